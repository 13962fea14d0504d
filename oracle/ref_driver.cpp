// ref_driver.cpp -- harness that runs the reference solver's own functions.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the reference
// headers where they lie (-I<reference dir>); nothing of the reference is
// copied.  It replaces the reference's main() (main.cpp:13-66), whose grid
// path is hard-coded, with: a partGrid reader equivalent to main.cpp:16-58
// (neighbour width taken from the file instead of the fixed 20), the same
// set-up calls, and a loop that calls the reference's fpi_solver() or, for
// the iterations asked for, its stage functions in fpi_solver's order
// (fpi_solver_mod.h:19-61) with a binary dump after each stage.
//
// The numerics globals that the reference keeps mutable (CFL, VL_CONST,
// inner_iterations; data_structure_mod.h:72,96,118) are set from the command
// line.  mach and aoa are compile-time constants there (parameter_mod.h:4-5);
// --mach/--aoa override the derived free-stream state (theta, q_init, q_inf,
// prim) after initial_conditions(), and Cl/Cd/Cm are rescaled from the
// constant Mach the reference's force routine divides by.
//
// Usage:
//   ref_driver run GRID OUTDIR --iters N [--cfl C] [--vl K] [--inner I]
//              [--mach M --aoa A] [--stages it1,it2,...] [--dump it1,...]
//              [--diag 1] [--print 1] [--restart FILE]
// --print keeps the reference's own per-iteration line on stdout
// (fpi_solver_mod.h:150).  --restart copies FILE to ./restart.dat and runs
// with the reference's restart = 1: initial_conditions() reads it with
// restart_sol() (initial_conditions_mod.h:7-21) and the loop continues from
// its iteration count, q_lskum_mod.h:29-37 (--mach/--aoa then set only the
// free stream, not the state read from the file).
// --diag runs the reference's diagnostics on the final state: entropy and
// enstrophy are dumped; stagnation_pressure() prints its own line (its last
// two fields are ints printed with %f, so only the first four mean anything).
//   ref_driver kat DATAFILE NROWS OUTFILE
#include "split_fluxes_mod.h"
#include "quadrant_fluxes_mod.h"
#include "data_structure_mod.h"
#include <fstream>
#include <iomanip>
#include <iostream>
#include "initial_conditions_mod.h"
#include "q_lskum_mod.h"
#include "compute_force_coeffs_mod.h"
#include "compute_entropy_mod.h"
#include "compute_enstrophy_mod.h"
#include "stagnation_values_mod.h"
#include "ref_grid.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>

static std::string g_out;

static void dump_raw(const std::string &name, const std::vector<double> &v) {
  std::string p = g_out + "/" + name + ".bin";
  FILE *f = std::fopen(p.c_str(), "wb");
  if (!f) { std::perror(p.c_str()); std::exit(3); }
  std::fwrite(v.data(), sizeof(double), v.size(), f);
  std::fclose(f);
}

// 1-based reference arrays -> 0-based, component-major
static void dump2(const std::string &name,
                  const std::vector<std::vector<double>> &a, int ncomp) {
  std::vector<double> v;
  v.reserve((size_t)ncomp * max_points);
  for (int r = 1; r <= ncomp; r++)
    for (int i = 1; i <= max_points; i++) v.push_back(a[r][i]);
  dump_raw(name, v);
}

static void dump3(const std::string &name,
                  const std::vector<std::vector<std::vector<double>>> &a) {
  std::vector<double> v;
  v.reserve((size_t)8 * max_points);
  for (int d = 1; d <= 2; d++)
    for (int r = 1; r <= 4; r++)
      for (int i = 1; i <= max_points; i++) v.push_back(a[d][r][i]);
  dump_raw(name, v);
}

static void dump_state(const std::string &tag, bool with_q) {
  dump2(tag + "_prim", point.prim, 4);
  if (with_q) {
    dump2(tag + "_q", point.q, 4);
    dump3(tag + "_dq", point.dq);
    dump3(tag + "_qm", point.qm);
    dump2(tag + "_flux_res", point.flux_res, 4);
    std::vector<double> d(point.delta.begin() + 1, point.delta.end());
    dump_raw(tag + "_delta", d);
  }
}

static std::set<int> parse_list(const char *s) {
  std::set<int> out;
  std::string t(s);
  size_t p = 0;
  while (p < t.size()) {
    size_t q = t.find(',', p);
    if (q == std::string::npos) q = t.size();
    if (q > p) out.insert(std::atoi(t.substr(p, q - p).c_str()));
    p = q + 1;
  }
  return out;
}

static int cmd_run(int argc, char **argv) {
  if (argc < 4) return 2;
  const char *grid = argv[2];
  g_out = argv[3];
  int iters = 1, diag = 0, print = 0;
  std::string restart_file;
  double mach_o = -1, aoa_o = 0;
  std::set<int> stages, dumps;
  for (int a = 4; a < argc; a++) {
    std::string k = argv[a];
    if (a + 1 >= argc) return 2;
    const char *v = argv[++a];
    if (k == "--iters") iters = std::atoi(v);
    else if (k == "--cfl") CFL = std::atof(v);
    else if (k == "--vl") VL_CONST = std::atof(v);
    else if (k == "--inner") inner_iterations = std::atoi(v);
    else if (k == "--mach") mach_o = std::atof(v);
    else if (k == "--aoa") aoa_o = std::atof(v);
    else if (k == "--stages") stages = parse_list(v);
    else if (k == "--dump") dumps = parse_list(v);
    else if (k == "--diag") diag = std::atoi(v);
    else if (k == "--print") print = std::atoi(v);
    else if (k == "--restart") restart_file = v;
    else return 2;
  }
  int width = read_grid(grid);
  if (width < 0) { std::fprintf(stderr, "cannot read %s\n", grid); return 3; }
  // main.cpp:59-65
  allocate_soln_wide(width);
  compute_normals();
  generate_connectivity();
  if (!restart_file.empty()) {
    std::ifstream src(restart_file, std::ios::binary);
    std::ofstream dst("restart.dat", std::ios::binary);
    if (!src || !dst) { std::fprintf(stderr, "cannot copy %s\n", restart_file.c_str()); return 3; }
    dst << src.rdbuf();
    dst.close();
    restart = 1;
  }
  initial_conditions();
  double force_scale = 1.0;
  if (mach_o > 0) {
    theta = aoa_o * pi / 180;
    q_init[1] = q_inf[1] = rho_inf;
    q_init[2] = q_inf[2] = mach_o * cos(theta);
    q_init[3] = q_inf[3] = mach_o * sin(theta);
    q_init[4] = q_inf[4] = pr_inf;
    if (restart == 0)
      for (int k = 1; k <= max_points; k++)
        for (int r = 1; r <= 4; r++) point.prim[r][k] = q_init[r];
    force_scale = (0.5 * rho_inf * mach * mach) / (0.5 * rho_inf * mach_o * mach_o);
  }
  {  // sub-stencil lists, 0-based, as counts then members
    std::vector<double> v;
    for (int i = 1; i <= max_points; i++) {
      v.push_back(point.xpos_nbhs[i]); v.push_back(point.xneg_nbhs[i]);
      v.push_back(point.ypos_nbhs[i]); v.push_back(point.yneg_nbhs[i]);
    }
    dump_raw("substencil_counts", v);
    v.clear();
    for (int i = 1; i <= max_points; i++) {
      for (int j = 1; j <= point.xpos_nbhs[i]; j++) v.push_back(point.xpos_conn[i][j] - 1);
      for (int j = 1; j <= point.xneg_nbhs[i]; j++) v.push_back(point.xneg_conn[i][j] - 1);
      for (int j = 1; j <= point.ypos_nbhs[i]; j++) v.push_back(point.ypos_conn[i][j] - 1);
      for (int j = 1; j <= point.yneg_nbhs[i]; j++) v.push_back(point.yneg_conn[i][j] - 1);
    }
    dump_raw("substencil_members", v);
    std::vector<double> nrm(point.nx.begin() + 1, point.nx.end());
    nrm.insert(nrm.end(), point.ny.begin() + 1, point.ny.end());
    dump_raw("normals", nrm);
  }
  // q_lskum_mod.h:28-37
  if (restart == 0) itr = 0;
  std::vector<double> hist;  // per iteration: res_new, residue, Cl, Cd, Cm
  std::streambuf *saved = std::cout.rdbuf();
  std::ofstream sink("/dev/null");
  if (!print) std::cout.rdbuf(sink.rdbuf());
  const auto t0 = std::chrono::steady_clock::now();
  for (it = itr + 1; it <= itr + iters; it++) {
    if (stages.count(it)) {
      char tag[64];
      // fpi_solver_mod.h:19-61, stage by stage
      for (int i = 1; i <= max_points; i++)
        for (int j = 1; j <= 4; j++) point.prim_old[j][i] = point.prim[j][i];
      func_delta();
      std::snprintf(tag, sizeof tag, "it%d_delta", it);
      dump_raw(tag, std::vector<double>(point.delta.begin() + 1, point.delta.end()));
      for (int rk = 1; rk <= rks; rk++) {
        std::snprintf(tag, sizeof tag, "it%d_rk%d", it, rk);
        std::string t(tag);
        dump2(t + "_prim_in", point.prim, 4);
        eval_q_variables();
        dump2(t + "_q", point.q, 4);
        eval_q_derivatives();
        dump3(t + "_dq0", point.dq);
        dump3(t + "_qm", point.qm);
        for (int i = 1; i <= inner_iterations; i++) {
          eval_q_inner_loop();
          eval_update_innerloop();
        }
        dump3(t + "_dq", point.dq);
        cal_flux_residual();
        dump2(t + "_flux_res", point.flux_res, 4);
        state_update(rk);
        dump2(t + "_prim", point.prim, 4);
        dump_raw(t + "_red", {sum_res_sqr, max_res, (double)(max_res_point - 1)});
      }
      // fpi_solver_mod.h:138-147
      res_new = sqrt(sum_res_sqr) / max_points;
      if (it <= 2 && restart == 0) {
        res_old = res_new;
        residue = 0;
      } else {
        residue = log10(res_new / res_old);
      }
    } else {
      fpi_solver(it);
    }
    compute_cl_cd_cm();
    hist.push_back(res_new);
    hist.push_back(residue);
    hist.push_back(Cl[1] * force_scale);
    hist.push_back(Cd[1] * force_scale);
    hist.push_back(Cm[1] * force_scale);
    if (dumps.count(it)) {
      char tag[32];
      std::snprintf(tag, sizeof tag, "it%d", it);
      dump_state(tag, true);
    }
  }
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cout.rdbuf(saved);
  dump_raw("history", hist);
  dump_raw("loop_seconds", {secs});
  if (diag) {
    compute_entropy();
    compute_enstrophy();
    dump_raw("diag_entropy", std::vector<double>(point.entropy.begin() + 1, point.entropy.end()));
    dump_raw("diag_vorticity",
             std::vector<double>(point.vorticity.begin() + 1, point.vorticity.end()));
    dump_raw("diag_vorticity_sqr",
             std::vector<double>(point.vorticity_sqr.begin() + 1, point.vorticity_sqr.end()));
    dump_raw("diag_total_enstrophy", {total_enstrophy});
    std::fflush(stdout);
    stagnation_pressure();
    std::printf("\n");
    std::fflush(stdout);
  }
  return 0;
}

// Known answers for the leaf physics on the reference's own data.dat rows
// (main.cpp:70-100 harness): per row the eight fluxes and qtilde_to_primitive
// of q~ = (c1, c2, c3, -c4).
static int cmd_kat(int argc, char **argv) {
  if (argc < 5) return 2;
  std::ifstream fin(argv[2]);
  int nrows = std::atoi(argv[3]);
  g_out = ".";
  std::vector<double> out;
  setup_case_parameters();
  for (int row = 0; row < nrows; row++) {
    double a[5], nx, ny, u1, u2, rho, pr;
    fin >> a[1] >> a[2] >> a[3] >> a[4] >> nx >> ny >> u1 >> u2 >> rho >> pr;
    if (!fin) return 3;
    std::vector<double> G(5);
    void (*fl[8])(std::vector<double> &, double, double, double, double, double, double) = {
        flux_Gxp, flux_Gxn, flux_Gyp, flux_Gyn,
        flux_quad_GxI, flux_quad_GxII, flux_quad_GxIII, flux_quad_GxIV};
    for (int f = 0; f < 8; f++) {
      fl[f](G, nx, ny, u1, u2, rho, pr);
      for (int r = 1; r <= 4; r++) out.push_back(G[r]);
    }
    std::vector<double> qt = {0.0, a[1], a[2], a[3], -a[4]};
    double pu1, pu2, prho, ppr;
    qtilde_to_primitive(qt, pu1, pu2, prho, ppr);
    out.push_back(pu1); out.push_back(pu2); out.push_back(prho); out.push_back(ppr);
  }
  FILE *f = std::fopen(argv[4], "wb");
  if (!f) return 3;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 2 && !std::strcmp(argv[1], "run")) return cmd_run(argc, argv);
  if (argc >= 2 && !std::strcmp(argv[1], "kat")) return cmd_kat(argc, argv);
  std::fprintf(stderr, "usage: ref_driver run|kat ...\n");
  return 2;
}
