// fpi_main.cpp -- lskum_fpi: the reference's main() (main.cpp:13-66) as a
// program over the C ABI.  It reads a partGrid file (main.cpp:16-58), sets
// the free stream (or a restart file's state, initial_conditions_mod.h:7-37)
// and runs q_lskum() with fpi_solver's stage calls one by one
// (lskum_fpi.hpp), printing the reference's line per iteration
// (fpi_solver_mod.h:150).
//
// Usage: lskum_fpi GRID [--iters N] [--mach M] [--aoa A] [--cfl C] [--vl K]
//                  [--power P] [--inner I] [--mixed 1] [--restart FILE]
//                  [--write-restart FILE] [--out PREFIX]
// --out writes PREFIX.prim.bin ([4][n] doubles, the file's point order),
// PREFIX.flux_res.bin (the same layout) and PREFIX.forces.bin (Cl, Cd, Cm).
// Exit status: 0, 2 usage, 3 I/O, 4 a failing entry point (its message on
// stderr).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "lskum.h"
#include "lskum_fpi.hpp"

namespace {

int fail(lskum_ctx *ctx, int rc, const char *what) {
  std::fprintf(stderr, "lskum_fpi: %s failed (%d): %s\n", what, rc,
               ctx ? lskum_last_error(ctx) : "");
  return 4;
}

bool write_bin(const std::string &path, const double *v, size_t n) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const size_t w = std::fwrite(v, sizeof(double), n, f);
  return std::fclose(f) == 0 && w == n;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: lskum_fpi GRID [--iters N] [--mach M] [--aoa A] [--cfl C] "
                         "[--vl K] [--power P] [--inner I] [--mixed 1] [--restart FILE] "
                         "[--write-restart FILE] [--out PREFIX]\n");
    return 2;
  }
  // the reference's defaults (parameter_mod.h:4-5, data_structure_mod.h:72-118)
  lskum_params p = {};
  p.mach = 0.63;
  p.aoa_deg = 2.0;
  p.cfl = 0.1;
  p.vl_const = 150.0;
  p.power = 0.0;
  p.inner_iterations = 3;
  p.precision = LSKUM_FP64;
  p.reorder = LSKUM_REORDER_HILBERT;
  p.lds_tiles = LSKUM_TILES_AUTO;
  int max_iters = 1000;
  std::string restart_in, restart_out, out;
  for (int a = 2; a < argc; a += 2) {
    if (a + 1 >= argc) return 2;
    const std::string k = argv[a];
    const char *v = argv[a + 1];
    if (k == "--iters") max_iters = std::atoi(v);
    else if (k == "--mach") p.mach = std::atof(v);
    else if (k == "--aoa") p.aoa_deg = std::atof(v);
    else if (k == "--cfl") p.cfl = std::atof(v);
    else if (k == "--vl") p.vl_const = std::atof(v);
    else if (k == "--power") p.power = std::atof(v);
    else if (k == "--inner") p.inner_iterations = std::atoi(v);
    else if (k == "--mixed") p.precision = std::atoi(v) ? LSKUM_MIXED : LSKUM_FP64;
    else if (k == "--restart") restart_in = v;
    else if (k == "--write-restart") restart_out = v;
    else if (k == "--out") out = v;
    else return 2;
  }
  lskum_cloud_h *cloud_h = nullptr;
  int rc = lskum_cloud_read_partgrid(argv[1], &cloud_h);
  if (rc) {
    std::fprintf(stderr, "lskum_fpi: cannot read %s (%d)\n", argv[1], rc);
    lskum_cloud_free(cloud_h);
    return 3;
  }
  lskum_cloud cloud;
  lskum_cloud_view(cloud_h, &cloud);
  const int32_t n = cloud.n_points;
  lskum_ctx *ctx = nullptr;
  rc = lskum_create(&cloud, &p, &ctx);
  if (rc) {
    const int code = fail(ctx, rc, "lskum_create");
    lskum_destroy(ctx);
    lskum_cloud_free(cloud_h);
    return code;
  }
  lskum_fpi::Globals g;
  std::vector<double> prim((size_t)4 * n);
  if (!restart_in.empty()) {  // initial_conditions_mod.h:7-21 with restart = 1
    int32_t itr = 0;
    double res_old = 0.0;
    rc = lskum_restart_read(restart_in.c_str(), n, &itr, &res_old, prim.data());
    if (!rc) rc = lskum_set_state(ctx, prim.data());
    if (rc) {
      const int code = fail(ctx, rc, "restart");
      lskum_destroy(ctx);
      lskum_cloud_free(cloud_h);
      return code;
    }
    g.restart = 1;
    g.itr = itr;
    g.res_old = res_old;
  }
  rc = lskum_fpi::q_lskum(ctx, max_iters, p.inner_iterations, n, g, stdout);
  std::fflush(stdout);
  int code = rc ? fail(ctx, rc, "q_lskum") : 0;
  double forces[3] = {0, 0, 0};
  if (!code && (rc = lskum_forces(ctx, &forces[0], &forces[1], &forces[2])))
    code = fail(ctx, rc, "lskum_forces");
  if (!code && (rc = lskum_get_state(ctx, prim.data()))) code = fail(ctx, rc, "lskum_get_state");
  if (!code && !restart_out.empty()) {
    int32_t it_done = 0;
    double res_old = 0.0;
    if ((rc = lskum_get_iteration(ctx, &it_done, &res_old)) ||
        (rc = lskum_restart_write(restart_out.c_str(), &cloud, it_done, res_old, prim.data())))
      code = fail(ctx, rc, "lskum_restart_write");
  }
  if (!code && !out.empty()) {
    std::vector<double> fres((size_t)4 * n);
    if ((rc = lskum_get_array(ctx, LSKUM_A_FLUX_RES, fres.data()))) {
      code = fail(ctx, rc, "lskum_get_array");
    } else if (!write_bin(out + ".prim.bin", prim.data(), prim.size()) ||
               !write_bin(out + ".flux_res.bin", fres.data(), fres.size()) ||
               !write_bin(out + ".forces.bin", forces, 3)) {
      std::fprintf(stderr, "lskum_fpi: cannot write %s.*\n", out.c_str());
      code = 3;
    }
  }
  lskum_destroy(ctx);
  lskum_cloud_free(cloud_h);
  return code;
}
