// fpi_dropin.cpp -- the reference's main() with q_lskum() replaced by the
// MI355X drop-in (include/lskum_fpi_reference.hpp).
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target dropin)
// against the reference headers where they lie and against liblskum; it is
// what a maintainer's main.cpp becomes with the drop-in, so a GPU test can
// run the reference's own reading, normals, sub-stencils and initial
// conditions followed by the GPU loop, and compare its printed lines and
// final state with the reference's.
//
// Usage: fpi_dropin GRID OUTDIR --iters N [--cfl C] [--vl K] [--inner I]
//                   [--restart FILE]
// Prints the reference's per-iteration line; writes OUTDIR/prim.bin
// ([4][n], 0-based) and OUTDIR/forces.bin (Cl, Cd, Cm).
#include "split_fluxes_mod.h"
#include "quadrant_fluxes_mod.h"
#include "data_structure_mod.h"
#include <fstream>
#include <iomanip>
#include <iostream>
#include "initial_conditions_mod.h"
#include "q_lskum_mod.h"
#include "ref_grid.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>

#include "lskum_fpi_reference.hpp"

static int write_bin(const std::string &path, const std::vector<double> &v) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return 3;
  const size_t w = std::fwrite(v.data(), sizeof(double), v.size(), f);
  std::fclose(f);
  return w == v.size() ? 0 : 3;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: fpi_dropin GRID OUTDIR --iters N [--cfl C] [--vl K] "
                         "[--inner I] [--restart FILE]\n");
    return 2;
  }
  const std::string out = argv[2];
  std::string restart_file;
  max_iters = 1;
  for (int a = 3; a + 1 < argc; a += 2) {
    const std::string k = argv[a];
    const char *v = argv[a + 1];
    if (k == "--iters") max_iters = std::atoi(v);
    else if (k == "--cfl") CFL = std::atof(v);
    else if (k == "--vl") VL_CONST = std::atof(v);
    else if (k == "--inner") inner_iterations = std::atoi(v);
    else if (k == "--restart") restart_file = v;
    else return 2;
  }
  const int width = read_grid(argv[1]);
  if (width < 0) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 3;
  }
  // main.cpp:59-64
  allocate_soln_wide(width);
  compute_normals();
  generate_connectivity();
  if (!restart_file.empty()) {
    std::ifstream src(restart_file, std::ios::binary);
    std::ofstream dst("restart.dat", std::ios::binary);
    if (!src || !dst) return 3;
    dst << src.rdbuf();
    dst.close();
    restart = 1;
  }
  initial_conditions();
  // main.cpp:65, q_lskum() -> the drop-in
  const int rc = lskum_fpi::q_lskum_mi355x(stdout);
  std::fflush(stdout);
  if (rc != LSKUM_OK) return 4;
  std::vector<double> prim;
  for (int r = 1; r <= 4; r++)
    for (int i = 1; i <= max_points; i++) prim.push_back(point.prim[r][i]);
  if (write_bin(out + "/prim.bin", prim)) return 3;
  return write_bin(out + "/forces.bin", {Cl[1], Cd[1], Cm[1]});
}
