// host_sanitize.cpp -- drives the library's host code (cloud generator, k-d
// tree kNN, visibility and balance passes, partGrid reader/writer, normals,
// sub-stencils, RCM, restart and output files, Hicks-Henne fields) for an
// AddressSanitizer / UndefinedBehaviorSanitizer build of those sources
// (tests/test_host_sanitize.py builds it with g++ and runs it).  Exits 0 and
// prints "host code OK" when every call behaved; the sanitizers abort on the
// first invalid access or undefined operation.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "lskum.h"

#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
      std::exit(1);                                                   \
    }                                                                 \
  } while (0)

static lskum_cloud_h *generate(int m, int p, int t, int n_wall, int n_radial, int n_circ,
                               double stretch, int k, uint64_t seed) {
  lskum_gen_params g = {};
  g.naca_m = m;
  g.naca_p = p;
  g.naca_t = t;
  g.n_wall = n_wall;
  g.n_radial = n_radial;
  g.n_circ = n_circ;
  g.r_far = 20.0;
  g.stretch = stretch;
  g.jitter = 0.1;
  g.k = k;
  g.radius_factor = 3.0;
  g.seed = seed;
  lskum_cloud_h *c = nullptr;
  CHECK(lskum_cloud_generate(&g, &c) == LSKUM_OK && c);
  return c;
}

static void exercise(lskum_cloud_h *h, const std::string &dir) {
  lskum_cloud v;
  CHECK(lskum_cloud_view(h, &v) == LSKUM_OK);
  const int n = v.n_points;
  const int nnz = v.conn_off[n];
  int32_t lo = 0, hi = 0;
  double dets[3];
  CHECK(lskum_cloud_stats(h, &lo, &hi, dets) == LSKUM_OK && lo >= 3 && hi >= lo);
  std::vector<double> nx(n), ny(n);
  std::vector<uint8_t> mask(nnz);
  CHECK(lskum_cloud_connectivity(h, nx.data(), ny.data(), mask.data()) == LSKUM_OK);
  for (int i = 0; i < n; i++) CHECK(std::fabs(nx[i] * nx[i] + ny[i] * ny[i] - 1.0) < 1e-12);
  std::vector<int32_t> perm(n), seen(n, 0);
  CHECK(lskum_cloud_rcm(h, perm.data()) == LSKUM_OK);
  for (int i = 0; i < n; i++) CHECK(perm[i] >= 0 && perm[i] < n && !seen[perm[i]]++);
  // partGrid round trip and a copy from arrays
  const std::string grid = dir + "/grid.txt";
  CHECK(lskum_cloud_write_partgrid(h, grid.c_str()) == LSKUM_OK);
  lskum_cloud_h *back = nullptr;
  CHECK(lskum_cloud_read_partgrid(grid.c_str(), &back) == LSKUM_OK);
  lskum_cloud bv;
  CHECK(lskum_cloud_view(back, &bv) == LSKUM_OK && bv.n_points == n);
  for (int i = 0; i < n; i++) CHECK(bv.x[i] == v.x[i] && bv.y[i] == v.y[i]);
  for (int e = 0; e < nnz; e++) CHECK(bv.conn[e] == v.conn[e]);
  lskum_cloud_h *copy = nullptr;
  CHECK(lskum_cloud_from_arrays(&bv, &copy) == LSKUM_OK);
  lskum_cloud_free(copy);
  lskum_cloud_free(back);
  // restart, solution and Cp files of a perturbed free stream
  std::vector<double> prim(4 * (size_t)n);
  for (int i = 0; i < n; i++) {
    prim[i] = 1.0 + 1e-3 * std::sin(0.1 * i);
    prim[n + i] = 0.6;
    prim[2 * n + i] = 0.02;
    prim[3 * n + i] = 1.0 / 1.4;
  }
  const std::string rst = dir + "/restart.dat";
  CHECK(lskum_restart_write(rst.c_str(), &v, 17, 1.25e-6, prim.data()) == LSKUM_OK);
  int32_t it = 0;
  double res_old = 0.0;
  std::vector<double> prim2(prim.size());
  CHECK(lskum_restart_read(rst.c_str(), n, &it, &res_old, prim2.data()) == LSKUM_OK);
  CHECK(it == 17 && res_old == 1.25e-6 && prim2 == prim);
  CHECK(lskum_restart_read(rst.c_str(), n + 1, &it, &res_old, prim2.data()) == LSKUM_EINVAL);
  const std::string sol = dir + "/solution.dat", cp = dir + "/cp.dat";
  CHECK(lskum_solution_write(sol.c_str(), &v, 17, 1.25e-6, prim.data(), nullptr) == LSKUM_OK);
  CHECK(lskum_cp_write(cp.c_str(), &v, prim.data(), 0.8) == LSKUM_OK);
  // Hicks-Henne design fields
  std::vector<double> dy(20 * (size_t)n);
  CHECK(lskum_hicks_henne_fields(&v, 10, 0.05, 0.95, 3, 0.5, dy.data()) == LSKUM_OK);
  for (double d : dy) CHECK(std::isfinite(d));
}

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  // the small symmetric cloud of the parity fixtures, a k = 24 cloud with a
  // matched first ring, and a cambered section (the many-case recipe's
  // rules: trailing-edge bisector, no point inside the body)
  lskum_cloud_h *a = generate(0, 0, 12, 40, 24, 60, 1.2, 20, 11);
  lskum_cloud_h *b = generate(0, 0, 12, 80, 30, 118, 0.0, 24, 12);
  lskum_cloud_h *c = generate(4, 6, 12, 60, 28, 80, 0.0, 20, 7);
  for (lskum_cloud_h *h : {a, b, c}) exercise(h, dir);
  for (lskum_cloud_h *h : {a, b, c}) lskum_cloud_free(h);
  // reader errors: a missing file and a truncated one
  lskum_cloud_h *bad = nullptr;
  CHECK(lskum_cloud_read_partgrid((dir + "/missing.txt").c_str(), &bad) == LSKUM_EIO);
  lskum_cloud_free(bad);
  const std::string trunc = dir + "/truncated.txt";
  FILE *f = std::fopen(trunc.c_str(), "w");
  CHECK(f);
  std::fprintf(f, "3\n0.5 0.25 0 0 1 0 0.1 2 2 3\n0.6\n");
  std::fclose(f);
  bad = nullptr;
  CHECK(lskum_cloud_read_partgrid(trunc.c_str(), &bad) == LSKUM_EIO);
  lskum_cloud_free(bad);
  int32_t it = 0;
  double ro = 0.0, p4[4];
  CHECK(lskum_restart_read(trunc.c_str(), 1, &it, &ro, p4) != LSKUM_OK);
  std::printf("host code OK\n");
  return 0;
}
