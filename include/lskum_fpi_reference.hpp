/*
 * lskum_fpi_reference.hpp -- the reference solver's q_lskum() on an MI355X,
 * on the reference's own globals.
 *
 * For a maintainer of the reference (Pathlessbark8/C-_serial): include this
 * header in main.cpp after the reference's modules (it reads and writes the
 * globals of data_structure_mod.h and parameter_mod.h) and replace
 *
 *     q_lskum();                                  // main.cpp:65
 * with
 *     if (lskum_fpi::q_lskum_mi355x(stdout) != LSKUM_OK) return 1;
 *
 * compute_normals(), generate_connectivity() and initial_conditions()
 * (main.cpp:62-64) stay where they are: the context derives the same normals
 * and sub-stencils from the cloud bit for bit, and takes its starting state
 * from point.prim, so a restart file read by initial_conditions() is used as
 * the reference would.  The loop is lskum_fpi::q_lskum, fpi_solver's stage
 * calls through lskum.h (lskum_fpi.hpp), printing the reference's line per
 * iteration.  Afterwards point.prim, the residual globals, it and Cl/Cd/Cm
 * hold what the reference's q_lskum() followed by compute_cl_cd_cm() leaves.
 * Build with -I<this include dir> and link -llskum.
 */
#ifndef LSKUM_FPI_REFERENCE_HPP
#define LSKUM_FPI_REFERENCE_HPP

#include <cstdint>
#include <cstdio>
#include <vector>

#include "lskum_fpi.hpp"

namespace lskum_fpi {

/* The reference's 1-based point struct (main.cpp:30-58) as the 0-based
 * arrays lskum_cloud points into: left/right 0 (none) become -1, conn[i][k]
 * becomes CSR. */
struct ReferenceCloud {
  std::vector<double> x, y, min_dist;
  std::vector<int32_t> left, right, flag_1, flag_2, conn_off, conn;
  lskum_cloud view() const {
    lskum_cloud c;
    c.n_points = (int32_t)x.size();
    c.x = x.data();
    c.y = y.data();
    c.left = left.data();
    c.right = right.data();
    c.flag_1 = flag_1.data();
    c.flag_2 = flag_2.data();
    c.min_dist = min_dist.data();
    c.conn_off = conn_off.data();
    c.conn = conn.data();
    return c;
  }
};

inline ReferenceCloud cloud_from_reference() {
  ReferenceCloud c;
  const int n = max_points;
  c.x.resize(n);
  c.y.resize(n);
  c.min_dist.resize(n);
  c.left.resize(n);
  c.right.resize(n);
  c.flag_1.resize(n);
  c.flag_2.resize(n);
  c.conn_off.assign(n + 1, 0);
  for (int i = 1; i <= n; i++) {
    c.x[i - 1] = point.x[i];
    c.y[i - 1] = point.y[i];
    c.min_dist[i - 1] = point.min_dist[i];
    c.left[i - 1] = point.left[i] - 1;
    c.right[i - 1] = point.right[i] - 1;
    c.flag_1[i - 1] = point.flag_1[i];
    c.flag_2[i - 1] = point.flag_2[i];
    for (int k = 1; k <= point.nbhs[i]; k++) c.conn.push_back(point.conn[i][k] - 1);
    c.conn_off[i] = (int32_t)c.conn.size();
  }
  return c;
}

/* The reference's numerics globals as lskum_params: mach and aoa
 * (parameter_mod.h:4-5), CFL, power, VL_CONST, inner_iterations
 * (data_structure_mod.h:72, 90, 96, 118). */
inline lskum_params params_from_reference(int32_t precision = LSKUM_FP64) {
  lskum_params p = {};
  p.mach = mach;
  p.aoa_deg = aoa;
  p.cfl = CFL;
  p.vl_const = VL_CONST;
  p.power = power;
  p.inner_iterations = inner_iterations;
  p.precision = precision;
  p.reorder = LSKUM_REORDER_HILBERT;
  p.lds_tiles = LSKUM_TILES_AUTO;
  p.tile_block = 0;
  p.points_per_thread = 0;
  return p;
}

/* point.prim[1..4][1..n] <-> the context's state ([4][n], 0-based). */
inline int state_to_context(lskum_ctx *ctx) {
  const int n = max_points;
  std::vector<double> prim((size_t)4 * n);
  for (int v = 0; v < 4; v++)
    for (int i = 1; i <= n; i++) prim[(size_t)v * n + i - 1] = point.prim[v + 1][i];
  return lskum_set_state(ctx, prim.data());
}

inline int state_from_context(lskum_ctx *ctx) {
  const int n = max_points;
  std::vector<double> prim((size_t)4 * n);
  const int rc = lskum_get_state(ctx, prim.data());
  if (rc) return rc;
  for (int v = 0; v < 4; v++)
    for (int i = 1; i <= n; i++) point.prim[v + 1][i] = prim[(size_t)v * n + i - 1];
  return LSKUM_OK;
}

/* q_lskum() (q_lskum_mod.h:13-38) with fpi_solver's stages on the GPU, then
 * compute_cl_cd_cm() (compute_force_coeffs_mod.h:7-69; Cl/Cd/Cm of shape 1,
 * without the Cp file).  Messages of a failing entry point go to stderr. */
inline int q_lskum_mi355x(std::FILE *out, int32_t precision = LSKUM_FP64) {
  const ReferenceCloud cloud = cloud_from_reference();
  const lskum_cloud view = cloud.view();
  const lskum_params params = params_from_reference(precision);
  lskum_ctx *ctx = nullptr;
  int rc = lskum_create(&view, &params, &ctx);
  if (!rc) rc = state_to_context(ctx);
  Globals g;
  g.restart = restart;
  g.itr = itr;
  g.res_old = res_old;
  if (!rc) rc = q_lskum(ctx, max_iters, inner_iterations, max_points, g, out);
  it = g.it;
  itr = g.itr;
  res_old = g.res_old;
  res_new = g.res_new;
  residue = g.residue;
  sum_res_sqr = g.sum_res_sqr;
  max_res = g.max_res;
  max_res_point = g.max_res_point;
  int rc2 = ctx ? state_from_context(ctx) : LSKUM_OK;
  if (!rc2 && ctx) rc2 = lskum_forces(ctx, &Cl[1], &Cd[1], &Cm[1]);
  if (rc || rc2)
    std::fprintf(stderr, "q_lskum_mi355x: status %d: %s\n", rc ? rc : rc2,
                 ctx ? lskum_last_error(ctx) : "context not created");
  lskum_destroy(ctx);
  return rc ? rc : rc2;
}

}  // namespace lskum_fpi

#endif /* LSKUM_FPI_REFERENCE_HPP */
