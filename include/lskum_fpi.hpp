/*
 * lskum_fpi.hpp -- the reference's host loop over the C ABI, stage by stage.
 *
 * The reference drives its hot path from two functions: q_lskum()
 * (q_lskum_mod.h:13-38) runs fpi_solver(it) max_iters times, and fpi_solver
 * (fpi_solver_mod.h:14-150) calls the stage functions in a fixed order, then
 * turns the last state_update's sum of squares into the residual and prints
 * one line per iteration.  This header keeps that loop, in C++, with every
 * stage function replaced by its lskum.h entry point:
 *
 *   prim_old copy loop (:19-25)           lskum_snapshot
 *   func_delta() (:32)                    lskum_func_delta
 *   per rk = 1..rks (:36-83):
 *     eval_q_variables()                  lskum_q_variables
 *     eval_q_derivatives()                lskum_q_derivatives
 *     eval_q_inner_loop();
 *     eval_update_innerloop() x inner     lskum_q_inner(ctx, inner)
 *     cal_flux_residual()                 lskum_flux_residual
 *     state_update(rk)                    lskum_state_update
 *   residual bookkeeping (:138-147), the line of :150
 *
 * Header only; it needs lskum.h and links against liblskum.  The calls are
 * stream-ordered: only the stage-4 state update, whose sums the residual
 * needs, synchronises (once per iteration).  lskum_iterate runs the same
 * sequence with one synchronisation per call; this loop is for callers
 * that keep the reference's structure, or put their own work between the
 * stages.
 */
#ifndef LSKUM_FPI_HPP
#define LSKUM_FPI_HPP

#include <cmath>
#include <cstdint>
#include <cstdio>

#include "lskum.h"

namespace lskum_fpi {

/* The globals fpi_solver and state_update leave behind
 * (data_structure_mod.h:56-64): the iteration counters it and itr, the
 * restart flag, the residual of the last iteration and the reductions of
 * its last state_update (max_res_point 1-based, as the reference numbers
 * points). */
struct Globals {
  int it = 0;
  int itr = 0;
  int restart = 0;
  double sum_res_sqr = 0.0;
  double max_res = 0.0;
  int max_res_point = 0;
  double res_old = 0.0;
  double res_new = 0.0;
  double residue = 0.0;
};

/* The line fpi_solver prints after every iteration (fpi_solver_mod.h:150,
 * after std::setprecision(13) << std::scientific at :87): the iteration as
 * an integer and the residue in scientific notation with 13 digits after
 * the point -- printf's %.13e gives the same text as that stream state. */
inline void print_iteration(std::FILE *out, int it, double residue) {
  std::fprintf(out, "iterations %d residue : %.13e\n", it, residue);
}

/* One fpi_solver(it) (fpi_solver_mod.h:14-150) on a single-case context.
 * n_points is max_points, the residual's denominator (:138); rks = 4.
 * Prints the reference's line to out when out is not NULL.  Returns the
 * first failing entry point's status, or LSKUM_ENONFINITE after printing
 * when the residual came out NaN or inf (the reference carries it on). */
inline int fpi_solver(lskum_ctx *ctx, int it, int inner_iterations, int32_t n_points,
                      Globals &g, std::FILE *out) {
  int rc = lskum_snapshot(ctx);
  if (!rc) rc = lskum_func_delta(ctx);
  for (int rk = 1; rk <= 4 && !rc; rk++) {
    rc = lskum_q_variables(ctx);
    if (!rc) rc = lskum_q_derivatives(ctx);
    if (!rc && inner_iterations > 0) rc = lskum_q_inner(ctx, inner_iterations);
    if (!rc) rc = lskum_flux_residual(ctx);
    if (rc) break;
    if (rk < 4) {
      /* every state_update resets the sums (state_update_mod.h:106-107):
       * only the last stage's survive, so the others are not fetched */
      rc = lskum_state_update(ctx, rk, nullptr, nullptr, nullptr);
    } else {
      int32_t point0 = 0;
      rc = lskum_state_update(ctx, rk, &g.sum_res_sqr, &g.max_res, &point0);
      g.max_res_point = point0 + 1;
    }
  }
  if (rc) return rc;
  g.res_new = std::sqrt(g.sum_res_sqr) / n_points;
  if (it <= 2 && g.restart == 0) {
    g.res_old = g.res_new;
    g.residue = 0;
  } else {
    g.residue = std::log10(g.res_new / g.res_old);
  }
  if (out) print_iteration(out, it, g.residue);
  return std::isfinite(g.res_new) ? LSKUM_OK : LSKUM_ENONFINITE;
}

/* q_lskum() (q_lskum_mod.h:28-37): iterations itr+1 .. itr+max_iters,
 * itr = 0 unless g.restart is set (then itr and res_old are those of a
 * restart file).  res_new / residue, when not NULL, receive max_iters values.
 * The context's own iteration count and res_old follow the loop, so
 * lskum_iterate or a restart file written afterwards continue from here. */
inline int q_lskum(lskum_ctx *ctx, int max_iters, int inner_iterations, int32_t n_points,
                   Globals &g, std::FILE *out, double *res_new = nullptr,
                   double *residue = nullptr) {
  if (g.restart == 0) g.itr = 0;
  int rc = LSKUM_OK;
  int done = 0;  /* iterations that reached the residual */
  for (g.it = g.itr + 1; g.it <= g.itr + max_iters; g.it++) {
    rc = fpi_solver(ctx, g.it, inner_iterations, n_points, g, out);
    if (rc && rc != LSKUM_ENONFINITE) break;
    if (res_new) res_new[done] = g.res_new;
    if (residue) residue[done] = g.residue;
    done++;
    if (rc) break;
  }
  const int rc2 = lskum_set_iteration(ctx, g.itr + done, g.res_old, g.restart);
  return rc ? rc : rc2;
}

}  // namespace lskum_fpi

#endif /* LSKUM_FPI_HPP */
