/*
 * lskum.h -- C ABI of the MI355X q-LSKUM iteration.
 *
 * The reference solver (Pathlessbark8/C-_serial) has no plugin or FFI layer: its
 * hot path is a set of argument-less void functions over one global
 * `points point` struct, called in a fixed order by fpi_solver()
 * (fpi_solver_mod.h:32-61).  This header replaces that set one for one with
 * explicit, stream-ordered entry points on an opaque context that owns the
 * device state.  Every entry point cites the reference function it replaces.
 *
 * Conventions
 *  - Status is int: 0 = ok, negative = error (see LSKUM_E*).  No exception
 *    crosses the ABI.  lskum_last_error() returns a message for the last
 *    failing call on a context.
 *  - Arrays are 0-based and component-major ("SoA"), as prim[var*n + point].
 *    Host arrays given to lskum_create() are copied; outputs go to caller-owned
 *    buffers.  Point order at the ABI is always the caller's order, even when
 *    the context reorders points internally (LSKUM_REORDER_RCM).
 *  - Calls on one context are ordered on its HIP stream.  Only calls that
 *    return host scalars or host arrays synchronise.
 *  - One context per host thread; contexts are independent.
 *  - A context runs n_cases independent cases: one from lskum_create, many
 *    from lskum_create_batch.  Per-case outputs are arrays of n_cases
 *    (case-major where they are per case and per iteration); per-point arrays
 *    hold the cases one after the other, each in its own [comp][n_case]
 *    layout and caller order.  With one case every array reduces to the
 *    reference's shape.
 */
#ifndef LSKUM_H
#define LSKUM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define LSKUM_OK 0
#define LSKUM_EINVAL (-1)     /* bad argument or inconsistent input        */
#define LSKUM_EHIP (-2)       /* a HIP runtime call failed                 */
#define LSKUM_ENOMEM (-3)     /* host or device allocation failed          */
#define LSKUM_EIO (-4)        /* file missing, unreadable or truncated     */
#define LSKUM_ENONFINITE (-5) /* residual or force came out NaN/Inf       */
#define LSKUM_ESTENCIL (-6)   /* degenerate sub-stencil (empty or det<=0)  */

/* --------------------------------------------------------------- clouds */
/*
 * Point cloud in the reference's "partGrid" content (main.cpp:30-58), 0-based.
 * flag_1: 0 wall, 1 interior, 2 outer (main.cpp:37-51).  flag_2 > 0: shape
 * index (main.cpp:53-56).  left/right: previous/next boundary point, -1 for
 * interior points.  conn is CSR: neighbours of point i are
 * conn[conn_off[i] .. conn_off[i+1]-1].
 */
typedef struct lskum_cloud {
  int32_t n_points;
  const double *x, *y;
  const int32_t *left, *right;
  const int32_t *flag_1, *flag_2;
  const double *min_dist;
  const int32_t *conn_off; /* n_points + 1 entries */
  const int32_t *conn;     /* conn_off[n_points] entries */
} lskum_cloud;

/* Host-owned cloud made by the generator or the partGrid reader. */
typedef struct lskum_cloud_h lskum_cloud_h;

/*
 * Synthetic NACA 4-digit meshfree cloud (no reference counterpart: the
 * reference reads a fixed file, main.cpp:17).  Wall points first (cosine
 * chord spacing, clockwise from the trailing edge), then n_radial-1 jittered
 * O-grid rings of interior points, then the far-field ring (flag_1 = 2).
 * Neighbours: the k nearest visible points within radius_factor x local
 * spacing, ties broken by index; a pair is not visible when its segment
 * crosses the body outline or when both are wall points on opposite
 * surfaces.  The selection is then balanced: every sub-stencil the flux
 * uses gets at least 3 members and a relative least-squares determinant of
 * at least 0.05, by swapping in the nearest qualifying points within twice
 * the radius for the farthest members no sub-stencil needs.  Neighbours are
 * listed by increasing distance; min_dist is the nearest one's distance.
 * Deterministic from seed.
 */
typedef struct lskum_gen_params {
  int32_t naca_m;   /* max camber, % chord (NACA 4-digit first digit)   */
  int32_t naca_p;   /* camber position, tenths of chord (second digit)  */
  int32_t naca_t;   /* thickness, % chord (last two digits)             */
  int32_t n_wall;   /* wall points                                      */
  int32_t n_radial; /* rings, the last one being the far field          */
  int32_t n_circ;   /* points per ring                                  */
  double r_far;     /* far-field radius, chords                         */
  double stretch;   /* radial growth ratio; <= 0: first ring spacing    */
                    /* matched to the first ring's circumferential one  */
  double jitter;    /* uniform jitter, fraction of local spacing        */
  int32_t k;        /* neighbours per point                             */
  double radius_factor; /* search radius in local spacings              */
  uint64_t seed;
  double wall_blend; /* 0: cosine chord spacing of the wall points; w in  */
                     /* (0, 1]: their arc-length positions blended, w to */
                     /* 1 - w, between uniform arc length and cosine     */
                     /* spacing (bounds the clustering at the edges)     */
} lskum_gen_params;

int lskum_cloud_generate(const lskum_gen_params *p, lskum_cloud_h **out);
/* Reader for the legacy partGrid text file (main.cpp:16-58). */
int lskum_cloud_read_partgrid(const char *path, lskum_cloud_h **out);
/* Writer in the same layout, doubles at 17 significant digits. */
int lskum_cloud_write_partgrid(const lskum_cloud_h *c, const char *path);
/* Copy caller arrays into a new host cloud. */
int lskum_cloud_from_arrays(const lskum_cloud *src, lskum_cloud_h **out);
/* Borrowed view of the arrays of c (valid until lskum_cloud_free). */
int lskum_cloud_view(const lskum_cloud_h *c, lskum_cloud *view);
/* Generator statistics: min/max neighbour count and min_rel_det[3]: per
 * point type (flag_1 0..2), the smallest sub-stencil least-squares
 * determinant relative to its scale; -1 if no point of the type. */
int lskum_cloud_stats(const lskum_cloud_h *c, int32_t *min_nbhs,
                      int32_t *max_nbhs, double *min_rel_det);
/* Host-only set-up of a cloud, as lskum_create does it: normals
 * (point_normals_mod.h:17-55) and, per conn entry, the sub-stencil bits
 * 1 xpos, 2 xneg, 4 ypos, 8 yneg (generate_connectivity_mod.h:5-249).
 * nx, ny: n_points doubles; mask: conn_off[n_points] bytes. */
int lskum_cloud_connectivity(const lskum_cloud_h *c, double *nx, double *ny,
                             uint8_t *mask);
/* Reverse Cuthill-McKee order of the symmetrised neighbour graph:
 * perm[new] = old (n_points entries). */
int lskum_cloud_rcm(const lskum_cloud_h *c, int32_t *perm);
void lskum_cloud_free(lskum_cloud_h *c);

/* ------------------------------------------------------------ parameters */
#define LSKUM_FP64 0  /* every operation in FP64, in the reference's order  */
/* Mixed: neighbour q, dq, qmax/qmin and eps^2 stored and gathered in FP32;
 * the limiter, q~ -> primitive, the split/quadrant fluxes and their
 * per-point sums, the least-squares right-hand sums and each neighbour's
 * local-step term in FP32; coordinates, the least-squares geometry (per-point
 * solve coefficients formed once in FP64 on the host, then rounded), the
 * state, the SSPRK update, the stored local time step and all reductions in
 * FP64.  q, dq and qm read back through lskum_get_array are the FP32 values
 * widened. */
#define LSKUM_MIXED 1

#define LSKUM_REORDER_NONE 0
#define LSKUM_REORDER_RCM 1 /* reverse Cuthill-McKee on the neighbour graph */
#define LSKUM_REORDER_HILBERT 2 /* Hilbert curve over the coordinates:     */
                                /* compact blocks, neighbour tiles in LDS  */

/*
 * Case and numerics parameters.  The reference hard-codes these as globals
 * (parameter_mod.h:4-10, data_structure_mod.h:57-121); here they are runtime
 * fields.  gamma = 1.4 is fixed, as the reference's 2.5 / 0.4 / 0.2 / 7 / 6
 * factors assume it (q_variables_mod.h:25, state_update_mod.h:171).
 */
typedef struct lskum_params {
  double mach;      /* free-stream Mach number (parameter_mod.h:4)          */
  double aoa_deg;   /* angle of attack, degrees (parameter_mod.h:5)         */
  double cfl;       /* CFL (data_structure_mod.h:72)                        */
  double vl_const;  /* Venkatakrishnan K (data_structure_mod.h:96)          */
  double power;     /* LS weight exponent (data_structure_mod.h:90)         */
  int32_t inner_iterations; /* (data_structure_mod.h:118)                  */
  int32_t precision;        /* LSKUM_FP64 or LSKUM_MIXED                    */
  int32_t reorder;          /* LSKUM_REORDER_*                              */
  int32_t lds_tiles;        /* kernels that stage neighbour data in per-   */
                            /* block LDS tiles: LSKUM_TILES_AUTO or a mask */
                            /* of LSKUM_TILE_* (used when the tiles fit)   */
  int32_t tile_block;       /* points per block of the stencil kernels (=  */
                            /* per LDS tile): 64, 128, 256; 0 = default    */
  int32_t points_per_thread; /* points per thread of the pointwise state   */
                            /* update (its blocks stay 256 points): 1, 2, */
                            /* 4; 0 = default (1).  Results do not depend */
                            /* on it, bit for bit                          */
} lskum_params;

#define LSKUM_TILES_AUTO (-1)
#define LSKUM_TILE_QDERIV 1
#define LSKUM_TILE_QINNER 2
#define LSKUM_TILE_DELTA 4
#define LSKUM_TILE_FLUX 8

/* --------------------------------------------------------------- context */
typedef struct lskum_ctx lskum_ctx;

/* One case of a batch: a cloud (index into the clouds array) and its flow
 * condition; the other numerics are the batch's lskum_params. */
typedef struct lskum_case {
  int32_t cloud;
  double mach;
  double aoa_deg;
} lskum_case;

/*
 * Copies the cloud to the device, computes normals (point_normals_mod.h:17-55)
 * and the x+/x-/y+/y- sub-stencils (generate_connectivity_mod.h:5-249) on the
 * host, validates every sub-stencil, and sets the free-stream state
 * (initial_conditions_mod.h:23-37, parameter_mod.h:13-30).
 * Replaces main.cpp:59-65.
 */
int lskum_create(const lskum_cloud *cloud, const lskum_params *params,
                 lskum_ctx **out);
/*
 * Many flow conditions over several clouds in one context (no reference
 * counterpart: the reference runs one case per process).  Every stage call
 * and lskum_iterate advance all cases together, one kernel launch per stage
 * for the whole batch; each case's arithmetic is that of a context of its
 * own (bit-identical results).  params->mach / aoa_deg are ignored in favour
 * of the cases'.
 */
int lskum_create_batch(const lskum_cloud *clouds, int32_t n_clouds,
                       const lskum_case *cases, int32_t n_cases,
                       const lskum_params *params, lskum_ctx **out);
/* Number of cases and each case's point count (points may be NULL). */
int lskum_case_count(lskum_ctx *ctx, int32_t *n_cases, int32_t *points);
/* Per case, the first iteration whose residual was non-finite (0: none). */
int lskum_case_status(lskum_ctx *ctx, int32_t *failed_at);
void lskum_destroy(lskum_ctx *ctx);
const char *lskum_last_error(const lskum_ctx *ctx);

/* prim is [4][n] per case: rho, u1, u2, p (point.prim,
 * data_structure_mod.h:26). */
int lskum_set_state(lskum_ctx *ctx, const double *prim);
int lskum_get_state(lskum_ctx *ctx, double *prim);

/* prim_old <- prim (fpi_solver_mod.h:19-25). */
int lskum_snapshot(lskum_ctx *ctx);
/* Local time step delta (state_update_mod.h:448 func_delta). */
int lskum_func_delta(lskum_ctx *ctx);
/* q from prim (q_variables_mod.h:9 eval_q_variables). */
int lskum_q_variables(lskum_ctx *ctx);
/* dq, qmax, qmin (q_variables_mod.h:36 eval_q_derivatives). */
int lskum_q_derivatives(lskum_ctx *ctx);
/* n_inner passes of eval_q_inner_loop + eval_update_innerloop
 * (q_variables_mod.h:120, :215). */
int lskum_q_inner(lskum_ctx *ctx, int32_t n_inner);
/* flux_res (flux_residual_mod.h:10 cal_flux_residual). */
int lskum_flux_residual(lskum_ctx *ctx);
/* SSPRK stage rk in 1..4 (state_update_mod.h:92 state_update); returns,
 * per case, the globals sum_res_sqr, max_res, max_res_point (0-based) it
 * sets.  Any pointer may be NULL (then the call does not synchronise). */
int lskum_state_update(lskum_ctx *ctx, int32_t rk, double *sum_res_sqr,
                       double *max_res, int32_t *max_res_point);
/* Cl, Cd, Cm of shape 1 per case (compute_force_coeffs_mod.h:7
 * compute_cl_cd_cm). */
int lskum_forces(lskum_ctx *ctx, double *cl, double *cd, double *cm);

/*
 * n_iter whole iterations, each one fpi_solver() (fpi_solver_mod.h:14-150):
 * snapshot, func_delta, then per rk = 1..4 q_variables, q_derivatives,
 * inner_iterations inner passes, flux_residual, state_update.  Iterations are
 * numbered it0+1 .. it0+n_iter (q_lskum_mod.h:33).  res_new[i] =
 * sqrt(sum_res_sqr)/n and residue[i] = log10(res_new/res_old) with res_old
 * latched at it <= 2 (fpi_solver_mod.h:138-146), unless restart is set.
 * Outputs are [n_cases][n_iter]; either may be NULL.  A non-finite residual
 * fails a single-case context with LSKUM_ENONFINITE; in a batch the case
 * goes on as NaN without touching the others (no case shares points,
 * stencils or reductions with another) and lskum_case_status reports it.
 */
int lskum_iterate(lskum_ctx *ctx, int32_t n_iter, double *res_new,
                  double *residue);
/*
 * lskum_iterate without any allocation (no reference counterpart: it is for
 * callers that cannot allocate device memory -- tight serving loops,
 * harnesses that forbid hipMalloc, code around graph capture).  The same
 * kernels as lskum_iterate, launched directly (no graph is instantiated);
 * the per-iteration residual sums go to a static device array of the
 * library, LSKUM_NOALLOC_HIST_MAX doubles (1 MiB) per device, held until
 * the process ends, instead of a buffer the context allocates.  Limits and
 * rules:
 *  - n_iter x n_cases <= LSKUM_NOALLOC_HIST_MAX, else LSKUM_EINVAL and
 *    nothing runs;
 *  - res_new is required (the sums are copied into it and turned into the
 *    residuals in place); residue may be NULL;
 *  - kernel timing (lskum_set_timing) must be off, else LSKUM_EINVAL;
 *  - every context on one device shares that device's array: calls that
 *    overlap in time on one device (two host threads, or two streams) race
 *    and give wrong residuals; the state itself is per context.
 * Results equal lskum_iterate's bit for bit.
 */
#define LSKUM_NOALLOC_HIST_MAX 131072
int lskum_iterate_noalloc(lskum_ctx *ctx, int32_t n_iter, double *res_new,
                          double *residue);
/*
 * Run whole iterations as captured HIP graphs of iters_per_graph iterations
 * each (0: launch every kernel from the host, the default).  Graphs are
 * captured on first use and replayed; kernel timing (lskum_set_timing)
 * runs the launches from the host instead.
 */
int lskum_set_graph(lskum_ctx *ctx, int32_t iters_per_graph);
/* Iteration counter and res_old (per case), for restart
 * (q_lskum_mod.h:29-37, fpi_solver_mod.h:139); set_iteration gives every
 * case the same res_old.  restart != 0 keeps res_old across it <= 2. */
int lskum_get_iteration(lskum_ctx *ctx, int32_t *it, double *res_old);
int lskum_set_iteration(lskum_ctx *ctx, int32_t it, double res_old,
                        int32_t restart);

/*
 * Restart files in the reference's restart.dat layout
 * (initial_conditions_mod.h:7-21 reads it): a header "n itr res_old", then
 * nine numbers per point of which the reader skips five and takes
 * prim[1..4].  The writer fills the skipped fields with x, y, flag_1,
 * flag_2 and min_dist (the reference has no writer for this layout).  prim
 * is [4][n] in the cloud's order; 17 significant digits round-trip exactly.
 * The reader fails with LSKUM_EIO on a missing or short file and
 * LSKUM_EINVAL when the header's point count is not n.  A context resumes
 * from a file with lskum_set_state(prim) and lskum_set_iteration(itr,
 * res_old, 1), the reference's restart = 1 (q_lskum_mod.h:29-33,
 * fpi_solver_mod.h:139).
 */
int lskum_restart_write(const char *path, const lskum_cloud *cloud, int32_t it,
                        double res_old, const double *prim);
int lskum_restart_read(const char *path, int32_t n, int32_t *it, double *res_old,
                       double *prim);

/* Arrays for tests and post-processing; sizes in doubles. */
#define LSKUM_A_PRIM 0     /* [4][n]                      */
#define LSKUM_A_PRIM_OLD 1 /* [4][n]                      */
#define LSKUM_A_Q 2        /* [4][n]                      */
#define LSKUM_A_DQ 3       /* [2][4][n]: d/dx, d/dy       */
#define LSKUM_A_QM 4       /* [2][4][n]: max, min         */
#define LSKUM_A_DELTA 5    /* [n]                         */
#define LSKUM_A_FLUX_RES 6 /* [4][n]                      */
#define LSKUM_A_NORMALS 7  /* [2][n]: nx, ny              */
int lskum_get_array(lskum_ctx *ctx, int32_t which, double *host_dst);
/* Sub-stencil lists as made by generate_connectivity, caller order (of the
 * first case's cloud).
 * which: 0 xpos, 1 xneg, 2 ypos, 3 yneg.  counts[n]; members are written
 * to conn_out in point order, conn order within a point. */
int lskum_get_substencil(lskum_ctx *ctx, int32_t which, int32_t *counts,
                         int32_t *members);

/*
 * Discrete adjoint and tangent of the flux residual (no reference
 * counterpart).  The map is one cal_flux_residual evaluation
 * (flux_residual_mod.h:10) as a function of the primitive state prim, after
 * eval_q_variables, eval_q_derivatives and the inner passes
 * (q_variables_mod.h:9-228), with the local time step delta held at its
 * last func_delta value.  Both recompute q, dq and qmax/qmin from the
 * context's current state (lskum_set_state), so nothing of a forward run is
 * stored.  Single-case FP64 contexts with power = 0 (else LSKUM_EINVAL).
 *   lskum_flux_vjp: grad = J^T lam, lam and grad [4][n];
 *   lskum_flux_jvp: out = J v, v and out [4][n].
 * The limiter's switches (|del_neg| <= 1e-5, the ratio reaching 1) and the
 * qmax/qmin selections are differentiated on the branch the state takes.
 */
int lskum_flux_vjp(lskum_ctx *ctx, const double *lam, double *grad);
int lskum_flux_jvp(lskum_ctx *ctx, const double *v, double *out);
/* Gradients of Cl, Cd, Cm (compute_force_coeffs_mod.h:7-69) with respect
 * to prim, each [4][n] (only the pressure row is non-zero); any output may
 * be NULL.  Single-case contexts. */
int lskum_forces_grad(lskum_ctx *ctx, double *dcl, double *dcd, double *dcm);

/*
 * Diagnostics of the current state (per case for a batch: per-point arrays
 * case after case in each case's caller order, totals one per case; any
 * output may be NULL):
 *   lskum_compute_entropy     compute_entropy_mod.h:5-24: entropy[k] =
 *       |ln(p/rho^gamma) - ln p_inf|, total_entropy = their sum (the
 *       reference forms the sum and discards it);
 *   lskum_compute_enstrophy   compute_enstrophy_mod.h:5-76: vorticity =
 *       du2/dx - du1/dy by the (power-weighted) least squares over each
 *       point's neighbours, vorticity_sqr, total_enstrophy = its sum;
 *   lskum_stagnation_pressure stagnation_values_mod.h:8-46: values[4] =
 *       pMin, pMax, pMin/p0_inf, pMax/p0_inf of p0 = p (1 + (gamma-1)/2
 *       M^2)^(gamma/(gamma-1)), index[2] = their points (0-based, the first
 *       in point order, as the reference keeps them; it prints them with a
 *       mismatched %f);
 *   lskum_objective_J         stagnation_values_mod.h:49-80: the loss
 *       sum (p0_inf - p0)^2 / (p0_inf^2 n).  The reference multiplies an
 *       uninitialised total_p0 instead of the sum it forms; this returns
 *       the sum's value.
 * Sums are reduced in a fixed order, so results are deterministic; they
 * differ from the reference's serial sums at the rounding level.
 */
int lskum_compute_entropy(lskum_ctx *ctx, double *entropy, double *total_entropy);
int lskum_compute_enstrophy(lskum_ctx *ctx, double *vorticity, double *vorticity_sqr,
                            double *total_enstrophy);
int lskum_stagnation_pressure(lskum_ctx *ctx, double *values, int32_t *index);
int lskum_objective_J(lskum_ctx *ctx, double *J);

/*
 * Output files of a state (host code; prim [4][n] caller order, as from
 * lskum_get_state), the reference's writers with their defects fixed:
 *   lskum_solution_write  post_processing_mod.h:7-27: a header line
 *       "n it res_old", then per point "x y flag_1 rho u1 u2 p Mach entropy",
 *       one point per line at 17 significant digits (the reference writes
 *       no separator between points, at 6 digits), Mach = |u| / sqrt(gamma
 *       p / rho) (the reference takes sqrt(|u|^2 / c)); entropy may be NULL
 *       (zeros, as the reference writes before compute_entropy has run);
 *   lskum_cp_write  the Cp file of compute_force_coeffs_mod.h:20-21,54:
 *       per shape point (the first shape_points points, as the force loop
 *       takes them) "flag_2 x cp", cp = -(p - p_inf) / (rho_inf mach^2 / 2)
 *       (the reference's `fout << a, b, c;` writes only flag_2).
 * LSKUM_EIO on a file that cannot be written.
 */
int lskum_solution_write(const char *path, const lskum_cloud *cloud, int32_t it,
                         double res_old, const double *prim, const double *entropy);
int lskum_cp_write(const char *path, const lskum_cloud *cloud, const double *prim,
                   double mach);

/*
 * Shape sensitivities (no reference counterpart).
 *
 * lskum_set_coordinates moves the points of a single-case context to
 * (x, y), caller order, and recomputes what the set-up derives from them:
 * the wall and outer normals (point_normals_mod.h:17-55) and the
 * sub-stencil least-squares coefficients.  The neighbour lists, their
 * sub-stencil split (generate_connectivity_mod.h) and min_dist stay as
 * built.  A move that leaves a sub-stencil degenerate is refused
 * (LSKUM_ESTENCIL) and the context keeps its coordinates.
 *
 * The objective is J = w[0] Cl + w[1] Cd + w[2] Cm + w[3] J_p0 of the state
 * that n_iter
 * whole iterations (fpi_solver_mod.h:14-150) reach from the context's
 * current state P0, as a function of the coordinates and of P0, with J_p0
 * the stagnation-pressure loss of lskum_objective_J:
 *   lskum_shape_vjp: J, its gradient x_bar, y_bar [n] (caller order) and,
 *     if prim_bar is non-NULL, [4][n] with respect to P0, by the discrete
 *     adjoint of the n_iter iterations -- one reverse sweep that keeps the
 *     state at the start of every iteration (32 bytes per point per
 *     iteration on the device) and recomputes each stage's q-derivatives,
 *     limiter and fluxes from it;
 *   lskum_shape_jvp: J and dJ = dJ/d(x, y) . (dx, dy) + dJ/dP0 . dprim
 *     (any of dx, dy, dprim may be NULL for zero), by a forward-mode run of
 *     the same iterations in dual numbers -- independent derivative code
 *     that checks the adjoint.
 * Neither changes the context's state or iteration counter.  min_dist (the
 * limiter's epsilon) and the sub-stencil split are held fixed, and the
 * limiter's switches, the qmax/qmin and func_delta minima are
 * differentiated on the branch the state takes.  Single-case FP64 contexts
 * with power = 0 (else LSKUM_EINVAL).
 */
int lskum_set_coordinates(lskum_ctx *ctx, const double *x, const double *y);
int lskum_shape_vjp(lskum_ctx *ctx, int32_t n_iter, const double w[4], double *J,
                    double *x_bar, double *y_bar, double *prim_bar);
int lskum_shape_jvp(lskum_ctx *ctx, int32_t n_iter, const double w[4], const double *dx,
                    const double *dy, const double *dprim, double *J, double *dJ);

/*
 * Hicks-Henne design variables of a cloud whose wall points come first in
 * the generator's order (clockwise from the trailing edge, lower surface
 * first): bump j has its peak at p_j, evenly spaced over [p_lo, p_hi],
 * b_j(x) = sin(pi x^(ln 0.5/ln p_j))^width_exponent, n_per_side bumps on
 * each surface, moving it outward (variables 0..n_per_side-1 the upper
 * surface, +y; the rest the lower, -y).  Other points follow by
 * inverse-distance weights (|p - w|^-3 over the wall points), faded by
 * (1 - d/decay_radius)^2 with d their distance from the wall.  Writes
 * dy[2 n_per_side][n]: the y displacement of every point per unit
 * amplitude (x does not move).  No reference counterpart.
 */
int lskum_hicks_henne_fields(const lskum_cloud *cloud, int32_t n_per_side, double p_lo,
                             double p_hi, int32_t width_exponent, double decay_radius,
                             double *dy);

/*
 * Autotuner (no reference counterpart): builds the given cases once per
 * candidate layout -- points per block of the stencil kernels 64, 128 or
 * 256 (the block shape; one point per thread, whose neighbour list the
 * thread walks), and the flux's neighbour data staged in LDS or read
 * through L2; then, on the fastest of those six, 2 and 4 points per thread
 * of the pointwise state update -- times one captured graph of `iters`
 * whole iterations after a warm-up replay, and returns the fastest in
 * *best.  Per-point arithmetic does not depend on the layout, so the choice
 * changes speed only.  trials (up to max_trials entries, 8 candidates)
 * receives every candidate; lds_tiles is the mask that fitted.  Pass a
 * representative subset of a large batch (for example one case per cloud):
 * the set-up runs once per candidate.
 */
typedef struct lskum_tune_result {
  int32_t tile_block;
  int32_t lds_tiles;
  int32_t points_per_thread;
  double ms_per_iter;
} lskum_tune_result;
int lskum_autotune(const lskum_cloud *clouds, int32_t n_clouds,
                   const lskum_case *cases, int32_t n_cases,
                   const lskum_params *params, int32_t iters,
                   lskum_tune_result *best, lskum_tune_result *trials,
                   int32_t max_trials, int32_t *n_trials);

/* Stream the context's kernels run on (a hipStream_t). */
void *lskum_stream(lskum_ctx *ctx);
/* Device layout chosen at create: tiled = mask of the LSKUM_TILE_* kernels
 * that stage neighbour data in per-block LDS tiles; tile_max = largest tile
 * (points);
 * order = the LSKUM_REORDER_* in use.  Any pointer may be NULL. */
int lskum_layout(lskum_ctx *ctx, int32_t *tiled, int32_t *tile_max, int32_t *order);

/* Tile entries the flux stages in LDS per block (no reference counterpart):
 * as many as let a CU hold the blocks the flux kernel's registers allow.  A
 * block whose tile is larger reads its further neighbours through L2; 0
 * when the flux reads every neighbour record through L2. */
int lskum_flux_tile(lskum_ctx *ctx, int32_t *cap);

/* Average duration in ms of the named kernel over the last lskum_iterate
 * call with timing enabled (0 disables).  Kernel ids: LSKUM_K_*. */
#define LSKUM_K_FLUX 0
#define LSKUM_K_QDERIV 1
#define LSKUM_K_QINNER 2
#define LSKUM_K_STATE 3
#define LSKUM_K_DELTA 4
#define LSKUM_K_QVAR 5
int lskum_set_timing(lskum_ctx *ctx, int32_t enable);
int lskum_kernel_time(lskum_ctx *ctx, int32_t kernel, double *avg_ms,
                      int32_t *launches);

/*
 * Device copies of the leaf physics, for known-answer tests: evaluates the
 * eight flux functions (split_fluxes_mod.h:6,45,86,127 and
 * quadrant_fluxes_mod.h:6,57,108,161) on n input rows
 * (nx, ny, u1, u2, rho, pr) -> out[n][8][4], order Gxp, Gxn, Gyp, Gyn, GxI,
 * GxII, GxIII, GxIV; and qtilde_to_primitive (q_variables_mod.h:230), in
 * the form the flux kernels evaluate it, on n rows of q~[4] ->
 * prim_out[n][4] as (u1, u2, rho, pr).  mixed != 0
 * evaluates the FP32 flux functions of LSKUM_MIXED on FP32-rounded inputs.
 */
int lskum_kat_fluxes(int32_t n, const double *in6, double *out, int32_t mixed);
int lskum_kat_qtilde(int32_t n, const double *qt4, double *prim_out);

#ifdef __cplusplus
}
#endif
#endif /* LSKUM_H */
