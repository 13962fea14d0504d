/*
 * lskum_oracle.h -- CPU restatement of the reference q-LSKUM iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP path: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 * The product library never links or calls it.
 *
 * Each function restates one reference function operation by operation (same
 * operand order, same constants, same libm calls) so that, built with gcc
 * without FMA contraction, it reproduces the reference bit for bit.  Arrays
 * are 0-based and component-major: prim[r*n + i], dq[(d*4 + r)*n + i],
 * qm[(m*4 + r)*n + i] with d = 0 x, 1 y and m = 0 max, 1 min.
 */
#ifndef LSKUM_ORACLE_H
#define LSKUM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx {
  int n;
  /* cloud, main.cpp:30-58 (0-based; left/right -1 when absent) */
  double *x, *y, *nx, *ny, *min_dist;
  int *left, *right, *flag1, *flag2;
  int *conn_off, *conn;
  /* sub-stencils, generate_connectivity_mod.h:5-249 (CSR, conn order) */
  int *sub_off[4]; /* 0 xpos, 1 xneg, 2 ypos, 3 yneg */
  int *sub[4];
  /* parameters, parameter_mod.h:4-10 and data_structure_mod.h:57-121 */
  double mach, aoa, cfl, vl_const, power, euler;
  int inner_iterations, rks, restart, shapes, shape_points;
  double theta, rho_inf, pr_inf, gamma_new, pi;
  double q_init[4], q_inf[4];
  /* state, data_structure_mod.h:10-47 */
  double *prim, *prim_old, *q, *dq, *qm, *temp, *delta, *flux_res;
  double *res_pt; /* per-point (dU1)^2 of the last state_update */
  /* globals set by state_update / fpi_solver / compute_cl_cd_cm */
  double sum_res_sqr, max_res, res_old, res_new, residue;
  int max_res_point, it;
  double cl, cd, cm;
  int nthreads;
} orc_ctx;

/* cloud arrays are copied */
orc_ctx *orc_create(int n, const double *x, const double *y, const int *left,
                    const int *right, const int *flag1, const int *flag2,
                    const double *min_dist, const int *conn_off,
                    const int *conn);
void orc_destroy(orc_ctx *c);
void orc_set_params(orc_ctx *c, double mach, double aoa_deg, double cfl,
                    double vl_const, double power, int inner_iterations);
void orc_set_threads(orc_ctx *c, int nthreads);

void orc_compute_normals(orc_ctx *c);
void orc_generate_connectivity(orc_ctx *c);
void orc_setup_case_parameters(orc_ctx *c);
void orc_initial_conditions(orc_ctx *c);

void orc_func_delta(orc_ctx *c);
void orc_eval_q_variables(orc_ctx *c);
void orc_eval_q_derivatives(orc_ctx *c);
void orc_eval_q_inner_loop(orc_ctx *c);
void orc_eval_update_innerloop(orc_ctx *c);
void orc_cal_flux_residual(orc_ctx *c);
void orc_state_update(orc_ctx *c, int rk);
void orc_compute_cl_cd_cm(orc_ctx *c);
void orc_fpi_solver(orc_ctx *c, int it);
/* iterations it+1 .. it+n_iter; res_new/residue per iteration (may be NULL) */
void orc_q_lskum(orc_ctx *c, int n_iter, double *res_new, double *residue);

/* leaf physics */
void orc_qtilde_to_primitive(const double qt[4], double *u1, double *u2,
                             double *rho, double *pr);
/* which: 0 Gxp 1 Gxn 2 Gyp 3 Gyn 4 GxI 5 GxII 6 GxIII 7 GxIV */
void orc_flux(int which, double nx, double ny, double u1, double u2,
              double rho, double pr, double G[4]);
void orc_venkat_limiter(const orc_ctx *c, const double qtilde[4],
                        double phi[4], int k);

/* diagnostics (compute_entropy_mod.h, compute_enstrophy_mod.h,
 * stagnation_values_mod.h); per-point outputs may be NULL */
double orc_compute_entropy(const orc_ctx *c, double *entropy);
double orc_compute_enstrophy(const orc_ctx *c, double *vorticity, double *vorticity_sqr);
void orc_stagnation_pressure(const orc_ctx *c, double out[4], int idx[2]);
double orc_objective_J(const orc_ctx *c);

void orc_scalars(const orc_ctx *c, double out[11]);
void orc_set_iteration(orc_ctx *c, int it, double res_old, int restart);

/* 0 prim 1 prim_old 2 q 3 dq 4 qm 5 delta 6 flux_res 7 nx|ny 8 res_pt */
double *orc_array(orc_ctx *c, int which, int *len);
int orc_sub_count(const orc_ctx *c, int which, int i);
const int *orc_sub_list(const orc_ctx *c, int which, int i);

#ifdef __cplusplus
}
#endif
#endif
