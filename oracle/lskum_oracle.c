/*
 * lskum_oracle.c -- CPU restatement of the reference q-LSKUM iteration.
 *
 * TEST INFRASTRUCTURE ONLY (see lskum_oracle.h).  Every function names the
 * reference file:line it restates.  Build: gcc -O2 -std=c99
 * -ffp-contract=off [-fopenmp]; per-point loops may run on several threads,
 * reductions always run serially in point order, so results do not depend on
 * the thread count.
 */
#include "lskum_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define Q(c, r, i) ((c)->q[(r) * (c)->n + (i)])
#define PRIM(c, r, i) ((c)->prim[(r) * (c)->n + (i)])
#define PRIMO(c, r, i) ((c)->prim_old[(r) * (c)->n + (i)])
#define DQ(c, d, r, i) ((c)->dq[((d) * 4 + (r)) * (c)->n + (i)])
#define QM(c, m, r, i) ((c)->qm[((m) * 4 + (r)) * (c)->n + (i)])
#define TMP(c, d, r, i) ((c)->temp[((d) * 4 + (r)) * (c)->n + (i)])
#define FRES(c, r, i) ((c)->flux_res[(r) * (c)->n + (i)])

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) abort();
  return p;
}

orc_ctx *orc_create(int n, const double *x, const double *y, const int *left,
                    const int *right, const int *flag1, const int *flag2,
                    const double *min_dist, const int *conn_off,
                    const int *conn) {
  orc_ctx *c = (orc_ctx *)xcalloc(1, sizeof(orc_ctx));
  int nnz = conn_off[n];
  c->n = n;
  c->x = xcalloc(n, sizeof(double));
  c->y = xcalloc(n, sizeof(double));
  c->nx = xcalloc(n, sizeof(double));
  c->ny = xcalloc(n, sizeof(double));
  c->min_dist = xcalloc(n, sizeof(double));
  c->left = xcalloc(n, sizeof(int));
  c->right = xcalloc(n, sizeof(int));
  c->flag1 = xcalloc(n, sizeof(int));
  c->flag2 = xcalloc(n, sizeof(int));
  c->conn_off = xcalloc(n + 1, sizeof(int));
  c->conn = xcalloc(nnz, sizeof(int));
  memcpy(c->x, x, n * sizeof(double));
  memcpy(c->y, y, n * sizeof(double));
  memcpy(c->min_dist, min_dist, n * sizeof(double));
  memcpy(c->left, left, n * sizeof(int));
  memcpy(c->right, right, n * sizeof(int));
  memcpy(c->flag1, flag1, n * sizeof(int));
  memcpy(c->flag2, flag2, n * sizeof(int));
  memcpy(c->conn_off, conn_off, (n + 1) * sizeof(int));
  memcpy(c->conn, conn, nnz * sizeof(int));
  for (int s = 0; s < 4; s++) {
    c->sub_off[s] = xcalloc(n + 1, sizeof(int));
    c->sub[s] = xcalloc(nnz, sizeof(int));
  }
  c->prim = xcalloc(4 * (size_t)n, sizeof(double));
  c->prim_old = xcalloc(4 * (size_t)n, sizeof(double));
  c->q = xcalloc(4 * (size_t)n, sizeof(double));
  c->dq = xcalloc(8 * (size_t)n, sizeof(double));
  c->qm = xcalloc(8 * (size_t)n, sizeof(double));
  c->temp = xcalloc(8 * (size_t)n, sizeof(double));
  c->delta = xcalloc(n, sizeof(double));
  c->flux_res = xcalloc(4 * (size_t)n, sizeof(double));
  c->res_pt = xcalloc(n, sizeof(double));
  /* main.cpp:53-56: shape points are those with flag_2 > 0 */
  c->shapes = 1;
  for (int i = 0; i < n; i++) {
    if (flag2[i] > 0) c->shape_points++;
    if (flag2[i] > c->shapes) c->shapes = flag2[i];
  }
  /* parameter_mod.h:6-10, data_structure_mod.h:69-118 defaults */
  c->rho_inf = 1.0;
  c->pr_inf = 1.0 / 1.4;
  c->gamma_new = 1.4;
  c->pi = 4 * atan(1.0);
  c->mach = 0.63;
  c->aoa = 2.0;
  c->cfl = 0.1;
  c->vl_const = 150;
  c->power = 0;
  c->inner_iterations = 3;
  c->rks = 4;
  c->euler = 1.0;
  c->restart = 0;
  c->nthreads = 1;
  return c;
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  free(c->x); free(c->y); free(c->nx); free(c->ny); free(c->min_dist);
  free(c->left); free(c->right); free(c->flag1); free(c->flag2);
  free(c->conn_off); free(c->conn);
  for (int s = 0; s < 4; s++) { free(c->sub_off[s]); free(c->sub[s]); }
  free(c->prim); free(c->prim_old); free(c->q); free(c->dq); free(c->qm);
  free(c->temp); free(c->delta); free(c->flux_res); free(c->res_pt);
  free(c);
}

void orc_set_params(orc_ctx *c, double mach, double aoa_deg, double cfl,
                    double vl_const, double power, int inner_iterations) {
  c->mach = mach;
  c->aoa = aoa_deg;
  c->cfl = cfl;
  c->vl_const = vl_const;
  c->power = power;
  c->inner_iterations = inner_iterations;
}

void orc_set_threads(orc_ctx *c, int nthreads) {
  c->nthreads = nthreads > 0 ? nthreads : 1;
}

#ifdef _OPENMP
#define PAR_FOR _Pragma("omp parallel for schedule(static) num_threads(c->nthreads)")
#else
#define PAR_FOR
#endif

/* point_normals_mod.h:17-55 */
void orc_compute_normals(orc_ctx *c) {
  for (int i = 0; i < c->n; i++) {
    if (c->flag1[i] == 0 || c->flag1[i] == 2) {
      int m = i, l = c->left[m], r = c->right[m];
      double lx = c->x[l], ly = c->y[l];
      double mx = c->x[m], my = c->y[m];
      double rx = c->x[r], ry = c->y[r];
      double nx1 = my - ly, nx2 = ry - my;
      double ny1 = mx - lx, ny2 = rx - mx;
      double nx = 0.5 * (nx1 + nx2);
      double ny = 0.5 * (ny1 + ny2);
      double det = sqrt(nx * nx + ny * ny);
      nx = -nx / det;
      ny = ny / det;
      c->nx[m] = nx;
      c->ny[m] = ny;
    } else {
      c->nx[i] = 0;
      c->ny[i] = 1;
    }
  }
}

/*
 * generate_connectivity_mod.h:221-249 with get_interior_neighbours :5-83,
 * get_wall_boundary_neighbours :91-153, get_outer_boundary_neighbours
 * :155-219.  Wall: y- gets every neighbour; outer: y+ gets every neighbour.
 */
void orc_generate_connectivity(orc_ctx *c) {
  int cnt[4] = {0, 0, 0, 0};
  for (int i = 0; i < c->n; i++) {
    double xi = c->x[i], yi = c->y[i];
    double nx = c->nx[i], ny = c->ny[i];
    double tx = ny, ty = -nx;
    int f = c->flag1[i];
    for (int s = 0; s < 4; s++) c->sub_off[s][i] = cnt[s];
    for (int r = c->conn_off[i]; r < c->conn_off[i + 1]; r++) {
      int nbh = c->conn[r];
      double xk = c->x[nbh], yk = c->y[nbh];
      double delx = xk - xi, dely = yk - yi;
      double dels = delx * tx + dely * ty;
      double deln = delx * nx + dely * ny;
      if (dels <= 0.0) c->sub[0][cnt[0]++] = nbh;
      if (dels >= 0.0) c->sub[1][cnt[1]++] = nbh;
      if (f == 0) {
        c->sub[3][cnt[3]++] = nbh;
      } else if (f == 1) {
        if (deln <= 0.0) c->sub[2][cnt[2]++] = nbh;
        if (deln >= 0.0) c->sub[3][cnt[3]++] = nbh;
      } else {
        c->sub[2][cnt[2]++] = nbh;
      }
    }
  }
  for (int s = 0; s < 4; s++) c->sub_off[s][c->n] = cnt[s];
}

/* parameter_mod.h:13-30 */
void orc_setup_case_parameters(orc_ctx *c) {
  c->theta = c->aoa * c->pi / 180;
  c->q_init[0] = c->rho_inf;
  c->q_init[1] = c->mach * cos(c->theta);
  c->q_init[2] = c->mach * sin(c->theta);
  c->q_init[3] = c->pr_inf;
  c->q_inf[0] = c->rho_inf;
  c->q_inf[1] = c->mach * cos(c->theta);
  c->q_inf[2] = c->mach * sin(c->theta);
  c->q_inf[3] = c->pr_inf;
}

/* initial_conditions_mod.h:23-37 (restart == 0 branch) */
void orc_initial_conditions(orc_ctx *c) {
  orc_setup_case_parameters(c);
  for (int k = 0; k < c->n; k++)
    for (int r = 0; r < 4; r++) PRIM(c, r, k) = c->q_init[r];
}

/* state_update_mod.h:448-496 */
void orc_func_delta(orc_ctx *c) {
  PAR_FOR
  for (int i = 0; i < c->n; i++) {
    double min_delt = 1.0;
    for (int r = c->conn_off[i]; r < c->conn_off[i + 1]; r++) {
      int k = c->conn[r];
      double rho = PRIM(c, 0, k), u1 = PRIM(c, 1, k), u2 = PRIM(c, 2, k),
             pr = PRIM(c, 3, k);
      double x_i = c->x[i], y_i = c->y[i], x_k = c->x[k], y_k = c->y[k];
      double dist = (x_k - x_i) * (x_k - x_i) + (y_k - y_i) * (y_k - y_i);
      dist = sqrt(dist);
      double mod_u = sqrt(u1 * u1 + u2 * u2);
      double delta_t = dist / (mod_u + 3.0 * sqrt(pr / rho));
      delta_t = c->cfl * delta_t;
      if (min_delt > delta_t) min_delt = delta_t;
    }
    c->delta[i] = min_delt;
  }
}

/* q_variables_mod.h:9-34 */
void orc_eval_q_variables(orc_ctx *c) {
  PAR_FOR
  for (int k = 0; k < c->n; k++) {
    double rho = PRIM(c, 0, k), u1 = PRIM(c, 1, k), u2 = PRIM(c, 2, k),
           pr = PRIM(c, 3, k);
    double beta = 0.5 * rho / pr;
    Q(c, 0, k) = log(rho) + (log(beta) * 2.5) - beta * (u1 * u1 + u2 * u2);
    double two_times_beta = 2.0 * beta;
    Q(c, 1, k) = two_times_beta * u1;
    Q(c, 2, k) = two_times_beta * u2;
    Q(c, 3, k) = -two_times_beta;
  }
}

/* q_variables_mod.h:36-118 */
void orc_eval_q_derivatives(orc_ctx *c) {
  PAR_FOR
  for (int i = 0; i < c->n; i++) {
    double x_i = c->x[i], y_i = c->y[i];
    double sum_delx_sqr = 0., sum_dely_sqr = 0., sum_delx_dely = 0.;
    double sum_delx_delq[4] = {0, 0, 0, 0}, sum_dely_delq[4] = {0, 0, 0, 0};
    for (int r = 0; r < 4; r++) {
      QM(c, 0, r, i) = Q(c, r, i);
      QM(c, 1, r, i) = Q(c, r, i);
    }
    for (int kk = c->conn_off[i]; kk < c->conn_off[i + 1]; kk++) {
      int nbh = c->conn[kk];
      for (int r = 0; r < 4; r++) {
        if (Q(c, r, nbh) > QM(c, 0, r, i)) QM(c, 0, r, i) = Q(c, r, nbh);
        if (Q(c, r, nbh) < QM(c, 1, r, i)) QM(c, 1, r, i) = Q(c, r, nbh);
      }
      double x_k = c->x[nbh], y_k = c->y[nbh];
      double delx = x_k - x_i, dely = y_k - y_i;
      double dist = sqrt(delx * delx + dely * dely);
      double weights = pow(dist, c->power);
      sum_delx_sqr = sum_delx_sqr + delx * delx * weights;
      sum_dely_sqr = sum_dely_sqr + dely * dely * weights;
      sum_delx_dely = sum_delx_dely + delx * dely * weights;
      for (int r = 0; r < 4; r++) {
        sum_delx_delq[r] =
            sum_delx_delq[r] + weights * delx * (Q(c, r, nbh) - Q(c, r, i));
        sum_dely_delq[r] =
            sum_dely_delq[r] + weights * dely * (Q(c, r, nbh) - Q(c, r, i));
      }
    }
    double det = sum_delx_sqr * sum_dely_sqr - sum_delx_dely * sum_delx_dely;
    double one_by_det = 1.0 / det;
    for (int r = 0; r < 4; r++) {
      DQ(c, 0, r, i) = (sum_delx_delq[r] * sum_dely_sqr -
                        sum_dely_delq[r] * sum_delx_dely) * one_by_det;
      DQ(c, 1, r, i) = (sum_dely_delq[r] * sum_delx_sqr -
                        sum_delx_delq[r] * sum_delx_dely) * one_by_det;
    }
  }
}

/* q_variables_mod.h:120-213 */
void orc_eval_q_inner_loop(orc_ctx *c) {
  PAR_FOR
  for (int i = 0; i < c->n; i++) {
    double x_i = c->x[i], y_i = c->y[i];
    double sum_delx_sqr = 0., sum_dely_sqr = 0., sum_delx_dely = 0.;
    double sx[4] = {0., 0., 0., 0.}, sy[4] = {0., 0., 0., 0.};
    double qi[4];
    for (int r = 0; r < 4; r++) qi[r] = Q(c, r, i);
    for (int kk = c->conn_off[i]; kk < c->conn_off[i + 1]; kk++) {
      int nbh = c->conn[kk];
      double x_k = c->x[nbh], y_k = c->y[nbh];
      double delx = x_k - x_i, dely = y_k - y_i;
      double dist = sqrt(delx * delx + dely * dely);
      double weights = pow(dist, c->power);
      sum_delx_sqr = sum_delx_sqr + delx * delx * weights;
      sum_dely_sqr = sum_dely_sqr + dely * dely * weights;
      sum_delx_dely = sum_delx_dely + delx * dely * weights;
      for (int r = 0; r < 4; r++) {
        double temp1 =
            qi[r] - 0.5 * (delx * DQ(c, 0, r, i) + dely * DQ(c, 1, r, i));
        double temp2 = Q(c, r, nbh) -
                       0.5 * (delx * DQ(c, 0, r, nbh) + dely * DQ(c, 1, r, nbh));
        sx[r] = sx[r] + (weights * delx * (temp2 - temp1));
        sy[r] = sy[r] + (weights * dely * (temp2 - temp1));
      }
    }
    double det = sum_delx_sqr * sum_dely_sqr - sum_delx_dely * sum_delx_dely;
    double one_by_det = 1.0 / det;
    for (int r = 0; r < 4; r++) {
      TMP(c, 0, r, i) =
          one_by_det * (sx[r] * sum_dely_sqr - sy[r] * sum_delx_dely);
      TMP(c, 1, r, i) =
          one_by_det * (sy[r] * sum_delx_sqr - sx[r] * sum_delx_dely);
    }
  }
}

/* q_variables_mod.h:215-228 */
void orc_eval_update_innerloop(orc_ctx *c) {
  memcpy(c->dq, c->temp, 8 * (size_t)c->n * sizeof(double));
}

/* q_variables_mod.h:230-253 */
void orc_qtilde_to_primitive(const double qt[4], double *u1, double *u2,
                             double *rho, double *pr) {
  const double gamma_new = 1.4;
  double q1 = qt[0], q2 = qt[1], q3 = qt[2], q4 = qt[3];
  double beta = -q4 * 0.5;
  double temp = 0.5 / beta;
  *u1 = q2 * temp;
  *u2 = q3 * temp;
  double temp1 = q1 + beta * (*u1 * *u1 + *u2 * *u2);
  double temp2 = temp1 - (log(beta) / (gamma_new - 1));
  *rho = exp(temp2);
  *pr = *rho * temp;
}

/*
 * limiters_mod.h:55-104.  phi keeps its previous value when del_neg is NaN,
 * as in the reference (neither branch of :66/:70 is taken).
 */
void orc_venkat_limiter(const orc_ctx *c, const double qtilde[4],
                        double phi[4], int k) {
  for (int r = 0; r < 4; r++) {
    double q = Q(c, r, k);
    double del_neg = qtilde[r] - q, del_pos = 0.0;
    if (fabs(del_neg) <= 10e-6) {
      phi[r] = 1.0;
    } else if (fabs(del_neg) > 10e-6) {
      if (del_neg > 0)
        del_pos = QM(c, 0, r, k) - q;
      else if (del_neg < 0)
        del_pos = QM(c, 1, r, k) - q;
      double epsi = c->vl_const * c->min_dist[k];
      epsi = pow(epsi, 3);
      double num = (del_pos * del_pos) + (epsi * epsi);
      num = num * del_neg + 2.0 * del_neg * del_neg * del_pos;
      double den = del_pos * del_pos + 2.0 * del_neg * del_neg;
      den = den + del_neg * del_pos + epsi * epsi;
      den = den * del_neg;
      double temp = num / den;
      if (temp < 1)
        phi[r] = temp;
      else
        phi[r] = 1;
    }
  }
}

/* split_fluxes_mod.h:6-43 (sgn = +1, Gxp) and :45-84 (sgn = -1, Gxn) */
static void flux_Gx(double G[4], double nx, double ny, double u1, double u2,
                    double rho, double pr, int pos, double pi) {
  double tx = ny, ty = -nx;
  double ut = u1 * tx + u2 * ty;
  double un = u1 * nx + u2 * ny;
  double beta = 0.5 * rho / pr;
  double S1 = ut * sqrt(beta);
  double B1 = 0.5 * exp(-S1 * S1) / sqrt(pi * beta);
  double pr_by_rho = pr / rho;
  double u_sqr = ut * ut + un * un;
  double temp1, temp2;
  if (pos) {
    double A1pos = 0.5 * (1.0 + erf(S1));
    G[0] = rho * (ut * A1pos + B1);
    temp1 = pr_by_rho + ut * ut;
    temp2 = temp1 * A1pos + ut * B1;
    G[1] = rho * temp2;
    temp1 = ut * un * A1pos + un * B1;
    G[2] = rho * temp1;
    temp1 = (7.0 * pr_by_rho) + u_sqr;
    temp2 = 0.5 * ut * temp1 * A1pos;
    temp1 = (6.0 * pr_by_rho) + u_sqr;
    G[3] = rho * (temp2 + 0.5 * temp1 * B1);
  } else {
    double A1neg = 0.5 * (1.0 - erf(S1));
    G[0] = rho * (ut * A1neg - B1);
    temp1 = pr_by_rho + ut * ut;
    temp2 = temp1 * A1neg - ut * B1;
    G[1] = rho * temp2;
    temp1 = ut * un * A1neg - un * B1;
    G[2] = rho * temp1;
    temp1 = (7.0 * pr_by_rho) + u_sqr;
    temp2 = 0.5 * ut * temp1 * A1neg;
    temp1 = (6.0 * pr_by_rho) + u_sqr;
    G[3] = rho * (temp2 - 0.5 * temp1 * B1);
  }
}

/* split_fluxes_mod.h:86-125 (Gyp) and :127-164 (Gyn) */
static void flux_Gy(double G[4], double nx, double ny, double u1, double u2,
                    double rho, double pr, int pos, double pi) {
  double tx = ny, ty = -nx;
  double ut = u1 * tx + u2 * ty;
  double un = u1 * nx + u2 * ny;
  double beta = 0.5 * rho / pr;
  double S2 = un * sqrt(beta);
  double B2 = 0.5 * exp(-S2 * S2) / sqrt(pi * beta);
  double pr_by_rho = pr / rho;
  double u_sqr = ut * ut + un * un;
  double temp1, temp2;
  if (pos) {
    double A2pos = 0.5 * (1.0 + erf(S2));
    G[0] = rho * (un * A2pos + B2);
    temp1 = pr_by_rho + un * un;
    temp2 = temp1 * A2pos + un * B2;
    G[2] = rho * temp2;
    temp1 = ut * un * A2pos + ut * B2;
    G[1] = rho * temp1;
    temp1 = (7.0 * pr_by_rho) + u_sqr;
    temp2 = 0.5 * un * temp1 * A2pos;
    temp1 = (6.0 * pr_by_rho) + u_sqr;
    G[3] = rho * (temp2 + 0.5 * temp1 * B2);
  } else {
    double A2neg = 0.5 * (1.0 - erf(S2));
    G[0] = rho * (un * A2neg - B2);
    temp1 = pr_by_rho + un * un;
    temp2 = temp1 * A2neg - un * B2;
    G[2] = rho * temp2;
    temp1 = ut * un * A2neg - ut * B2;
    G[1] = rho * temp1;
    temp1 = (7.0 * pr_by_rho) + u_sqr;
    temp2 = 0.5 * un * temp1 * A2neg;
    temp1 = (6.0 * pr_by_rho) + u_sqr;
    G[3] = rho * (temp2 - 0.5 * temp1 * B2);
  }
}

/*
 * quadrant_fluxes_mod.h:6-55 GxI (s1 -, s2 -), :57-106 GxII (s1 +, s2 -),
 * :108-159 GxIII (s1 +, s2 +), :161-210 GxIV (s1 -, s2 +).
 */
static void flux_quad(double G[4], double nx, double ny, double u1, double u2,
                      double rho, double pr, int s1pos, int s2pos, double pi) {
  double tx = ny, ty = -nx;
  double ut = u1 * tx + u2 * ty;
  double un = u1 * nx + u2 * ny;
  double beta = 0.5 * rho / pr;
  double S1 = ut * sqrt(beta);
  double S2 = un * sqrt(beta);
  double B1 = 0.5 * exp(-S1 * S1) / sqrt(pi * beta);
  double B2 = 0.5 * exp(-S2 * S2) / sqrt(pi * beta);
  double A1 = s1pos ? 0.5 * (1.0 + erf(S1)) : 0.5 * (1.0 - erf(S1));
  double A2 = s2pos ? 0.5 * (1.0 + erf(S2)) : 0.5 * (1.0 - erf(S2));
  double pr_by_rho = pr / rho;
  double u_sqr = ut * ut + un * un;
  double temp1, temp2, temp3, temp4;
  /* ut*A1 + B1 (s1 +) or ut*A1 - B1 (s1 -) */
  double a1b1 = s1pos ? ut * A1 + B1 : ut * A1 - B1;
  G[0] = rho * A2 * a1b1;
  temp1 = pr_by_rho + ut * ut;
  temp2 = s1pos ? temp1 * A1 + ut * B1 : temp1 * A1 - ut * B1;
  G[1] = rho * A2 * temp2;
  temp1 = a1b1;
  temp2 = s2pos ? un * A2 + B2 : un * A2 - B2;
  G[2] = rho * temp1 * temp2;
  temp1 = (7.0 * pr_by_rho) + u_sqr;
  temp2 = 0.5 * ut * temp1 * A1;
  temp1 = (6.0 * pr_by_rho) + u_sqr;
  temp3 = 0.5 * B1 * temp1;
  temp1 = a1b1;
  temp4 = 0.5 * rho * un * B2 * temp1;
  double inner = s1pos ? temp2 + temp3 : temp2 - temp3;
  G[3] = s2pos ? rho * A2 * inner + temp4 : rho * A2 * inner - temp4;
}

void orc_flux(int which, double nx, double ny, double u1, double u2,
              double rho, double pr, double G[4]) {
  const double pi = 4 * atan(1.0);
  switch (which) {
    case 0: flux_Gx(G, nx, ny, u1, u2, rho, pr, 1, pi); break;
    case 1: flux_Gx(G, nx, ny, u1, u2, rho, pr, 0, pi); break;
    case 2: flux_Gy(G, nx, ny, u1, u2, rho, pr, 1, pi); break;
    case 3: flux_Gy(G, nx, ny, u1, u2, rho, pr, 0, pi); break;
    case 4: flux_quad(G, nx, ny, u1, u2, rho, pr, 0, 0, pi); break;
    case 5: flux_quad(G, nx, ny, u1, u2, rho, pr, 1, 0, pi); break;
    case 6: flux_quad(G, nx, ny, u1, u2, rho, pr, 1, 1, pi); break;
    default: flux_quad(G, nx, ny, u1, u2, rho, pr, 0, 1, pi); break;
  }
}

/*
 * One sub-stencil flux derivative: interior_dGx_pos (interior_fluxes_mod.h:
 * 12-132) and its nine siblings (:134, :237, :333; wall_fluxes_mod.h:11, 103,
 * 195; outer_fluxes_mod.h:11, 103, 195) differ only in the list walked, the
 * flux called and the final 2x2 solve (x-form :110-113 / y-form :325-328).
 * phi_i / phi_k persist across the neighbour loop as in the reference.
 */
static void orc_dG(const orc_ctx *c, int i, int sub, int flux, int yform,
                   double G[4]) {
  double sum_delx_sqr = 0.0, sum_dely_sqr = 0.0, sum_delx_dely = 0.0;
  double sum_delx_delf[4] = {0, 0, 0, 0}, sum_dely_delf[4] = {0, 0, 0, 0};
  double qtilde_i[4], qtilde_k[4], phi_i[4] = {0, 0, 0, 0},
         phi_k[4] = {0, 0, 0, 0};
  double G_i[4], G_k[4];
  double x_i = c->x[i], y_i = c->y[i];
  double nx = c->nx[i], ny = c->ny[i];
  double tx = ny, ty = -nx;
  const int *list = c->sub[sub];
  for (int j = c->sub_off[sub][i]; j < c->sub_off[sub][i + 1]; j++) {
    int k = list[j];
    double x_k = c->x[k], y_k = c->y[k];
    double delx = x_k - x_i, dely = y_k - y_i;
    double dels = delx * tx + dely * ty;
    double deln = delx * nx + dely * ny;
    double dist = sqrt(dels * dels + deln * deln);
    double weights = pow(dist, c->power);
    double dels_weights = dels * weights;
    double deln_weights = deln * weights;
    sum_delx_sqr = sum_delx_sqr + dels * dels_weights;
    sum_dely_sqr = sum_dely_sqr + deln * deln_weights;
    sum_delx_dely = sum_delx_dely + dels * deln_weights;
    for (int r = 0; r < 4; r++) {
      qtilde_i[r] = Q(c, r, i) - 0.5 * (delx * DQ(c, 0, r, i) + dely * DQ(c, 1, r, i));
      qtilde_k[r] = Q(c, r, k) - 0.5 * (delx * DQ(c, 0, r, k) + dely * DQ(c, 1, r, k));
    }
    orc_venkat_limiter(c, qtilde_i, phi_i, i);
    orc_venkat_limiter(c, qtilde_k, phi_k, k);
    for (int r = 0; r < 4; r++) {
      qtilde_i[r] = Q(c, r, i) - 0.5 * phi_i[r] * (delx * DQ(c, 0, r, i) + dely * DQ(c, 1, r, i));
      qtilde_k[r] = Q(c, r, k) - 0.5 * phi_k[r] * (delx * DQ(c, 0, r, k) + dely * DQ(c, 1, r, k));
    }
    double u1, u2, rho, pr;
    orc_qtilde_to_primitive(qtilde_i, &u1, &u2, &rho, &pr);
    orc_flux(flux, nx, ny, u1, u2, rho, pr, G_i);
    orc_qtilde_to_primitive(qtilde_k, &u1, &u2, &rho, &pr);
    orc_flux(flux, nx, ny, u1, u2, rho, pr, G_k);
    for (int r = 0; r < 4; r++) {
      sum_delx_delf[r] = sum_delx_delf[r] + (G_k[r] - G_i[r]) * dels_weights;
      sum_dely_delf[r] = sum_dely_delf[r] + (G_k[r] - G_i[r]) * deln_weights;
    }
  }
  double det = sum_delx_sqr * sum_dely_sqr - sum_delx_dely * sum_delx_dely;
  double one_by_det = 1.0 / det;
  for (int r = 0; r < 4; r++) {
    if (!yform)
      G[r] = (sum_delx_delf[r] * sum_dely_sqr - sum_dely_delf[r] * sum_delx_dely) * one_by_det;
    else
      G[r] = (sum_dely_delf[r] * sum_delx_sqr - sum_delx_delf[r] * sum_delx_dely) * one_by_det;
  }
}

/* flux_residual_mod.h:10-89 */
void orc_cal_flux_residual(orc_ctx *c) {
  PAR_FOR
  for (int i = 0; i < c->n; i++) {
    double Gxp[4], Gxn[4], Gyp[4], Gyn[4];
    if (c->flag1[i] == 0) {
      orc_dG(c, i, 0, 5, 0, Gxp); /* wall_dGx_pos: GxII */
      orc_dG(c, i, 1, 4, 0, Gxn); /* wall_dGx_neg: GxI  */
      orc_dG(c, i, 3, 3, 1, Gyn); /* wall_dGy_neg: Gyn  */
      for (int j = 0; j < 4; j++)
        FRES(c, j, i) = (2.0 * c->delta[i]) * (Gxp[j] + Gxn[j] + Gyn[j]);
    } else if (c->flag1[i] == 1) {
      orc_dG(c, i, 0, 0, 0, Gxp);
      orc_dG(c, i, 1, 1, 0, Gxn);
      orc_dG(c, i, 2, 2, 1, Gyp);
      orc_dG(c, i, 3, 3, 1, Gyn);
      for (int j = 0; j < 4; j++)
        FRES(c, j, i) = c->delta[i] * (Gxp[j] + Gxn[j] + Gyp[j] + Gyn[j]);
    } else {
      orc_dG(c, i, 0, 6, 0, Gxp); /* outer_dGx_pos: GxIII */
      orc_dG(c, i, 1, 7, 0, Gxn); /* outer_dGx_neg: GxIV  */
      orc_dG(c, i, 2, 2, 1, Gyp); /* outer_dGy_pos: Gyp   */
      for (int j = 0; j < 4; j++)
        FRES(c, j, i) = c->delta[i] * (Gxp[j] + Gxn[j] + Gyp[j]);
    }
  }
}

/* state_update_mod.h:18-33 */
static void primitive_to_conserved(const double prim[4], double nx, double ny,
                                   double U[4]) {
  double rho = prim[0];
  U[0] = rho;
  double temp1 = rho * prim[1];
  double temp2 = rho * prim[2];
  U[3] = 2.5 * prim[3] + 0.5 * (temp1 * temp1 + temp2 * temp2) / rho;
  U[1] = temp1 * ny - temp2 * nx;
  U[2] = temp1 * nx + temp2 * ny;
}

/* state_update_mod.h:35-90 */
static void conserved_vector_Ubar(const orc_ctx *c, const double prim[4],
                                  double Ubar[4], double nx, double ny) {
  const double rho_inf = c->rho_inf, pr_inf = c->pr_inf, gamma_new = 1.4;
  const double pi = c->pi;
  double u1_inf = c->q_inf[1], u2_inf = c->q_inf[2];
  double tx = ny, ty = -nx;
  double u1_inf_rot = u1_inf * tx + u2_inf * ty;
  double u2_inf_rot = u1_inf * nx + u2_inf * ny;
  double temp1 = (u1_inf_rot * u1_inf_rot + u2_inf_rot * u2_inf_rot);
  double e_inf = pr_inf / (rho_inf * (gamma_new - 1.0)) + 0.5 * (temp1);
  double beta = (0.5 * rho_inf) / pr_inf;
  double S2 = u2_inf_rot * sqrt(beta);
  double B2_inf = exp(-S2 * S2) / (2.0 * sqrt(pi * beta));
  double A2n_inf = 0.5 * (1.0 - erf(S2));
  double rho = prim[0], u1 = prim[1], u2 = prim[2], pr = prim[3];
  double u1_rot = u1 * tx + u2 * ty;
  double u2_rot = u1 * nx + u2 * ny;
  temp1 = (u1_rot * u1_rot + u2_rot * u2_rot);
  double e = pr / (rho * (gamma_new - 1.0)) + 0.5 * (temp1);
  beta = (rho) / (2.0 * pr);
  S2 = u2_rot * sqrt(beta);
  double B2 = exp(-S2 * S2) / (2.0 * sqrt(pi * beta));
  double A2p = 0.5 * (1.0 + erf(S2));
  Ubar[0] = (rho_inf * A2n_inf) + (rho * A2p);
  Ubar[1] = (rho_inf * u1_inf_rot * A2n_inf) + (rho * u1_rot * A2p);
  temp1 = rho_inf * (u2_inf_rot * A2n_inf - B2_inf);
  double temp2 = rho * (u2_rot * A2p + B2);
  Ubar[2] = temp1 + temp2;
  temp1 = (rho_inf * A2n_inf * e_inf - 0.5 * rho_inf * u2_inf_rot * B2_inf);
  temp2 = (rho * A2p * e + 0.5 * rho * u2_rot * B2);
  Ubar[3] = temp1 + temp2;
}

/* state_update_mod.h:92-283 */
void orc_state_update(orc_ctx *c, int rk) {
  const double obt = 1.0 / 3.0;
  const double tbt = 2.0 / 3.0;
  PAR_FOR
  for (int k = 0; k < c->n; k++) {
    double nx = c->nx[k], ny = c->ny[k];
    double p1[4], p2[4], U[4], U_old[4];
    for (int r = 0; r < 4; r++) {
      p1[r] = PRIM(c, r, k);
      p2[r] = PRIMO(c, r, k);
    }
    int f = c->flag1[k];
    if (f == 0 || f == 1) {
      primitive_to_conserved(p1, nx, ny, U);
      primitive_to_conserved(p2, nx, ny, U_old);
    } else {
      conserved_vector_Ubar(c, p1, U, nx, ny);
      conserved_vector_Ubar(c, p2, U_old, nx, ny);
    }
    double temp = U[0];
    if (rk != 3) {
      for (int j = 0; j < 4; j++)
        U[j] = U[j] - (0.5 * c->euler * FRES(c, j, k));
    } else {
      for (int j = 0; j < 4; j++)
        U[j] = tbt * U_old[j] + obt * (U[j] - 0.5 * FRES(c, j, k));
    }
    if (f == 0) U[2] = 0.;
    double U2_rot = U[1], U3_rot = U[2];
    U[1] = U2_rot * ny + U3_rot * nx;
    U[2] = U3_rot * ny - U2_rot * nx;
    c->res_pt[k] = (f == 2) ? 0.0 : (U[0] - temp) * (U[0] - temp);
    PRIM(c, 0, k) = U[0];
    temp = 1.0 / U[0];
    PRIM(c, 1, k) = U[1] * temp;
    PRIM(c, 2, k) = U[2] * temp;
    PRIM(c, 3, k) = 0.4 * U[3] - (0.2 * temp) * (U[1] * U[1] + U[2] * U[2]);
  }
  /* :106-107, :157-165, :215-223 -- serial, in point order */
  c->max_res = 0.0;
  c->sum_res_sqr = 0.0;
  for (int k = 0; k < c->n; k++) {
    if (c->flag1[k] == 2) continue;
    double res_sqr = c->res_pt[k];
    if (res_sqr > c->max_res) {
      c->max_res = res_sqr;
      c->max_res_point = k;
    }
    c->sum_res_sqr = c->sum_res_sqr + res_sqr;
  }
}

/* compute_force_coeffs_mod.h:7-69 (shape 1 reported) */
void orc_compute_cl_cd_cm(orc_ctx *c) {
  int ns = c->shapes;
  double *H = xcalloc(ns + 1, sizeof(double));
  double *V = xcalloc(ns + 1, sizeof(double));
  double *P = xcalloc(ns + 1, sizeof(double));
  double temp = 0.5 * c->rho_inf * c->mach * c->mach;
  for (int j = 0; j < c->shape_points; j++) {
    int m = j, r = c->right[m], l = c->left[m];
    double lx = c->x[l], ly = c->y[l];
    double mx = c->x[m], my = c->y[m];
    double rx = c->x[r], ry = c->y[r];
    double ds1 = pow((mx - lx), 2) + pow((my - ly), 2);
    ds1 = sqrt(ds1);
    double ds2 = pow((rx - mx), 2) + pow((ry - my), 2);
    ds2 = sqrt(ds2);
    double ds = 0.5 * (ds1 + ds2);
    double nx = c->nx[m], ny = c->ny[m];
    double cp = PRIM(c, 3, m) - c->pr_inf;
    cp = -cp / temp;
    int s = c->flag2[m];
    H[s] = H[s] + cp * nx * ds;
    V[s] = V[s] + cp * ny * ds;
    P[s] = P[s] + (-cp * ny * ds * (mx - 0.25) + cp * nx * ds * (my));
  }
  c->cl = V[1] * cos(c->theta) - H[1] * sin(c->theta);
  c->cd = H[1] * cos(c->theta) + V[1] * sin(c->theta);
  c->cm = P[1];
  free(H); free(V); free(P);
}

/* fpi_solver_mod.h:14-150 */
void orc_fpi_solver(orc_ctx *c, int it) {
  memcpy(c->prim_old, c->prim, 4 * (size_t)c->n * sizeof(double));
  orc_func_delta(c);
  for (int rk = 1; rk <= c->rks; rk++) {
    orc_eval_q_variables(c);
    orc_eval_q_derivatives(c);
    for (int i = 1; i <= c->inner_iterations; i++) {
      orc_eval_q_inner_loop(c);
      orc_eval_update_innerloop(c);
    }
    orc_cal_flux_residual(c);
    orc_state_update(c, rk);
  }
  c->res_new = sqrt(c->sum_res_sqr) / c->n;
  if (it <= 2 && c->restart == 0) {
    c->res_old = c->res_new;
    c->residue = 0;
  } else {
    c->residue = log10(c->res_new / c->res_old);
  }
  c->it = it;
}

/* q_lskum_mod.h:28-37 */
void orc_q_lskum(orc_ctx *c, int n_iter, double *res_new, double *residue) {
  int itr = c->it;
  for (int it = itr + 1; it <= itr + n_iter; it++) {
    orc_fpi_solver(c, it);
    if (res_new) res_new[it - itr - 1] = c->res_new;
    if (residue) residue[it - itr - 1] = c->residue;
  }
}

/* ---------------------------------------------------------------- diagnostics */

/* compute_entropy_mod.h:5-24: per point |log(p / rho^gamma) - log(pr_inf)|;
 * returns the sum (the reference's total_entropy) */
double orc_compute_entropy(const orc_ctx *c, double *entropy) {
  const int n = c->n;
  double total_entropy = 0.;
  const double temp2 = log(c->pr_inf);
  for (int k = 0; k < n; k++) {
    double temp1 = pow(PRIM(c, 0, k), c->gamma_new);
    temp1 = PRIM(c, 3, k) / temp1;
    temp1 = log(temp1);
    if (entropy) entropy[k] = fabs(temp1 - temp2);
    total_entropy = total_entropy + fabs(temp1 - temp2);
  }
  return total_entropy;
}

/* compute_enstrophy_mod.h:5-76: vorticity du2/dx - du1/dy from the weighted
 * least squares over the point's neighbours, its square; returns the sum */
double orc_compute_enstrophy(const orc_ctx *c, double *vorticity, double *vorticity_sqr) {
  const int n = c->n;
  double total_enstrophy = 0;
  for (int i = 0; i < n; i++) {
    const double x_i = c->x[i], y_i = c->y[i];
    double sum_delx_sqr = 0., sum_dely_sqr = 0., sum_delx_dely = 0.;
    double sum_delx_delu1 = 0., sum_dely_delu1 = 0., sum_delx_delu2 = 0., sum_dely_delu2 = 0.;
    for (int e = c->conn_off[i]; e < c->conn_off[i + 1]; e++) {
      const int nbh = c->conn[e];
      const double delx = c->x[nbh] - x_i, dely = c->y[nbh] - y_i;
      const double dist = sqrt(delx * delx + dely * dely);
      const double weights = pow(dist, c->power);
      sum_delx_sqr = sum_delx_sqr + delx * delx * weights;
      sum_dely_sqr = sum_dely_sqr + dely * dely * weights;
      sum_delx_dely = sum_delx_dely + delx * dely * weights;
      sum_delx_delu1 = sum_delx_delu1 + weights * delx * (PRIM(c, 1, nbh) - PRIM(c, 1, i));
      sum_delx_delu2 = sum_delx_delu2 + weights * delx * (PRIM(c, 2, nbh) - PRIM(c, 2, i));
      sum_dely_delu1 = sum_dely_delu1 + weights * dely * (PRIM(c, 1, nbh) - PRIM(c, 1, i));
      sum_dely_delu2 = sum_dely_delu2 + weights * dely * (PRIM(c, 2, nbh) - PRIM(c, 2, i));
    }
    const double det = sum_delx_sqr * sum_dely_sqr - sum_delx_dely * sum_delx_dely;
    const double one_by_det = 1.0 / det;
    const double du2_dx =
        (sum_delx_delu2 * sum_dely_sqr - sum_dely_delu2 * sum_delx_dely) * one_by_det;
    const double du1_dy =
        (sum_dely_delu1 * sum_delx_sqr - sum_delx_delu1 * sum_delx_dely) * one_by_det;
    const double temp = du2_dx - du1_dy;
    if (vorticity) vorticity[i] = temp;
    if (vorticity_sqr) vorticity_sqr[i] = temp * temp;
    total_enstrophy = total_enstrophy + temp * temp;
  }
  return total_enstrophy;
}

/* stagnation pressure of one state, stagnation_values_mod.h:29-31 */
static double stag_p0(const orc_ctx *c, int i, double gammaPower) {
  const double g = c->gamma_new;
  const double angle = sqrt(g * PRIM(c, 3, i) / PRIM(c, 0, i));
  const double mach_t = sqrt(pow(PRIM(c, 1, i), 2) + pow(PRIM(c, 2, i), 2)) / angle;
  return PRIM(c, 3, i) * (pow((1 + ((g - 1) / 2) * mach_t * mach_t), gammaPower));
}

static double stag_p0_inf(const orc_ctx *c, double gammaPower) {
  const double g = c->gamma_new;
  return c->pr_inf * (pow((1 + ((g - 1) / 2) * c->mach * c->mach), gammaPower));
}

/* stagnation_values_mod.h:8-46: out = pMin, pMax, pMin/p0_inf, pMax/p0_inf;
 * idx = the (0-based) points of the first minimum and maximum */
void orc_stagnation_pressure(const orc_ctx *c, double out[4], int idx[2]) {
  const double gammaPower = c->gamma_new / (c->gamma_new - 1);
  const double p0_inf = stag_p0_inf(c, gammaPower);
  double pMin = 0, pMax = 0;
  int indexMin = 0, indexMax = 0;
  for (int i = 0; i < c->n; i++) {
    const double p0 = stag_p0(c, i, gammaPower);
    if (i == 0) {
      pMin = p0;
      indexMin = i;
      indexMax = i;
      pMax = p0;
    } else if (p0 < pMin) {
      pMin = p0;
      indexMin = i;
    } else if (p0 > pMax) {
      pMax = p0;
      indexMax = i;
    }
  }
  out[0] = pMin;
  out[1] = pMax;
  out[2] = pMin / p0_inf;
  out[3] = pMax / p0_inf;
  idx[0] = indexMin;
  idx[1] = indexMax;
}

/* stagnation_values_mod.h:49-80, the loss sum (p0_inf - p0)^2 / (p0_inf^2
 * plen) with plen = max_points (meshfree_solver_mod.h:58).  The reference
 * forms p0_sum and then multiplies an uninitialised total_p0 by the
 * constant; this returns p0_sum times the constant. */
double orc_objective_J(const orc_ctx *c) {
  const double gammaPower = c->gamma_new / (c->gamma_new - 1);
  const double p0_inf = stag_p0_inf(c, gammaPower);
  const double constant = 1 / (pow(p0_inf, 2) * c->n);
  double p0_sum = 0.0;
  for (int i = 0; i < c->n; i++) {
    const double p0 = stag_p0(c, i, gammaPower);
    p0_sum = p0_sum + pow((p0_inf - p0), 2);
  }
  return p0_sum * constant;
}

double *orc_array(orc_ctx *c, int which, int *len) {
  int n = c->n;
  switch (which) {
    case 0: *len = 4 * n; return c->prim;
    case 1: *len = 4 * n; return c->prim_old;
    case 2: *len = 4 * n; return c->q;
    case 3: *len = 8 * n; return c->dq;
    case 4: *len = 8 * n; return c->qm;
    case 5: *len = n; return c->delta;
    case 6: *len = 4 * n; return c->flux_res;
    case 7: *len = n; return c->nx;
    case 8: *len = n; return c->ny;
    case 9: *len = n; return c->res_pt;
    default: *len = 0; return 0;
  }
}

int orc_sub_count(const orc_ctx *c, int which, int i) {
  return c->sub_off[which][i + 1] - c->sub_off[which][i];
}

const int *orc_sub_list(const orc_ctx *c, int which, int i) {
  return c->sub[which] + c->sub_off[which][i];
}

/* out: sum_res_sqr, max_res, res_old, res_new, residue, max_res_point, it,
 * cl, cd, cm, theta */
void orc_scalars(const orc_ctx *c, double out[11]) {
  out[0] = c->sum_res_sqr;
  out[1] = c->max_res;
  out[2] = c->res_old;
  out[3] = c->res_new;
  out[4] = c->residue;
  out[5] = c->max_res_point;
  out[6] = c->it;
  out[7] = c->cl;
  out[8] = c->cd;
  out[9] = c->cm;
  out[10] = c->theta;
}

void orc_set_iteration(orc_ctx *c, int it, double res_old, int restart) {
  c->it = it;
  c->res_old = res_old;
  c->restart = restart;
}
