// shape_api.hip -- shape sensitivities of the force coefficients (no
// reference counterpart): the coordinates of a context's points as inputs
// (lskum_set_coordinates), the discrete adjoint of whole iterations with
// respect to them (lskum_shape_vjp) and the tangent-linear run it is
// checked against (lskum_shape_jvp).
//
// The reverse sweep keeps the state at the start of every iteration (the
// checkpoints) and recomputes the rest: per iteration, its func_delta and the
// four stages' input states and flux residuals; per stage, q, the
// q-derivatives of every least-squares pass, qmax/qmin and the limiter.  The
// reverse kernels accumulate the adjoints of the geometry they read -- per
// pair d/d(delx, dely), per point d/d(nx, ny) and d/d of the sub-stencil
// coefficients -- and the host maps them onto the coordinates once, at the
// end: coefficients -> pairs and normals, normals -> the wall and outer
// points' left and right neighbours (point_normals_mod.h:17-55), pairs ->
// their two points.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "context.hpp"

using namespace lskum;

namespace {

// device arrays released together
struct Scratch {
  std::vector<void *> ptrs;
  ~Scratch() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  bool get(T **p, size_t count) {
    if (dalloc(p, count) != hipSuccess) return false;
    ptrs.push_back((void *)*p);
    return true;
  }
};

int shape_ok(lskum_ctx *c) {
  if (c->n_cases() != 1 || c->mixed || c->prm.power != 0.0)
    return set_err(c, LSKUM_EINVAL,
                   "shape sensitivities need a single-case FP64 context with power 0");
  return LSKUM_OK;
}

// J = w0 Cl + w1 Cd + w2 Cm with Cl = V cos th - H sin th, Cd = H cos th +
// V sin th, Cm = P (compute_force_coeffs_mod.h:62-67): the weights of H, V, P
void force_weights(double theta, const double w[3], double &wH, double &wV, double &wP) {
  const double ct = std::cos(theta), st = std::sin(theta);
  wH = -w[0] * st + w[1] * ct;
  wV = w[0] * ct + w[1] * st;
  wP = w[2];
}

// The stagnation-pressure loss of a state (stagnation_values_mod.h:49-80,
// the sum the reference forms times its constant 1/(p0_inf^2 n)) and, with
// grad, weight times its gradient added to grad [4][n]: p0 = p A^g with
// A = 1 + (gamma-1)/2 M^2, M^2 = |u|^2 rho / (gamma p), g = gamma/(gamma-1).
double stagnation_loss(const double *prim, int32_t n, double mach, double pr_inf, double weight,
                       double *grad) {
  const double gm = 1.4, gp = gm / (gm - 1), c1 = (gm - 1) / 2;
  const double p0_inf = pr_inf * std::pow(1 + c1 * mach * mach, gp);
  const double constant = 1 / (std::pow(p0_inf, 2) * n);
  double sum = 0.0;
  for (int32_t i = 0; i < n; i++) {
    const double rho = prim[i], u1 = prim[(size_t)n + i], u2 = prim[(size_t)2 * n + i];
    const double pr = prim[(size_t)3 * n + i];
    const double uu = u1 * u1 + u2 * u2, M2 = uu * rho / (gm * pr);
    const double A = 1 + c1 * M2, Ag = std::pow(A, gp), p0 = pr * Ag;
    sum += (p0_inf - p0) * (p0_inf - p0);
    if (!grad) continue;
    const double dp0 = weight * (-2.0 * (p0_inf - p0) * constant);
    const double dM2 = dp0 * pr * gp * std::pow(A, gp - 1) * c1;  // d/dM2 of weight J
    grad[i] += dM2 * (uu / (gm * pr));
    grad[(size_t)n + i] += dM2 * (2.0 * u1 * rho / (gm * pr));
    grad[(size_t)2 * n + i] += dM2 * (2.0 * u2 * rho / (gm * pr));
    grad[(size_t)3 * n + i] += dp0 * Ag + dM2 * (-uu * rho / (gm * pr * pr));
  }
  return sum * constant;
}

// one fpi_solver iteration without the residual history
int plain_iteration(lskum_ctx *c) {
  const int inner = c->prm.inner_iterations;
  int rc = do_snapshot(c);
  if (!rc) rc = do_func_delta(c);
  for (int rk = 1; rk <= 4 && !rc; rk++) {
    rc = do_q_derivatives(c, inner, inner > 0);
    if (!rc) rc = do_q_inner(c, inner);
    if (!rc) rc = do_flux(c);
    if (!rc) rc = do_state_update(c, rk, false);
  }
  return rc;
}

// J of the context's current state (the forces kernel, finished on the host)
int objective(lskum_ctx *c, const double w[3], double *J) {
  launch_forces(1, c->soff, c->f_m, c->f_l, c->f_r, c->f_s, c->base, c->nx, c->ny, c->prim,
                c->n, c->pr_inf, c->qdyn, c->forces, c->stream);
  int rc = check_launch(c, "forces");
  if (rc) return rc;
  double h[3];
  HIP_TRY(c, hipMemcpyAsync(h, c->forces, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const double th = c->cases[0].theta;
  const double l = h[1] * std::cos(th) - h[0] * std::sin(th);
  const double d = h[0] * std::cos(th) + h[1] * std::sin(th);
  *J = w[0] * l + w[1] * d + w[2] * h[2];
  if (!std::isfinite(*J)) return set_err(c, LSKUM_ENONFINITE, "objective is not finite");
  return LSKUM_OK;
}

// device indices of the left and right neighbours of every wall and outer
// point (0 for interior points)
std::vector<int32_t> left_right(const CloudLayout &L) {
  std::vector<int32_t> lr((size_t)2 * L.n, 0);
  for (int32_t d = 0; d < L.n; d++) {
    if (L.flag[d] == 1) continue;
    const int32_t v = L.perm[d];
    lr[2 * (size_t)d] = L.inv[L.left_h[v]];
    lr[2 * (size_t)d + 1] = L.inv[L.right_h[v]];
  }
  return lr;
}

// a value and one directional derivative, for the host end of the tangent
struct HD {
  double v = 0.0, d = 0.0;
};
HD operator+(HD a, HD b) { return {a.v + b.v, a.d + b.d}; }
HD operator-(HD a, HD b) { return {a.v - b.v, a.d - b.d}; }
HD operator*(HD a, HD b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
HD operator*(double a, HD b) { return {a * b.v, a * b.d}; }
HD hsqrt(HD a) {
  const double r = std::sqrt(a.v);
  return {r, 0.5 * a.d / r};
}
HD operator/(HD a, HD b) { return {a.v / b.v, (a.d - a.v / b.v * b.d) / b.v}; }
HD hpow(HD a, double e) { return {std::pow(a.v, e), e * std::pow(a.v, e - 1) * a.d}; }

}  // namespace

extern "C" {

int lskum_set_coordinates(lskum_ctx *ctx, const double *x, const double *y) {
  CTX_OK(ctx);
  if (!x || !y) return set_err(ctx, LSKUM_EINVAL, "null coordinates");
  if (ctx->n_cases() != 1) return set_err(ctx, LSKUM_EINVAL, "single-case contexts only");
  CloudLayout &L = ctx->clouds[0];
  const int32_t n = L.n;
  for (int32_t i = 0; i < n; i++)
    if (!std::isfinite(x[i]) || !std::isfinite(y[i]))
      return set_err(ctx, LSKUM_EINVAL, "non-finite coordinate at point " + std::to_string(i));
  lskum_cloud view;
  view.n_points = n;
  view.x = x;
  view.y = y;
  view.left = L.left_h.data();
  view.right = L.right_h.data();
  view.flag_1 = L.flag1_h.data();
  view.flag_2 = L.flag2_h.data();
  view.min_dist = L.min_dist_h.data();
  view.conn_off = L.conn_off_h.data();
  view.conn = L.conn_h.data();
  std::vector<double> nxh, nyh, coef;
  compute_normals(view, nxh, nyh);
  for (int32_t i = 0; i < n; i++) {
    if (!(substencil_min_rel_det(view, nxh, nyh, L.mask_h, i) > 0.0))
      return set_err(ctx, LSKUM_ESTENCIL,
                     "degenerate sub-stencil at point " + std::to_string(i) +
                         " after the move (the neighbour lists and their split are kept)");
  }
  compute_coef(L, x, y, nxh.data(), nyh.data(), ctx->prm.power, coef);
  std::vector<double> base_h;
  int rc = fetch_plain(ctx, ctx->base, B_STRIDE, base_h);  // [n][B_STRIDE] records
  if (rc) return rc;
  std::vector<double> xy_h((size_t)2 * n);
  for (int32_t d = 0; d < n; d++) {
    const int32_t v = L.perm[d];
    L.x[d] = x[v];
    L.y[d] = y[v];
    L.nx[d] = nxh[v];
    L.ny[d] = nyh[v];
    base_h[(size_t)B_STRIDE * d + B_X] = x[v];
    base_h[(size_t)B_STRIDE * d + B_Y] = y[v];
    xy_h[2 * (size_t)d] = x[v];
    xy_h[2 * (size_t)d + 1] = y[v];
  }
  L.nx_h = nxh;
  L.ny_h = nyh;
  L.coef = coef;
  hipStream_t s = ctx->stream;
  HIP_TRY(ctx, hipMemcpyAsync(ctx->base, base_h.data(), base_h.size() * sizeof(double),
                              hipMemcpyHostToDevice, s));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->nx, L.nx.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->ny, L.ny.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->coef, coef.data(), coef.size() * sizeof(double),
                              hipMemcpyHostToDevice, s));
  if (ctx->xy)
    HIP_TRY(ctx, hipMemcpyAsync(ctx->xy, xy_h.data(), xy_h.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  return LSKUM_OK;
}

int lskum_shape_vjp(lskum_ctx *ctx, int32_t n_iter, const double w[4], double *J,
                    double *x_bar, double *y_bar, double *prim_bar) {
  CTX_OK(ctx);
  if (!w || n_iter < 0) return set_err(ctx, LSKUM_EINVAL, "null weights or n_iter < 0");
  int rc = shape_ok(ctx);
  if (!rc) rc = prepare_adjoint(ctx);
  if (rc) return rc;
  const CloudLayout &L = ctx->clouds[0];
  const CaseInfo &cs = ctx->cases[0];
  const int32_t n = ctx->n;
  const size_t n4 = (size_t)4 * n, E = (size_t)L.off[n];
  const int inner = ctx->prm.inner_iterations;
  hipStream_t s = ctx->stream;
  const DevCloud dc = ctx->dc();
  AdjointBuffers &a = ctx->adj;
  Scratch S;
  double *ckpt = nullptr, *stage = nullptr, *sres = nullptr, *dqs = nullptr, *pbar = nullptr,
         *pold = nullptr, *rbar = nullptr, *amin_v = nullptr;
  int32_t *amin_e = nullptr;
  GeoAdjoint g{};
  if (!S.get(&ckpt, n4 * (size_t)std::max(n_iter, 1)) || !S.get(&stage, 4 * n4) ||
      !S.get(&sres, 4 * n4) || !S.get(&dqs, (size_t)DQ_STRIDE * n * (inner + 1)) ||
      !S.get(&pbar, n4) || !S.get(&pold, n4) || !S.get(&rbar, n4) || !S.get(&amin_v, n) ||
      !S.get(&amin_e, n) || !S.get(&g.edge, 2 * E) || !S.get(&g.nrm, (size_t)2 * n) ||
      !S.get(&g.coef, (size_t)8 * n) || !S.get(&g.delta, n))
    return set_err(ctx, LSKUM_ENOMEM, "shape adjoint: device allocation failed");
  auto dq_of = [&](int p) { return dqs + (size_t)DQ_STRIDE * n * p; };
  auto copy = [&](const double *src, double *dst, size_t count) {
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToDevice, s));
    return LSKUM_OK;
  };
  // forward: P_0 .. P_{N-1} kept, P_N left in prim
  rc = copy(ctx->prim, ckpt, n4);
  if (!rc) rc = do_q_variables(ctx);
  if (!rc) rc = do_speed(ctx);
  for (int it = 0; it < n_iter && !rc; it++) {
    rc = copy(ctx->prim, ckpt + n4 * it, n4);
    if (!rc) rc = plain_iteration(ctx);
  }
  double Jv = 0.0;
  if (!rc) rc = objective(ctx, w, &Jv);
  std::vector<double> pN;
  if (!rc) rc = fetch_plain(ctx, ctx->prim, 4, pN);
  if (rc) return rc;
  // the objective's partials: the wall pressures of P_N, and the geometry of
  // its panels (compute_force_coeffs_mod.h:25-66)
  std::vector<double> pb((size_t)n4, 0.0), xb(n, 0.0), yb(n, 0.0), nb((size_t)2 * n, 0.0);
  double wH, wV, wP;
  force_weights(cs.theta, w, wH, wV, wP);
  const double qdyn = 0.5 * ctx->rho_inf * cs.mach * cs.mach;
  for (size_t j = 0; j < L.fm.size(); j++) {
    if (L.fs[j] != 1) continue;
    const int32_t m = L.fm[j], l = L.fl[j], r = L.fr[j];
    const double ax = L.x[m] - L.x[l], ay = L.y[m] - L.y[l];
    const double bx = L.x[r] - L.x[m], by = L.y[r] - L.y[m];
    const double ds1 = std::sqrt(ax * ax + ay * ay), ds2 = std::sqrt(bx * bx + by * by);
    const double ds = 0.5 * (ds1 + ds2);
    const double nx = L.nx[m], ny = L.ny[m], xm = L.x[m] - 0.25, ym = L.y[m];
    const double cp = -(pN[(size_t)3 * n + m] - ctx->pr_inf) / qdyn;
    const double f = wH * nx + wV * ny + wP * (-ny * xm + nx * ym);  // per unit cp ds
    pb[(size_t)3 * n + m] += f * ds * (-1.0 / qdyn);
    nb[2 * (size_t)m] += cp * ds * (wH + wP * ym);
    nb[2 * (size_t)m + 1] += cp * ds * (wV - wP * xm);
    xb[m] += -wP * cp * ny * ds;
    yb[m] += wP * cp * nx * ds;
    const double dsb = 0.5 * cp * f;
    xb[m] += dsb * (ax / ds1 - bx / ds2);
    yb[m] += dsb * (ay / ds1 - by / ds2);
    xb[l] -= dsb * ax / ds1;
    yb[l] -= dsb * ay / ds1;
    xb[r] += dsb * bx / ds2;
    yb[r] += dsb * by / ds2;
  }
  // the stagnation-pressure loss: pointwise in P_N, no geometry
  if (w[3] != 0.0)
    Jv += w[3] * stagnation_loss(pN.data(), n, cs.mach, ctx->pr_inf, w[3], pb.data());
  HIP_TRY(ctx, hipMemcpyAsync(pbar, pb.data(), n4 * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(ctx, hipMemsetAsync(g.edge, 0, 2 * E * sizeof(double), s));
  HIP_TRY(ctx, hipMemsetAsync(g.nrm, 0, (size_t)2 * n * sizeof(double), s));
  HIP_TRY(ctx, hipMemsetAsync(g.coef, 0, (size_t)8 * n * sizeof(double), s));
  // reverse sweep
  for (int it = n_iter - 1; it >= 0 && !rc; it--) {
    const double *P0 = ckpt + n4 * it;
    // recompute the iteration: each stage's input state and flux residual
    rc = copy(P0, ctx->prim, n4);
    if (!rc) rc = do_q_variables(ctx);
    if (!rc) rc = do_speed(ctx);
    if (!rc) rc = do_snapshot(ctx);
    if (!rc) rc = do_func_delta(ctx);
    for (int rk = 1; rk <= 4 && !rc; rk++) {
      rc = copy(ctx->prim, stage + n4 * (rk - 1), n4);
      if (!rc) rc = do_q_derivatives(ctx, inner, inner > 0);
      if (!rc) rc = do_q_inner(ctx, inner);
      if (!rc) rc = do_flux(ctx);
      if (!rc) rc = copy(ctx->flux_res, sres + n4 * (rk - 1), n4);
      if (!rc) rc = do_state_update(ctx, rk, false);
    }
    if (rc) break;
    HIP_TRY(ctx, hipMemsetAsync(g.delta, 0, (size_t)n * sizeof(double), s));
    HIP_TRY(ctx, hipMemsetAsync(pold, 0, n4 * sizeof(double), s));
    for (int rk = 4; rk >= 1; rk--) {
      const double *Ps = stage + n4 * (rk - 1);
      launch_su_vjp(dc, rk, Ps, P0, sres + n4 * (rk - 1), ctx->fcs, pbar, pold, rbar, g.nrm, s);
      // the stage's q, derivatives of every pass, qmax/qmin and their origins
      launch_q_variables(dc, Ps, ctx->base, s);
      launch_q_derivatives_arg(dc, ctx->base, dq_of(0), ctx->qm, ctx->qsum, a.arg, a.lsum, s);
      for (int p = 0; p < inner; p++)
        launch_q_inner(dc, ctx->qsum, dq_of(p), dq_of(p + 1), 0.0, s);
      launch_flux_vjp(dc, a, ctx->base, dq_of(inner), ctx->qm, ctx->delta, rbar, &g, s);
      int b = 0;
      for (int p = inner - 1; p >= 0; p--) {
        launch_ls_vjp(dc, a, ctx->base, a.ldq[b], a.ldq[1 - b], false, dq_of(p + 1), dq_of(p),
                      g.edge, s);
        b = 1 - b;
      }
      launch_ls_vjp(dc, a, ctx->base, a.ldq[b], nullptr, true, dq_of(0), nullptr, g.edge, s);
      launch_qvar_vjp(dc, a, Ps, pbar, true, s);
    }
    launch_axpy(n4, pold, pbar, s);
    launch_speed(dc, P0, ctx->base, s);  // the speed this iteration's func_delta read
    launch_delta_vjp(dc, a, ctx->base, P0, ctx->prm.cfl, g.delta, amin_e, amin_v, g.edge, pbar,
                     s);
    rc = check_launch(ctx, "shape adjoint");
  }
  if (rc) return rc;
  // back to P_0
  rc = copy(ckpt, ctx->prim, n4);
  ctx->q_fresh = ctx->speed_fresh = false;
  std::vector<double> ge, gn, gc, pbh;
  if (!rc) rc = fetch_plain(ctx, pbar, 4, pbh);
  if (rc) return rc;
  ge.resize(2 * E);
  HIP_TRY(ctx, hipMemcpyAsync(ge.data(), g.edge, 2 * E * sizeof(double), hipMemcpyDeviceToHost, s));
  rc = fetch_plain(ctx, g.nrm, 2, gn);
  if (!rc) rc = fetch_plain(ctx, g.coef, 8, gc);
  if (rc) return rc;
  for (size_t j = 0; j < 2 * (size_t)n; j++) gn[j] += nb[j];
  // sub-stencil coefficients (coef = (syy, -sxy)/det or (sxx, -sxy)/det of
  // sums over the members) -> pairs and normals
  coef_adjoint(n, L.x.data(), L.y.data(), L.nx.data(), L.ny.data(), L.off.data(),
               L.conn.data(), gc.data(), ge.data(), gn.data());
  // pairs -> their points
  for (int32_t d = 0; d < n; d++)
    for (int32_t e = L.off[d]; e < L.off[d + 1]; e++) {
      const int32_t k = (int32_t)(L.conn[e] & kIndexMask);
      xb[k] += ge[2 * (size_t)e];
      yb[k] += ge[2 * (size_t)e + 1];
      xb[d] -= ge[2 * (size_t)e];
      yb[d] -= ge[2 * (size_t)e + 1];
    }
  // normals of wall and outer points -> their left and right neighbours:
  // n = (-a, b)/|(a, b)| with a = (y_r - y_l)/2, b = (x_r - x_l)/2
  const std::vector<int32_t> lr = left_right(L);
  for (int32_t d = 0; d < n; d++) {
    if (L.flag[d] == 1) continue;
    const int32_t l = lr[2 * (size_t)d], r = lr[2 * (size_t)d + 1];
    const double av = 0.5 * ((L.y[d] - L.y[l]) + (L.y[r] - L.y[d]));
    const double bv = 0.5 * ((L.x[d] - L.x[l]) + (L.x[r] - L.x[d]));
    const double D = std::sqrt(av * av + bv * bv), D3 = D * D * D;
    const double gx = gn[2 * (size_t)d], gy = gn[2 * (size_t)d + 1];
    const double ab = gx * (-1.0 / D + av * av / D3) + gy * (-av * bv / D3);
    const double bb = gx * (av * bv / D3) + gy * (1.0 / D - bv * bv / D3);
    yb[l] -= 0.5 * ab;
    yb[r] += 0.5 * ab;
    xb[l] -= 0.5 * bb;
    xb[r] += 0.5 * bb;
  }
  for (int32_t d = 0; d < n; d++) {
    if (x_bar) x_bar[L.perm[d]] = xb[d];
    if (y_bar) y_bar[L.perm[d]] = yb[d];
  }
  if (prim_bar) permute_out(ctx, pbh.data(), prim_bar, 4);
  if (J) *J = Jv;
  return LSKUM_OK;
}

int lskum_shape_jvp(lskum_ctx *ctx, int32_t n_iter, const double w[4], const double *dx,
                    const double *dy, const double *dprim, double *J, double *dJ) {
  CTX_OK(ctx);
  if (!w || n_iter < 0) return set_err(ctx, LSKUM_EINVAL, "null weights or n_iter < 0");
  int rc = shape_ok(ctx);
  if (rc) return rc;
  const CloudLayout &L = ctx->clouds[0];
  const CaseInfo &cs = ctx->cases[0];
  const int32_t n = ctx->n;
  hipStream_t s = ctx->stream;
  const DevCloud dc = ctx->dc();
  Scratch S;
  TanState t{};
  int32_t *lr_d = nullptr;
  if (!S.get(&t.x, n) || !S.get(&t.y, n) || !S.get(&t.nx, n) || !S.get(&t.ny, n) ||
      !S.get(&t.coef, (size_t)8 * n) || !S.get(&t.prim, (size_t)4 * n) ||
      !S.get(&t.prim_old, (size_t)4 * n) || !S.get(&t.fres, (size_t)4 * n) ||
      !S.get(&t.q, (size_t)4 * n) || !S.get(&t.speed, n) || !S.get(&t.delta, n) ||
      !S.get(&t.dq[0], (size_t)8 * n) || !S.get(&t.dq[1], (size_t)8 * n) ||
      !S.get(&t.qm, (size_t)8 * n) || !S.get(&lr_d, (size_t)2 * n))
    return set_err(ctx, LSKUM_ENOMEM, "shape tangent: device allocation failed");
  std::vector<double> p0, dp((size_t)4 * n, 0.0);
  rc = fetch_plain(ctx, ctx->prim, 4, p0);
  if (rc) return rc;
  if (dprim) permute_in(ctx, dprim, dp.data(), 4);
  std::vector<Tan> xh(n), yh(n), ph((size_t)4 * n);
  for (int32_t d = 0; d < n; d++) {
    const int32_t v = L.perm[d];
    xh[d] = {L.x[d], dx ? dx[v] : 0.0};
    yh[d] = {L.y[d], dy ? dy[v] : 0.0};
  }
  for (size_t j = 0; j < (size_t)4 * n; j++) ph[j] = {p0[j], dp[j]};
  const std::vector<int32_t> lr = left_right(L);
  HIP_TRY(ctx, hipMemcpyAsync(t.x, xh.data(), n * sizeof(Tan), hipMemcpyHostToDevice, s));
  HIP_TRY(ctx, hipMemcpyAsync(t.y, yh.data(), n * sizeof(Tan), hipMemcpyHostToDevice, s));
  HIP_TRY(ctx, hipMemcpyAsync(t.prim, ph.data(), ph.size() * sizeof(Tan), hipMemcpyHostToDevice, s));
  HIP_TRY(ctx, hipMemcpyAsync(lr_d, lr.data(), lr.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  launch_tan_geometry(dc, lr_d, t, s);
  for (int it = 0; it < n_iter; it++)
    launch_tan_iteration(dc, ctx->base, ctx->fcs, ctx->prm.cfl, ctx->prm.inner_iterations, t, s);
  rc = check_launch(ctx, "shape tangent");
  if (rc) return rc;
  std::vector<Tan> nxh(n), nyh(n);
  HIP_TRY(ctx, hipMemcpyAsync(ph.data(), t.prim, ph.size() * sizeof(Tan), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipMemcpyAsync(nxh.data(), t.nx, n * sizeof(Tan), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipMemcpyAsync(nyh.data(), t.ny, n * sizeof(Tan), hipMemcpyDeviceToHost, s));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  // compute_force_coeffs_mod.h:25-67 on (value, derivative) pairs
  auto hd = [](const Tan &a) { return HD{a.v, a.d}; };
  const double qdyn = 0.5 * ctx->rho_inf * cs.mach * cs.mach;
  HD H, V, P;
  for (size_t j = 0; j < L.fm.size(); j++) {
    if (L.fs[j] != 1) continue;
    const int32_t m = L.fm[j], l = L.fl[j], r = L.fr[j];
    const HD xm = hd(xh[m]), ym = hd(yh[m]), xl = hd(xh[l]), yl = hd(yh[l]);
    const HD xr = hd(xh[r]), yr = hd(yh[r]);
    const HD ds1 = hsqrt((xm - xl) * (xm - xl) + (ym - yl) * (ym - yl));
    const HD ds2 = hsqrt((xr - xm) * (xr - xm) + (yr - ym) * (yr - ym));
    const HD ds = 0.5 * (ds1 + ds2);
    const HD nx = hd(nxh[m]), ny = hd(nyh[m]);
    const HD cp = (-1.0 / qdyn) * (hd(ph[(size_t)3 * n + m]) - HD{ctx->pr_inf, 0.0});
    H = H + cp * nx * ds;
    V = V + cp * ny * ds;
    P = P + ((-1.0) * cp * ny * ds * (xm - HD{0.25, 0.0}) + cp * nx * ds * ym);
  }
  double wH, wV, wP;
  force_weights(cs.theta, w, wH, wV, wP);
  const double th = cs.theta;
  const double l = V.v * std::cos(th) - H.v * std::sin(th);
  const double d = H.v * std::cos(th) + V.v * std::sin(th);
  double Jv = w[0] * l + w[1] * d + w[2] * P.v, dJv = wH * H.d + wV * V.d + wP * P.d;
  if (w[3] != 0.0) {  // the stagnation-pressure loss on (value, derivative) pairs
    const double gm = 1.4, gp = gm / (gm - 1), c1 = (gm - 1) / 2;
    const double p0_inf = ctx->pr_inf * std::pow(1 + c1 * cs.mach * cs.mach, gp);
    HD sum;
    for (int32_t i = 0; i < n; i++) {
      const HD rho = hd(ph[i]), u1 = hd(ph[(size_t)n + i]), u2 = hd(ph[(size_t)2 * n + i]);
      const HD pr = hd(ph[(size_t)3 * n + i]);
      const HD M2 = (u1 * u1 + u2 * u2) * rho / (gm * pr);
      const HD p0 = pr * hpow(HD{1.0, 0.0} + c1 * M2, gp);
      const HD t = HD{p0_inf, 0.0} - p0;
      sum = sum + t * t;
    }
    const double constant = 1 / (std::pow(p0_inf, 2) * n);
    Jv += w[3] * sum.v * constant;
    dJv += w[3] * sum.d * constant;
  }
  if (J) *J = Jv;
  if (dJ) *dJ = dJv;
  return LSKUM_OK;
}

}  // extern "C"
