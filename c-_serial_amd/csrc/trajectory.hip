// trajectory.hip -- derivatives of whole iterations (fpi_solver_mod.h:14-150),
// FP64, one case, power = 0.
//
// Reverse: the kernels the flux-residual adjoint (adjoint.hip) lacks for a
// whole iteration -- the stage update (state_update_mod.h:92-283), whose
// pointwise Jacobians come from dual numbers through update.hpp, and the
// local time step (state_update_mod.h:448-496), a minimum over the pairs of
// a point: its adjoint goes to the pair that set it.
//
// Forward: a tangent-linear run of the same iterations in which every
// quantity is a dual number {value, derivative along one direction} --
// coordinates, normals and sub-stencil coefficients included -- through the
// templated leaf physics.  It shares no derivative code with the reverse
// pass, which it checks.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"
#include "physics.hpp"
#include "update.hpp"
// after physics.hpp and update.hpp (see adjoint.hip)
#include "dual.hpp"

namespace lskum {

namespace {

using D1 = Dual<1>;
static_assert(sizeof(D1) == sizeof(Tan), "Tan and Dual<1> share one layout");

int blocks_of(int n) { return (n + 255) / 256; }

// ------------------------------------------------------------ stage update
__global__ __launch_bounds__(256) void k_su_vjp(int n, int rk, const double *__restrict__ prim,
                                                const double *__restrict__ prim_old,
                                                const double *__restrict__ fres,
                                                const double *__restrict__ nxa,
                                                const double *__restrict__ nya,
                                                const uint8_t *__restrict__ flag,
                                                const FlowConsts *__restrict__ fcs,
                                                double *__restrict__ pbar,
                                                double *__restrict__ pold_bar,
                                                double *__restrict__ rbar,
                                                double *__restrict__ nrm_bar) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const FlowConsts fc = fcs[0];
  const int type = flag[i];
  double p1[4], p2[4], R[4], a[4], gb[4][4];
  for (int r = 0; r < 4; r++) {
    p1[r] = prim[(size_t)r * n + i];
    p2[r] = prim_old[(size_t)r * n + i];
    R[r] = fres[(size_t)r * n + i];
    a[r] = pbar[(size_t)r * n + i];
  }
  // one pass per input group: the stage's state, the iteration's start
  // state, the flux residual, the normal
  for (int g = 0; g < 4; g++) {
    Dual<4> d1[4], d2[4], dr[4], out[4], dnx(nxa[i]), dny(nya[i]);
    for (int r = 0; r < 4; r++) {
      d1[r] = Dual<4>(p1[r]);
      d2[r] = Dual<4>(p2[r]);
      dr[r] = Dual<4>(R[r]);
    }
    if (g == 0) {
      for (int r = 0; r < 4; r++) d1[r].d[r] = 1.0;
    } else if (g == 1) {
      for (int r = 0; r < 4; r++) d2[r].d[r] = 1.0;
    } else if (g == 2) {
      for (int r = 0; r < 4; r++) dr[r].d[r] = 1.0;
    } else {
      dnx.d[0] = 1.0;
      dny.d[1] = 1.0;
    }
    su_point(rk, type, d1, d2, dr, dnx, dny, fc, out);
    for (int c = 0; c < 4; c++) {
      double t = 0.0;
      for (int r = 0; r < 4; r++) t += a[r] * out[r].d[c];
      gb[g][c] = t;
    }
  }
  for (int c = 0; c < 4; c++) {
    pbar[(size_t)c * n + i] = gb[0][c];
    pold_bar[(size_t)c * n + i] += gb[1][c];
    rbar[(size_t)c * n + i] = gb[2][c];
  }
  nrm_bar[2 * (size_t)i] += gb[3][0];
  nrm_bar[2 * (size_t)i + 1] += gb[3][1];
}

// ------------------------------------------------------------ local time step
// delta_i = min(1, min over pairs of cfl |x_k - x_i| / speed_k), the first
// strict minimum in conn order as in the forward kernel: its pair receives
// the geometry adjoint, and amin_v[i] = d delta_i / d speed_k times dbar_i
// waits for the gather at k.
__global__ __launch_bounds__(256) void k_delta_vjp(int n, const int32_t *__restrict__ off,
                                                   const uint32_t *__restrict__ conn,
                                                   const double *__restrict__ base, double cfl,
                                                   const double *__restrict__ dbar,
                                                   int32_t *__restrict__ amin_e,
                                                   double *__restrict__ amin_v,
                                                   double *__restrict__ geo_edge) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x_i = base[(size_t)B_STRIDE * i + B_X], y_i = base[(size_t)B_STRIDE * i + B_Y];
  double min_delt = 1.0, bx = 0.0, by = 0.0, bd = 1.0, bs = 1.0;
  int es = -1;
  for (int e = off[i]; e < off[i + 1]; e++) {
    const int k = (int)(conn[e] & kIndexMask);
    const double *hk = base + (size_t)B_STRIDE * k;
    const double delx = hk[B_X] - x_i, dely = hk[B_Y] - y_i, speed = hk[B_SPEED];
    double dist = delx * delx + dely * dely;
    dist = sqrt(dist);
    double delta_t = dist / speed;
    delta_t = cfl * delta_t;
    if (min_delt > delta_t) {
      min_delt = delta_t;
      es = e;
      bx = delx;
      by = dely;
      bd = dist;
      bs = speed;
    }
  }
  amin_e[i] = es;
  amin_v[i] = 0.0;
  if (es < 0) return;
  const double b = dbar[i];
  const double g = b * cfl / (bd * bs);
  geo_edge[2 * (size_t)es] += g * bx;
  geo_edge[2 * (size_t)es + 1] += g * by;
  amin_v[i] = -b * cfl * bd / (bs * bs);
}

// per point j: the speed adjoint of every pair that set an owner's delta
// with j as its neighbour (owners ascending), through speed = |u| +
// 3 sqrt(p/rho) (state_update_mod.h:483-485) into pbar
__global__ __launch_bounds__(256) void k_delta_vjp_gather(
    int n, const int32_t *__restrict__ tr_off, const int32_t *__restrict__ tr_edge,
    const int32_t *__restrict__ tr_owner, const int32_t *__restrict__ amin_e,
    const double *__restrict__ amin_v, const double *__restrict__ prim,
    double *__restrict__ pbar) {
  const int j = xcd_block() * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double sb = 0.0;
  for (int t = tr_off[j]; t < tr_off[j + 1]; t++) {
    const int i = tr_owner[t];
    if (amin_e[i] == tr_edge[t]) sb += amin_v[i];
  }
  if (sb == 0.0) return;
  const double rho = prim[j], u1 = prim[(size_t)n + j], u2 = prim[(size_t)2 * n + j];
  const double pr = prim[(size_t)3 * n + j];
  const double mod_u = sqrt(u1 * u1 + u2 * u2), c = sqrt(pr / rho);
  pbar[j] += sb * (-1.5 * c / rho);
  if (mod_u > 0.0) {
    pbar[(size_t)n + j] += sb * (u1 / mod_u);
    pbar[(size_t)2 * n + j] += sb * (u2 / mod_u);
  }
  pbar[(size_t)3 * n + j] += sb * (1.5 / (rho * c));
}

__global__ __launch_bounds__(256) void k_axpy(size_t count, const double *__restrict__ x,
                                              double *__restrict__ y) {
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < count;
       j += (size_t)gridDim.x * blockDim.x)
    y[j] += x[j];
}

// ------------------------------------------------------------ tangent run
// point_normals_mod.h:17-55
__global__ __launch_bounds__(256) void t_normals(int n, const uint8_t *__restrict__ flag,
                                                 const int32_t *__restrict__ lr,
                                                 const D1 *__restrict__ x,
                                                 const D1 *__restrict__ y, D1 *__restrict__ nx,
                                                 D1 *__restrict__ ny) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i] == 1) {
    nx[i] = D1(0.0);
    ny[i] = D1(1.0);
    return;
  }
  const int l = lr[2 * i], r = lr[2 * i + 1];
  const D1 nx1 = y[i] - y[l], nx2 = y[r] - y[i];
  const D1 ny1 = x[i] - x[l], ny2 = x[r] - x[i];
  const D1 a = D1(0.5) * (nx1 + nx2), b = D1(0.5) * (ny1 + ny2);
  const D1 det = sqrt(a * a + b * b);
  nx[i] = -a / det;
  ny[i] = b / det;
}

// the sub-stencil least-squares coefficients (power 0): per sub-stencil the
// sums of dels^2, deln^2, dels deln over its members, then (syy, -sxy)/det
// (x forms) or (sxx, -sxy)/det (y forms)
__global__ __launch_bounds__(256) void t_coef(int n, const int32_t *__restrict__ off,
                                              const uint32_t *__restrict__ conn,
                                              const D1 *__restrict__ x, const D1 *__restrict__ y,
                                              const D1 *__restrict__ nxa,
                                              const D1 *__restrict__ nya, D1 *__restrict__ coef) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const D1 nx = nxa[i], ny = nya[i], tx = ny, ty = -nx;
  D1 sxx[4], syy[4], sxy[4];
  for (int e = off[i]; e < off[i + 1]; e++) {
    const uint32_t w = conn[e];
    const int k = (int)(w & kIndexMask);
    const uint32_t m = w >> kMaskShift;
    const D1 delx = x[k] - x[i], dely = y[k] - y[i];
    const D1 dels = delx * tx + dely * ty, deln = delx * nx + dely * ny;
    for (int sb = 0; sb < 4; sb++) {
      if (!(m & (1u << sb))) continue;
      sxx[sb] = sxx[sb] + dels * dels;
      syy[sb] = syy[sb] + deln * deln;
      sxy[sb] = sxy[sb] + dels * deln;
    }
  }
  for (int sb = 0; sb < 4; sb++) {
    const D1 det = sxx[sb] * syy[sb] - sxy[sb] * sxy[sb];
    const D1 inv = det != D1(0.0) ? D1(1.0) / det : D1(0.0);
    coef[8 * (size_t)i + 2 * sb] = (sb < 2 ? syy[sb] : sxx[sb]) * inv;
    coef[8 * (size_t)i + 2 * sb + 1] = -sxy[sb] * inv;
  }
}

// q_variables_mod.h:9-34
__global__ __launch_bounds__(256) void t_qvar(int n, const D1 *__restrict__ prim,
                                              D1 *__restrict__ q) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const D1 rho = prim[i], u1 = prim[(size_t)n + i], u2 = prim[(size_t)2 * n + i];
  const D1 pr = prim[(size_t)3 * n + i];
  const D1 beta = D1(0.5) * rho / pr;
  q[4 * (size_t)i] = log(rho) + (log(beta) * D1(2.5)) - beta * (u1 * u1 + u2 * u2);
  const D1 ttb = D1(2.0) * beta;
  q[4 * (size_t)i + 1] = ttb * u1;
  q[4 * (size_t)i + 2] = ttb * u2;
  q[4 * (size_t)i + 3] = -ttb;
}

// state_update_mod.h:483-485
__global__ __launch_bounds__(256) void t_speed(int n, const D1 *__restrict__ prim,
                                               D1 *__restrict__ speed) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const D1 rho = prim[i], u1 = prim[(size_t)n + i], u2 = prim[(size_t)2 * n + i];
  const D1 pr = prim[(size_t)3 * n + i];
  const D1 uu = u1 * u1 + u2 * u2;
  const D1 mod_u = uu.v > 0.0 ? sqrt(uu) : D1(0.0);
  speed[i] = mod_u + D1(3.0) * sqrt(pr / rho);
}

// state_update_mod.h:448-496
__global__ __launch_bounds__(256) void t_delta(int n, const int32_t *__restrict__ off,
                                               const uint32_t *__restrict__ conn,
                                               const D1 *__restrict__ x, const D1 *__restrict__ y,
                                               const D1 *__restrict__ speed, double cfl,
                                               D1 *__restrict__ delta) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  D1 min_delt(1.0);
  for (int e = off[i]; e < off[i + 1]; e++) {
    const int k = (int)(conn[e] & kIndexMask);
    D1 dist = (x[k] - x[i]) * (x[k] - x[i]) + (y[k] - y[i]) * (y[k] - y[i]);
    dist = sqrt(dist);
    D1 delta_t = dist / speed[k];
    delta_t = D1(cfl) * delta_t;
    if (min_delt > delta_t) min_delt = delta_t;
  }
  delta[i] = min_delt;
}

// q_variables_mod.h:36-118 (power 0)
__global__ __launch_bounds__(256) void t_qder(int n, const int32_t *__restrict__ off,
                                              const uint32_t *__restrict__ conn,
                                              const D1 *__restrict__ x, const D1 *__restrict__ y,
                                              const D1 *__restrict__ q, D1 *__restrict__ dq,
                                              D1 *__restrict__ qm) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  D1 qi[4], qmax[4], qmin[4], sdx[4], sdy[4], sxx, syy, sxy;
  for (int r = 0; r < 4; r++) qi[r] = qmax[r] = qmin[r] = q[4 * (size_t)i + r];
  for (int e = off[i]; e < off[i + 1]; e++) {
    const int k = (int)(conn[e] & kIndexMask);
    const D1 delx = x[k] - x[i], dely = y[k] - y[i];
    sxx = sxx + delx * delx;
    syy = syy + dely * dely;
    sxy = sxy + delx * dely;
    for (int r = 0; r < 4; r++) {
      const D1 qk = q[4 * (size_t)k + r];
      if (qk > qmax[r]) qmax[r] = qk;
      if (qk < qmin[r]) qmin[r] = qk;
      sdx[r] = sdx[r] + delx * (qk - qi[r]);
      sdy[r] = sdy[r] + dely * (qk - qi[r]);
    }
  }
  const D1 det = sxx * syy - sxy * sxy;
  const D1 one_by_det = det != D1(0.0) ? D1(1.0) / det : D1(0.0);
  for (int r = 0; r < 4; r++) {
    dq[8 * (size_t)i + r] = (sdx[r] * syy - sdy[r] * sxy) * one_by_det;
    dq[8 * (size_t)i + 4 + r] = (sdy[r] * sxx - sdx[r] * sxy) * one_by_det;
    qm[8 * (size_t)i + r] = qmax[r];
    qm[8 * (size_t)i + 4 + r] = qmin[r];
  }
}

// q_variables_mod.h:120-228 (power 0): one Jacobi pass src -> dst
__global__ __launch_bounds__(256) void t_inner(int n, const int32_t *__restrict__ off,
                                               const uint32_t *__restrict__ conn,
                                               const D1 *__restrict__ x, const D1 *__restrict__ y,
                                               const D1 *__restrict__ q,
                                               const D1 *__restrict__ src, D1 *__restrict__ dst) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  D1 sxx, syy, sxy, sx[4], sy[4];
  for (int e = off[i]; e < off[i + 1]; e++) {
    const int k = (int)(conn[e] & kIndexMask);
    const D1 delx = x[k] - x[i], dely = y[k] - y[i];
    sxx = sxx + delx * delx;
    syy = syy + dely * dely;
    sxy = sxy + delx * dely;
    for (int r = 0; r < 4; r++) {
      const D1 temp1 = q[4 * (size_t)i + r] -
                       D1(0.5) * (delx * src[8 * (size_t)i + r] + dely * src[8 * (size_t)i + 4 + r]);
      const D1 temp2 = q[4 * (size_t)k + r] -
                       D1(0.5) * (delx * src[8 * (size_t)k + r] + dely * src[8 * (size_t)k + 4 + r]);
      sx[r] = sx[r] + delx * (temp2 - temp1);
      sy[r] = sy[r] + dely * (temp2 - temp1);
    }
  }
  const D1 det = sxx * syy - sxy * sxy;
  const D1 one_by_det = det != D1(0.0) ? D1(1.0) / det : D1(0.0);
  for (int r = 0; r < 4; r++) {
    dst[8 * (size_t)i + r] = one_by_det * (sx[r] * syy - sy[r] * sxy);
    dst[8 * (size_t)i + 4 + r] = one_by_det * (sy[r] * sxx - sx[r] * sxy);
  }
}

// flux_residual_mod.h:10-89 with the sub-stencil routines of
// interior_fluxes_mod.h, wall_fluxes_mod.h and outer_fluxes_mod.h, the
// geometry sums taken from the coefficients as in the FP64 kernel
__global__ __launch_bounds__(256) void t_flux(
    int n, const int32_t *__restrict__ off, const uint32_t *__restrict__ conn,
    const D1 *__restrict__ x, const D1 *__restrict__ y, const D1 *__restrict__ q,
    const D1 *__restrict__ dq, const D1 *__restrict__ qm, const D1 *__restrict__ coef,
    const D1 *__restrict__ nxa, const D1 *__restrict__ nya, const uint8_t *__restrict__ flag,
    const double *__restrict__ base, const D1 *__restrict__ delta, D1 *__restrict__ fres) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  D1 qi[4], dxi[4], dyi[4], qmxi[4], qmni[4], cf[8], acc[4];
  for (int r = 0; r < 4; r++) {
    qi[r] = q[4 * (size_t)i + r];
    dxi[r] = dq[8 * (size_t)i + r];
    dyi[r] = dq[8 * (size_t)i + 4 + r];
    qmxi[r] = qm[8 * (size_t)i + r];
    qmni[r] = qm[8 * (size_t)i + 4 + r];
  }
  for (int j = 0; j < 8; j++) cf[j] = coef[8 * (size_t)i + j];
  const D1 eps2_i(base[(size_t)B_STRIDE * i + B_EPS2]);
  const D1 nx = nxa[i], ny = nya[i], tx = ny, ty = -nx;
  const int type = flag[i];
  const D1 s2b(type == 0 ? -1.0 : 1.0);
  for (int e = off[i]; e < off[i + 1]; e++) {
    const uint32_t w = conn[e];
    const int k = (int)(w & kIndexMask);
    const uint32_t m = w >> kMaskShift;
    const D1 eps2_k(base[(size_t)B_STRIDE * k + B_EPS2]);
    const D1 delx = x[k] - x[i], dely = y[k] - y[i];
    const D1 dels = delx * tx + dely * ty, deln = delx * nx + dely * ny;
    D1 qti[4], qtk[4];
    for (int r = 0; r < 4; r++) {
      const D1 qk = q[4 * (size_t)k + r];
      const D1 di = delx * dxi[r] + dely * dyi[r];
      const D1 dk = delx * dq[8 * (size_t)k + r] + dely * dq[8 * (size_t)k + 4 + r];
      const D1 phi_i = venkat_phi(qi[r] - D1(0.5) * di, qi[r], qmxi[r], qmni[r], eps2_i);
      const D1 phi_k = venkat_phi(qk - D1(0.5) * dk, qk, qm[8 * (size_t)k + r],
                                  qm[8 * (size_t)k + 4 + r], eps2_k);
      qti[r] = qi[r] - D1(0.5) * phi_i * di;
      qtk[r] = qk - D1(0.5) * phi_k * dk;
    }
    D1 u1, u2, rho, sb, temp;
    qtilde_state(qti, u1, u2, rho, sb, temp);
    const EndTerms<D1> ei = end_terms_sb(nx, ny, u1, u2, rho, sb, temp);
    qtilde_state(qtk, u1, u2, rho, sb, temp);
    const EndTerms<D1> ek = end_terms_sb(nx, ny, u1, u2, rho, sb, temp);
    D1 Gi[4], Gk[4];
    const bool xp = (m & XPOS) != 0;
    const D1 sx(xp ? 1.0 : -1.0);
    if (type == 1) {
      split_x(ei, sx, Gi);
      split_x(ek, sx, Gk);
    } else {
      quad_x(ei, sx, s2b, Gi);
      quad_x(ek, sx, s2b, Gk);
    }
    const D1 wx = xp ? cf[0] * dels + cf[1] * deln : cf[2] * dels + cf[3] * deln;
    for (int r = 0; r < 4; r++) acc[r] = acc[r] + (Gk[r] - Gi[r]) * wx;
    if (xp && (m & XNEG)) {
      if (type == 1) {
        split_x(ei, D1(-1.0), Gi);
        split_x(ek, D1(-1.0), Gk);
      } else {
        quad_x(ei, D1(-1.0), s2b, Gi);
        quad_x(ek, D1(-1.0), s2b, Gk);
      }
      const D1 wn = cf[2] * dels + cf[3] * deln;
      for (int r = 0; r < 4; r++) acc[r] = acc[r] + (Gk[r] - Gi[r]) * wn;
    }
    const bool yp = type == 1 ? (m & YPOS) != 0 : type == 2;
    const D1 sy(yp ? 1.0 : -1.0);
    split_y(ei, sy, Gi);
    split_y(ek, sy, Gk);
    const D1 wy = yp ? cf[4] * deln + cf[5] * dels : cf[6] * deln + cf[7] * dels;
    for (int r = 0; r < 4; r++) acc[r] = acc[r] + (Gk[r] - Gi[r]) * wy;
    if (type == 1 && yp && (m & YNEG)) {
      split_y(ei, D1(-1.0), Gi);
      split_y(ek, D1(-1.0), Gk);
      const D1 wn = cf[6] * deln + cf[7] * dels;
      for (int r = 0; r < 4; r++) acc[r] = acc[r] + (Gk[r] - Gi[r]) * wn;
    }
  }
  const D1 d = type == 0 ? D1(2.0) * delta[i] : delta[i];
  for (int r = 0; r < 4; r++) fres[(size_t)r * n + i] = d * acc[r];
}

__global__ __launch_bounds__(256) void t_state_update(int n, int rk, D1 *__restrict__ prim,
                                                      const D1 *__restrict__ prim_old,
                                                      const D1 *__restrict__ fres,
                                                      const D1 *__restrict__ nxa,
                                                      const D1 *__restrict__ nya,
                                                      const uint8_t *__restrict__ flag,
                                                      const FlowConsts *__restrict__ fcs) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  D1 p1[4], p2[4], R[4], out[4];
  for (int r = 0; r < 4; r++) {
    p1[r] = prim[(size_t)r * n + i];
    p2[r] = prim_old[(size_t)r * n + i];
    R[r] = fres[(size_t)r * n + i];
  }
  su_point(rk, (int)flag[i], p1, p2, R, nxa[i], nya[i], fcs[0], out);
  for (int r = 0; r < 4; r++) prim[(size_t)r * n + i] = out[r];
}

D1 *as_d1(Tan *p) { return reinterpret_cast<D1 *>(p); }

}  // namespace

void launch_su_vjp(const DevCloud &c, int rk, const double *prim, const double *prim_old,
                   const double *fres, const FlowConsts *fc, double *pbar, double *pold_bar,
                   double *rbar, double *nrm_bar, hipStream_t s) {
  hipLaunchKernelGGL(k_su_vjp, dim3(blocks_of(c.n)), dim3(256), 0, s, c.n, rk, prim, prim_old,
                     fres, c.nx, c.ny, c.flag, fc, pbar, pold_bar, rbar, nrm_bar);
}

void launch_delta_vjp(const DevCloud &c, const AdjointBuffers &a, const double *base,
                      const double *prim, double cfl, const double *dbar, int32_t *amin_e,
                      double *amin_v, double *geo_edge, double *pbar, hipStream_t s) {
  hipLaunchKernelGGL(k_delta_vjp, dim3(blocks_of(c.n)), dim3(256), 0, s, c.n, c.conn_off,
                     c.conn, base, cfl, dbar, amin_e, amin_v, geo_edge);
  hipLaunchKernelGGL(k_delta_vjp_gather, dim3(blocks_of(c.n)), dim3(256), 0, s, c.n, a.tr_off,
                     a.tr_edge, a.tr_owner, amin_e, amin_v, prim, pbar);
}

void launch_axpy(size_t count, const double *x, double *y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(1024), dim3(256), 0, s, count, x, y);
}

void launch_tan_geometry(const DevCloud &c, const int32_t *lr, const TanState &t,
                         hipStream_t s) {
  const dim3 g(blocks_of(c.n)), b(256);
  hipLaunchKernelGGL(t_normals, g, b, 0, s, c.n, c.flag, lr, as_d1(t.x), as_d1(t.y),
                     as_d1(t.nx), as_d1(t.ny));
  hipLaunchKernelGGL(t_coef, g, b, 0, s, c.n, c.conn_off, c.conn, as_d1(t.x), as_d1(t.y),
                     as_d1(t.nx), as_d1(t.ny), as_d1(t.coef));
}

void launch_tan_iteration(const DevCloud &c, const double *base, const FlowConsts *fc,
                          double cfl, int inner, const TanState &t, hipStream_t s) {
  const int n = c.n;
  const dim3 g(blocks_of(n)), b(256);
  (void)hipMemcpyAsync(t.prim_old, t.prim, sizeof(Tan) * 4 * (size_t)n,
                       hipMemcpyDeviceToDevice, s);
  hipLaunchKernelGGL(t_speed, g, b, 0, s, n, as_d1(t.prim), as_d1(t.speed));
  hipLaunchKernelGGL(t_delta, g, b, 0, s, n, c.conn_off, c.conn, as_d1(t.x), as_d1(t.y),
                     as_d1(t.speed), cfl, as_d1(t.delta));
  for (int rk = 1; rk <= 4; rk++) {
    hipLaunchKernelGGL(t_qvar, g, b, 0, s, n, as_d1(t.prim), as_d1(t.q));
    hipLaunchKernelGGL(t_qder, g, b, 0, s, n, c.conn_off, c.conn, as_d1(t.x), as_d1(t.y),
                       as_d1(t.q), as_d1(t.dq[0]), as_d1(t.qm));
    int cur = 0;
    for (int p = 0; p < inner; p++) {
      hipLaunchKernelGGL(t_inner, g, b, 0, s, n, c.conn_off, c.conn, as_d1(t.x), as_d1(t.y),
                         as_d1(t.q), as_d1(t.dq[cur]), as_d1(t.dq[1 - cur]));
      cur = 1 - cur;
    }
    hipLaunchKernelGGL(t_flux, g, b, 0, s, n, c.conn_off, c.conn, as_d1(t.x), as_d1(t.y),
                       as_d1(t.q), as_d1(t.dq[cur]), as_d1(t.qm), as_d1(t.coef), as_d1(t.nx),
                       as_d1(t.ny), c.flag, base, as_d1(t.delta), as_d1(t.fres));
    hipLaunchKernelGGL(t_state_update, g, b, 0, s, n, rk, as_d1(t.prim), as_d1(t.prim_old),
                       as_d1(t.fres), as_d1(t.nx), as_d1(t.ny), c.flag, fc);
  }
}

}  // namespace lskum
