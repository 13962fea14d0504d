// adjoint.hip -- discrete adjoint (reverse) and tangent-linear kernels of the
// flux residual with respect to the primitive state, FP64, power = 0.
//
// The map differentiated is one cal_flux_residual evaluation as a function
// of prim, with the local time step delta held fixed (flux_residual_mod.h:
// 10-89 after q_variables_mod.h:9-228):
//   prim -> q (eval_q_variables) -> dq, qmax/qmin (eval_q_derivatives)
//        -> n inner passes (eval_q_inner_loop) -> flux_res.
// The reverse pass recomputes q, dq and qmax/qmin from prim (the checkpoint
// is prim alone; nothing of a forward run is stored) and then runs each
// stage backwards:
//  - flux: per (point, neighbour) pair, each end's summed sub-stencil
//    fluxes run forward and then in reverse (the split and quadrant fluxes,
//    the kinetic terms and q~ -> primitive, each reversed by hand), giving
//    the adjoint of q~ at both ends and of the flux frame; then the limiter
//    and reconstruction partials.  The neighbour's share of each pair goes
//    to a per-edge buffer and a transpose-CSR gather sums it per point in a
//    fixed order (no atomics, deterministic).
//  - qmax/qmin: routed to the point each extremum came from.
//  - inner passes and q_derivatives: linear least-squares stencils, whose
//    transposes are applied by the same gather pattern.
//  - q_variables: the pointwise 4x4 Jacobian transposed.
// The tangent-linear map runs the same chain forwards (the linear stencils
// through the forward kernels themselves) for the duality check.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"
#include "physics.hpp"
// after physics.hpp: the physics templates look up exp/log/sqrt/erf for
// double at their definition and find the Dual overloads by argument
#include "dual.hpp"

namespace lskum {

namespace {

constexpr int COEF_STRIDE = 8;

// limiters_mod.h:55-104 for one component (del_neg outside the threshold)
// with d phi/d del_neg (pn) and d phi/d del_pos (pp); phi = 1 and zero
// partials where the ratio reaches 1.
__device__ __forceinline__ double venkat_phi_d(double N, double P, double eps2, double &pn,
                                               double &pp) {
  pn = pp = 0.0;
  const double num = (P * P + eps2) * N + 2.0 * N * N * P;
  const double den = (P * P + 2.0 * N * N + N * P + eps2) * N;
  const double t = num / den;
  if (!(t < 1.0)) return 1.0;
  const double num_n = P * P + eps2 + 4.0 * N * P;
  const double den_n = 6.0 * N * N + 2.0 * N * P + P * P + eps2;
  const double num_p = 2.0 * P * N + 2.0 * N * N;
  const double den_p = (2.0 * P + N) * N;
  const double inv2 = 1.0 / (den * den);
  pn = (num_n * den - num * den_n) * inv2;
  pp = (num_p * den - num * den_p) * inv2;
  return t;
}

// q~ = q - 0.5 phi d (interior_fluxes_mod.h:72-83) with its partials with
// respect to q, the increment d = dx.dq, and the qmax or qmin it used.
struct Recon {
  double qt, dq, dd, dm;
  bool use_max;
};

__device__ __forceinline__ Recon recon(double q, double d, double qmax, double qmin,
                                       double eps2) {
  Recon o;
  const double N = (q - 0.5 * d) - q;
  double phi = 1.0, pn = 0.0, pp = 0.0;
  o.use_max = N > 0.0;
  if (fabs(N) > 10e-6) phi = venkat_phi_d(N, (N > 0.0 ? qmax : qmin) - q, eps2, pn, pp);
  o.qt = q - 0.5 * phi * d;
  o.dq = 1.0 + 0.5 * d * pp;           // through del_pos = qm - q
  o.dd = -0.5 * phi + 0.25 * d * pn;   // del_neg = -0.5 d
  o.dm = -0.5 * d * pp;
  return o;
}

// Sum over the sub-stencil fluxes a (point, neighbour) pair feeds of
// w_s G_s(e), for one end's terms e: the selection of the FP64 flux kernel
// (x+ or x-, both on a tie; interior y+ / y-, wall y-, outer y+).
template <typename T>
__device__ __forceinline__ void end_sum(const EndTerms<T> &e, int type, uint32_t m, double s2b,
                                        const double cf[8], double dsw, double dnw, T out[4]) {
  T G[4];
  const bool xp = (m & XPOS) != 0;
  if (type == 1)
    split_x(e, T(xp ? 1.0 : -1.0), G);
  else
    quad_x(e, T(xp ? 1.0 : -1.0), T(s2b), G);
  const double wx = xp ? cf[0] * dsw + cf[1] * dnw : cf[2] * dsw + cf[3] * dnw;
  for (int r = 0; r < 4; r++) out[r] = G[r] * wx;
  if (xp && (m & XNEG)) {
    if (type == 1)
      split_x(e, T(-1.0), G);
    else
      quad_x(e, T(-1.0), T(s2b), G);
    const double wn = cf[2] * dsw + cf[3] * dnw;
    for (int r = 0; r < 4; r++) out[r] = out[r] + G[r] * wn;
  }
  const bool yp = type == 1 ? (m & YPOS) != 0 : type == 2;
  split_y(e, T(yp ? 1.0 : -1.0), G);
  const double wy = yp ? cf[4] * dnw + cf[5] * dsw : cf[6] * dnw + cf[7] * dsw;
  for (int r = 0; r < 4; r++) out[r] = out[r] + G[r] * wy;
  if (type == 1 && yp && (m & YNEG)) {
    split_y(e, T(-1.0), G);
    const double wn = cf[6] * dnw + cf[7] * dsw;
    for (int r = 0; r < 4; r++) out[r] = out[r] + G[r] * wn;
  }
}

// end_sum for the state q~ with N seeded directions: the value and, per
// direction, its derivative along the seed (the tangent kernel's Dual<1>).
template <int N>
__device__ __forceinline__ void end_sum_dual(const Dual<N> qt[4], double nx, double ny, int type,
                                             uint32_t m, double s2b, const double cf[8],
                                             double dsw, double dnw, Dual<N> out[4]) {
  Dual<N> u1, u2, rho, sb, temp;
  qtilde_state(qt, u1, u2, rho, sb, temp);
  const EndTerms<Dual<N>> e = end_terms_sb(Dual<N>(nx), Dual<N>(ny), u1, u2, rho, sb, temp);
  end_sum(e, type, m, s2b, cf, dsw, dnw, out);
}

// ------------------------------------------------------------ leaf physics, reverse
// Adjoints of the per-end terms (EndTerms fields), accumulated by the
// reverse flux functions below.
struct EndBar {
  double rho = 0.0, ut = 0.0, un = 0.0, prr = 0.0, usq = 0.0, B1 = 0.0, B2 = 0.0, e1 = 0.0,
         e2 = 0.0;
};

// split_x (split_fluxes_mod.h:6-84) reversed: gb is the adjoint of G[4].
// G_r = rho P_r with P0 = ut A1 + s B1, P1 = (p/rho + ut^2) A1 + s ut B1,
// P2 = ut un A1 + s un B1, P3 = ut (7p/rho + |u|^2) A1 / 2 + s (6p/rho +
// |u|^2) B1 / 2, A1 = (1 + s erf S1) / 2.
__device__ __forceinline__ void split_x_adj(const EndTerms<double> &e, double s,
                                            const double gb[4], EndBar &b) {
  const double A1 = 0.5 * (1.0 + s * e.e1);
  const double t1 = e.pr_by_rho + e.ut * e.ut;
  const double t7 = 7.0 * e.pr_by_rho + e.u_sqr, t6 = 6.0 * e.pr_by_rho + e.u_sqr;
  const double P0 = e.ut * A1 + s * e.B1;
  const double P1 = t1 * A1 + s * (e.ut * e.B1);
  const double P2 = e.ut * e.un * A1 + s * (e.un * e.B1);
  const double P3 = 0.5 * e.ut * t7 * A1 + s * (0.5 * t6 * e.B1);
  b.rho += gb[0] * P0 + gb[1] * P1 + gb[2] * P2 + gb[3] * P3;
  const double R0 = e.rho * gb[0], R1 = e.rho * gb[1], R2 = e.rho * gb[2], R3 = e.rho * gb[3];
  const double A1b = R0 * e.ut + R1 * t1 + R2 * (e.ut * e.un) + R3 * (0.5 * e.ut * t7);
  b.ut += R0 * A1 + R1 * (2.0 * e.ut * A1 + s * e.B1) + R2 * (e.un * A1) + R3 * (0.5 * t7 * A1);
  b.un += R2 * (e.ut * A1 + s * e.B1);
  b.prr += R1 * A1 + R3 * (3.5 * e.ut * A1 + 3.0 * s * e.B1);
  b.usq += R3 * (0.5 * e.ut * A1 + 0.5 * s * e.B1);
  b.B1 += s * (R0 + R1 * e.ut + R2 * e.un + R3 * (0.5 * t6));
  b.e1 += A1b * (0.5 * s);
}

// split_y (split_fluxes_mod.h:86-164) reversed: the same with (un, A2, B2)
// and the momentum rows exchanged.
__device__ __forceinline__ void split_y_adj(const EndTerms<double> &e, double s,
                                            const double gb[4], EndBar &b) {
  const double A2 = 0.5 * (1.0 + s * e.e2);
  const double t1 = e.pr_by_rho + e.un * e.un;
  const double t7 = 7.0 * e.pr_by_rho + e.u_sqr, t6 = 6.0 * e.pr_by_rho + e.u_sqr;
  const double P0 = e.un * A2 + s * e.B2;
  const double P1 = e.ut * e.un * A2 + s * (e.ut * e.B2);
  const double P2 = t1 * A2 + s * (e.un * e.B2);
  const double P3 = 0.5 * e.un * t7 * A2 + s * (0.5 * t6 * e.B2);
  b.rho += gb[0] * P0 + gb[1] * P1 + gb[2] * P2 + gb[3] * P3;
  const double R0 = e.rho * gb[0], R1 = e.rho * gb[1], R2 = e.rho * gb[2], R3 = e.rho * gb[3];
  const double A2b = R0 * e.un + R1 * (e.ut * e.un) + R2 * t1 + R3 * (0.5 * e.un * t7);
  b.un += R0 * A2 + R1 * (e.ut * A2) + R2 * (2.0 * e.un * A2 + s * e.B2) + R3 * (0.5 * t7 * A2);
  b.ut += R1 * (e.un * A2 + s * e.B2);
  b.prr += R2 * A2 + R3 * (3.5 * e.un * A2 + 3.0 * s * e.B2);
  b.usq += R3 * (0.5 * e.un * A2 + 0.5 * s * e.B2);
  b.B2 += s * (R0 + R1 * e.ut + R2 * e.un + R3 * (0.5 * t6));
  b.e2 += A2b * (0.5 * s);
}

// quad_x (quadrant_fluxes_mod.h:6-212) reversed.
__device__ __forceinline__ void quad_x_adj(const EndTerms<double> &e, double s1, double s2,
                                           const double gb[4], EndBar &b) {
  const double A1 = 0.5 * (1.0 + s1 * e.e1), A2 = 0.5 * (1.0 + s2 * e.e2);
  const double a1b1 = e.ut * A1 + s1 * e.B1;
  const double t1 = e.pr_by_rho + e.ut * e.ut;
  const double t2 = t1 * A1 + s1 * (e.ut * e.B1);
  const double t2b = e.un * A2 + s2 * e.B2;
  const double t7 = 7.0 * e.pr_by_rho + e.u_sqr, t6 = 6.0 * e.pr_by_rho + e.u_sqr;
  const double t2c = 0.5 * e.ut * t7 * A1, t3 = 0.5 * e.B1 * t6;
  const double M = t2c + s1 * t3;
  // G0 = rho A2 a1b1, G1 = rho A2 t2, G2 = rho a1b1 t2b,
  // G3 = rho A2 M + s2 (rho un B2 a1b1) / 2
  double rhob = gb[0] * A2 * a1b1 + gb[1] * A2 * t2 + gb[2] * a1b1 * t2b + gb[3] * A2 * M;
  double A2b = e.rho * (gb[0] * a1b1 + gb[1] * t2 + gb[3] * M);
  double a1b1b = e.rho * (gb[0] * A2 + gb[2] * t2b);
  const double t2bar = gb[1] * e.rho * A2, t2bb = gb[2] * e.rho * a1b1;
  const double t2cb = gb[3] * e.rho * A2, t3b = t2cb * s1, t4b = gb[3] * s2;
  double A1b = t2bar * t1, utb = t2bar * (s1 * e.B1), B1b = t2bar * (s1 * e.ut);
  const double t1b = t2bar * A1;
  double prrb = t1b, unb = t2bb * A2, B2b = t2bb * s2, usqb = 0.0;
  utb += t1b * 2.0 * e.ut;
  A2b += t2bb * e.un;
  utb += t2cb * (0.5 * t7 * A1);
  const double t7b = t2cb * (0.5 * e.ut * A1);
  A1b += t2cb * (0.5 * e.ut * t7);
  prrb += 7.0 * t7b;
  usqb += t7b;
  B1b += t3b * (0.5 * t6);
  const double t6b = t3b * (0.5 * e.B1);
  prrb += 6.0 * t6b;
  usqb += t6b;
  rhob += t4b * (0.5 * e.un * e.B2 * a1b1);
  unb += t4b * (0.5 * e.rho * e.B2 * a1b1);
  B2b += t4b * (0.5 * e.rho * e.un * a1b1);
  a1b1b += t4b * (0.5 * e.rho * e.un * e.B2);
  utb += a1b1b * A1;
  A1b += a1b1b * e.ut;
  B1b += a1b1b * s1;
  b.rho += rhob;
  b.ut += utb;
  b.un += unb;
  b.prr += prrb;
  b.usq += usqb;
  b.B1 += B1b;
  b.B2 += B2b;
  b.e1 += A1b * (0.5 * s1);
  b.e2 += A2b * (0.5 * s2);
}

// One end of a pair run forward and then in reverse.  From its limited
// state q~, the sum S[r] of the fluxes of end_sum (with the same
// selection), per flux j the contraction g[j] = sum_r la_r G_j,r, and from
// the adjoint weights sgn * la_r * w_j of those fluxes (sgn = -1 at the
// point's own end, +1 at the neighbour's) the adjoint of q~ (qb) and of the
// normal through the flux frame (nb).
__device__ __forceinline__ void end_adj(const double qt[4], double nx, double ny, int type,
                                        uint32_t m, double s2b, const double cf[8], double dels,
                                        double deln, const double la[4], double sgn,
                                        double qb[4], double nb[2], double g[4], double S[4]) {
  double u1, u2, rho, sb, temp;
  qtilde_state(qt, u1, u2, rho, sb, temp);
  const EndTerms<double> e = end_terms_sb(nx, ny, u1, u2, rho, sb, temp);
  EndBar b;
  for (int r = 0; r < 4; r++) S[r] = g[r] = 0.0;
  // kind 0 split_x, 1 quad_x, 2 split_y
  auto flux = [&](int j, int kind, double s1, double w) {
    double G[4], gb[4], t = 0.0;
    if (kind == 0)
      split_x(e, s1, G);
    else if (kind == 1)
      quad_x(e, s1, s2b, G);
    else
      split_y(e, s1, G);
    for (int r = 0; r < 4; r++) {
      S[r] += G[r] * w;
      t += la[r] * G[r];
      gb[r] = sgn * la[r] * w;
    }
    g[j] = t;
    if (kind == 0)
      split_x_adj(e, s1, gb, b);
    else if (kind == 1)
      quad_x_adj(e, s1, s2b, gb, b);
    else
      split_y_adj(e, s1, gb, b);
  };
  const bool xp = (m & XPOS) != 0;
  const int kx = type == 1 ? 0 : 1;
  flux(0, kx, xp ? 1.0 : -1.0, xp ? cf[0] * dels + cf[1] * deln : cf[2] * dels + cf[3] * deln);
  if (xp && (m & XNEG)) flux(1, kx, -1.0, cf[2] * dels + cf[3] * deln);
  const bool yp = type == 1 ? (m & YPOS) != 0 : type == 2;
  flux(2, 2, yp ? 1.0 : -1.0, yp ? cf[4] * deln + cf[5] * dels : cf[6] * deln + cf[7] * dels);
  if (type == 1 && yp && (m & YNEG)) flux(3, 2, -1.0, cf[6] * deln + cf[7] * dels);
  // end_terms_sb reversed: S1 = ut sb, S2 = un sb, h = sb (p/rho) / sqrt(pi),
  // B = exp(-S^2) h, e = erf(S), |u|^2 = ut^2 + un^2
  const double h = kInvSqrtPi * sb * temp;
  const double S1 = e.ut * sb, S2 = e.un * sb;
  const double X1 = e.B1 / h, X2 = e.B2 / h;  // exp(-S1^2), exp(-S2^2)
  constexpr double kTwoBySqrtPi = 1.1283791670955126;
  const double hb = b.B1 * X1 + b.B2 * X2;
  const double S1b = X1 * (-2.0 * S1 * (b.B1 * h) + kTwoBySqrtPi * b.e1);
  const double S2b = X2 * (-2.0 * S2 * (b.B2 * h) + kTwoBySqrtPi * b.e2);
  const double utb = b.ut + 2.0 * e.ut * b.usq + S1b * sb;
  const double unb = b.un + 2.0 * e.un * b.usq + S2b * sb;
  const double sbb = S1b * e.ut + S2b * e.un + hb * kInvSqrtPi * temp;
  double tempb = b.prr + hb * kInvSqrtPi * sb;
  // ut = u1 ny - u2 nx, un = u1 nx + u2 ny
  double u1b = utb * ny + unb * nx, u2b = -utb * nx + unb * ny;
  nb[0] = -utb * u2 + unb * u1;
  nb[1] = utb * u1 + unb * u2;
  // qtilde_state reversed: beta = -q3/2, p/rho = 1/(2 beta), u = q1..2 p/rho,
  // rho = exp(q0 + beta |u|^2) beta^-2.5 (sb = sqrt(beta))
  const double beta = -qt[3] * 0.5;
  const double T1b = b.rho * rho;
  double betab = T1b * (-2.5 / beta) + T1b * (u1 * u1 + u2 * u2) + sbb * 0.5 / sb;
  u1b += T1b * 2.0 * beta * u1;
  u2b += T1b * 2.0 * beta * u2;
  qb[0] = T1b;
  qb[1] = u1b * temp;
  qb[2] = u2b * temp;
  tempb += u1b * qt[1] + u2b * qt[2];
  betab += -tempb * temp / beta;
  qb[3] = -0.5 * betab;
}

__device__ __forceinline__ void load8(const double *p, double a[4], double b[4]) {
  load4(p, a);
  load4(p + 4, b);
}

// ------------------------------------------------------------ flux, reverse
// Per point i, over its neighbours: the adjoint of q~ at both ends of each
// pair, pushed through the reconstruction.  i's own share accumulates in
// registers (own_q, own_dq, own_qm); the neighbour's goes to the edge's
// record: c_q[4], c_d[4] (per unit of d, times dx/dy at the gather) and
// c_m[4], the adjoint of the extremum each limiter read, with a byte whose
// bit r says whether that was qmax or qmin, both stored at the edge's
// position in the transpose lists so that the gather reads each point's
// records contiguously (96 + 1 bytes a pair).
//
// GEO adds the adjoint of the geometry the flux reads, accumulated (+=)
// into per-edge and per-point arrays owned by i alone:
//   geo_edge[e] (d/d delx, d/d dely of pair e): through the increments
//     d = delx dqx + dely dqy of both reconstructions and through the
//     weights' dels = delx ny - dely nx, deln = delx nx + dely ny;
//   geo_n[i] (d/d nx, d/d ny): through dels, deln and the flux frame;
//   geo_cf[i] (d/d of the eight sub-stencil coefficients);
//   geo_delta[i] (d/d delta).
//
// TILED stages the block's neighbour records (base, dq, qm: 192 B per tile
// entry) in LDS first.  The kernel's registers allow one wave per SIMD, one
// block per CU, so the tile costs no occupancy, and it takes the three
// dependent gathers of every neighbour off the critical path.
template <bool GEO, bool TILED>
__global__ __launch_bounds__(256) void k_flux_vjp(
    int n, const int32_t *__restrict__ off, const uint32_t *__restrict__ conn,
    const uint16_t *__restrict__ conn16, const int32_t *__restrict__ tile_off,
    const int32_t *__restrict__ tile_idx,
    const double *__restrict__ base, const double *__restrict__ dq,
    const double *__restrict__ qm, const double *__restrict__ coef,
    const double *__restrict__ nxa, const double *__restrict__ nya,
    const uint8_t *__restrict__ flag, const double *__restrict__ delta,
    const double *__restrict__ lam, double *__restrict__ own_q, double *__restrict__ own_dq,
    double *__restrict__ own_qm, double *__restrict__ edge, uint8_t *__restrict__ edge_mx,
    const int32_t *__restrict__ tr_pos, double *__restrict__ geo_edge,
    double *__restrict__ geo_n, double *__restrict__ geo_cf, double *__restrict__ geo_delta) {
  extern __shared__ double lds[];
  constexpr int S = 24;  // base (8), dq (8), qm (8)
  const int b = xcd_block();
  if (TILED) {
    load_tile<S>(lds, tile_off, tile_idx, b, base, B_STRIDE, 0, 4, 0);
    load_tile<S>(lds, tile_off, tile_idx, b, dq, DQ_STRIDE, 0, 4, 8);
    load_tile<S>(lds, tile_off, tile_idx, b, qm, QM_STRIDE, 0, 4, 16);
    __syncthreads();
  }
  const int i = b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int type = flag[i];
  // flux_res = c delta acc with c = 2 on the wall (flux_residual_mod.h:29)
  const double cw = type == 0 ? 2.0 : 1.0;
  const double sc = cw * delta[i];
  double la[4], lc[4];
  for (int r = 0; r < 4; r++) {
    la[r] = sc * lam[(size_t)r * n + i];
    lc[r] = cw * lam[(size_t)r * n + i];
  }
  double gcf[8] = {0, 0, 0, 0, 0, 0, 0, 0}, gnx = 0.0, gny = 0.0, gdel = 0.0;
  const double *hi = TILED ? lds + S * threadIdx.x : base + (size_t)B_STRIDE * i;
  const double x_i = hi[B_X], y_i = hi[B_Y], eps2_i = hi[B_EPS2];
  double qi[4], dxi[4], dyi[4], qmxi[4], qmni[4], cf[8];
  load4(hi + B_Q, qi);
  load8(TILED ? hi + 8 : dq + (size_t)DQ_STRIDE * i, dxi, dyi);
  load8(TILED ? hi + 16 : qm + (size_t)QM_STRIDE * i, qmxi, qmni);
  load8(coef + (size_t)COEF_STRIDE * i, cf, cf + 4);
  const double nx = nxa[i], ny = nya[i], tx = ny, ty = -nx;
  const double s2b = type == 0 ? -1.0 : 1.0;
  double aq[4] = {0, 0, 0, 0}, adx[4] = {0, 0, 0, 0}, ady[4] = {0, 0, 0, 0};
  double amx[4] = {0, 0, 0, 0}, amn[4] = {0, 0, 0, 0};
  for (int e = off[i]; e < off[i + 1]; e++) {
    const uint32_t w = conn_raw(conn, conn16, e, TILED);
    const int k = conn_index(w, TILED);
    const uint32_t m = conn_mask(w, TILED);
    const double *hk = TILED ? lds + S * k : base + (size_t)B_STRIDE * k;
    double qk[4], dxk[4], dyk[4], qmxk[4], qmnk[4];
    load4(hk + B_Q, qk);
    load8(TILED ? hk + 8 : dq + (size_t)DQ_STRIDE * k, dxk, dyk);
    load8(TILED ? hk + 16 : qm + (size_t)QM_STRIDE * k, qmxk, qmnk);
    const double delx = hk[B_X] - x_i, dely = hk[B_Y] - y_i;
    const double dels = delx * tx + dely * ty, deln = delx * nx + dely * ny;
    Recon Ri[4], Rk[4];
    for (int r = 0; r < 4; r++) {
      Ri[r] = recon(qi[r], delx * dxi[r] + dely * dyi[r], qmxi[r], qmni[r], eps2_i);
      Rk[r] = recon(qk[r], delx * dxk[r] + dely * dyk[r], qmxk[r], qmnk[r], hk[B_EPS2]);
    }
    // acc += end_sum(k end) - end_sum(i end): the adjoint of q~ at each end
    double a_i[4], a_k[4], g_i[4], g_k[4], S_i[4], S_k[4], n_i[2], n_k[2];
    {
      double qt[4];
      for (int r = 0; r < 4; r++) qt[r] = Ri[r].qt;
      end_adj(qt, nx, ny, type, m, s2b, cf, dels, deln, la, -1.0, a_i, n_i, g_i, S_i);
      for (int r = 0; r < 4; r++) qt[r] = Rk[r].qt;
      end_adj(qt, nx, ny, type, m, s2b, cf, dels, deln, la, 1.0, a_k, n_k, g_k, S_k);
    }
    if (GEO) {
      for (int r = 0; r < 4; r++) gdel += lc[r] * (S_k[r] - S_i[r]);
      // weights: x flux a dels + b deln, y flux a deln + b dels, with (a, b)
      // the coefficients of the sub-stencil the flux belongs to
      double g[4], bs = 0.0, bn = 0.0;
      for (int j = 0; j < 4; j++) g[j] = g_k[j] - g_i[j];
      const bool xp = (m & XPOS) != 0;
      const bool yp = type == 1 ? (m & YPOS) != 0 : type == 2;
      const int cx = xp ? 0 : 2, cy = yp ? 4 : 6;
      gcf[cx] += g[0] * dels;
      gcf[cx + 1] += g[0] * deln;
      bs += g[0] * cf[cx];
      bn += g[0] * cf[cx + 1];
      gcf[2] += g[1] * dels;
      gcf[3] += g[1] * deln;
      bs += g[1] * cf[2];
      bn += g[1] * cf[3];
      gcf[cy] += g[2] * deln;
      gcf[cy + 1] += g[2] * dels;
      bn += g[2] * cf[cy];
      bs += g[2] * cf[cy + 1];
      gcf[6] += g[3] * deln;
      gcf[7] += g[3] * dels;
      bn += g[3] * cf[6];
      bs += g[3] * cf[7];
      gnx += -bs * dely + bn * delx + (n_k[0] + n_i[0]);
      gny += bs * delx + bn * dely + (n_k[1] + n_i[1]);
      double gx = bs * ny + bn * nx, gy = -bs * nx + bn * ny;
      for (int r = 0; r < 4; r++) {
        const double ai = a_i[r] * Ri[r].dd, ak = a_k[r] * Rk[r].dd;
        gx += ai * dxi[r] + ak * dxk[r];
        gy += ai * dyi[r] + ak * dyk[r];
      }
      geo_edge[2 * (size_t)e] += gx;
      geo_edge[2 * (size_t)e + 1] += gy;
    }
    const size_t tp = (size_t)tr_pos[e];
    double *rec = edge + EDGE_REC * tp;
    uint32_t mx = 0;
    for (int r = 0; r < 4; r++) {
      aq[r] += a_i[r] * Ri[r].dq;
      const double ad = a_i[r] * Ri[r].dd;
      adx[r] += ad * delx;
      ady[r] += ad * dely;
      if (Ri[r].use_max)
        amx[r] += a_i[r] * Ri[r].dm;
      else
        amn[r] += a_i[r] * Ri[r].dm;
      rec[r] = a_k[r] * Rk[r].dq;
      rec[4 + r] = a_k[r] * Rk[r].dd;
      rec[8 + r] = a_k[r] * Rk[r].dm;
      mx |= (Rk[r].use_max ? 1u : 0u) << r;
    }
    edge_mx[tp] = (uint8_t)mx;
  }
  store4(own_q + 4 * (size_t)i, aq);
  store4(own_dq + 8 * (size_t)i, adx);
  store4(own_dq + 8 * (size_t)i + 4, ady);
  store4(own_qm + 8 * (size_t)i, amx);
  store4(own_qm + 8 * (size_t)i + 4, amn);
  if (GEO) {
    for (int j = 0; j < 8; j++) geo_cf[8 * (size_t)i + j] += gcf[j];
    geo_n[2 * (size_t)i] += gnx;
    geo_n[2 * (size_t)i + 1] += gny;
    geo_delta[i] += gdel;
  }
}

// Per point j: its own share plus the edge records of every pair in which
// it is the neighbour (owners ascending), its r-th at slot
// sbase[j / 256] + 256 r + j % 256.
__global__ __launch_bounds__(256) void k_flux_vjp_gather(
    int n, const int32_t *__restrict__ tr_off, const int32_t *__restrict__ sbase,
    const int32_t *__restrict__ owner_s, const double2 *__restrict__ xy,
    const double *__restrict__ edge,
    const uint8_t *__restrict__ edge_mx, const double *__restrict__ own_q,
    const double *__restrict__ own_dq, const double *__restrict__ own_qm,
    double *__restrict__ lq, double *__restrict__ ldq, double *__restrict__ lqm) {
  const int j = xcd_block() * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double aq[4], adx[4], ady[4], amx[4], amn[4];
  load4(own_q + 4 * (size_t)j, aq);
  load8(own_dq + 8 * (size_t)j, adx, ady);
  load8(own_qm + 8 * (size_t)j, amx, amn);
  const double x_j = xy[j].x, y_j = xy[j].y;
  // record r + 1 with its owner's coordinates, and owner r + 2's index, load
  // while record r is added
  const int s0 = sbase[j >> 8] + (j & 255), rlast = tr_off[j + 1] - tr_off[j] - 1;
  auto slot = [&](int r) { return (size_t)s0 + 256 * (size_t)r; };
  auto fetch = [&](int r, int i, double2 &p, double r12[EDGE_REC], uint32_t &mx) {
    p = xy[i];
    const double *rec = edge + EDGE_REC * slot(r);
    load8(rec, r12, r12 + 4);
    load4(rec + 8, r12 + 8);
    mx = edge_mx[slot(r)];
  };
  double2 po = {0.0, 0.0};
  double rc[EDGE_REC] = {};
  uint32_t mc = 0;
  int i_n = 0;
  if (rlast >= 0) {
    fetch(0, owner_s[slot(0)], po, rc, mc);
    i_n = owner_s[slot(min(1, rlast))];
  }
  for (int t = 0; t <= rlast; t++) {
    double2 po_n;
    double rn[EDGE_REC];
    uint32_t mn;
    fetch(min(t + 1, rlast), i_n, po_n, rn, mn);
    i_n = owner_s[slot(min(t + 2, rlast))];
    const double delx = x_j - po.x;
    const double dely = y_j - po.y;
    for (int r = 0; r < 4; r++) {
      aq[r] += rc[r];
      adx[r] += rc[4 + r] * delx;
      ady[r] += rc[4 + r] * dely;
      // the other extremum gains an exact 0, as it did from a record of both
      const bool use_max = (mc >> r) & 1u;
      amx[r] += use_max ? rc[8 + r] : 0.0;
      amn[r] += use_max ? 0.0 : rc[8 + r];
    }
    po = po_n;
    for (int r = 0; r < EDGE_REC; r++) rc[r] = rn[r];
    mc = mn;
  }
  store4(lq + 4 * (size_t)j, aq);
  store4(ldq + 8 * (size_t)j, adx);
  store4(ldq + 8 * (size_t)j + 4, ady);
  store4(lqm + 8 * (size_t)j, amx);
  store4(lqm + 8 * (size_t)j + 4, amn);
}

// ------------------------------------------------------------ LS stencils, reverse
// ls_i = M_i^-T lo_i: the adjoint of the sums (sx, sy) a least-squares pass
// solves, from the adjoint of its output (q_variables_mod.h:94-112 and
// :176-210 solve with the same 2x2 matrix of the point's geometry, whose
// sums lsum the recomputed q_derivatives stored).  The pairs (i, k) of i's
// own list then give i its share t = dx.ls_i per component: -t to its q
// adjoint (lq, accumulated) and, for an inner pass (with_dq), +0.5 dx t to
// the adjoint of the derivatives the pass read (ldq_in, written); the
// gather adds the shares of the pairs in which i is the neighbour.
//
// GEO adds, per pair e = (i, k) with rho = (delx, dely), the derivative of
// sum_r lo_r . out_r with respect to rho (accumulated into geo_edge[e]):
// out_r = M^-1 s_r with M = sum rho rho^T and s_r = sum rho dq_r, so
//   d/d rho = sum_r ls_r (dq_r - rho . out_r) - (ls_r . rho) (out_r + 0.5 h_r)
// where dq_r is the pair's difference the pass sums (q_k - q_i, less
// 0.5 rho . h_r for an inner pass) and h_r = dq_in_k,r - dq_in_i,r the
// difference of the derivatives an inner pass reads (0 for q_derivatives,
// dq_in = NULL).  TILED (GEO with dq_in = NULL only) first stages the
// coordinates and q of the block's tile in LDS (64 B per entry) and reads
// the neighbours' from there.
template <bool GEO, bool TILED = false>
__global__ __launch_bounds__(256) void k_ls_adj(int n, const int32_t *__restrict__ off,
                                                const uint32_t *__restrict__ conn,
                                                const uint16_t *__restrict__ conn16,
                                                const int32_t *__restrict__ tile_off,
                                                const int32_t *__restrict__ tile_idx,
                                                const double *__restrict__ base,
                                                const double *__restrict__ lsum,
                                                const double *__restrict__ lo,
                                                double *__restrict__ ls,
                                                const double *__restrict__ out,
                                                const double *__restrict__ dq_in,
                                                double *__restrict__ geo_edge,
                                                double *__restrict__ lq, int with_dq,
                                                double *__restrict__ ldq_in) {
  extern __shared__ double lds[];
  constexpr int S = 8;  // x, y, q[4], pad
  const int b = xcd_block();
  if (TILED) {
    load_tile<S>(lds, tile_off, tile_idx, b, base, B_STRIDE, 0, 3, 0);
    __syncthreads();
  }
  const int i = b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x_i = base[(size_t)B_STRIDE * i + B_X], y_i = base[(size_t)B_STRIDE * i + B_Y];
  const int e0 = off[i], last = off[i + 1] - 1;
  // the geometry sums the recomputed q_derivatives formed
  const double sxx = lsum[4 * (size_t)i], syy = lsum[4 * (size_t)i + 1],
               sxy = lsum[4 * (size_t)i + 2];
  const double det = sxx * syy - sxy * sxy;
  const double inv = det != 0.0 ? 1.0 / det : 0.0;
  double ox[4], oy[4], sx[4], sy[4];
  load8(lo + 8 * (size_t)i, ox, oy);
  for (int r = 0; r < 4; r++) {
    sx[r] = (syy * ox[r] - sxy * oy[r]) * inv;
    sy[r] = (sxx * oy[r] - sxy * ox[r]) * inv;
  }
  store4(ls + 8 * (size_t)i, sx);
  store4(ls + 8 * (size_t)i + 4, sy);
  double aq[4], adx[4] = {0, 0, 0, 0}, ady[4] = {0, 0, 0, 0};
  load4(lq + 4 * (size_t)i, aq);
  // one pair's share to i (the same t as the gather's)
  auto own_share = [&](double delx, double dely) {
    for (int r = 0; r < 4; r++) {
      const double t = delx * sx[r] + dely * sy[r];
      aq[r] -= t;
      adx[r] += 0.5 * delx * t;
      ady[r] += 0.5 * dely * t;
    }
  };
  if (!GEO) {
    // neighbour e + 1's coordinates and e + 2's conn entry load while e is used
    auto xy_of = [&](uint32_t w) {
      return reinterpret_cast<const double2 *>(base + (size_t)B_STRIDE * (w & kIndexMask))[0];
    };
    double2 pk = {0.0, 0.0};
    uint32_t w_n = 0;
    if (e0 <= last) {
      pk = xy_of(conn[e0]);
      w_n = conn[min(e0 + 1, last)];
    }
    for (int e = e0; e <= last; e++) {
      const double2 pk_n = xy_of(w_n);
      w_n = conn[min(e + 2, last)];
      own_share(pk.x - x_i, pk.y - y_i);
      pk = pk_n;
    }
  } else {
    double qi[4], px[4], py[4], hxi[4] = {0, 0, 0, 0}, hyi[4] = {0, 0, 0, 0};
    load4(base + (size_t)B_STRIDE * i + B_Q, qi);
    load8(out + (size_t)DQ_STRIDE * i, px, py);
    if (dq_in) load8(dq_in + (size_t)DQ_STRIDE * i, hxi, hyi);
    // neighbour e + 1's base record and e + 2's conn entry load while e is
    // used (its dq_in is read when it is used)
    auto rec_of = [&](uint32_t w, double2 &p, double q[4]) {
      const double *h = TILED ? lds + S * conn_index(w, true)
                              : base + (size_t)B_STRIDE * conn_index(w, false);
      p = reinterpret_cast<const double2 *>(h)[0];
      load4(h + B_Q, q);
    };
    double2 pk = {0.0, 0.0};
    double qk[4] = {0, 0, 0, 0};
    uint32_t w_c = 0, w_n = 0;
    if (e0 <= last) {
      w_c = conn_raw(conn, conn16, e0, TILED);
      rec_of(w_c, pk, qk);
      w_n = conn_raw(conn, conn16, min(e0 + 1, last), TILED);
    }
    for (int e = e0; e <= last; e++) {
      double2 pk_n;
      double qk_n[4];
      rec_of(w_n, pk_n, qk_n);
      const uint32_t w_nn = conn_raw(conn, conn16, min(e + 2, last), TILED);
      const double delx = pk.x - x_i;
      const double dely = pk.y - y_i;
      double hxk[4] = {0, 0, 0, 0}, hyk[4] = {0, 0, 0, 0};
      if (!TILED && dq_in) load8(dq_in + (size_t)DQ_STRIDE * conn_index(w_c, false), hxk, hyk);
      double gx = 0.0, gy = 0.0;
      for (int r = 0; r < 4; r++) {
        const double hx = hxk[r] - hxi[r], hy = hyk[r] - hyi[r];
        const double d = (qk[r] - qi[r]) - 0.5 * (delx * hx + dely * hy);
        const double lr = delx * sx[r] + dely * sy[r];
        const double ro = delx * px[r] + dely * py[r];
        gx += sx[r] * (d - ro) - lr * (px[r] + 0.5 * hx);
        gy += sy[r] * (d - ro) - lr * (py[r] + 0.5 * hy);
      }
      geo_edge[2 * (size_t)e] += gx;
      geo_edge[2 * (size_t)e + 1] += gy;
      own_share(delx, dely);
      pk = pk_n;
      for (int r = 0; r < 4; r++) qk[r] = qk_n[r];
      w_c = w_n;
      w_n = w_nn;
    }
  }
  store4(lq + 4 * (size_t)i, aq);
  if (with_dq) {
    store4(ldq_in + 8 * (size_t)i, adx);
    store4(ldq_in + 8 * (size_t)i + 4, ady);
  }
}

// Per point j, from the sums' adjoints ls of every point: the adjoint of
// q (added to lq) and, for an inner pass (with_dq), of the derivatives the
// pass read (added to ldq_in).  Each pair (i, k) contributes t = dx.ls_i per
// component to q_k and -t to q_i, and for an inner pass -0.5 dx t to dq_k
// and +0.5 dx t to dq_i; k_ls_adj has added the shares of j's own pairs,
// this kernel those of the pairs in which j is the neighbour.  For
// q_derivatives (with_ext) the adjoints of qmax/qmin are routed to the
// point each came from.
__global__ __launch_bounds__(256) void k_ls_adj_gather(
    int n, const int32_t *__restrict__ tr_off, const int32_t *__restrict__ tr_owner,
    const double2 *__restrict__ xy, const double *__restrict__ ls, int with_dq,
    double *__restrict__ ldq_in, const int32_t *__restrict__ arg,
    const double *__restrict__ lqm, double *__restrict__ lq) {
  const int j = xcd_block() * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double x_j = xy[j].x, y_j = xy[j].y;
  double aq[4], adx[4] = {0, 0, 0, 0}, ady[4] = {0, 0, 0, 0};
  load4(lq + 4 * (size_t)j, aq);
  if (with_dq) load8(ldq_in + 8 * (size_t)j, adx, ady);
  // pairs (i, j) in which j is the neighbour; owner t + 1's coordinates and
  // sums' adjoints and t + 2's index load while t is used
  const int t0 = tr_off[j], tlast = tr_off[j + 1] - 1;
  auto own_of = [&](int i, double2 &p, double s8[8]) {
    p = xy[i];
    load8(ls + 8 * (size_t)i, s8, s8 + 4);
  };
  double2 po = {0.0, 0.0};
  double so[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int i_c = 0, i_n = 0;
  if (t0 <= tlast) {
    i_c = tr_owner[t0];
    own_of(i_c, po, so);
    i_n = tr_owner[min(t0 + 1, tlast)];
  }
  int prev = -1;
  for (int tt = t0; tt <= tlast; tt++) {
    double2 po_n;
    double so_n[8];
    own_of(i_n, po_n, so_n);
    const int i_nn = tr_owner[min(tt + 2, tlast)];
    const int i = i_c;
    const double delx = x_j - po.x;
    const double dely = y_j - po.y;
    double si[4], ti[4];
    for (int r = 0; r < 4; r++) {
      si[r] = so[r];
      ti[r] = so[4 + r];
    }
    po = po_n;
    for (int r = 0; r < 8; r++) so[r] = so_n[r];
    i_c = i_n;
    i_n = i_nn;
    for (int r = 0; r < 4; r++) {
      const double t = delx * si[r] + dely * ti[r];
      aq[r] += t;
      adx[r] -= 0.5 * delx * t;
      ady[r] -= 0.5 * dely * t;
    }
    if (arg && i != prev) {  // a repeated neighbour entry routes an extremum once
      for (int r = 0; r < 8; r++)
        if (arg[8 * (size_t)i + r] == j) aq[r & 3] += lqm[8 * (size_t)i + r];
    }
    prev = i;
  }
  if (arg)
    for (int r = 0; r < 8; r++)
      if (arg[8 * (size_t)j + r] == j) aq[r & 3] += lqm[8 * (size_t)j + r];
  store4(lq + 4 * (size_t)j, aq);
  if (with_dq) {
    store4(ldq_in + 8 * (size_t)j, adx);
    store4(ldq_in + 8 * (size_t)j + 4, ady);
  }
}

// ------------------------------------------------------------ q variables
// d q / d prim (q_variables_mod.h:9-34) at one state, prim order (rho, u1,
// u2, p): J[r][c] = d q_r / d prim_c.
__device__ __forceinline__ void qvar_jac(const double p[4], double J[4][4]) {
  const double rho = p[0], u1 = p[1], u2 = p[2], pr = p[3];
  const double beta = 0.5 * rho / pr, uu = u1 * u1 + u2 * u2;
  J[0][0] = (3.5 - beta * uu) / rho;
  J[0][1] = -2.0 * beta * u1;
  J[0][2] = -2.0 * beta * u2;
  J[0][3] = (beta * uu - 2.5) / pr;
  J[1][0] = 2.0 * beta * u1 / rho;
  J[1][1] = 2.0 * beta;
  J[1][2] = 0.0;
  J[1][3] = -2.0 * beta * u1 / pr;
  J[2][0] = 2.0 * beta * u2 / rho;
  J[2][1] = 0.0;
  J[2][2] = 2.0 * beta;
  J[2][3] = -2.0 * beta * u2 / pr;
  J[3][0] = -2.0 * beta / rho;
  J[3][1] = 0.0;
  J[3][2] = 0.0;
  J[3][3] = 2.0 * beta / pr;
}

// grad = J^T lq (accumulate: grad += J^T lq)
__global__ __launch_bounds__(256) void k_qvar_vjp(int n, const double *__restrict__ prim,
                                                  const double *__restrict__ lq,
                                                  double *__restrict__ grad, int accumulate) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p[4], J[4][4], a[4];
  for (int c = 0; c < 4; c++) p[c] = prim[(size_t)c * n + i];
  qvar_jac(p, J);
  load4(lq + 4 * (size_t)i, a);
  for (int c = 0; c < 4; c++) {
    const double g = J[0][c] * a[0] + J[1][c] * a[1] + J[2][c] * a[2] + J[3][c] * a[3];
    grad[(size_t)c * n + i] = accumulate ? grad[(size_t)c * n + i] + g : g;
  }
}

// ------------------------------------------------------------ tangent-linear
// q tangent into the q slots of a copy of the base records (so that the
// forward least-squares kernels apply the linear stencils to it)
__global__ __launch_bounds__(256) void k_qvar_jvp(int n, const double *__restrict__ prim,
                                                  const double *__restrict__ v,
                                                  double *__restrict__ tbase) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p[4], J[4][4], t[4];
  for (int c = 0; c < 4; c++) p[c] = prim[(size_t)c * n + i];
  qvar_jac(p, J);
  for (int r = 0; r < 4; r++) {
    t[r] = 0.0;
    for (int c = 0; c < 4; c++) t[r] += J[r][c] * v[(size_t)c * n + i];
  }
  store4(tbase + (size_t)B_STRIDE * i + B_Q, t);
}

// qmax/qmin tangents: the tangent of the point each extremum came from
__global__ __launch_bounds__(256) void k_ext_jvp(int n, const int32_t *__restrict__ arg,
                                                 const double *__restrict__ tbase,
                                                 double *__restrict__ tqm) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int r = 0; r < 8; r++)
    tqm[8 * (size_t)i + r] = tbase[(size_t)B_STRIDE * arg[8 * (size_t)i + r] + B_Q + (r & 3)];
}

// flux tangent: per pair, the tangent of q~ at each end through the
// reconstruction, then one dual direction through the summed fluxes
__global__ __launch_bounds__(256) void k_flux_jvp(
    int n, const int32_t *__restrict__ off, const uint32_t *__restrict__ conn,
    const double *__restrict__ base, const double *__restrict__ dq,
    const double *__restrict__ qm, const double *__restrict__ coef,
    const double *__restrict__ nxa, const double *__restrict__ nya,
    const uint8_t *__restrict__ flag, const double *__restrict__ delta,
    const double *__restrict__ tbase, const double *__restrict__ tdq,
    const double *__restrict__ tqm, double *__restrict__ out) {
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int type = flag[i];
  const double *hi = base + (size_t)B_STRIDE * i;
  const double x_i = hi[B_X], y_i = hi[B_Y], eps2_i = hi[B_EPS2];
  double qi[4], dxi[4], dyi[4], qmxi[4], qmni[4], cf[8];
  double tqi[4], tdxi[4], tdyi[4], tmxi[4], tmni[4];
  load4(hi + B_Q, qi);
  load8(dq + (size_t)DQ_STRIDE * i, dxi, dyi);
  load8(qm + (size_t)QM_STRIDE * i, qmxi, qmni);
  load8(coef + (size_t)COEF_STRIDE * i, cf, cf + 4);
  load4(tbase + (size_t)B_STRIDE * i + B_Q, tqi);
  load8(tdq + (size_t)DQ_STRIDE * i, tdxi, tdyi);
  load8(tqm + 8 * (size_t)i, tmxi, tmni);
  const double nx = nxa[i], ny = nya[i], tx = ny, ty = -nx;
  const double s2b = type == 0 ? -1.0 : 1.0;
  double acc[4] = {0, 0, 0, 0};
  for (int e = off[i]; e < off[i + 1]; e++) {
    const uint32_t w = conn[e];
    const int k = (int)(w & kIndexMask);
    const uint32_t m = w >> kMaskShift;
    const double *hk = base + (size_t)B_STRIDE * k;
    double qk[4], dxk[4], dyk[4], qmxk[4], qmnk[4], tqk[4], tdxk[4], tdyk[4], tmxk[4], tmnk[4];
    load4(hk + B_Q, qk);
    load8(dq + (size_t)DQ_STRIDE * k, dxk, dyk);
    load8(qm + (size_t)QM_STRIDE * k, qmxk, qmnk);
    load4(tbase + (size_t)B_STRIDE * k + B_Q, tqk);
    load8(tdq + (size_t)DQ_STRIDE * k, tdxk, tdyk);
    load8(tqm + 8 * (size_t)k, tmxk, tmnk);
    const double delx = hk[B_X] - x_i, dely = hk[B_Y] - y_i;
    const double dels = delx * tx + dely * ty, deln = delx * nx + dely * ny;
    Dual<1> qti[4], qtk[4];
    for (int r = 0; r < 4; r++) {
      const Recon a = recon(qi[r], delx * dxi[r] + dely * dyi[r], qmxi[r], qmni[r], eps2_i);
      const Recon b = recon(qk[r], delx * dxk[r] + dely * dyk[r], qmxk[r], qmnk[r], hk[B_EPS2]);
      qti[r] = Dual<1>(a.qt);
      qti[r].d[0] = a.dq * tqi[r] + a.dd * (delx * tdxi[r] + dely * tdyi[r]) +
                    a.dm * (a.use_max ? tmxi[r] : tmni[r]);
      qtk[r] = Dual<1>(b.qt);
      qtk[r].d[0] = b.dq * tqk[r] + b.dd * (delx * tdxk[r] + dely * tdyk[r]) +
                    b.dm * (b.use_max ? tmxk[r] : tmnk[r]);
    }
    Dual<1> Si[4], Sk[4];
    end_sum_dual(qti, nx, ny, type, m, s2b, cf, dels, deln, Si);
    end_sum_dual(qtk, nx, ny, type, m, s2b, cf, dels, deln, Sk);
    for (int r = 0; r < 4; r++) acc[r] += Sk[r].d[0] - Si[r].d[0];
  }
  const double sc = (type == 0 ? 2.0 : 1.0) * delta[i];
  for (int r = 0; r < 4; r++) out[(size_t)r * n + i] = sc * acc[r];
}

int blocks_of(int n) { return (n + 255) / 256; }

}  // namespace

size_t lds_bytes_flux_vjp(int tile_max) { return sizeof(double) * 24 * (size_t)tile_max; }

bool prepare_flux_vjp_tiles(const DevCloud &c, size_t lds_limit) {
  if (!c.tiles || c.tile_max <= 0) return false;
  const size_t L = lds_bytes_flux_vjp(c.tile_max);
  if (L > lds_limit) return false;
  const void *fs[2] = {(const void *)k_flux_vjp<true, true>,
                       (const void *)k_flux_vjp<false, true>};
  for (const void *f : fs)
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L) != hipSuccess)
      return false;
  return true;
}

void launch_flux_vjp(const DevCloud &c, const AdjointBuffers &a, const double *base,
                     const double *dq, const double *qm, const double *delta, const double *lam,
                     const GeoAdjoint *geo, hipStream_t s) {
#define VJP_ARGS(G)                                                                       \
  c.n, c.conn_off, c.conn, c.conn16, c.tile_off, c.tile_idx, base, dq, qm, c.coef, c.nx,  \
      c.ny, c.flag, delta, lam, a.own_q, a.own_dq, a.own_qm, a.edge, a.edge_mx, a.tr_pos, \
      G ? geo->edge : nullptr, G ? geo->nrm : nullptr, G ? geo->coef : nullptr,           \
      G ? geo->delta : nullptr
  if (a.vjp_tiled) {
    const dim3 g(blocks_for(c.n, c.tile_block)), b(c.tile_block);
    const size_t L = lds_bytes_flux_vjp(c.tile_max);
    if (geo)
      hipLaunchKernelGGL((k_flux_vjp<true, true>), g, b, L, s, VJP_ARGS(true));
    else
      hipLaunchKernelGGL((k_flux_vjp<false, true>), g, b, L, s, VJP_ARGS(false));
  } else if (geo) {
    hipLaunchKernelGGL((k_flux_vjp<true, false>), dim3(blocks_of(c.n)), dim3(256), 0, s,
                       VJP_ARGS(true));
  } else {
    hipLaunchKernelGGL((k_flux_vjp<false, false>), dim3(blocks_of(c.n)), dim3(256), 0, s,
                       VJP_ARGS(false));
  }
#undef VJP_ARGS
  hipLaunchKernelGGL(k_flux_vjp_gather, dim3(blocks_of(c.n)), dim3(256), 0, s, c.n, a.tr_off,
                     a.tr_sbase, a.tr_owner_s, reinterpret_cast<const double2 *>(c.xy), a.edge,
                     a.edge_mx, a.own_q,
                     a.own_dq, a.own_qm, a.lq, a.ldq[0], a.lqm);
}

size_t lds_bytes_ls_adj(int tile_max) { return (size_t)tile_max * 8 * sizeof(double); }

bool prepare_ls_adj_tiles(const DevCloud &c, size_t lds_limit) {
  if (!c.tiles || c.tile_max <= 0) return false;
  const size_t L = lds_bytes_ls_adj(c.tile_max);
  return L <= lds_limit &&
         hipFuncSetAttribute((const void *)k_ls_adj<true, true>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)L) == hipSuccess;
}

void launch_ls_vjp(const DevCloud &c, const AdjointBuffers &a, const double *base,
                   const double *lo, double *ldq_in, bool extrema, const double *out,
                   const double *dq_in, double *geo_edge, hipStream_t s) {
  if (geo_edge && a.ls_tiled && !dq_in)
    hipLaunchKernelGGL((k_ls_adj<true, true>), dim3(blocks_for(c.n, c.tile_block)),
                       dim3(c.tile_block), lds_bytes_ls_adj(c.tile_max), s, c.n, c.conn_off,
                       c.conn, c.conn16, c.tile_off, c.tile_idx, base, a.lsum, lo, a.ls, out,
                       nullptr, geo_edge, a.lq, extrema ? 0 : 1, ldq_in);
  else if (geo_edge)
    hipLaunchKernelGGL((k_ls_adj<true, false>), dim3(blocks_of(c.n)), dim3(256), 0, s, c.n,
                       c.conn_off, c.conn, nullptr, nullptr, nullptr, base, a.lsum, lo, a.ls,
                       out, dq_in, geo_edge, a.lq, extrema ? 0 : 1, ldq_in);
  else
    hipLaunchKernelGGL((k_ls_adj<false, false>), dim3(blocks_of(c.n)), dim3(256), 0, s, c.n,
                       c.conn_off, c.conn, nullptr, nullptr, nullptr, base, a.lsum, lo, a.ls,
                       nullptr, nullptr, nullptr, a.lq, extrema ? 0 : 1, ldq_in);
  hipLaunchKernelGGL(k_ls_adj_gather, dim3(blocks_of(c.n)), dim3(256), 0, s, c.n, a.tr_off,
                     a.tr_owner, reinterpret_cast<const double2 *>(c.xy), a.ls, extrema ? 0 : 1,
                     ldq_in, extrema ? a.arg : nullptr, a.lqm, a.lq);
}

void launch_qvar_vjp(const DevCloud &c, const AdjointBuffers &a, const double *prim,
                     double *grad, bool accumulate, hipStream_t s) {
  hipLaunchKernelGGL(k_qvar_vjp, dim3(blocks_of(c.n)), dim3(256), 0, s, c.n, prim, a.lq, grad,
                     accumulate ? 1 : 0);
}

void launch_qvar_jvp(const DevCloud &c, const double *prim, const double *v, double *tbase,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_qvar_jvp, dim3(blocks_of(c.n)), dim3(256), 0, s, c.n, prim, v, tbase);
}

void launch_ext_jvp(const DevCloud &c, const int32_t *arg, const double *tbase, double *tqm,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_ext_jvp, dim3(blocks_of(c.n)), dim3(256), 0, s, c.n, arg, tbase, tqm);
}

void launch_flux_jvp(const DevCloud &c, const double *base, const double *dq, const double *qm,
                     const double *delta, const double *tbase, const double *tdq,
                     const double *tqm, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_flux_jvp, dim3(blocks_of(c.n)), dim3(256), 0, s, c.n, c.conn_off, c.conn,
                     base, dq, qm, c.coef, c.nx, c.ny, c.flag, delta, tbase, tdq, tqm, out);
}

}  // namespace lskum
