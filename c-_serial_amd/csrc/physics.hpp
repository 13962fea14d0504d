// physics.hpp -- q-LSKUM leaf physics as __device__ functions for gfx950.
//
// The FP64 form keeps the reference's formulas and operation order except
// where noted; the build contracts a*b+c into FMAs, so results differ from
// the reference's at the rounding level (the ROCm device exp/log/erf also
// round differently from glibc by about an ulp).
//
// The split/quadrant fluxes are restructured around one set of per-end
// terms (EndTerms): every flux the reference evaluates for a (point,
// neighbour) pair -- two sub-stencils per neighbour, so two to four fluxes
// per end -- shares beta, sqrt(beta), 0.5/sqrt(pi*beta), exp(-S^2) and
// erf(S), which are computed once.  Quotients in the limiter and in
// q~ -> primitive use qdiv (a refined hardware reciprocal), within an ulp or
// two of the IEEE quotient.  A flux's +/- variant is selected by a sign s =
// +-1.0: 0.5*(1 + s*erf) and x + s*y are bit-identical to the reference's
// 0.5*(1 - erf) and x - y because multiplying by -1.0 is exact.
#pragma once

#include <hip/hip_runtime.h>

namespace lskum {

constexpr double kPi = 3.14159265358979323846;  // == 4*atan(1.0) in double
constexpr double kGammaM1 = 1.4 - 1.0;          // gamma_new - 1 as evaluated
constexpr double kInvGammaM1 = 1.0 / kGammaM1;
constexpr double kInvSqrtPi = 0.56418958354775628695;  // 1/sqrt(pi)

// 1/b for normal b from the hardware estimate r0 (relative error e0 well
// under 2^-22): with e = 1 - b r0, 1/b = r0 (1 + e + e^2 + e^3 + ...), and one
// third-order step r0 + r0 (e + e^2) leaves e0^3 (2^-66 and below) plus its
// own rounding -- the accuracy of two Newton steps in three FMAs instead of
// four.
__device__ __forceinline__ double rcp_refined(double b) {
  const double r = __builtin_amdgcn_rcp(b);
  const double e = fma(-b, r, 1.0);
  return fma(r, fma(e, e, e), r);
}

// a / b from the refined reciprocal and a final residual correction (the
// refinement the correctly rounded division also performs, without its
// scaling for operands near the ends of the exponent range); within an ulp
// or two of a / b for the normal-range operands it gets here.  Other types
// (the dual numbers of the derivative kernels) divide plainly.
template <typename T>
__device__ __forceinline__ T qdiv(T a, T b) {
  return a / b;
}
template <>
__device__ __forceinline__ double qdiv(double a, double b) {
  const double r = rcp_refined(b);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}
template <>
__device__ __forceinline__ float qdiv(float a, float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  const float q = a * r;
  return fmaf(fmaf(-b, q, a), r, q);
}

// e^a as the fluxes use it (other types: the library exp).  Floats take the
// hardware base-2 exponential of a log2(e): the rounding of the product
// costs |a| 6e-8 in relative terms, as much as a has already lost in FP32,
// in place of the library's range-reduced expf.
template <typename T>
__device__ __forceinline__ T exp_k(T a) {
  return exp(a);
}
template <>
__device__ __forceinline__ float exp_k(float a) {
  return __builtin_amdgcn_exp2f(a * 1.44269504088896341f);
}
// Doubles: a = n ln2 + r with n = rint(a / ln2) and |r| <= ln2/2 (r by two
// FMAs against ln 2 split into a double and its remainder), e^r from
// p(r) = 1 + r + c2 r^2 + ... + c10 r^10, a fit in relative error (3.3e-16;
// exp_fit.py beside this file derives it), then scaled by 2^n: within two
// or three ulps.  n comes from the shifter 1.5 * 2^52: nb = a / ln2 +
// 1.5 * 2^52, for |a / ln2| < 2^51, lies in [2^52, 2^53) and is therefore
// rounded to an integer, n = nb - 1.5 * 2^52 exactly, and the low 32 bits
// of nb are n itself (2^51 = 0 mod 2^32) -- n and its integer without
// rint or a conversion.  n may differ from rint(a * (1/ln2)) by one at a
// tie of the rounded product, which leaves |r| at ln2/2 to within
// rounding.  The callers first raise a to -708, which keeps 2^n normal in
// place of the library's underflow tests (e^-708 = 3e-308 stands for 0).
// e^r on |r| <= ln2/2 (the fit above)
__device__ __forceinline__ double exp_poly(double r) {
  double p = 2.743938836864164e-07;
  p = fma(p, r, 2.7635062228304872e-06);
  p = fma(p, r, 2.4802048439065005e-05);
  p = fma(p, r, 0.00019841185049797322);
  p = fma(p, r, 0.0013888888387934915);
  p = fma(p, r, 0.008333333371024567);
  p = fma(p, r, 0.04166666666871237);
  p = fma(p, r, 0.16666666666611155);
  p = fma(p, r, 0.49999999999997463);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return p;
}

constexpr double kExpShift = 6755399441055744.0;  // 1.5 * 2^52

// e^a from nb = 1.5 * 2^52 + n
__device__ __forceinline__ double exp_shifted(double a, double nb) {
  const double n = nb - kExpShift;
  double r = fma(-n, 0.6931471805599453, a);
  r = fma(-n, 2.3190468138462996e-17, r);
  return ldexp(exp_poly(r), (int)(uint32_t)__double_as_longlong(nb));
}

// Any a >= -708.  nb is capped at n = 1100 (the fmin, which also takes a
// NaN nb there): for a past 762 the capped n leaves r large, and every
// coefficient being positive, p(r) and then 2^n p overflow to inf, as the
// library's exp does past 709.78 -- for huge finite a and for +inf
// (r = inf - 1100 ln2 = inf) too.  A NaN a gives a NaN r and result.
__device__ __forceinline__ double exp_reduced(double a) {
  return exp_shifted(a, fmin(fma(a, 1.4426950408889634, kExpShift), kExpShift + 1100.0));
}

// The same for a in [-708, 0] (or NaN), where nb needs no cap.
__device__ __forceinline__ double exp_reduced_nonpos(double a) {
  return exp_shifted(a, fma(a, 1.4426950408889634, kExpShift));
}

// The clamp compares, so a NaN a stays NaN.
template <>
__device__ __forceinline__ double exp_k(double a) {
  return exp_reduced(a < -708.0 ? -708.0 : a);
}

// e^a for the kinetic terms' a = -S^2 <= 0.  The clamp is one fmax, which
// maps a NaN a to -708; a NaN S still reaches every flux through the
// velocity component or the sqrt(beta) it is the product of.
template <typename T>
__device__ __forceinline__ T exp_neg_sq(T a) {
  return exp_k(a);
}
template <>
__device__ __forceinline__ double exp_neg_sq(double a) {
  return exp_reduced_nonpos(fmax(a, -708.0));
}

// a / b without qdiv's last correction: the refined reciprocal times a,
// within two ulps.  For quotients whose last bits do not matter downstream
// (the limiter ratio, erf's rational).
template <typename T>
__device__ __forceinline__ T qdiv_loose(T a, T b) {
  return qdiv(a, b);
}
template <>
__device__ __forceinline__ double qdiv_loose(double a, double b) {
  return a * rcp_refined(b);
}
template <>
__device__ __forceinline__ float qdiv_loose(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}

// 1/sqrt(b) for normal b > 0 from the hardware estimate y0: with
// e = 1 - b y0^2, 1/sqrt(b) = y0 (1 - e)^(-1/2) = y0 (1 + e/2 + 3e^2/8 + ...),
// and the step y0 + y0 e (1/2 + 3e/8) leaves 5e^3/16 (the estimate is good
// to 2^-24 here, so about 2^-70) plus its own rounding: within an ulp or
// two, the accuracy of two Newton steps in five instructions instead of
// seven.
__device__ __forceinline__ double rsqrt_nr(double b) {
  const double y = __builtin_amdgcn_rsq(b);
  const double e = fma(-b * y, y, 1.0);
  return fma(y, e * fma(0.375, e, 0.5), y);
}
// In FP32 the estimate is within an ulp or so; one step rounds it off.
__device__ __forceinline__ float rsqrt_nr(float b) {
  float y = __builtin_amdgcn_rsqf(b);
  return fmaf(y, fmaf(-0.5f * b * y, y, 0.5f), y);
}

// qtilde_to_primitive (q_variables_mod.h:230-253) that also returns
// sqrt(beta) and temp = 0.5/beta = p/rho, which the fluxes need again (split_fluxes_mod.h:20-25 recompute
// beta = 0.5*rho/p and p/rho; here they are reused, equal up to rounding).
// rho = exp(temp1 - log(beta)/(gamma-1)) is evaluated as exp(temp1) /
// (beta^2 sqrt(beta)): 1/(gamma-1) = 2.5 for gamma = 1.4, so the logarithm
// becomes a quotient of the square root the end terms need anyway.
template <typename T>
__device__ __forceinline__ void qtilde_state(const T qt[4], T &u1, T &u2, T &rho, T &sb,
                                             T &temp) {
  const T beta = -qt[3] * T(0.5);
  temp = qdiv(T(0.5), beta);
  u1 = qt[1] * temp;
  u2 = qt[2] * temp;
  const T temp1 = qt[0] + beta * (u1 * u1 + u2 * u2);
  sb = sqrt(beta);
  rho = exp(temp1) * qdiv(T(1.0), beta * beta * sb);
}
// For doubles and floats, one y = 1/sqrt(beta) gives all three: 0.5/beta = y^2/2,
// sqrt(beta) = beta y and beta^-2.5 = y^5 (a few ulps in all), in place of
// a square root and two divisions.
template <typename T>
__device__ __forceinline__ void qtilde_state_rsqrt(const T qt[4], T &u1, T &u2, T &rho, T &sb,
                                                   T &temp) {
  const T beta = -qt[3] * T(0.5);
  const T y = rsqrt_nr(beta);
  const T y2 = y * y;
  temp = T(0.5) * y2;
  u1 = qt[1] * temp;
  u2 = qt[2] * temp;
  const T temp1 = qt[0] + beta * (u1 * u1 + u2 * u2);
  sb = beta * y;
  rho = exp_k(temp1) * (y2 * y2 * y);
}
template <>
__device__ __forceinline__ void qtilde_state(const double qt[4], double &u1, double &u2,
                                             double &rho, double &sb, double &temp) {
  qtilde_state_rsqrt(qt, u1, u2, rho, sb, temp);
}
template <>
__device__ __forceinline__ void qtilde_state(const float qt[4], float &u1, float &u2,
                                             float &rho, float &sb, float &temp) {
  qtilde_state_rsqrt(qt, u1, u2, rho, sb, temp);
}

// erf(s) given x = exp(-s^2), which the kinetic terms compute anyway.  For
// doubles, erf(s) = sign(s) (1 - x R(|s|)) with R a rational (8, 8)
// approximation of erfcx(t) = exp(t^2) erfc(t) on [0, 6], fitted in relative
// error (max 2.9e-16; erfcx_fit.py beside this file derives the
// coefficients).  Evaluated in double the result is within 6e-16 of erf in
// absolute terms, the precision the fluxes keep anyway (they use erf only
// through A = (1 +- erf)/2), and it has no branches.  Past |s| = 6, x R(6) <
// 3e-17 and the result rounds to +-1.  Other types (FP32, the dual numbers
// of the derivative kernels) use the library erf.
template <typename T>
__device__ __forceinline__ T erf_x(T s, T /*x*/) {
  return erf(s);
}
template <>
__device__ __forceinline__ double erf_x(double s, double x) {
  const double t = fmin(fabs(s), 6.0);
  double p = 1.0833939560829976e-11;
  p = fma(p, t, 0.00045073903614461776);
  p = fma(p, t, 0.007157824410058684);
  p = fma(p, t, 0.053534056393831096);
  p = fma(p, t, 0.24246846990245913);
  p = fma(p, t, 0.7171759137974627);
  p = fma(p, t, 1.3930631679029237);
  p = fma(p, t, 1.6689634151475463);
  p = fma(p, t, 0.9999999999999997);
  double q = 0.0007989152115252066;
  q = fma(q, t, 0.012686885026544848);
  q = fma(q, t, 0.09528658017053802);
  q = fma(q, t, 0.4361018725716152);
  q = fma(q, t, 1.3182577689390016);
  q = fma(q, t, 2.6772975954680143);
  q = fma(q, t, 3.5495262609381824);
  q = fma(q, t, 2.797342582242985);
  q = fma(q, t, 1.0);
  // e >= 0 (x R(t) <= R(0) < 1), so the sign of s is copied onto it in one
  // instruction; erf(-0) comes out as -e(0), i.e. -3e-16, as erf(+0) is +3e-16
  const double e = fma(-x, qdiv_loose(p, q), 1.0);
  return copysign(e, s);
}

// Terms shared by all split and quadrant fluxes of one state in the frame
// t = (ny, -nx), n = (nx, ny): split_fluxes_mod.h:14-26 and
// quadrant_fluxes_mod.h:17-32.
template <typename T>
struct EndTerms {
  T rho, ut, un, pr_by_rho, u_sqr, B1, B2, e1, e2;
};

// From (sqrt(beta), p/rho), beta = rho/2p.  B1 and B2 share h =
// 0.5/sqrt(pi beta), formed without a square root or a division: p/rho =
// 0.5/beta, so h = (1/sqrt(pi)) sqrt(beta) (p/rho).
template <typename T>
__device__ __forceinline__ EndTerms<T> end_terms_sb(T nx, T ny, T u1, T u2, T rho, T sb,
                                                    T pr_by_rho) {
  EndTerms<T> e;
  const T tx = ny, ty = -nx;
  e.rho = rho;
  e.ut = u1 * tx + u2 * ty;
  e.un = u1 * nx + u2 * ny;
  const T S1 = e.ut * sb;
  const T S2 = e.un * sb;
  const T h = T(kInvSqrtPi) * sb * pr_by_rho;
  const T X1 = exp_neg_sq(-S1 * S1), X2 = exp_neg_sq(-S2 * S2);
  e.B1 = X1 * h;
  e.B2 = X2 * h;
  e.e1 = erf_x(S1, X1);
  e.e2 = erf_x(S2, X2);
  e.pr_by_rho = pr_by_rho;
  e.u_sqr = e.ut * e.ut + e.un * e.un;
  return e;
}

// q~ -> primitive and the end terms in one, for the FP64 and FP32 fluxes:
// the numbers of qtilde_state followed by end_terms_sb, up to rounding, in
// fewer operations -- |u|^2 is the u1^2 + u2^2 that q~ -> primitive forms
// (the frame's ut^2 + un^2 in end_terms_sb; a rotation leaves it
// unchanged), and h = sqrt(beta) (p/rho) / sqrt(pi) = y / (2 sqrt(pi)) with
// y = beta^-1/2, since beta y^2 = 1.
template <typename T>
__device__ __forceinline__ EndTerms<T> qtilde_end_terms(const T qt[4], T nx, T ny) {
  const T beta = -qt[3] * T(0.5);
  const T y = rsqrt_nr(beta);
  const T y2 = y * y;
  const T temp = T(0.5) * y2;
  const T u1 = qt[1] * temp, u2 = qt[2] * temp;
  const T usq = u1 * u1 + u2 * u2;
  const T sb = beta * y;
  EndTerms<T> e;
  e.rho = exp_k(qt[0] + beta * usq) * (y2 * y2 * y);
  const T tx = ny, ty = -nx;
  e.ut = u1 * tx + u2 * ty;
  e.un = u1 * nx + u2 * ny;
  const T S1 = e.ut * sb, S2 = e.un * sb;
  const T h = T(0.5 * kInvSqrtPi) * y;
  const T X1 = exp_neg_sq(-S1 * S1), X2 = exp_neg_sq(-S2 * S2);
  e.B1 = X1 * h;
  e.B2 = X2 * h;
  e.e1 = erf_x(S1, X1);
  e.e2 = erf_x(S2, X2);
  e.pr_by_rho = temp;
  e.u_sqr = usq;
  return e;
}

// The same from a primitive state (u1, u2, rho, pr), as the flux functions
// of split_fluxes_mod.h / quadrant_fluxes_mod.h take it.
template <typename T>
__device__ __forceinline__ EndTerms<T> end_terms(T nx, T ny, T u1, T u2, T rho, T pr) {
  return end_terms_sb(nx, ny, u1, u2, rho, sqrt(T(0.5) * rho / pr), pr / rho);
}

// split_fluxes_mod.h:6-43 (s = +1, Gxp) and :45-84 (s = -1, Gxn).
// Regrouped, equal up to rounding: G[2] = rho (ut un A1 + s un B1) is
// un G[0], formed as such.  The rest is the same numbers by fewer
// operations, since products by s = +-1 and by powers of two are exact:
//  - the sign s is applied once, to B1 (s (x B1) and x (s B1) are equal);
//  - A1 = (1 + s erf)/2 is fma(s/2, erf, 1/2);
//  - 0.5 ut (7 p/rho + |u|^2) A1 and 0.5 (6 p/rho + |u|^2) s B1 take the
//    halves inside: (3.5 p/rho + |u|^2/2) and (3 p/rho + |u|^2/2), which
//    the x and y fluxes of one end share.
// split_y likewise with the roles of ut and un exchanged.
template <typename T>
__device__ __forceinline__ void split_x(const EndTerms<T> &e, T s, T G[4]) {
  const T sB1 = s * e.B1;
  const T A1 = (T(0.5) * s) * e.e1 + T(0.5);
  G[0] = e.rho * (e.ut * A1 + sB1);
  const T temp1 = e.pr_by_rho + e.ut * e.ut;
  const T temp2 = temp1 * A1 + e.ut * sB1;
  G[1] = e.rho * temp2;
  G[2] = e.un * G[0];
  const T hu = T(0.5) * e.u_sqr;
  const T t7h = T(3.5) * e.pr_by_rho + hu, t6h = T(3.0) * e.pr_by_rho + hu;
  G[3] = e.rho * (e.ut * t7h * A1 + t6h * sB1);
}

// split_fluxes_mod.h:86-125 (s = +1, Gyp) and :127-164 (s = -1, Gyn)
template <typename T>
__device__ __forceinline__ void split_y(const EndTerms<T> &e, T s, T G[4]) {
  const T sB2 = s * e.B2;
  const T A2 = (T(0.5) * s) * e.e2 + T(0.5);
  G[0] = e.rho * (e.un * A2 + sB2);
  const T temp1 = e.pr_by_rho + e.un * e.un;
  const T temp2 = temp1 * A2 + e.un * sB2;
  G[2] = e.rho * temp2;
  G[1] = e.ut * G[0];
  const T hu = T(0.5) * e.u_sqr;
  const T t7h = T(3.5) * e.pr_by_rho + hu, t6h = T(3.0) * e.pr_by_rho + hu;
  G[3] = e.rho * (e.un * t7h * A2 + t6h * sB2);
}

// quadrant_fluxes_mod.h: GxI :6 (s1 -, s2 -), GxII :57 (+, -),
// GxIII :108 (+, +), GxIV :161 (-, +); the signs and halves placed as in
// split_x (the same numbers as the reference's grouping).
template <typename T>
__device__ __forceinline__ void quad_x(const EndTerms<T> &e, T s1, T s2, T G[4]) {
  const T sB1 = s1 * e.B1, sB2 = s2 * e.B2;
  const T A1 = (T(0.5) * s1) * e.e1 + T(0.5);
  const T A2 = (T(0.5) * s2) * e.e2 + T(0.5);
  const T a1b1 = e.ut * A1 + sB1;
  G[0] = e.rho * A2 * a1b1;
  const T temp1 = e.pr_by_rho + e.ut * e.ut;
  T temp2 = temp1 * A1 + e.ut * sB1;
  G[1] = e.rho * A2 * temp2;
  temp2 = e.un * A2 + sB2;
  G[2] = e.rho * a1b1 * temp2;
  const T hu = T(0.5) * e.u_sqr;
  const T t7h = T(3.5) * e.pr_by_rho + hu, t6h = T(3.0) * e.pr_by_rho + hu;
  temp2 = e.ut * t7h * A1;
  const T temp3 = sB1 * t6h;
  const T temp4 = T(0.5) * e.rho * e.un * sB2 * a1b1;
  G[3] = e.rho * A2 * (temp2 + temp3) + temp4;
}

// The Venkatakrishnan ratio of limiters_mod.h:85-99.  There num and den
// both carry a factor del_neg, which the callers have already shown to be
// non-zero (|del_neg| > 1e-5); cancelled, with a = del_pos^2 + eps2,
//   phi = min(1, (a + 2 dn dp) / (a + dn (2 dn + dp))).
// dp has the sign of dn (qmax - q >= 0 for dn > 0, qmin - q <= 0 for
// dn < 0), so every term is non-negative and nothing cancels.
template <typename T>
__device__ __forceinline__ T venkat_ratio(T dn, T dp, T eps2) {
  const T a = dp * dp + eps2;
  const T num = a + T(2.0) * dn * dp;
  const T den = a + dn * (T(2.0) * dn + dp);
  const T temp = qdiv_loose(num, den);
  return temp < T(1) ? temp : T(1);
}

// limiters_mod.h:55-104 for one component; eps2 = (pow(K*min_dist,3))^2 is
// precomputed per point on the host with the reference's pow call.
template <typename T>
__device__ __forceinline__ T venkat_phi(T qtilde, T q, T qmax, T qmin, T eps2) {
  const T del_neg = qtilde - q;
  if (fabs(del_neg) <= T(10e-6)) return T(1.0);
  return venkat_ratio(del_neg, (del_neg > T(0) ? qmax : qmin) - q, eps2);
}

// The same limiter from the increments themselves: del_neg = q~ - q =
// -0.5*(dq . dx) and dmax = qmax - q, dmin = qmin - q.  In FP32, forming q~
// first and subtracting q again would lose the increment's low bits to the
// size of q; this form keeps them.
template <typename T>
__device__ __forceinline__ T venkat_phi_delta(T del_neg, T dmax, T dmin, T eps2) {
  if (fabs(del_neg) <= T(10e-6)) return T(1.0);
  return venkat_ratio(del_neg, del_neg > T(0) ? dmax : dmin, eps2);
}

}  // namespace lskum
