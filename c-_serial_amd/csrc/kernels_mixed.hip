// kernels_mixed.hip -- q-LSKUM stage kernels, mixed-precision form (gfx950).
//
// Neighbour data are gathered in FP32: q (float4), dq/dx and dq/dy (two
// float4), qmax/qmin (two float4) and the limiter's eps^2 (float).  The
// coordinates stay FP64 so that a neighbour offset dx = x_k - x_i is exact
// before it is rounded to FP32: at a cosine-clustered leading edge |dx| is
// ~1e-6 chord, and an FP32 difference of absolute coordinates would keep two
// digits of it.  The limiter, q~ -> primitive and the split/quadrant fluxes
// run in FP32, and so do the flux's sub-stencil weights, its per-point sums
// of the FP32 flux differences, the least-squares right-hand sides of
// q_derivatives and the inner passes (sums of w dx dq over FP32 neighbour
// data) and each neighbour's local-step term.  FP64 is kept for
//  - the least-squares geometry: every kernel applies per-point
//    coefficients of its 2x2 solves, formed in FP64 once on the host from
//    the geometry sums and rounded once (the full stencil's for
//    q_derivatives and the inner passes, the sub-stencils' for the flux),
//  - the state, the SSPRK update and the residual and force reductions
//    (kernels.hip, shared with the FP64 form).
// Differences are formed directly where FP32 would otherwise cancel: the
// limiter takes del_neg = -0.5*(dq . dx) rather than q~ - q, and an inner
// pass sums (q_k - q_i) - 0.5*(dx . (dq_k - dq_i)) rather than the
// difference of two reconstructed values.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"
#include "physics.hpp"

namespace lskum {

namespace {

// One tile's records gathered from up to three arrays in one round of loads
// per thread: entry r takes C0 16-byte chunks of s0 (s0[idx[r] C0 + c])
// into d0[r C0 + c], C1 of s1 into d1 and C2 of s2 into d2.  Every load is
// issued before any LDS store, so a block waits for memory once rather than
// once per array.
template <int C0, int C1, int C2>
__device__ __forceinline__ void tile_gather(uint4 *d0, const uint4 *__restrict__ s0, uint4 *d1,
                                            const uint4 *__restrict__ s1, uint4 *d2,
                                            const uint4 *__restrict__ s2,
                                            const int32_t *__restrict__ idx, int count) {
  constexpr int CH = C0 + C1 + C2, U = 8;
  const int total = count * CH;
  for (int base = threadIdx.x; base < total; base += U * blockDim.x) {
    uint4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      int t = base + u * blockDim.x;
      t = t < total ? t : total - 1;
      const int r = t / CH, c = t - r * CH;
      const size_t g = (size_t)idx[r];
      const uint4 *p = s0 + C0 * g + c;
      if constexpr (C2 > 0) {
        if (c >= C0 + C1) p = s2 + C2 * g + (c - C0 - C1);
      }
      if (c >= C0 && c < C0 + C1) p = s1 + C1 * g + (c - C0);
      v[u] = *p;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int t = base + u * blockDim.x;
      if (t < total) {
        const int r = t / CH, c = t - r * CH;
        uint4 *q = d0 + C0 * r + c;
        if constexpr (C2 > 0) {
          if (c >= C0 + C1) q = d2 + C2 * r + (c - C0 - C1);
        }
        if (c >= C0 && c < C0 + C1) q = d1 + C1 * r + (c - C0);
        *q = v[u];
      }
    }
  }
}

// The flux's tile: per entry q, dq/dx, dq/dy as three 16-byte chunks
// (dst[3r + ch]) and eps2 (eps[r]), in one round of loads.
__device__ __forceinline__ void tile_copy_flux(uint4 *dst, float *eps, const uint4 *__restrict__ q,
                                               const uint4 *__restrict__ dq,
                                               const float *__restrict__ eps32,
                                               const int32_t *__restrict__ idx, int count) {
  constexpr int U = 10;  // a 605-entry tile's 2,420 loads in one round
  const int total = count * 4;
  for (int base = threadIdx.x; base < total; base += U * blockDim.x) {
    uint4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      int t = base + u * blockDim.x;
      t = t < total ? t : total - 1;
      const int r = t >> 2, ch = t & 3;
      const size_t g = (size_t)idx[r];
      if (ch == 3)
        v[u] = make_uint4(__float_as_uint(eps32[g]), 0u, 0u, 0u);
      else
        v[u] = ch == 0 ? q[g] : dq[2 * g + (ch - 1)];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int t = base + u * blockDim.x;
      if (t < total) {
        const int r = t >> 2, ch = t & 3;
        if (ch == 3)
          eps[r] = __uint_as_float(v[u].x);
        else
          dst[3 * r + ch] = v[u];
      }
    }
  }
}


// ------------------------------------------------------------ q variables
// q_variables_mod.h:9-34 in FP64, rounded once to FP32
__global__ __launch_bounds__(256) void k_q_variables_m(int n, const double *__restrict__ prim,
                                                       float4 *__restrict__ q32) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double q[4];
  q_of_state(prim[k], prim[n + k], prim[2 * n + k], prim[3 * n + k], q);
  q32[k] = make_float4((float)q[0], (float)q[1], (float)q[2], (float)q[3]);
}

// ------------------------------------------------------------ local dt
// One neighbour's step cfl |x_k - x_i| / speed_k in FP32 (the offset formed
// in FP64, then rounded) and the running minimum with the reference's
// strict comparison; k_func_delta_m and stage 1's q_derivatives share it,
// so the two give the same bits.
__device__ __forceinline__ void delta_step_m(double2 xy, double2 xyk, float speed, float cfl,
                                             float &min_delt) {
  const float fx = (float)(xyk.x - xy.x), fy = (float)(xyk.y - xy.y);
  const float dist = sqrtf(fmaf(fx, fx, fy * fy));
  const float delta_t = cfl * (dist * __builtin_amdgcn_rcpf(speed));
  if (min_delt > delta_t) min_delt = delta_t;
}

// state_update_mod.h:448-496 for the mixed form: the neighbours' coordinates
// and speeds gathered as k_func_delta does (FP64 records; the offsets
// x_k - x_i formed in FP64 before they are rounded), the step
// cfl |x_k - x_i| / speed_k of each neighbour in FP32 -- a neighbour term,
// like the fluxes it scales -- and the minimum over the neighbours, with the
// reference's strict comparison from 1, stored as a double.
template <bool TILED>
__global__ __launch_bounds__(256) void k_func_delta_m(
    int n, const int32_t *__restrict__ off, const int32_t *__restrict__ slice_off,
    const uint32_t *__restrict__ conn, const uint16_t *__restrict__ conn16,
    const int32_t *__restrict__ tile_off, const int32_t *__restrict__ tile_idx,
    const double *__restrict__ base, float cfl, double *__restrict__ delta) {
  extern __shared__ double lds[];
  constexpr int S = 4;  // x, y, speed, pad
  const int b = xcd_block();
  if (TILED) {
    const int t0 = tile_off[b], tn = tile_off[b + 1] - t0;
    for (int t = threadIdx.x; t < tn; t += blockDim.x) {
      const double *h = base + (size_t)B_STRIDE * tile_idx[t0 + t];
      reinterpret_cast<double2 *>(lds + S * t)[0] = reinterpret_cast<const double2 *>(h)[0];
      lds[S * t + 2] = h[B_SPEED];
    }
    __syncthreads();
  }
  const int i = b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *hi = TILED ? lds + S * threadIdx.x : base + (size_t)B_STRIDE * i;
  const double2 xy = reinterpret_cast<const double2 *>(hi)[0];
  float min_delt = 1.0f;
  // neighbour e + 1's record and e + 2's conn entry load while e is used
  auto fetch = [&](uint32_t w, double2 &xyk, double &speed) {
    const int k = conn_index(w, TILED);
    const double *hk = TILED ? lds + S * k : base + (size_t)B_STRIDE * k;
    xyk = reinterpret_cast<const double2 *>(hk)[0];
    speed = TILED ? hk[2] : hk[B_SPEED];
  };
  const Slice L = slice_of(slice_off, conn, conn16, b);
  const int e0 = 0, last = off[i + 1] - off[i] - 1;
  double2 xyk = {0.0, 0.0};
  double speed = 1.0;
  uint32_t w_n = 0;
  if (e0 <= last) {
    fetch(L.raw(e0, TILED), xyk, speed);
    w_n = L.raw(min(e0 + 1, last), TILED);
  }
  for (int e = e0; e <= last; e++) {
    double2 xy_n;
    double speed_n;
    fetch(w_n, xy_n, speed_n);
    w_n = L.raw(min(e + 2, last), TILED);
    delta_step_m(xy, xyk, (float)speed, cfl, min_delt);
    xyk = xy_n;
    speed = speed_n;
  }
  delta[i] = (double)min_delt;
}

// ------------------------------------------------------------ q derivatives
// q_variables_mod.h:36-118: FP32 gathers of q, FP32 right-hand sums
// w dx (q_k - q_i) with the offsets formed in FP64, the solve from the
// point's coefficients; qmax/qmin over FP32 values (rounding is monotone,
// so they are the rounded FP64 extrema).
// A neighbour's weighted offset (w dx, w dy) in FP32: the difference of the
// FP64 coordinates, rounded once.
template <bool POW>
__device__ __forceinline__ void ls_offset(double2 pi, double2 pk, double power, float &fx,
                                          float &fy) {
  const double delx = pk.x - pi.x, dely = pk.y - pi.y;
  if (POW) {
    const double weights = pow(sqrt(fma(delx, delx, dely * dely)), power);
    fx = (float)(weights * delx);
    fy = (float)(weights * dely);
  } else {
    fx = (float)delx;
    fy = (float)dely;
  }
}

// The 2x2 solve of q_variables_mod.h:98-104 from the point's coefficients
// c = (syy, -sxy, sxx) / det of its geometry sums.  A batch's padding points
// have c = 0, so their derivatives are 0.
__device__ __forceinline__ void ls_solve(float4 c, float sx, float sy, float &ox, float &oy) {
  ox = fmaf(c.x, sx, c.y * sy);
  oy = fmaf(c.z, sy, c.y * sx);
}

// One neighbour's terms of the mixed q_derivatives: qmax/qmin and the
// right-hand sums, every fused product written out (the kernels that
// inline this then agree to the last bit).
template <bool POW>
__device__ __forceinline__ void qdm_term(double2 pi, const float qi[4], double2 pk,
                                         const float qk[4], double power, float qmax[4],
                                         float qmin[4], float sdx[4], float sdy[4]) {
  for (int r = 0; r < 4; r++) {
    qmax[r] = fmaxf(qmax[r], qk[r]);  // the reference's strict comparisons for finite q
    qmin[r] = fminf(qmin[r], qk[r]);
  }
  float fwx, fwy;
  ls_offset<POW>(pi, pk, power, fwx, fwy);
  for (int r = 0; r < 4; r++) {
    const float d = qk[r] - qi[r];
    sdx[r] = fmaf(fwx, d, sdx[r]);
    sdy[r] = fmaf(fwy, d, sdy[r]);
  }
}

// The 2x2 solve and the stores of point i's dq and qmax/qmin.
__device__ __forceinline__ void qdm_finish(float4 ls, const float sdx[4], const float sdy[4],
                                           const float qmax[4], const float qmin[4],
                                           float4 *dq_i, float4 *qm_i) {
  float dqx[4], dqy[4];
  for (int r = 0; r < 4; r++) ls_solve(ls, sdx[r], sdy[r], dqx[r], dqy[r]);
  dq_i[0] = make_float4(dqx[0], dqx[1], dqx[2], dqx[3]);
  dq_i[1] = make_float4(dqy[0], dqy[1], dqy[2], dqy[3]);
  qm_i[0] = make_float4(qmax[0], qmax[1], qmax[2], qmax[3]);
  qm_i[1] = make_float4(qmin[0], qmin[1], qmin[2], qmin[3]);
}

// DELTA (stage 1 of an iteration) also takes func_delta's minimum over the
// same neighbours, from the speeds in the base records (tiled: staged as
// floats after the tile's q).
template <bool TILED, bool POW, bool DELTA = false>
__global__ __launch_bounds__(256) void k_q_derivatives_m(
    int n, const int32_t *__restrict__ off, const int32_t *__restrict__ slice_off,
    const uint32_t *__restrict__ conn,
    const uint16_t *__restrict__ conn16, const int32_t *__restrict__ tile_off,
    const int32_t *__restrict__ tile_idx, int tile_max, const double2 *__restrict__ xy,
    const float4 *__restrict__ q32, const float4 *__restrict__ ls, float4 *__restrict__ dq_out,
    float4 *__restrict__ qm_out, double power, const double *__restrict__ base, float cfl,
    double *__restrict__ delta) {
  extern __shared__ uint4 lds4[];
  double2 *lxy = reinterpret_cast<double2 *>(lds4);
  float4 *lq = reinterpret_cast<float4 *>(lds4 + tile_max);
  float *lsp = reinterpret_cast<float *>(lds4 + 2 * tile_max);
  const int b = xcd_block();
  if (TILED) {
    const int t0 = tile_off[b], tn = tile_off[b + 1] - t0;
    tile_gather<1, 1, 0>(lds4, reinterpret_cast<const uint4 *>(xy), lds4 + tile_max,
                         reinterpret_cast<const uint4 *>(q32), nullptr, nullptr,
                         tile_idx + t0, tn);
    if (DELTA)
      for (int t = threadIdx.x; t < tn; t += blockDim.x)
        lsp[t] = (float)base[(size_t)B_STRIDE * tile_idx[t0 + t] + B_SPEED];
    __syncthreads();
  }
  const int i = b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 *gxy = TILED ? lxy : xy;
  const float4 *gq = TILED ? lq : q32;
  const int li = TILED ? (int)threadIdx.x : i;
  const double2 pi = gxy[li];
  float qi[4], qmax[4], qmin[4];
  unpack4(gq[li], qi);
  const float4 lsi = ls[i];
  float sdx[4] = {0.0f, 0.0f, 0.0f, 0.0f}, sdy[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int r = 0; r < 4; r++) qmax[r] = qmin[r] = qi[r];
  // neighbour e + 1's record and e + 2's conn entry load while e is used
  const Slice L = slice_of(slice_off, conn, conn16, b);
  const int e0 = 0, last = off[i + 1] - off[i] - 1;
  auto speed_of = [&](int k) {
    return TILED ? lsp[k] : (float)base[(size_t)B_STRIDE * k + B_SPEED];
  };
  double2 pk = {0.0, 0.0};
  float4 qk4 = {0.0f, 0.0f, 0.0f, 0.0f};
  float spk = 1.0f, min_delt = 1.0f;
  uint32_t w_n = 0;
  if (e0 <= last) {
    const int k0 = conn_index(L.raw(e0, TILED), TILED);
    pk = gxy[k0];
    qk4 = gq[k0];
    if (DELTA) spk = speed_of(k0);
    w_n = L.raw(min(e0 + 1, last), TILED);
  }
  for (int e = e0; e <= last; e++) {
    const int kn = conn_index(w_n, TILED);
    const double2 pk_n = gxy[kn];
    const float4 qk4_n = gq[kn];
    const float spk_n = DELTA ? speed_of(kn) : 1.0f;
    w_n = L.raw(min(e + 2, last), TILED);
    float qk[4];
    unpack4(qk4, qk);
    qdm_term<POW>(pi, qi, pk, qk, power, qmax, qmin, sdx, sdy);
    if (DELTA) delta_step_m(pi, pk, spk, cfl, min_delt);
    pk = pk_n;
    qk4 = qk4_n;
    spk = spk_n;
  }
  qdm_finish(lsi, sdx, sdy, qmax, qmin, dq_out + 2 * (size_t)i, qm_out + 2 * (size_t)i);
  if (DELTA) delta[i] = (double)min_delt;
}

// ------------------------------------------------------------ inner passes
// q_variables_mod.h:120-228, one Jacobi pass (src -> dst), FP32 gathers,
// corrected differences and right-hand sums, the solve from the point's
// coefficients.
// One neighbour's terms, every fused product written out (the kernels that
// inline this then agree to the last bit).
template <bool POW>
__device__ __forceinline__ void qim_term(double2 pi, const float qi[4], const float dxi[4],
                                         const float dyi[4], double2 pk, const float qk[4],
                                         const float dxk[4], const float dyk[4], double power,
                                         float sx[4], float sy[4]) {
  float fx, fy, fwx, fwy;
  ls_offset<false>(pi, pk, 0.0, fx, fy);
  if (POW)
    ls_offset<true>(pi, pk, power, fwx, fwy);
  else
    fwx = fx, fwy = fy;
  for (int r = 0; r < 4; r++) {
    const float t1 = fmaf(fx, dxk[r] - dxi[r], fy * (dyk[r] - dyi[r]));
    const float d = fmaf(-0.5f, t1, qk[r] - qi[r]);
    sx[r] = fmaf(fwx, d, sx[r]);
    sy[r] = fmaf(fwy, d, sy[r]);
  }
}

__device__ __forceinline__ void qim_finish(float4 ls, const float sx[4], const float sy[4],
                                           float4 *d_i) {
  float ox[4], oy[4];
  for (int r = 0; r < 4; r++) ls_solve(ls, sx[r], sy[r], ox[r], oy[r]);
  d_i[0] = make_float4(ox[0], ox[1], ox[2], ox[3]);
  d_i[1] = make_float4(oy[0], oy[1], oy[2], oy[3]);
}

template <bool TILED, bool POW>
__global__ __launch_bounds__(256) void k_q_inner_m(
    int n, const int32_t *__restrict__ off, const int32_t *__restrict__ slice_off,
    const uint32_t *__restrict__ conn,
    const uint16_t *__restrict__ conn16, const int32_t *__restrict__ tile_off,
    const int32_t *__restrict__ tile_idx, int tile_max, const double2 *__restrict__ xy,
    const float4 *__restrict__ q32, const float4 *__restrict__ ls,
    const float4 *__restrict__ src, float4 *__restrict__ dst, double power) {
  extern __shared__ uint4 lds4[];
  double2 *lxy = reinterpret_cast<double2 *>(lds4);
  float4 *lq = reinterpret_cast<float4 *>(lds4 + tile_max);
  float4 *ld = reinterpret_cast<float4 *>(lds4 + 2 * tile_max);
  const int b = xcd_block();
  if (TILED) {
    const int t0 = tile_off[b], tn = tile_off[b + 1] - t0;
    tile_gather<1, 1, 2>(lds4, reinterpret_cast<const uint4 *>(xy), lds4 + tile_max,
                         reinterpret_cast<const uint4 *>(q32), lds4 + 2 * tile_max,
                         reinterpret_cast<const uint4 *>(src), tile_idx + t0, tn);
    __syncthreads();
  }
  const int i = b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 *gxy = TILED ? lxy : xy;
  const float4 *gq = TILED ? lq : q32;
  const float4 *gd = TILED ? ld : src;
  const int li = TILED ? (int)threadIdx.x : i;
  const double2 pi = gxy[li];
  float qi[4], dxi[4], dyi[4];
  unpack4(gq[li], qi);
  unpack4(gd[2 * li], dxi);
  unpack4(gd[2 * li + 1], dyi);
  const float4 lsi = ls[i];
  float sx[4] = {0.0f, 0.0f, 0.0f, 0.0f}, sy[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  // neighbour e + 1's record loads while e is used (conn entries further ahead)
  const Slice L = slice_of(slice_off, conn, conn16, b);
  const int e0 = 0, last = off[i + 1] - off[i] - 1;
  double2 pk = {0.0, 0.0};
  float4 qk4 = {}, dxk4 = {}, dyk4 = {};
  ConnAhead<kAheadQInner> ahead;
  if (e0 <= last) {
    const int k0 = conn_index(L.raw(e0, TILED), TILED);
    pk = gxy[k0];
    qk4 = gq[k0];
    dxk4 = gd[2 * k0];
    dyk4 = gd[2 * k0 + 1];
    ahead.start(L, TILED, e0 + 1, last);
  }
#pragma unroll kAheadQInner
  for (int e = e0; e <= last; e++) {
    const int kn = conn_index(ahead.pop(L, TILED, e + 1 + kAheadQInner, last), TILED);
    const double2 pk_n = gxy[kn];
    const float4 qk4_n = gq[kn], dxk4_n = gd[2 * kn], dyk4_n = gd[2 * kn + 1];
    float qk[4], dxk[4], dyk[4];
    unpack4(qk4, qk);
    unpack4(dxk4, dxk);
    unpack4(dyk4, dyk);
    qim_term<POW>(pi, qi, dxi, dyi, pk, qk, dxk, dyk, power, sx, sy);
    pk = pk_n;
    qk4 = qk4_n;
    dxk4 = dxk4_n;
    dyk4 = dyk4_n;
  }
  qim_finish(lsi, sx, sy, dst + 2 * (size_t)i);
}

// ------------------------------------------------------------ flux residual
// flux_residual_mod.h:10-89 and the *_dG* sub-stencil routines, as the FP64
// kernel (one pass over the neighbour list, shared end terms, per-point
// least-squares coefficients), with the reconstruction, limiter, q~ ->
// primitive, fluxes, coefficient products and sums in FP32 (the sums
// were FP64 before: 0.505 -> 0.494 ms at 2.5M points, |dCl| and |dCd|
// against the FP64 form unchanged at about 2e-9 after 55 iterations).
constexpr int COEF_STRIDE = 8;

template <bool TILED, bool POW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(POW ? 1 : 4))) void k_flux_m(
    int n, const int32_t *__restrict__ off, const int32_t *__restrict__ slice_off,
    const uint32_t *__restrict__ conn,
    const uint16_t *__restrict__ conn16, const int32_t *__restrict__ tile_off,
    const int32_t *__restrict__ tile_idx, int cap, const double2 *__restrict__ xy,
    const float4 *__restrict__ q32, const float4 *__restrict__ dq32,
    const float4 *__restrict__ qm32, const float *__restrict__ eps32,
    const double *__restrict__ coef, const double *__restrict__ nxa,
    const double *__restrict__ nya, const uint8_t *__restrict__ flag,
    const double *__restrict__ delta, double power, double *__restrict__ fres) {
  // LDS tile: q, dq (48 B) and eps2 (4 B) per entry; the FP64 coordinates
  // (16 B) and qmax/qmin (32 B) are read through L2.  cap entries (LDS for
  // cap entries per block) let a CU hold the four blocks the kernel's
  // registers allow.  Tiled, a launch takes the blocks whose tile has at
  // most cap entries; untiled with cap > 0, the others; untiled with cap =
  // 0, every block.
  extern __shared__ uint4 lds4[];
  const float4 *lt = reinterpret_cast<const float4 *>(lds4);
  float *le = reinterpret_cast<float *>(lds4 + 3 * cap);
  const int b = xcd_block();
  if (TILED || cap > 0) {
    const int tn = tile_off[b + 1] - tile_off[b];
    if (TILED ? tn > cap : tn <= cap) return;
  }
  if (TILED) {
    const int t0 = tile_off[b], tn = tile_off[b + 1] - t0;
    const int32_t *idx = tile_idx + t0;
    tile_copy_flux(lds4, le, reinterpret_cast<const uint4 *>(q32),
                   reinterpret_cast<const uint4 *>(dq32), eps32, idx, tn);
    __syncthreads();
  }
  const int i = b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 pi = xy[i];
  float qi[4], dxi[4], dyi[4], dmaxi[4], dmini[4];
  float eps2_i;
  if (TILED) {
    const int li = 3 * (int)threadIdx.x;
    unpack4(lt[li], qi);
    unpack4(lt[li + 1], dxi);
    unpack4(lt[li + 2], dyi);
    unpack4(qm32[2 * i], dmaxi);
    unpack4(qm32[2 * i + 1], dmini);
    eps2_i = le[threadIdx.x];
  } else {
    unpack4(q32[i], qi);
    unpack4(dq32[2 * i], dxi);
    unpack4(dq32[2 * i + 1], dyi);
    unpack4(qm32[2 * i], dmaxi);
    unpack4(qm32[2 * i + 1], dmini);
    eps2_i = eps32[i];
  }
  for (int r = 0; r < 4; r++) {
    dmaxi[r] -= qi[r];
    dmini[r] -= qi[r];
  }
  // The sub-stencil weights and the point's flux sums are FP32: their terms
  // are FP32 flux differences G_k - G_i already, whose rounding (a few ulps
  // of |G| over a neighbour distance) outweighs what FP32 products and sums
  // of the terms add; the coefficients come from the FP64 solves of the
  // set-up and are rounded once.
  float cf[8];
  {
    double c8[8];
    load4(coef + (size_t)COEF_STRIDE * i, c8);
    load4(coef + (size_t)COEF_STRIDE * i + 4, c8 + 4);
    for (int j = 0; j < 8; j++) cf[j] = (float)c8[j];
  }
  const double nx = nxa[i], ny = nya[i];
  const double tx = ny, ty = -nx;
  const float nxf = (float)nx, nyf = (float)ny;
  const float txf = nyf, tyf = -nxf;
  const int type = flag[i];
  const float s2b = type == 0 ? -1.0f : 1.0f;
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  // neighbour e + 1's record and e + 2's conn entry load while e is used
  struct Rec {
    double2 p;
    float4 q, dx, dy, mx, mn;
    float eps2;
  };
  // the global entry (coordinates) and, tiled, the local one (the rest)
  auto fetch = [&](uint32_t wg, uint32_t wl) -> Rec {
    const int kg = conn_index(wg, false);
    if (TILED) {
      const int kl = conn_index(wl, true);
      const float4 *t = lt + 3 * kl;
      return {xy[kg], t[0], t[1], t[2], qm32[2 * kg], qm32[2 * kg + 1], le[kl]};
    }
    return {xy[kg], q32[kg], dq32[2 * kg], dq32[2 * kg + 1], qm32[2 * kg], qm32[2 * kg + 1],
            eps32[kg]};
  };
  const Slice L = slice_of(slice_off, conn, conn16, b);
  const int e0 = 0, last = off[i + 1] - off[i] - 1;
  Rec rc = {};
  uint32_t w_c = 0, w_n = 0, l_n = 0;
  if (e0 <= last) {
    w_c = L.w32(e0);
    rc = fetch(w_c, TILED ? L.w16(e0) : 0u);
    w_n = L.w32(min(e0 + 1, last));
    l_n = TILED ? L.w16(min(e0 + 1, last)) : 0u;
  }
  for (int e = e0; e <= last; e++) {
    const Rec rn = fetch(w_n, l_n);
    const int enn = min(e + 2, last);
    const uint32_t w_nn = L.w32(enn);
    const uint32_t l_nn = TILED ? L.w16(enn) : 0u;
    const uint32_t m = conn_mask(w_c, false);
    const double2 pk = rc.p;
    float qk[4], dxk[4], dyk[4], dmaxk[4], dmink[4];
    unpack4(rc.q, qk);
    unpack4(rc.dx, dxk);
    unpack4(rc.dy, dyk);
    unpack4(rc.mx, dmaxk);
    unpack4(rc.mn, dmink);
    const float eps2_k = rc.eps2;
    const double delx = pk.x - pi.x, dely = pk.y - pi.y;
    const float fx = (float)delx, fy = (float)dely;
    float dsw, dnw;
    if (POW) {
      const double dels = delx * tx + dely * ty;
      const double deln = delx * nx + dely * ny;
      const double weights = pow(sqrt(dels * dels + deln * deln), power);
      dsw = (float)(dels * weights);
      dnw = (float)(deln * weights);
    } else {
      dsw = fx * txf + fy * tyf;
      dnw = fx * nxf + fy * nyf;
    }
    // limited q~ at both ends (interior_fluxes_mod.h:72-83, limiters_mod.h:55-104)
    float qti[4], qtk[4];
    for (int r = 0; r < 4; r++) {
      const float di = fx * dxi[r] + fy * dyi[r];
      const float dk = fx * dxk[r] + fy * dyk[r];
      const float phi_i = venkat_phi_delta(-0.5f * di, dmaxi[r], dmini[r], eps2_i);
      const float phi_k =
          venkat_phi_delta(-0.5f * dk, dmaxk[r] - qk[r], dmink[r] - qk[r], eps2_k);
      qti[r] = qi[r] - 0.5f * phi_i * di;
      qtk[r] = qk[r] - 0.5f * phi_k * dk;
    }
    const EndTerms<float> ei = qtilde_end_terms(qti, nxf, nyf);
    const EndTerms<float> ek = qtilde_end_terms(qtk, nxf, nyf);
    float Gi[4], Gk[4];
    const bool xp = (m & XPOS) != 0;
    const float sx = xp ? 1.0f : -1.0f;
    if (type == 1) {
      split_x(ei, sx, Gi);
      split_x(ek, sx, Gk);
    } else {
      quad_x(ei, sx, s2b, Gi);
      quad_x(ek, sx, s2b, Gk);
    }
    const float wx = xp ? cf[0] * dsw + cf[1] * dnw : cf[2] * dsw + cf[3] * dnw;
    for (int r = 0; r < 4; r++) acc[r] += diff_unfused(Gk[r], Gi[r]) * wx;
    if (xp && (m & XNEG)) {  // dels == 0: the neighbour is in both lists
      if (type == 1) {
        split_x(ei, -1.0f, Gi);
        split_x(ek, -1.0f, Gk);
      } else {
        quad_x(ei, -1.0f, s2b, Gi);
        quad_x(ek, -1.0f, s2b, Gk);
      }
      const float wn = cf[2] * dsw + cf[3] * dnw;
      for (int r = 0; r < 4; r++) acc[r] += diff_unfused(Gk[r], Gi[r]) * wn;
    }
    const bool yp = type == 1 ? (m & YPOS) != 0 : type == 2;
    const float sy = yp ? 1.0f : -1.0f;
    split_y(ei, sy, Gi);
    split_y(ek, sy, Gk);
    const float wy = yp ? cf[4] * dnw + cf[5] * dsw : cf[6] * dnw + cf[7] * dsw;
    for (int r = 0; r < 4; r++) acc[r] += diff_unfused(Gk[r], Gi[r]) * wy;
    if (type == 1 && yp && (m & YNEG)) {  // deln == 0
      split_y(ei, -1.0f, Gi);
      split_y(ek, -1.0f, Gk);
      const float wn = cf[6] * dnw + cf[7] * dsw;
      for (int r = 0; r < 4; r++) acc[r] += diff_unfused(Gk[r], Gi[r]) * wn;
    }
    rc = rn;
    w_c = w_n;
    w_n = w_nn;
    l_n = l_nn;
  }
  const double d = type == 0 ? 2.0 * delta[i] : delta[i];
  for (int j = 0; j < 4; j++) fres[(size_t)j * n + i] = d * (double)acc[j];
}

// ------------------------------------------------------------ KAT
// The eight flux functions in FP32 on FP32-rounded inputs.
__global__ void k_kat_fluxes_m(int n, const double *__restrict__ in, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *v = in + 6 * (size_t)i;
  const EndTerms<float> e = end_terms((float)v[0], (float)v[1], (float)v[2], (float)v[3],
                                      (float)v[4], (float)v[5]);
  float G[8][4];
  split_x(e, 1.0f, G[0]);
  split_x(e, -1.0f, G[1]);
  split_y(e, 1.0f, G[2]);
  split_y(e, -1.0f, G[3]);
  quad_x(e, -1.0f, -1.0f, G[4]);
  quad_x(e, 1.0f, -1.0f, G[5]);
  quad_x(e, 1.0f, 1.0f, G[6]);
  quad_x(e, -1.0f, 1.0f, G[7]);
  for (int f = 0; f < 8; f++)
    for (int r = 0; r < 4; r++) out[32 * (size_t)i + 4 * f + r] = (double)G[f][r];
}

}  // namespace

size_t lds_bytes_qderiv_m(int tile_max) { return 32 * (size_t)tile_max; }
// with func_delta: the tile's speeds after its coordinates and q
size_t lds_bytes_qderiv_delta_m(int tile_max) { return 36 * (size_t)tile_max; }

size_t lds_bytes_qinner_m(int tile_max) { return 64 * (size_t)tile_max; }
size_t lds_bytes_flux_m(int tile_max) { return 52 * (size_t)tile_max; }

int flux_tile_cap_m(int block, int tile_max, size_t lds_limit, size_t lds_per_cu) {
  int nb_regs = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_regs, k_flux_m<true, false>, block, 0) !=
          hipSuccess ||
      nb_regs <= 0)
    return 0;
  // per block: the CU's LDS shared by nb_regs blocks, less 2 KB of
  // allocation granularity, and at most what one block may allocate
  size_t per_block = lds_per_cu / (size_t)nb_regs;
  per_block = per_block > 2048 ? per_block - 2048 : 0;
  if (per_block > lds_limit) per_block = lds_limit;
  const size_t entry = lds_bytes_flux_m(1);
  int cap = (int)(per_block / entry);
  if (cap > tile_max) cap = tile_max;
  return cap >= block ? cap : 0;
}

unsigned prepare_tiled_m(unsigned tiles, int tile_max, int flux_cap, size_t lds_limit) {
  struct K { unsigned bit; const void *f; size_t bytes; };
  const K ks[] = {
      {TILE_DELTA, (const void *)k_func_delta_m<true>, lds_bytes_delta(tile_max)},
      {TILE_QDERIV, (const void *)k_q_derivatives_m<true, false>, lds_bytes_qderiv_m(tile_max)},
      {TILE_QDERIV, (const void *)k_q_derivatives_m<true, true>, lds_bytes_qderiv_m(tile_max)},
      {TILE_QDERIV, (const void *)k_q_derivatives_m<true, false, true>,
       lds_bytes_qderiv_delta_m(tile_max)},
      {TILE_QINNER, (const void *)k_q_inner_m<true, false>, lds_bytes_qinner_m(tile_max)},
      {TILE_QINNER, (const void *)k_q_inner_m<true, true>, lds_bytes_qinner_m(tile_max)},
      {TILE_FLUX, (const void *)k_flux_m<true, false>, lds_bytes_flux_m(flux_cap)},
      {TILE_FLUX, (const void *)k_flux_m<true, true>, lds_bytes_flux_m(flux_cap)}};
  unsigned ok = tiles;
  for (const K &k : ks) {
    if (!(tiles & k.bit)) continue;
    if (k.bytes > lds_limit ||
        hipFuncSetAttribute(k.f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)k.bytes) !=
            hipSuccess)
      ok &= ~k.bit;
  }
  return ok;
}

#define TILE_ARGS_M c.conn_off, c.slice_off, c.connS, c.conn16S, c.tile_off, c.tile_idx, c.tile_max

void launch_q_variables_m(const DevCloud &c, const double *prim, float *q32, hipStream_t s) {
  hipLaunchKernelGGL(k_q_variables_m, dim3(blocks_for(c.n, 256)), dim3(256), 0, s, c.n, prim,
                     reinterpret_cast<float4 *>(q32));
}

void launch_func_delta_m(const DevCloud &c, const double *base, double cfl, double *delta,
                         hipStream_t s) {
  dim3 g(blocks_for(c.n, c.tile_block)), b(c.tile_block);
  if (c.tiles & TILE_DELTA)
    hipLaunchKernelGGL(k_func_delta_m<true>, g, b, lds_bytes_delta(c.tile_max), s, c.n,
                       c.conn_off, c.slice_off, c.connS, c.conn16S, c.tile_off, c.tile_idx, base,
                       (float)cfl, delta);
  else
    hipLaunchKernelGGL(k_func_delta_m<false>, g, b, 0, s, c.n, c.conn_off, c.slice_off, c.connS,
                       c.conn16S, c.tile_off, c.tile_idx, base, (float)cfl, delta);
}

void launch_q_derivatives_m(const DevCloud &c, const MixedState &m, float *dq, double power,
                            hipStream_t s, const double *base, double cfl, double *delta) {
  dim3 g(blocks_for(c.n, c.tile_block)), b(c.tile_block);
  const bool t = (c.tiles & TILE_QDERIV) != 0;
  const size_t L = t ? lds_bytes_qderiv_m(c.tile_max) : 0;
#define QD_ARGS                                                                           \
  c.n, TILE_ARGS_M, reinterpret_cast<const double2 *>(m.xy),                              \
      reinterpret_cast<const float4 *>(m.q32), reinterpret_cast<const float4 *>(m.ls32),   \
      reinterpret_cast<float4 *>(dq), reinterpret_cast<float4 *>(m.qm32), power, base,     \
      (float)cfl, delta
  if (delta && t)  // power 0 (the caller's condition for fusing)
    hipLaunchKernelGGL((k_q_derivatives_m<true, false, true>), g, b,
                       lds_bytes_qderiv_delta_m(c.tile_max), s, QD_ARGS);
  else if (delta)
    hipLaunchKernelGGL((k_q_derivatives_m<false, false, true>), g, b, 0, s, QD_ARGS);
  else if (t && power == 0.0)
    hipLaunchKernelGGL((k_q_derivatives_m<true, false>), g, b, L, s, QD_ARGS);
  else if (t)
    hipLaunchKernelGGL((k_q_derivatives_m<true, true>), g, b, L, s, QD_ARGS);
  else if (power == 0.0)
    hipLaunchKernelGGL((k_q_derivatives_m<false, false>), g, b, 0, s, QD_ARGS);
  else
    hipLaunchKernelGGL((k_q_derivatives_m<false, true>), g, b, 0, s, QD_ARGS);
#undef QD_ARGS
}

void launch_q_inner_m(const DevCloud &c, const MixedState &m, const float *src, float *dst,
                      double power, hipStream_t s) {
  dim3 g(blocks_for(c.n, c.tile_block)), b(c.tile_block);
  const bool t = (c.tiles & TILE_QINNER) != 0;
  const size_t L = t ? lds_bytes_qinner_m(c.tile_max) : 0;
#define QI_ARGS                                                                           \
  c.n, TILE_ARGS_M, reinterpret_cast<const double2 *>(m.xy),                              \
      reinterpret_cast<const float4 *>(m.q32), reinterpret_cast<const float4 *>(m.ls32),   \
      reinterpret_cast<const float4 *>(src), reinterpret_cast<float4 *>(dst), power
  if (t && power == 0.0)
    hipLaunchKernelGGL((k_q_inner_m<true, false>), g, b, L, s, QI_ARGS);
  else if (t)
    hipLaunchKernelGGL((k_q_inner_m<true, true>), g, b, L, s, QI_ARGS);
  else if (power == 0.0)
    hipLaunchKernelGGL((k_q_inner_m<false, false>), g, b, 0, s, QI_ARGS);
  else
    hipLaunchKernelGGL((k_q_inner_m<false, true>), g, b, 0, s, QI_ARGS);
#undef QI_ARGS
}

void launch_flux_m(const DevCloud &c, const MixedState &m, const float *dq, const double *delta,
                   double power, double *flux_res, hipStream_t s) {
  dim3 g(blocks_for(c.n, c.tile_block)), b(c.tile_block);
  const bool t = (c.tiles & TILE_FLUX) != 0;
  const size_t L = t ? lds_bytes_flux_m(c.flux_cap) : 0;
  const int cap = t ? c.flux_cap : 0;
#define FLUX_ARGS                                                                          \
  c.n, c.conn_off, c.slice_off, c.connS, c.conn16S, c.tile_off, c.tile_idx, cap,          \
      reinterpret_cast<const double2 *>(m.xy),                                             \
      reinterpret_cast<const float4 *>(m.q32), reinterpret_cast<const float4 *>(dq),       \
      reinterpret_cast<const float4 *>(m.qm32), m.eps32, c.coef, c.nx, c.ny, c.flag, delta, \
      power, flux_res
  // tiled: the blocks whose tile fits the cap, then (if any tile is larger)
  // the rest through L2
  if (t && power == 0.0)
    hipLaunchKernelGGL((k_flux_m<true, false>), g, b, L, s, FLUX_ARGS);
  else if (t)
    hipLaunchKernelGGL((k_flux_m<true, true>), g, b, L, s, FLUX_ARGS);
  if (!t || c.tile_max > cap) {
    if (power == 0.0)
      hipLaunchKernelGGL((k_flux_m<false, false>), g, b, 0, s, FLUX_ARGS);
    else
      hipLaunchKernelGGL((k_flux_m<false, true>), g, b, 0, s, FLUX_ARGS);
  }
#undef FLUX_ARGS
}

void launch_kat_fluxes_m(int n, const double *in6, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_kat_fluxes_m, dim3(blocks_for(n, 128)), dim3(128), 0, s, n, in6, out);
}

}  // namespace lskum
