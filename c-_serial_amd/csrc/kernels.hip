// kernels.hip -- q-LSKUM stage kernels for gfx950 (FP64 form).
//
// One thread per point, 256 points per block.  Stencil kernels walk the
// point's CSR neighbour list in the reference's conn order, so every
// per-point sum is accumulated in the same order as the reference's serial
// loops.  Neighbour data come from the
// block's LDS tile (its own points and every point its lists reach, loaded
// once per block) for the kernels configured so, else straight from the
// 64-byte per-point records through L2.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"
#include "pipeline.hpp"
#include "physics.hpp"

namespace lskum {

namespace {

// ------------------------------------------------------------ q variables
// q_variables_mod.h:9-34
__global__ __launch_bounds__(256) void k_q_variables(int n, const double *__restrict__ prim,
                                                     double *__restrict__ base) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double q[4];
  q_of_state(prim[k], prim[n + k], prim[2 * n + k], prim[3 * n + k], q);
  store4(base + (size_t)B_STRIDE * k + B_Q, q);
}

// The denominator of state_update_mod.h:485 for the point's own state:
// func_delta divides each neighbour distance by its neighbour's value.
__global__ __launch_bounds__(256) void k_speed(int n, const double *__restrict__ prim,
                                               double *__restrict__ base) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  base[(size_t)B_STRIDE * k + B_SPEED] =
      speed_of_state(prim[k], prim[n + k], prim[2 * n + k], prim[3 * n + k]);
}

// ------------------------------------------------------------ local dt
// state_update_mod.h:448-496: min over the neighbours' states only.
template <bool TILED>
__global__ __launch_bounds__(256) void k_func_delta(int n, const int32_t *__restrict__ off,
                                                    const int32_t *__restrict__ slice_off,
                                                    const uint32_t *__restrict__ conn,
                                                    const uint16_t *__restrict__ conn16,
                                                    const int32_t *__restrict__ tile_off,
                                                    const int32_t *__restrict__ tile_idx,
                                                    const double *__restrict__ base, double cfl,
                                                    double *__restrict__ delta) {
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
  double min_delt = 1.0;
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
    const double delta_t = delta_step(xy, xyk, speed, cfl);
    if (min_delt > delta_t) min_delt = delta_t;
    xyk = xy_n;
    speed = speed_n;
  }
  delta[i] = min_delt;
}

// ------------------------------------------------------------ q derivatives
// q_variables_mod.h:36-118: unweighted (power = 0) or pow(dist, power)
// weighted least squares over the full stencil, and qmax/qmin over the
// point and its neighbours.
// ARG (the adjoint's recomputation) also records, per qmax/qmin
// component, the point the extremum came from (q_variables_mod.h:60-80:
// the point itself, replaced by a neighbour only on a strictly larger /
// smaller value, in conn order; device index), which the reverse sweep
// routes the extrema's adjoints to, and the geometry sums (sxx, syy, sxy)
// the least-squares reverse needs.
// One neighbour's terms of the unweighted least squares
// (q_variables_mod.h:83-101, weights = pow(dist, 0) = 1 exactly).
__device__ __forceinline__ void q_deriv_term(double x_i, double y_i, const double qi[4],
                                             double2 xyk, const double qk[4],
                                             double &sum_delx_sqr, double &sum_dely_sqr,
                                             double &sum_delx_dely, double sdx[4],
                                             double sdy[4]) {
  const double delx = xyk.x - x_i, dely = xyk.y - y_i;
  sum_delx_sqr = sum_delx_sqr + delx * delx;
  sum_dely_sqr = sum_dely_sqr + dely * dely;
  sum_delx_dely = sum_delx_dely + delx * dely;
  for (int r = 0; r < 4; r++) {
    sdx[r] = sdx[r] + delx * (qk[r] - qi[r]);
    sdy[r] = sdy[r] + dely * (qk[r] - qi[r]);
  }
}

// The 2x2 solve (q_variables_mod.h:103-115), its one fused product written
// out (the kernels that inline this then agree to the last bit), and the
// stores of dq, qmax/qmin and, when qs is set, the q sums.
__device__ __forceinline__ void q_deriv_finish(double sum_delx_sqr, double sum_dely_sqr,
                                               double sum_delx_dely, const double sdx[4],
                                               const double sdy[4], const double qmax[4],
                                               const double qmin[4], double *dq, double *qm,
                                               double *qs) {
  const double det = fma(sum_delx_sqr, sum_dely_sqr, -(sum_delx_dely * sum_delx_dely));
  // det > 0 for every validated point; 0 only for a batch's padding points,
  // which have no neighbours and must not produce NaNs
  const double one_by_det = det != 0.0 ? 1.0 / det : 0.0;
  double dqx[4], dqy[4];
  for (int r = 0; r < 4; r++) {
    dqx[r] = fma(sdx[r], sum_dely_sqr, -(sdy[r] * sum_delx_dely)) * one_by_det;
    dqy[r] = fma(sdy[r], sum_delx_sqr, -(sdx[r] * sum_delx_dely)) * one_by_det;
  }
  store4(dq, dqx);
  store4(dq + 4, dqy);
  store4(qm, qmax);
  store4(qm + 4, qmin);
  if (qs) {  // the q sums the inner passes reuse
    store4(qs, sdx);
    store4(qs + 4, sdy);
  }
}

template <bool TILED, bool POW, bool ARG = false>
__global__ __launch_bounds__(256) void k_q_derivatives(
    int n, const int32_t *__restrict__ off, const int32_t *__restrict__ slice_off,
    const uint32_t *__restrict__ conn,
    const uint16_t *__restrict__ conn16, const int32_t *__restrict__ tile_off,
    const int32_t *__restrict__ tile_idx, const double *__restrict__ base,
    double *__restrict__ dq_out, double *__restrict__ qm_out, double *__restrict__ qs_out,
    double power, int32_t *__restrict__ arg = nullptr, double *__restrict__ lsum = nullptr) {
  extern __shared__ double lds[];
  constexpr int S = 6;  // x, y, q[4]
  const int b = xcd_block();
  if (TILED) {
    load_tile<S, 8>(lds, tile_off, tile_idx, b, base, B_STRIDE, 0, 3, 0);  // 682 entries a round
    __syncthreads();
  }
  const int i = b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *hi = TILED ? lds + S * threadIdx.x : base + (size_t)B_STRIDE * i;
  const double2 xy = reinterpret_cast<const double2 *>(hi)[0];
  const double x_i = xy.x, y_i = xy.y;
  double qi[4], qmax[4], qmin[4];
  load4(hi + B_Q, qi);
  double sum_delx_sqr = 0., sum_dely_sqr = 0., sum_delx_dely = 0.;
  double sdx[4] = {0.0, 0.0, 0.0, 0.0}, sdy[4] = {0.0, 0.0, 0.0, 0.0};
  for (int r = 0; r < 4; r++) qmax[r] = qmin[r] = qi[r];
  int am[4] = {i, i, i, i}, an[4] = {i, i, i, i};
  const int32_t *tix = TILED ? tile_idx + tile_off[b] : nullptr;
  // neighbour e + 1's record loads while e is used (conn entries further ahead)
  auto fetch = [&](uint32_t w, double2 &xyk, double qk[4]) {
    const int k = conn_index(w, TILED);
    const double *hk = TILED ? lds + S * k : base + (size_t)B_STRIDE * k;
    xyk = reinterpret_cast<const double2 *>(hk)[0];
    load4(hk + B_Q, qk);
  };
  const Slice L = slice_of(slice_off, conn, conn16, b);
  const int e0 = 0, last = off[i + 1] - off[i] - 1;
  double2 xyk = {0.0, 0.0};
  double qk[4] = {0.0, 0.0, 0.0, 0.0};
  ConnAhead<kAheadQDeriv> ahead;
  uint32_t w_c = 0;
  if (e0 <= last) {
    w_c = L.raw(e0, TILED);
    fetch(w_c, xyk, qk);
    ahead.start(L, TILED, e0 + 1, last);
  }
#pragma unroll kAheadQDeriv
  for (int e = e0; e <= last; e++) {
    double2 xy_n;
    double q_n[4];
    const uint32_t w_n = ahead.pop(L, TILED, e + 1 + kAheadQDeriv, last);
    fetch(w_n, xy_n, q_n);
    const uint32_t w_k = w_c;
    w_c = w_n;
    if (ARG) {
      const int kl = conn_index(w_k, TILED);
      const int kg = TILED ? tix[kl] : kl;
      for (int r = 0; r < 4; r++) {
        if (qk[r] > qmax[r]) { qmax[r] = qk[r]; am[r] = kg; }
        if (qk[r] < qmin[r]) { qmin[r] = qk[r]; an[r] = kg; }
      }
    } else {
      // one max / min instruction each in place of a compare and two
      // selects: the same value as the reference's strict comparisons for
      // every finite q (on equal values, +0 and -0 included, either is the
      // result); only a NaN q, which no finite state has, would differ
      for (int r = 0; r < 4; r++) {
        qmax[r] = fmax(qmax[r], qk[r]);
        qmin[r] = fmin(qmin[r], qk[r]);
      }
    }
    const double delx = xyk.x - x_i, dely = xyk.y - y_i;
    if (POW) {
      const double weights = pow(sqrt(delx * delx + dely * dely), power);
      sum_delx_sqr = sum_delx_sqr + delx * delx * weights;
      sum_dely_sqr = sum_dely_sqr + dely * dely * weights;
      sum_delx_dely = sum_delx_dely + delx * dely * weights;
      for (int r = 0; r < 4; r++) {
        sdx[r] = sdx[r] + weights * delx * (qk[r] - qi[r]);
        sdy[r] = sdy[r] + weights * dely * (qk[r] - qi[r]);
      }
    } else {  // weights == 1.0 exactly: x*1.0 == x
      q_deriv_term(x_i, y_i, qi, xyk, qk, sum_delx_sqr, sum_dely_sqr, sum_delx_dely, sdx, sdy);
    }
    xyk = xy_n;
    for (int r = 0; r < 4; r++) qk[r] = q_n[r];
  }
  q_deriv_finish(sum_delx_sqr, sum_dely_sqr, sum_delx_dely, sdx, sdy, qmax, qmin,
                 dq_out + (size_t)DQ_STRIDE * i, qm_out + (size_t)QM_STRIDE * i,
                 qs_out ? qs_out + 8 * (size_t)i : nullptr);
  if (ARG) {
    for (int r = 0; r < 4; r++) {
      arg[8 * (size_t)i + r] = am[r];
      arg[8 * (size_t)i + 4 + r] = an[r];
    }
    lsum[4 * (size_t)i] = sum_delx_sqr;
    lsum[4 * (size_t)i + 1] = sum_dely_sqr;
    lsum[4 * (size_t)i + 2] = sum_delx_dely;
  }
}

// ------------------------------------------------------------ inner passes
// One neighbour's terms of an inner pass (q_variables_mod.h:150-195): the
// least-squares geometry sums and w dx (dx.(dq_k - dq_i)).  The reference's
// factor -0.5 of each term is applied to the sums once, in q_inner_finish:
// scaling by -0.5 is exact, so the sums are the same numbers.
template <bool POW>
__device__ __forceinline__ void q_inner_term(double x_i, double y_i, const double dxi[4],
                                             const double dyi[4], double2 xyk,
                                             const double dxk[4], const double dyk[4],
                                             double power, double &sum_delx_sqr,
                                             double &sum_dely_sqr, double &sum_delx_dely,
                                             double sx[4], double sy[4]) {
  const double delx = xyk.x - x_i, dely = xyk.y - y_i;
  const double weights = POW ? pow(sqrt(delx * delx + dely * dely), power) : 1.0;
  if (POW) {
    sum_delx_sqr = sum_delx_sqr + delx * delx * weights;
    sum_dely_sqr = sum_dely_sqr + dely * dely * weights;
    sum_delx_dely = sum_delx_dely + delx * dely * weights;
  } else {
    sum_delx_sqr = sum_delx_sqr + delx * delx;
    sum_dely_sqr = sum_dely_sqr + dely * dely;
    sum_delx_dely = sum_delx_dely + delx * dely;
  }
  for (int r = 0; r < 4; r++) {
    // the sum of two products with one of them fused, written out: left to
    // the compiler, which product it fuses differed between the kernels
    // that inline this (a last-bit difference between layouts)
    const double c = fma(delx, dxk[r] - dxi[r], dely * (dyk[r] - dyi[r]));
    if (POW) {
      sx[r] = sx[r] + (weights * delx * c);
      sy[r] = sy[r] + (weights * dely * c);
    } else {
      sx[r] = sx[r] + (delx * c);
      sy[r] = sy[r] + (dely * c);
    }
  }
}

// The pass's result (q_variables_mod.h:197-213): the sums scaled by -0.5
// and q_derivatives' q sums (qx, qy) added, the 2x2 solve, the point's new
// derivatives stored at d.
__device__ __forceinline__ void q_inner_finish(const double qx[4], const double qy[4],
                                               double sum_delx_sqr, double sum_dely_sqr,
                                               double sum_delx_dely, double sx[4],
                                               double sy[4], double *d) {
  for (int r = 0; r < 4; r++) {
    sx[r] = qx[r] + (-0.5 * sx[r]);
    sy[r] = qy[r] + (-0.5 * sy[r]);
  }
  const double det = fma(sum_delx_sqr, sum_dely_sqr, -(sum_delx_dely * sum_delx_dely));
  // det > 0 for every validated point; 0 only for a batch's padding points,
  // which have no neighbours and must not produce NaNs
  const double one_by_det = det != 0.0 ? 1.0 / det : 0.0;
  double ox[4], oy[4];
  for (int r = 0; r < 4; r++) {
    ox[r] = one_by_det * fma(sx[r], sum_dely_sqr, -(sy[r] * sum_delx_dely));
    oy[r] = one_by_det * fma(sy[r], sum_delx_sqr, -(sx[r] * sum_delx_dely));
  }
  store4(d, ox);
  store4(d + 4, oy);
}

// q_variables_mod.h:120-228: one defect-corrected LS pass, Jacobi (reads
// src, writes dst: the reference's temp array + copy).  Per pair the pass
// sums w dx ((q_k - q_i) - 0.5 dx.(dq_k - dq_i)) (dx = (delx, dely)); the
// first part, w dx (q_k - q_i), is q_derivatives' sum, which does not
// change between passes and is read from qs, so a pass reads only the
// neighbours' coordinates and derivatives (80 B per tile entry, against
// 112 B with q, which lets three blocks share a CU instead of two).
template <bool TILED, bool POW>
__global__ __launch_bounds__(256) void k_q_inner(
    int n, const int32_t *__restrict__ off, const int32_t *__restrict__ slice_off,
    const uint32_t *__restrict__ conn,
    const uint16_t *__restrict__ conn16, const int32_t *__restrict__ tile_off,
    const int32_t *__restrict__ tile_idx, const double *__restrict__ xy,
    const double *__restrict__ qs, const double *__restrict__ src, double *__restrict__ dst,
    double power) {
  extern __shared__ double lds[];
  constexpr int S = 10;  // x, y, dqx[4], dqy[4]
  const int b = xcd_block();
  if (TILED) {
    // one round each for tiles of up to 768 entries (3 and 10 chunks per
    // thread at 256 threads); the coordinates come from the compact xy
    // array, four points to a 64-byte line
    load_tile<S, 3>(lds, tile_off, tile_idx, b, xy, 2, 0, 1, 0);
    load_tile<S, 10>(lds, tile_off, tile_idx, b, src, DQ_STRIDE, 0, 4, 2);
    __syncthreads();
  }
  const int i = b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *hi = TILED ? lds + S * threadIdx.x : xy + 2 * (size_t)i;
  const double *di = TILED ? hi + 2 : src + (size_t)DQ_STRIDE * i;
  const double2 pi = reinterpret_cast<const double2 *>(hi)[0];
  const double x_i = pi.x, y_i = pi.y;
  double dxi[4], dyi[4];
  load4(di, dxi);
  load4(di + 4, dyi);
  double sum_delx_sqr = 0., sum_dely_sqr = 0., sum_delx_dely = 0.;
  double sx[4] = {0.0, 0.0, 0.0, 0.0}, sy[4] = {0.0, 0.0, 0.0, 0.0};
  // neighbour e + 1's record loads while e is used (conn entries further ahead)
  auto fetch = [&](uint32_t w, double2 &xyk, double dxk[4], double dyk[4]) {
    const int k = conn_index(w, TILED);
    const double *hk = TILED ? lds + S * k : xy + 2 * (size_t)k;
    const double *dk = TILED ? hk + 2 : src + (size_t)DQ_STRIDE * k;
    xyk = reinterpret_cast<const double2 *>(hk)[0];
    load4(dk, dxk);
    load4(dk + 4, dyk);
  };
  const Slice L = slice_of(slice_off, conn, conn16, b);
  const int e0 = 0, last = off[i + 1] - off[i] - 1;
  double2 xyk = {0.0, 0.0};
  double dxk[4] = {}, dyk[4] = {};
  ConnAhead<kAheadQInner> ahead;
  if (e0 <= last) {
    fetch(L.raw(e0, TILED), xyk, dxk, dyk);
    ahead.start(L, TILED, e0 + 1, last);
  }
#pragma unroll kAheadQInner
  for (int e = e0; e <= last; e++) {
    double2 xy_n;
    double dx_n[4], dy_n[4];
    fetch(ahead.pop(L, TILED, e + 1 + kAheadQInner, last), xy_n, dx_n, dy_n);
    q_inner_term<POW>(x_i, y_i, dxi, dyi, xyk, dxk, dyk, power, sum_delx_sqr, sum_dely_sqr,
                      sum_delx_dely, sx, sy);
    xyk = xy_n;
    for (int r = 0; r < 4; r++) {
      dxk[r] = dx_n[r];
      dyk[r] = dy_n[r];
    }
  }
  double qx[4], qy[4];
  load4(qs + 8 * (size_t)i, qx);
  load4(qs + 8 * (size_t)i + 4, qy);
  q_inner_finish(qx, qy, sum_delx_sqr, sum_dely_sqr, sum_delx_dely, sx, sy,
                 dst + (size_t)DQ_STRIDE * i);
}

// The inner pass again, pipelined.  A workgroup of tile_block threads walks
// a sequence of stencil blocks (persistent: a few workgroups per CU, each
// block of an XCD's range taken by a workgroup of that XCD).  While it
// computes block j from LDS, block j+1's gather -- its tile entries'
// coordinates and derivatives and its neighbour-list slice, as 16-byte
// chunks -- is in flight to registers, and block j+2's tile indices before
// it; the arithmetic reads its lists from LDS, so none of it waits on a
// global load.  Each thread holds E entry slots (tile entry t + u TB) with
// their chunks and NC slots of the list slice (chunk t + u TB); a slot past
// the end takes the last entry or chunk again and writes the same bytes to
// the same place, so every thread issues the same loads and the compiler's
// memory-operation counts stay exact.  The kind of every slot is fixed at
// compile time, so addresses take a few integer operations.  The per-point
// arithmetic is k_q_inner's (q_inner_term, q_inner_finish), so results are
// bit-identical.
// q_inner: entry slot u holds the coordinates (from xy) and the four
// derivative chunks (from src) of its tile entry, stored at LDS chunk 5 e
template <int E, int NC>
__device__ __forceinline__ void qi_load_data(qi_chunk (&v)[5 * E + NC], const int (&idx)[E],
                                             const QiBlock &q, int t, int TB,
                                             const double *__restrict__ xy,
                                             const double *__restrict__ src,
                                             const uint16_t *__restrict__ conn16) {
  const qi_chunk *xy2 = reinterpret_cast<const qi_chunk *>(xy);
  const qi_chunk *src2 = reinterpret_cast<const qi_chunk *>(src);
#pragma unroll
  for (int u = 0; u < E; u++) {
    const size_t g = (size_t)idx[u];
    v[5 * u] = xy2[g];
#pragma unroll
    for (int j = 0; j < 4; j++) v[5 * u + 1 + j] = src2[(DQ_STRIDE / 2) * g + j];
  }
  pf_load_conn<NC>(v, 5 * E, q, t, TB, conn16, xy2 + idx[0]);
}

template <int E, int NC>
__device__ __forceinline__ void qi_store_lds(qi_chunk *l2, const qi_chunk (&v)[5 * E + NC],
                                             const QiBlock &q, int t, int TB, int lds_conn2) {
#pragma unroll
  for (int u = 0; u < E; u++) {
    const int e = min(t + u * TB, q.tn - 1);
#pragma unroll
    for (int j = 0; j < 5; j++) l2[5 * e + j] = v[5 * u + j];
  }
  pf_store_conn<NC>(l2, v, 5 * E, q, t, TB, lds_conn2);
}

template <int E, int NC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_q_inner_pf(
    int n, int nblk, int lds_conn, const int32_t *__restrict__ off,
    const int32_t *__restrict__ slice_off, const uint16_t *__restrict__ conn16,
    const int32_t *__restrict__ tile_off, const int32_t *__restrict__ tile_idx,
    const double *__restrict__ xy, const double *__restrict__ qs,
    const double *__restrict__ src, double *__restrict__ dst) {
  extern __shared__ double lds[];
  constexpr int S = 10;  // x, y, dqx[4], dqy[4]
  const uint16_t *cl = reinterpret_cast<const uint16_t *>(lds + lds_conn);
  qi_chunk *l2 = reinterpret_cast<qi_chunk *>(lds);
  const int t = threadIdx.x, TB = blockDim.x;
  const int lg_tb = __ffs(TB) - 1;  // tile_block is 64, 128 or 256: row e of the slice at e << lg_tb
  const int G = gridDim.x;
  int idx[E];
  qi_chunk v[5 * E + NC];
  int vb = blockIdx.x;
  if (vb >= nblk) return;  // uniform per workgroup: no barrier is reached
  int b = xcd_map(vb, nblk);
  QiBlock q = qi_block(tile_off, slice_off, b);
  pf_load_idx<E>(idx, q, t, TB, tile_idx);
  qi_load_data<E, NC>(v, idx, q, t, TB, xy, src, conn16);
  // past the end of the sequence the next block is the current one again:
  // its loads are redundant but keep every pass through the loop alike
  int vn = vb + G;
  QiBlock qn = vn < nblk ? qi_block(tile_off, slice_off, xcd_map(vn, nblk)) : q;
  pf_load_idx<E>(idx, qn, t, TB, tile_idx);
  for (;;) {
    __syncthreads();  // the previous block's arithmetic is done with the LDS
    qi_store_lds<E, NC>(l2, v, q, t, TB, lds_conn / 2);
    const int i = b * TB + t;
    const int ii = min(i, n - 1);
    const int e_off = off[ii], deg = off[ii + 1] - e_off;
    double qx[4], qy[4];
    load4(qs + 8 * (size_t)ii, qx);
    load4(qs + 8 * (size_t)ii + 4, qy);
    const int vnn = vn + G;
    const QiBlock qnn = vnn < nblk ? qi_block(tile_off, slice_off, xcd_map(vnn, nblk)) : qn;
    qi_load_data<E, NC>(v, idx, qn, t, TB, xy, src, conn16);
    pf_load_idx<E>(idx, qnn, t, TB, tile_idx);
    __syncthreads();  // this block's tile and lists are in the LDS
    if (i < n) {
      const double *hi = lds + S * t;
      const double2 pi = reinterpret_cast<const double2 *>(hi)[0];
      const double x_i = pi.x, y_i = pi.y;
      double dxi[4], dyi[4];
      load4(hi + 2, dxi);
      load4(hi + 6, dyi);
      double sum_delx_sqr = 0., sum_dely_sqr = 0., sum_delx_dely = 0.;
      double sx[4] = {0.0, 0.0, 0.0, 0.0}, sy[4] = {0.0, 0.0, 0.0, 0.0};
      auto fetch = [&](int e, double2 &xyk, double dxk[4], double dyk[4]) {
        const double *hk = lds + S * (int)(cl[(e << lg_tb) + t] & kLocalMask);
        xyk = reinterpret_cast<const double2 *>(hk)[0];
        load4(hk + 2, dxk);
        load4(hk + 6, dyk);
      };
      const int last = deg - 1;
      double2 xyk = {0.0, 0.0};
      double dxk[4] = {}, dyk[4] = {};
      if (last >= 0) fetch(0, xyk, dxk, dyk);
      for (int e = 0; e <= last; e++) {
        double2 xy_n;
        double dx_n[4], dy_n[4];
        fetch(min(e + 1, last), xy_n, dx_n, dy_n);
        q_inner_term<false>(x_i, y_i, dxi, dyi, xyk, dxk, dyk, 0.0, sum_delx_sqr,
                            sum_dely_sqr, sum_delx_dely, sx, sy);
        xyk = xy_n;
        for (int r = 0; r < 4; r++) {
          dxk[r] = dx_n[r];
          dyk[r] = dy_n[r];
        }
      }
      q_inner_finish(qx, qy, sum_delx_sqr, sum_dely_sqr, sum_delx_dely, sx, sy,
                     dst + (size_t)DQ_STRIDE * i);
    }
    if (vn >= nblk) break;
    vb = vn;
    vn = vnn;
    b = xcd_map(vb, nblk);
    q = qn;
    qn = qnn;
  }
}

// q_derivatives pipelined the same way (power 0, without the extremum
// origins): entry slot u holds C chunks of its tile entry's base record,
// x, y and q (C = 3), or the whole record with eps^2 and the speed (C = 4)
// where the kernel also computes func_delta, which walks the same lists.
template <int E, int C, int NC>
__device__ __forceinline__ void qd_load_data(qi_chunk (&v)[C * E + NC], const int (&idx)[E],
                                             const QiBlock &q, int t, int TB,
                                             const double *__restrict__ base,
                                             const uint16_t *__restrict__ conn16) {
  const qi_chunk *b2 = reinterpret_cast<const qi_chunk *>(base);
#pragma unroll
  for (int u = 0; u < E; u++) {
    const size_t g = (size_t)idx[u];
#pragma unroll
    for (int j = 0; j < C; j++) v[C * u + j] = b2[(B_STRIDE / 2) * g + j];
  }
  pf_load_conn<NC>(v, C * E, q, t, TB, conn16, b2 + (B_STRIDE / 2) * (size_t)idx[0]);
}

template <int E, int C, int NC>
__device__ __forceinline__ void qd_store_lds(qi_chunk *l2, const qi_chunk (&v)[C * E + NC],
                                             const QiBlock &q, int t, int TB, int lds_conn2) {
#pragma unroll
  for (int u = 0; u < E; u++) {
    const int e = min(t + u * TB, q.tn - 1);
#pragma unroll
    for (int j = 0; j < C; j++) l2[C * e + j] = v[C * u + j];
  }
  pf_store_conn<NC>(l2, v, C * E, q, t, TB, lds_conn2);
}

// Three waves per SIMD (168 VGPRs) fit without spills except for the
// layouts with func_delta and more than three entry or list slots a thread.
template <int E, int NC, bool DELTA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    DELTA && (E > 3 || NC > 3) ? 2 : 3, DELTA && (E > 3 || NC > 3) ? 2 : 3))) void k_q_derivatives_pf(
    int n, int nblk, int lds_conn, const int32_t *__restrict__ off,
    const int32_t *__restrict__ slice_off, const uint16_t *__restrict__ conn16,
    const int32_t *__restrict__ tile_off, const int32_t *__restrict__ tile_idx,
    const double *__restrict__ base, double *__restrict__ dq_out, double *__restrict__ qm_out,
    double *__restrict__ qs_out, double cfl, double *__restrict__ delta) {
  extern __shared__ double lds[];
  constexpr int C = DELTA ? 4 : 3;
  constexpr int S = 2 * C;  // x, y, q[4] (, eps^2, speed)
  const uint16_t *cl = reinterpret_cast<const uint16_t *>(lds + lds_conn);
  qi_chunk *l2 = reinterpret_cast<qi_chunk *>(lds);
  const int t = threadIdx.x, TB = blockDim.x;
  const int lg_tb = __ffs(TB) - 1;  // tile_block is 64, 128 or 256: row e of the slice at e << lg_tb
  const int G = gridDim.x;
  int idx[E];
  qi_chunk v[C * E + NC];
  int vb = blockIdx.x;
  if (vb >= nblk) return;  // uniform per workgroup: no barrier is reached
  int b = xcd_map(vb, nblk);
  QiBlock q = qi_block(tile_off, slice_off, b);
  pf_load_idx<E>(idx, q, t, TB, tile_idx);
  qd_load_data<E, C, NC>(v, idx, q, t, TB, base, conn16);
  int vn = vb + G;
  QiBlock qn = vn < nblk ? qi_block(tile_off, slice_off, xcd_map(vn, nblk)) : q;
  pf_load_idx<E>(idx, qn, t, TB, tile_idx);
  for (;;) {
    __syncthreads();  // the previous block's arithmetic is done with the LDS
    qd_store_lds<E, C, NC>(l2, v, q, t, TB, lds_conn / 2);
    const int i = b * TB + t;
    const int ii = min(i, n - 1);
    const int deg = off[ii + 1] - off[ii];
    const int vnn = vn + G;
    const QiBlock qnn = vnn < nblk ? qi_block(tile_off, slice_off, xcd_map(vnn, nblk)) : qn;
    qd_load_data<E, C, NC>(v, idx, qn, t, TB, base, conn16);
    pf_load_idx<E>(idx, qnn, t, TB, tile_idx);
    __syncthreads();  // this block's tile and lists are in the LDS
    if (i < n) {
      const double *hi = lds + S * t;
      const double2 xy = reinterpret_cast<const double2 *>(hi)[0];
      double qi[4], qmax[4], qmin[4];
      load4(hi + 2, qi);
      double sum_delx_sqr = 0., sum_dely_sqr = 0., sum_delx_dely = 0.;
      double sdx[4] = {0.0, 0.0, 0.0, 0.0}, sdy[4] = {0.0, 0.0, 0.0, 0.0};
      double min_delt = 1.0;
      for (int r = 0; r < 4; r++) qmax[r] = qmin[r] = qi[r];
      auto fetch = [&](int e, double2 &xyk, double qk[4], double &speed) {
        const double *hk = lds + S * (int)(cl[(e << lg_tb) + t] & kLocalMask);
        xyk = reinterpret_cast<const double2 *>(hk)[0];
        load4(hk + 2, qk);
        speed = DELTA ? hk[7] : 1.0;
      };
      const int last = deg - 1;
      double2 xyk = {0.0, 0.0};
      double qk[4] = {0.0, 0.0, 0.0, 0.0};
      double speed = 1.0;
      if (last >= 0) fetch(0, xyk, qk, speed);
      for (int e = 0; e <= last; e++) {
        double2 xy_n;
        double q_n[4], speed_n;
        fetch(min(e + 1, last), xy_n, q_n, speed_n);
        for (int r = 0; r < 4; r++) {
          qmax[r] = fmax(qmax[r], qk[r]);
          qmin[r] = fmin(qmin[r], qk[r]);
        }
        q_deriv_term(xy.x, xy.y, qi, xyk, qk, sum_delx_sqr, sum_dely_sqr, sum_delx_dely, sdx,
                     sdy);
        if (DELTA) {
          const double delta_t = delta_step(xy, xyk, speed, cfl);
          if (min_delt > delta_t) min_delt = delta_t;
        }
        xyk = xy_n;
        for (int r = 0; r < 4; r++) qk[r] = q_n[r];
        speed = speed_n;
      }
      q_deriv_finish(sum_delx_sqr, sum_dely_sqr, sum_delx_dely, sdx, sdy, qmax, qmin,
                     dq_out + (size_t)DQ_STRIDE * i, qm_out + (size_t)QM_STRIDE * i,
                     qs_out ? qs_out + 8 * (size_t)i : nullptr);
      if (DELTA) delta[i] = min_delt;
    }
    if (vn >= nblk) break;
    vb = vn;
    vn = vnn;
    b = xcd_map(vb, nblk);
    q = qn;
    qn = qnn;
  }
}

// ------------------------------------------------------------ flux residual
// flux_residual_mod.h:10-89 with the ten *_dG* sub-stencil routines of
// interior_fluxes_mod.h, wall_fluxes_mod.h and outer_fluxes_mod.h fused into
// one pass over the neighbour list.
//  - The limited q~ at both ends and their primitive states depend only on
//    the (point, neighbour) pair, so they are computed once per neighbour
//    and feed the x- and y-direction fluxes of the (usually two)
//    sub-stencils the neighbour belongs to.
//  - Each sub-stencil's result G = (Sx*syy - Sy*sxy)/det (x-form,
//    interior_fluxes_mod.h:104-113) or (Sy*sxx - Sx*sxy)/det (y-form,
//    :324-328) is linear in the per-neighbour flux differences, and the
//    geometry sums are constant: with the precomputed coefficients
//    (a, b) = (syy, -sxy)/det or (sxx, -sxy)/det each neighbour adds
//    dG * (a*dels_w + b*deln_w) straight into the point's flux residual.
// coef per point: x+ (a,b), x- (a,b), y+ (a,b), y- (a,b).
constexpr int COEF_STRIDE = 8;

// The flux's LDS tile holds, per tile entry, the point's dq (8 doubles) and
// qmax/qmin (8 doubles) as eight 16-byte chunks; its 64-byte base record
// (x, y, q, eps2) is read through L2.  The tile then takes 128 bytes per
// entry, so two blocks of a 605-entry tile share a CU's LDS and the kernel
// keeps the two waves per SIMD its registers allow, while the bytes it
// gathers through L2 per neighbour fall from 192 to 64.  The eight chunks
// of an entry are consecutive, so the reads of one entry are eight
// ds_read_b128 at immediate offsets from one address; an XOR swizzle of
// the chunk positions, which spread the lanes' reads over more LDS banks,
// cost more in address arithmetic (16 instructions per neighbour) than the
// bank conflicts it saved (flux 1.093 -> 1.110 ms at 2.5M points): the
// flux keeps its LDS busy less than a fifth of the time.
constexpr int FLUX_TILE_CHUNKS = 8;

__device__ __forceinline__ void load_flux_tile(double2 *lds, const int32_t *__restrict__ tile_off,
                                               const int32_t *__restrict__ tile_idx, int b,
                                               const double *__restrict__ dq,
                                               const double *__restrict__ qm) {
  constexpr int U = 4;
  const int t0 = tile_off[b], total = (tile_off[b + 1] - t0) * FLUX_TILE_CHUNKS;
  for (int base = threadIdx.x; base < total; base += U * blockDim.x) {
    double2 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      int t = base + u * blockDim.x;
      t = t < total ? t : total - 1;
      const int r = t >> 3, ch = t & 7;
      const size_t g = (size_t)tile_idx[t0 + r];
      const double *src = ch < 4 ? dq + DQ_STRIDE * g : qm + QM_STRIDE * g;
      v[u] = reinterpret_cast<const double2 *>(src)[ch & 3];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int t = base + u * blockDim.x;
      if (t < total) {
        const int r = t >> 3, ch = t & 7;
        // dq halved (exact): the flux only uses d/2 = (dx . dq)/2
        const double sc = ch < 4 ? 0.5 : 1.0;
        lds[FLUX_TILE_CHUNKS * r + ch] = {v[u].x * sc, v[u].y * sc};
      }
    }
  }
}

// A point's base record (x, y, q[4], eps2, speed) as four 16-byte loads.
struct BaseRec {
  double2 xy, q01, q23, eps_speed;
};

__device__ __forceinline__ BaseRec load_base(const double *__restrict__ base, uint32_t k) {
  const double2 *p = reinterpret_cast<const double2 *>(base + (size_t)B_STRIDE * k);
  return {p[0], p[1], p[2], p[3]};
}

// entry r's dq/dx / 2, dq/dy / 2, qmax, qmin from the flux tile
__device__ __forceinline__ void flux_tile_entry(const double2 *lds, int r, double dx[4],
                                                double dy[4], double qmx[4], double qmn[4]) {
  double2 c[FLUX_TILE_CHUNKS];
#pragma unroll
  for (int j = 0; j < FLUX_TILE_CHUNKS; j++) c[j] = lds[FLUX_TILE_CHUNKS * r + j];
  dx[0] = c[0].x; dx[1] = c[0].y; dx[2] = c[1].x; dx[3] = c[1].y;
  dy[0] = c[2].x; dy[1] = c[2].y; dy[2] = c[3].x; dy[3] = c[3].y;
  qmx[0] = c[4].x; qmx[1] = c[4].y; qmx[2] = c[5].x; qmx[3] = c[5].y;
  qmn[0] = c[6].x; qmn[1] = c[6].y; qmn[2] = c[7].x; qmn[3] = c[7].y;
}

template <bool TILED, bool POW>
__global__ __launch_bounds__(256) void k_flux(
    int n, const int32_t *__restrict__ off, const int32_t *__restrict__ slice_off,
    const uint32_t *__restrict__ conn,
    const uint16_t *__restrict__ conn16, const int32_t *__restrict__ tile_off,
    const int32_t *__restrict__ tile_idx, int cap, const double *__restrict__ base,
    const double *__restrict__ dq, const double *__restrict__ qm,
    const double *__restrict__ coef, const double *__restrict__ nxa,
    const double *__restrict__ nya, const uint8_t *__restrict__ flag,
    const double *__restrict__ delta, double power, double *__restrict__ fres) {
  // Tiled, a launch takes the blocks whose tile has at most cap entries;
  // untiled with cap > 0, the others (their neighbours through L2); untiled
  // with cap = 0, every block.
  extern __shared__ double2 lds2[];
  const int b = xcd_block();
  if (TILED || cap > 0) {
    const int tn = tile_off[b + 1] - tile_off[b];
    if (TILED ? tn > cap : tn <= cap) return;
  }
  if (TILED) {
    load_flux_tile(lds2, tile_off, tile_idx, b, dq, qm);
    __syncthreads();
  }
  const int i = b * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *hi = base + (size_t)B_STRIDE * i;
  const double2 xy = reinterpret_cast<const double2 *>(hi)[0];
  const double x_i = xy.x, y_i = xy.y;
  // dxi, dyi, dxk, dyk: the derivatives halved (exact; the tile holds them
  // so); dmxi, dmni: the point's qmax - q and qmin - q, the candidates for
  // its limiters' del_pos, formed once (the reference forms the same
  // differences per pair, limiters_mod.h:72-78)
  double qi[4], dxi[4], dyi[4], dmxi[4], dmni[4], cf[8];
  load4(hi + B_Q, qi);
  if (TILED) {
    flux_tile_entry(lds2, threadIdx.x, dxi, dyi, dmxi, dmni);
    for (int r = 0; r < 4; r++) {
      dmxi[r] -= qi[r];
      dmni[r] -= qi[r];
    }
  } else {
    const double *di = dq + (size_t)DQ_STRIDE * i;
    const double *mi = qm + (size_t)QM_STRIDE * i;
    load4(di, dxi);
    load4(di + 4, dyi);
    load4(mi, dmxi);
    load4(mi + 4, dmni);
    for (int r = 0; r < 4; r++) {
      dxi[r] *= 0.5;
      dyi[r] *= 0.5;
      dmxi[r] -= qi[r];
      dmni[r] -= qi[r];
    }
  }
  load4(coef + (size_t)COEF_STRIDE * i, cf);
  load4(coef + (size_t)COEF_STRIDE * i + 4, cf + 4);
  const double eps2_i = hi[B_EPS2];
  const double nx = nxa[i], ny = nya[i];
  const double tx = ny, ty = -nx;
  const int type = flag[i];
  // wall: quadrant fluxes with A2neg and Gyn; outer: A2pos and Gyp
  const double s2b = type == 0 ? -1.0 : 1.0;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  // Software pipeline over the neighbour list: while neighbour e is
  // computed, the base record of e + 1 and the conn entries of e + 2 are in
  // flight (indices clamped to the last entry), so their L2 latency hides
  // behind a whole neighbour's arithmetic.
  const Slice L = slice_of(slice_off, conn, conn16, b);
  const int e0 = 0, last = off[i + 1] - off[i] - 1;
  uint32_t w_c = 0, w_n = 0;
  uint16_t l_c = 0, l_n = 0;
  BaseRec r_c = {};
  if (e0 <= last) {
    w_c = L.w32(e0);
    w_n = L.w32(min(e0 + 1, last));
    if (TILED) {
      l_c = L.w16(e0);
      l_n = L.w16(min(e0 + 1, last));
    }
    r_c = load_base(base, w_c & kIndexMask);
  }
  for (int e = e0; e <= last; e++) {
    const BaseRec r_n = load_base(base, w_n & kIndexMask);
    const int enn = min(e + 2, last);
    const uint32_t w_nn = L.w32(enn);
    const uint16_t l_nn = TILED ? L.w16(enn) : 0;
    const uint32_t m = w_c >> kMaskShift;
    const int k = (int)(w_c & kIndexMask);
    const double2 xyk = r_c.xy;
    double qk[4] = {r_c.q01.x, r_c.q01.y, r_c.q23.x, r_c.q23.y};
    double dxk[4], dyk[4], qmxk[4], qmnk[4];
    if (TILED) {
      flux_tile_entry(lds2, l_c & kLocalMask, dxk, dyk, qmxk, qmnk);
    } else {
      const double *dk = dq + (size_t)DQ_STRIDE * k;
      const double *mk = qm + (size_t)QM_STRIDE * k;
      load4(dk, dxk);
      load4(dk + 4, dyk);
      load4(mk, qmxk);
      load4(mk + 4, qmnk);
      for (int r = 0; r < 4; r++) {
        dxk[r] *= 0.5;
        dyk[r] *= 0.5;
      }
    }
    const double eps2_k = r_c.eps_speed.x;
    const double delx = xyk.x - x_i, dely = xyk.y - y_i;
    const double dels = delx * tx + dely * ty;
    const double deln = delx * nx + dely * ny;
    double dsw = dels, dnw = deln;
    if (POW) {
      const double weights = pow(sqrt(dels * dels + deln * deln), power);
      dsw = dels * weights;
      dnw = deln * weights;
    }
    // limited q~ at both ends (interior_fluxes_mod.h:72-83) from the halved
    // derivatives: d/2 = dx dq/dx/2 + dy dq/dy/2, and q - 0.5 phi d is
    // q - phi (d/2)
    double qti[4], qtk[4];
    for (int r = 0; r < 4; r++) {
      const double hi_r = fma(delx, dxi[r], dely * dyi[r]);
      const double hk_r = fma(delx, dxk[r], dely * dyk[r]);
      // del_neg = q~ - q as the reference rounds it: (q - d/2) - q
      const double phi_i = venkat_phi_delta((qi[r] - hi_r) - qi[r], dmxi[r], dmni[r], eps2_i);
      const double phi_k = venkat_phi(qk[r] - hk_r, qk[r], qmxk[r], qmnk[r], eps2_k);
      qti[r] = qi[r] - phi_i * hi_r;
      qtk[r] = qk[r] - phi_k * hk_r;
    }
    const EndTerms<double> ei = qtilde_end_terms(qti, nx, ny);
    const EndTerms<double> ek = qtilde_end_terms(qtk, nx, ny);
    double Gi[4], Gk[4];
    // x direction: x+ (dels <= 0) or x- (dels >= 0)
    const bool xp = (m & XPOS) != 0;
    const double sx = xp ? 1.0 : -1.0;
    if (type == 1) {
      split_x(ei, sx, Gi);
      split_x(ek, sx, Gk);
    } else {
      quad_x(ei, sx, s2b, Gi);
      quad_x(ek, sx, s2b, Gk);
    }
    const double wx = xp ? cf[0] * dsw + cf[1] * dnw : cf[2] * dsw + cf[3] * dnw;
    for (int r = 0; r < 4; r++) acc[r] += diff_unfused(Gk[r], Gi[r]) * wx;
    if (xp && (m & XNEG)) {  // dels == 0: the neighbour is in both lists
      if (type == 1) {
        split_x(ei, -1.0, Gi);
        split_x(ek, -1.0, Gk);
      } else {
        quad_x(ei, -1.0, s2b, Gi);
        quad_x(ek, -1.0, s2b, Gk);
      }
      const double wn = cf[2] * dsw + cf[3] * dnw;
      for (int r = 0; r < 4; r++) acc[r] += diff_unfused(Gk[r], Gi[r]) * wn;
    }
    // y direction: interior y+ / y-, wall y- (all), outer y+ (all)
    const bool yp = type == 1 ? (m & YPOS) != 0 : type == 2;
    const double sy = yp ? 1.0 : -1.0;
    split_y(ei, sy, Gi);
    split_y(ek, sy, Gk);
    const double wy = yp ? cf[4] * dnw + cf[5] * dsw : cf[6] * dnw + cf[7] * dsw;
    for (int r = 0; r < 4; r++) acc[r] += diff_unfused(Gk[r], Gi[r]) * wy;
    if (type == 1 && yp && (m & YNEG)) {  // deln == 0
      split_y(ei, -1.0, Gi);
      split_y(ek, -1.0, Gk);
      const double wn = cf[6] * dnw + cf[7] * dsw;
      for (int r = 0; r < 4; r++) acc[r] += diff_unfused(Gk[r], Gi[r]) * wn;
    }
    w_c = w_n;
    w_n = w_nn;
    l_c = l_n;
    l_n = l_nn;
    r_c = r_n;
  }
  // flux_residual_mod.h:29 (wall: 2 delta), :45, :85
  const double d = type == 0 ? 2.0 * delta[i] : delta[i];
  for (int j = 0; j < 4; j++) fres[(size_t)j * n + i] = d * acc[j];
}

// ------------------------------------------------------------ state update
// The state update's arithmetic is not contracted into FMAs: left to the
// compiler, which products it fuses could differ between the kernel's
// points-per-thread instances, and their results must agree to the bit.
// state_update_mod.h:18-33
__device__ __forceinline__ void primitive_to_conserved(const double p[4], double nx, double ny,
                                                       double U[4]) {
#pragma clang fp contract(off)
  const double rho = p[0];
  U[0] = rho;
  const double temp1 = rho * p[1];
  const double temp2 = rho * p[2];
  U[3] = 2.5 * p[3] + 0.5 * (temp1 * temp1 + temp2 * temp2) / rho;
  U[1] = temp1 * ny - temp2 * nx;
  U[2] = temp1 * nx + temp2 * ny;
}

// state_update_mod.h:35-90
__device__ __forceinline__ void conserved_vector_Ubar(const double p[4], double nx, double ny,
                                                      const FlowConsts &fc, double Ubar[4]) {
#pragma clang fp contract(off)
  const double rho_inf = fc.rho_inf, pr_inf = fc.pr_inf, gamma_new = 1.4;
  const double tx = ny, ty = -nx;
  const double u1_inf_rot = fc.u1_inf * tx + fc.u2_inf * ty;
  const double u2_inf_rot = fc.u1_inf * nx + fc.u2_inf * ny;
  double temp1 = (u1_inf_rot * u1_inf_rot + u2_inf_rot * u2_inf_rot);
  const double e_inf = pr_inf / (rho_inf * (gamma_new - 1.0)) + 0.5 * (temp1);
  double beta = (0.5 * rho_inf) / pr_inf;
  double S2 = u2_inf_rot * sqrt(beta);
  const double B2_inf = exp(-S2 * S2) / (2.0 * sqrt(kPi * beta));
  const double A2n_inf = 0.5 * (1.0 - erf(S2));
  const double rho = p[0], u1 = p[1], u2 = p[2], pr = p[3];
  const double u1_rot = u1 * tx + u2 * ty;
  const double u2_rot = u1 * nx + u2 * ny;
  temp1 = (u1_rot * u1_rot + u2_rot * u2_rot);
  const double e = pr / (rho * (gamma_new - 1.0)) + 0.5 * (temp1);
  beta = (rho) / (2.0 * pr);
  S2 = u2_rot * sqrt(beta);
  const double B2 = exp(-S2 * S2) / (2.0 * sqrt(kPi * beta));
  const double A2p = 0.5 * (1.0 + erf(S2));
  Ubar[0] = (rho_inf * A2n_inf) + (rho * A2p);
  Ubar[1] = (rho_inf * u1_inf_rot * A2n_inf) + (rho * u1_rot * A2p);
  temp1 = rho_inf * (u2_inf_rot * A2n_inf - B2_inf);
  double temp2 = rho * (u2_rot * A2p + B2);
  Ubar[2] = temp1 + temp2;
  temp1 = (rho_inf * A2n_inf * e_inf - 0.5 * rho_inf * u2_inf_rot * B2_inf);
  temp2 = (rho * A2p * e + 0.5 * rho * u2_rot * B2);
  Ubar[3] = temp1 + temp2;
}

// (sum, max, argmax) of the block, fixed tree order; argmax ties go to the
// smaller caller index, as the reference's strict '>' over points in order.
__device__ void block_reduce3(double &s, double &mx, int &arg) {
  __shared__ double ss[256], sm[256];
  __shared__ int sa[256];
  const int t = threadIdx.x;
  ss[t] = s;
  sm[t] = mx;
  sa[t] = arg;
  __syncthreads();
  for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
    if (t < w) {
      ss[t] = ss[t] + ss[t + w];
      const double a = sm[t], b = sm[t + w];
      const int ia = sa[t], ib = sa[t + w];
      if (b > a || (b == a && ib < ia)) {
        sm[t] = b;
        sa[t] = ib;
      }
    }
    __syncthreads();
  }
  s = ss[0];
  mx = sm[0];
  arg = sa[0];
}

// One point's inputs to the state update, loaded before any point of the
// thread is computed (so that a thread's PPT points have their loads in
// flight together).
struct StateIn {
  double p1[4], p2[4], fr[4], nx, ny;
  int type;
};

__device__ __forceinline__ void state_load(int n, int i, const double *__restrict__ prim,
                                           const double *__restrict__ prim_old,
                                           const double *__restrict__ fres,
                                           const double *__restrict__ nxa,
                                           const double *__restrict__ nya,
                                           const uint8_t *__restrict__ flag, StateIn &in) {
  for (int r = 0; r < 4; r++) {
    in.p1[r] = prim[(size_t)r * n + i];
    in.p2[r] = prim_old[(size_t)r * n + i];
    in.fr[r] = fres[(size_t)r * n + i];
  }
  in.nx = nxa[i];
  in.ny = nya[i];
  in.type = flag[i];
}

// state_update_mod.h:92-283 for point i.  It also writes, from the new
// state, the q-variables of the next stage (q_variables_mod.h:9-34) and,
// when `speed` is set, the denominator func_delta divides by
// (state_update_mod.h:483-485): both are pointwise functions of the state,
// so computing them here gives the values the separate kernels would.
// Returns the point's residual term (0 for outer points) and sets arg to
// its caller index (outer points: left unchanged).
__device__ __forceinline__ double state_point(int n, int i, int rk, const StateIn &in,
                                              const FlowConsts &fc, double *__restrict__ prim,
                                              const int32_t *__restrict__ perm,
                                              double *__restrict__ base,
                                              float4 *__restrict__ q32, int speed, int &arg) {
#pragma clang fp contract(off)
  const double obt = 1.0 / 3.0, tbt = 2.0 / 3.0;
  const double nx = in.nx, ny = in.ny;
  const int type = in.type;
  double U[4], U_old[4];
  double res_sqr = 0.0;
  if (type != 2) {
    primitive_to_conserved(in.p1, nx, ny, U);
    primitive_to_conserved(in.p2, nx, ny, U_old);
  } else {
    conserved_vector_Ubar(in.p1, nx, ny, fc, U);
    conserved_vector_Ubar(in.p2, nx, ny, fc, U_old);
  }
  const double temp0 = U[0];
  if (rk != 3) {
    for (int j = 0; j < 4; j++) U[j] = U[j] - (0.5 * in.fr[j]);
  } else {
    for (int j = 0; j < 4; j++) U[j] = tbt * U_old[j] + obt * (U[j] - 0.5 * in.fr[j]);
  }
  if (type == 0) U[2] = 0.;
  const double U2_rot = U[1], U3_rot = U[2];
  U[1] = U2_rot * ny + U3_rot * nx;
  U[2] = U3_rot * ny - U2_rot * nx;
  if (type != 2) {
    res_sqr = (U[0] - temp0) * (U[0] - temp0);
    arg = perm[i];
  }
  const double temp = 1.0 / U[0];
  const double rho = U[0], u1 = U[1] * temp, u2 = U[2] * temp;
  const double pr = 0.4 * U[3] - (0.2 * temp) * (U[1] * U[1] + U[2] * U[2]);
  prim[i] = rho;
  prim[(size_t)n + i] = u1;
  prim[(size_t)2 * n + i] = u2;
  prim[(size_t)3 * n + i] = pr;
  if (base || q32) {
    double q[4];
    q_of_state(rho, u1, u2, pr, q);
    if (q32)  // mixed form: q rounded once, from the FP64 state
      q32[i] = make_float4((float)q[0], (float)q[1], (float)q[2], (float)q[3]);
    else
      store4(base + (size_t)B_STRIDE * i + B_Q, q);
  }
  if (base && speed) base[(size_t)B_STRIDE * i + B_SPEED] = speed_of_state(rho, u1, u2, pr);
  return res_sqr;
}

// The state update over 256-point blocks, PPT points per thread (256 / PPT
// threads a block; thread t takes points t, t + 256/PPT, ...).  Each point's
// residual term goes to its slot in the block and the slots are reduced in
// one fixed tree, so the partial sums do not depend on PPT: every choice
// gives the same bits.  Ties of the maximum go to the smaller caller index,
// as the reference's strict '>' over points in order.
template <int PPT>
__global__ __launch_bounds__(256) void k_state_update(int n, int rk, double *__restrict__ prim,
                                                      const double *__restrict__ prim_old,
                                                      const double *__restrict__ fres,
                                                      const double *__restrict__ nxa,
                                                      const double *__restrict__ nya,
                                                      const uint8_t *__restrict__ flag,
                                                      const int32_t *__restrict__ perm,
                                                      const FlowConsts *__restrict__ fcs,
                                                      const uint16_t *__restrict__ block_case,
                                                      double *__restrict__ psum,
                                                      double *__restrict__ pmax,
                                                      int32_t *__restrict__ parg,
                                                      double *__restrict__ base,
                                                      float4 *__restrict__ q32, int speed) {
  constexpr int T = 256 / PPT;
  __shared__ double ss[256], sm[256];
  __shared__ int sa[256];
  const int t = threadIdx.x;
  const int i0 = blockIdx.x * 256 + t;
  // a batch pads every case to whole 256-point blocks: one case per block
  const FlowConsts fc = fcs[block_case ? block_case[blockIdx.x] : 0];
  StateIn in[PPT];
#pragma unroll
  for (int j = 0; j < PPT; j++) {
    const int i = i0 + j * T;
    if (i < n) state_load(n, i, prim, prim_old, fres, nxa, nya, flag, in[j]);
  }
#pragma unroll
  for (int j = 0; j < PPT; j++) {
    const int i = i0 + j * T;
    double res_sqr = 0.0;
    int arg = 0x7fffffff;
    if (i < n) res_sqr = state_point(n, i, rk, in[j], fc, prim, perm, base, q32, speed, arg);
    ss[t + j * T] = res_sqr;
    sm[t + j * T] = arg == 0x7fffffff ? -1.0 : res_sqr;
    sa[t + j * T] = arg;
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    for (int u = t; u < w; u += T) {
      ss[u] = ss[u] + ss[u + w];
      const double a = sm[u], b = sm[u + w];
      const int ia = sa[u], ib = sa[u + w];
      if (b > a || (b == a && ib < ia)) {
        sm[u] = b;
        sa[u] = ib;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    psum[blockIdx.x] = ss[0];
    pmax[blockIdx.x] = sm[0];
    parg[blockIdx.x] = sa[0];
  }
}

// One block per case: reduces the case's partials [seg[c], seg[c+1]) in a
// fixed order into out[3c..3c+2] = (sum, max, argmax as a double) and, when
// hist is given, hist[c*hist_stride + slot] = sum, with slot = *slot_ptr
// when slot_ptr is given (a graph replays the same launch every iteration).
__global__ __launch_bounds__(256) void k_reduce(const int32_t *__restrict__ seg, int nb,
                                                const double *__restrict__ psum,
                                                const double *__restrict__ pmax,
                                                const int32_t *__restrict__ parg,
                                                double *__restrict__ out, double *__restrict__ hist,
                                                int hist_stride, int slot,
                                                const int32_t *__restrict__ slot_ptr) {
  const int c = blockIdx.x;
  const int b0 = seg ? seg[c] : 0, b1 = seg ? seg[c + 1] : nb;
  double s = 0.0, mx = -1.0;
  int arg = 0x7fffffff;
  // eight partials a thread loaded at once (indices clamped to the last),
  // then taken in the same order as one at a time
  constexpr int U = 8;
  const int stride = (int)blockDim.x;
  for (int b = b0 + threadIdx.x; b < b1; b += U * stride) {
    double ps[U], pm[U];
    int pa[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int bu = min(b + u * stride, b1 - 1);
      ps[u] = psum[bu];
      pm[u] = pmax[bu];
      pa[u] = parg[bu];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (b + u * stride >= b1) break;
      s = s + ps[u];
      if (pm[u] > mx || (pm[u] == mx && pa[u] < arg)) {
        mx = pm[u];
        arg = pa[u];
      }
    }
  }
  block_reduce3(s, mx, arg);
  if (threadIdx.x == 0) {
    out[3 * c] = s;
    out[3 * c + 1] = mx;
    out[3 * c + 2] = (double)arg;
    if (hist) hist[(size_t)c * hist_stride + (slot_ptr ? *slot_ptr : slot)] = s;
  }
}

__global__ void k_tick(int32_t *counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += 1;
}

__global__ void k_set_int(int32_t *p, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

// dst <- src (graph-safe replacement of a device-to-device memcpy)
__global__ __launch_bounds__(256) void k_copy(size_t n, const double2 *__restrict__ src,
                                              double2 *__restrict__ dst) {
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (size_t)gridDim.x * blockDim.x)
    dst[j] = src[j];
}

// ------------------------------------------------------------ forces
// compute_force_coeffs_mod.h:25-60 for shape 1: H, V and the pitching
// moment about (0.25, 0), reduced in a fixed order by one block.
__global__ __launch_bounds__(256) void k_forces(const int32_t *__restrict__ soff,
                                                const int32_t *__restrict__ mi,
                                                const int32_t *__restrict__ li,
                                                const int32_t *__restrict__ ri,
                                                const int32_t *__restrict__ shape,
                                                const double *__restrict__ base,
                                                const double *__restrict__ nxa,
                                                const double *__restrict__ nya,
                                                const double *__restrict__ prim, int n,
                                                double pr_inf,
                                                const double *__restrict__ qdyns,
                                                double *__restrict__ out) {
  // one block per case: its shape points are [soff[c], soff[c+1])
  __shared__ double sh[3][256];
  const int c = blockIdx.x;
  const double qdyn = qdyns[c];
  double H = 0.0, V = 0.0, P = 0.0;
  for (int j = soff[c] + threadIdx.x; j < soff[c + 1]; j += blockDim.x) {
    if (shape[j] != 1) continue;
    const int m = mi[j], l = li[j], r = ri[j];
    const double2 L = reinterpret_cast<const double2 *>(base + (size_t)B_STRIDE * l)[0];
    const double2 Mp = reinterpret_cast<const double2 *>(base + (size_t)B_STRIDE * m)[0];
    const double2 Rp = reinterpret_cast<const double2 *>(base + (size_t)B_STRIDE * r)[0];
    double ds1 = (Mp.x - L.x) * (Mp.x - L.x) + (Mp.y - L.y) * (Mp.y - L.y);
    ds1 = sqrt(ds1);
    double ds2 = (Rp.x - Mp.x) * (Rp.x - Mp.x) + (Rp.y - Mp.y) * (Rp.y - Mp.y);
    ds2 = sqrt(ds2);
    const double ds = 0.5 * (ds1 + ds2);
    const double nx = nxa[m], ny = nya[m];
    double cp = prim[(size_t)3 * n + m] - pr_inf;
    cp = -cp / qdyn;
    H = H + cp * nx * ds;
    V = V + cp * ny * ds;
    P = P + (-cp * ny * ds * (Mp.x - 0.25) + cp * nx * ds * (Mp.y));
  }
  const int t = threadIdx.x;
  sh[0][t] = H;
  sh[1][t] = V;
  sh[2][t] = P;
  __syncthreads();
  for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
    if (t < w)
      for (int c = 0; c < 3; c++) sh[c][t] = sh[c][t] + sh[c][t + w];
    __syncthreads();
  }
  if (t < 3) out[3 * c + t] = sh[t][0];
}

// ------------------------------------------------------------ KAT
__global__ void k_kat_fluxes(int n, const double *__restrict__ in, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *v = in + 6 * (size_t)i;
  const EndTerms<double> e = end_terms(v[0], v[1], v[2], v[3], v[4], v[5]);
  double G[8][4];
  split_x(e, 1.0, G[0]);
  split_x(e, -1.0, G[1]);
  split_y(e, 1.0, G[2]);
  split_y(e, -1.0, G[3]);
  quad_x(e, -1.0, -1.0, G[4]);
  quad_x(e, 1.0, -1.0, G[5]);
  quad_x(e, 1.0, 1.0, G[6]);
  quad_x(e, -1.0, 1.0, G[7]);
  for (int f = 0; f < 8; f++)
    for (int r = 0; r < 4; r++) out[32 * (size_t)i + 4 * f + r] = G[f][r];
}

__global__ void k_kat_qtilde(int n, const double *__restrict__ in, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double qt[4];
  for (int r = 0; r < 4; r++) qt[r] = in[4 * (size_t)i + r];
  // the form the flux kernels use
  double u1, u2, rho, sb, temp;
  qtilde_state(qt, u1, u2, rho, sb, temp);
  out[4 * (size_t)i + 0] = u1;
  out[4 * (size_t)i + 1] = u2;
  out[4 * (size_t)i + 2] = rho;
  out[4 * (size_t)i + 3] = rho * temp;
}

}  // namespace

int blocks_for(int n, int block) { return (n + block - 1) / block; }

size_t lds_bytes_qderiv(int tile_max) { return sizeof(double) * 6 * (size_t)tile_max; }
size_t lds_bytes_qinner(int tile_max) { return sizeof(double) * 10 * (size_t)tile_max; }
size_t lds_bytes_delta(int tile_max) { return sizeof(double) * 4 * (size_t)tile_max; }
size_t lds_bytes_flux(int tile_max) {
  return sizeof(double2) * FLUX_TILE_CHUNKS * (size_t)tile_max;
}

int flux_tile_cap(int block, int tile_max, size_t lds_limit, size_t lds_per_cu) {
  int nb_regs = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_regs, k_flux<true, false>, block, 0) !=
          hipSuccess ||
      nb_regs <= 0)
    return 0;
  // per block: the CU's LDS shared by nb_regs blocks, less 2 KB of
  // allocation granularity, and at most what one block may allocate
  size_t per_block = lds_per_cu / (size_t)nb_regs;
  per_block = per_block > 2048 ? per_block - 2048 : 0;
  if (per_block > lds_limit) per_block = lds_limit;
  const size_t entry = lds_bytes_flux(1);
  int cap = (int)(per_block / entry);
  if (cap > tile_max) cap = tile_max;
  return cap >= block ? cap : 0;
}

unsigned prepare_tiled(unsigned tiles, int tile_max, int flux_cap, size_t lds_limit) {
  struct K { unsigned bit; const void *f; size_t bytes; };
  const K ks[] = {
      {TILE_QDERIV, (const void *)k_q_derivatives<true, false>, lds_bytes_qderiv(tile_max)},
      {TILE_QDERIV, (const void *)k_q_derivatives<true, true>, lds_bytes_qderiv(tile_max)},
      {TILE_QDERIV, (const void *)k_q_derivatives<true, false, true>, lds_bytes_qderiv(tile_max)},
      {TILE_QINNER, (const void *)k_q_inner<true, false>, lds_bytes_qinner(tile_max)},
      {TILE_QINNER, (const void *)k_q_inner<true, true>, lds_bytes_qinner(tile_max)},
      {TILE_DELTA, (const void *)k_func_delta<true>, lds_bytes_delta(tile_max)},
      {TILE_FLUX, (const void *)k_flux<true, false>, lds_bytes_flux(flux_cap)},
      {TILE_FLUX, (const void *)k_flux<true, true>, lds_bytes_flux(flux_cap)}};
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

// The pipelined inner pass's LDS: the tile, then the block's list slice
// (double offset lds_conn).
int prepare_qinner_pf(int tile_block, int tile_max, int slice_width, int nblk, int n_cu,
                      size_t lds_limit, int *grid) {
  *grid = 0;
  if (tile_block < 64 || (tile_block & (tile_block - 1)) || n_cu <= 0 || nblk <= 0) return 0;
  const int code = pf_layout(tile_block, tile_max, slice_width);
  const size_t bytes = pf_lds_bytes(tile_block, tile_max, slice_width, 10);
  const void *f = nullptr;
#define PF_F(e, nc) if (code == 10 * e + nc) f = (const void *)k_q_inner_pf<e, nc>;
  PF_LAYOUTS(PF_F)
#undef PF_F
  if (!f) return 0;
  *grid = pf_grid(f, tile_block, bytes, lds_limit, nblk, n_cu);
  return *grid ? code : 0;
}

int prepare_qderiv_pf(int tile_block, int tile_max, int slice_width, int nblk, int n_cu,
                      size_t lds_limit, bool with_delta, int *grid) {
  *grid = 0;
  if (tile_block < 64 || (tile_block & (tile_block - 1)) || n_cu <= 0 || nblk <= 0) return 0;
  const int code = pf_layout(tile_block, tile_max, slice_width);
  const size_t bytes = pf_lds_bytes(tile_block, tile_max, slice_width, with_delta ? 8 : 6);
  const void *f = nullptr;
#define PF_F(e, nc)                                                                   \
  if (code == 10 * e + nc)                                                            \
    f = with_delta ? (const void *)k_q_derivatives_pf<e, nc, true>                    \
                   : (const void *)k_q_derivatives_pf<e, nc, false>;
  PF_LAYOUTS(PF_F)
#undef PF_F
  if (!f) return 0;
  *grid = pf_grid(f, tile_block, bytes, lds_limit, nblk, n_cu);
  return *grid ? code : 0;
}

#define TILE_ARGS c.conn_off, c.slice_off, c.connS, c.conn16S, c.tile_off, c.tile_idx

void launch_q_variables(const DevCloud &c, const double *prim, double *base, hipStream_t s) {
  hipLaunchKernelGGL(k_q_variables, dim3(blocks_for(c.n, 256)), dim3(256), 0, s, c.n, prim,
                     base);
}

void launch_speed(const DevCloud &c, const double *prim, double *base, hipStream_t s) {
  hipLaunchKernelGGL(k_speed, dim3(blocks_for(c.n, 256)), dim3(256), 0, s, c.n, prim, base);
}

void launch_func_delta(const DevCloud &c, const double *base, double cfl, double *delta,
                       hipStream_t s) {
  dim3 g(blocks_for(c.n, c.tile_block)), b(c.tile_block);
  if (c.tiles & TILE_DELTA)
    hipLaunchKernelGGL(k_func_delta<true>, g, b, lds_bytes_delta(c.tile_max), s, c.n, TILE_ARGS,
                       base, cfl, delta);
  else
    hipLaunchKernelGGL(k_func_delta<false>, g, b, 0, s, c.n, TILE_ARGS, base, cfl, delta);
}

void launch_q_derivatives(const DevCloud &c, const double *base, double *dq, double *qm,
                          double *qs, double power, hipStream_t s) {
  dim3 g(blocks_for(c.n, c.tile_block)), b(c.tile_block);
  const bool t = (c.tiles & TILE_QDERIV) != 0;
  const size_t L = t ? lds_bytes_qderiv(c.tile_max) : 0;
#define QD_ARGS c.n, TILE_ARGS, base, dq, qm, qs, power
  if (t && power == 0.0 && c.qd_chunks) {
    const int nblk = blocks_for(c.n, c.tile_block);
    const size_t Lp = pf_lds_bytes(c.tile_block, c.tile_max, c.slice_width, 6);
    const int lc = 6 * c.tile_max;
#define PF_L(e, nc)                                                                        \
  if (c.qd_chunks == 10 * e + nc)                                                          \
    hipLaunchKernelGGL((k_q_derivatives_pf<e, nc, false>), dim3(c.qd_grid), b, Lp, s, c.n, \
                       nblk, lc, c.conn_off, c.slice_off, c.conn16S, c.tile_off, c.tile_idx, \
                       base, dq, qm, qs, 0.0, nullptr);
    PF_LAYOUTS(PF_L)
#undef PF_L
  } else if (t && power == 0.0)
    hipLaunchKernelGGL((k_q_derivatives<true, false>), g, b, L, s, QD_ARGS);
  else if (t)
    hipLaunchKernelGGL((k_q_derivatives<true, true>), g, b, L, s, QD_ARGS);
  else if (power == 0.0)
    hipLaunchKernelGGL((k_q_derivatives<false, false>), g, b, 0, s, QD_ARGS);
  else
    hipLaunchKernelGGL((k_q_derivatives<false, true>), g, b, 0, s, QD_ARGS);
#undef QD_ARGS
}

void launch_q_derivatives_delta(const DevCloud &c, const double *base, double *dq, double *qm,
                                double *qs, double cfl, double *delta, hipStream_t s) {
  const int nblk = blocks_for(c.n, c.tile_block);
  const size_t Lp = pf_lds_bytes(c.tile_block, c.tile_max, c.slice_width, 8);
  const int lc = 8 * c.tile_max;
  const dim3 g(c.qdd_grid), b(c.tile_block);
#define PF_L(e, nc)                                                                        \
  if (c.qdd_chunks == 10 * e + nc)                                                         \
    hipLaunchKernelGGL((k_q_derivatives_pf<e, nc, true>), g, b, Lp, s, c.n, nblk, lc,      \
                       c.conn_off, c.slice_off, c.conn16S, c.tile_off, c.tile_idx, base, dq, \
                       qm, qs, cfl, delta);
  PF_LAYOUTS(PF_L)
#undef PF_L
}

void launch_q_derivatives_arg(const DevCloud &c, const double *base, double *dq, double *qm,
                              double *qs, int32_t *arg, double *lsum, hipStream_t s) {
  dim3 g(blocks_for(c.n, c.tile_block)), b(c.tile_block);
  if (c.tiles & TILE_QDERIV)
    hipLaunchKernelGGL((k_q_derivatives<true, false, true>), g, b, lds_bytes_qderiv(c.tile_max),
                       s, c.n, TILE_ARGS, base, dq, qm, qs, 0.0, arg, lsum);
  else
    hipLaunchKernelGGL((k_q_derivatives<false, false, true>), g, b, 0, s, c.n, TILE_ARGS, base,
                       dq, qm, qs, 0.0, arg, lsum);
}

void launch_q_inner(const DevCloud &c, const double *qs, const double *dq_src, double *dq_dst,
                    double power, hipStream_t s) {
  dim3 g(blocks_for(c.n, c.tile_block)), b(c.tile_block);
  const bool t = (c.tiles & TILE_QINNER) != 0;
  const size_t L = t ? lds_bytes_qinner(c.tile_max) : 0;
#define QI_ARGS c.n, TILE_ARGS, c.xy, qs, dq_src, dq_dst, power
  if (t && power == 0.0 && c.qi_chunks) {
    const int nblk = blocks_for(c.n, c.tile_block);
    const size_t Lp = pf_lds_bytes(c.tile_block, c.tile_max, c.slice_width, 10);
    const int lc = 10 * c.tile_max;
#define PF_L(e, nc)                                                                          \
  if (c.qi_chunks == 10 * e + nc)                                                            \
    hipLaunchKernelGGL((k_q_inner_pf<e, nc>), dim3(c.qi_grid), b, Lp, s, c.n, nblk, lc,      \
                       c.conn_off, c.slice_off, c.conn16S, c.tile_off, c.tile_idx, c.xy, qs, \
                       dq_src, dq_dst);
    PF_LAYOUTS(PF_L)
#undef PF_L
  } else if (t && power == 0.0)
    hipLaunchKernelGGL((k_q_inner<true, false>), g, b, L, s, QI_ARGS);
  else if (t)
    hipLaunchKernelGGL((k_q_inner<true, true>), g, b, L, s, QI_ARGS);
  else if (power == 0.0)
    hipLaunchKernelGGL((k_q_inner<false, false>), g, b, 0, s, QI_ARGS);
  else
    hipLaunchKernelGGL((k_q_inner<false, true>), g, b, 0, s, QI_ARGS);
#undef QI_ARGS
}

void launch_flux(const DevCloud &c, const double *base, const double *dq, const double *qm,
                 const double *delta, double power, double *flux_res, hipStream_t s) {
  dim3 g(blocks_for(c.n, c.tile_block)), b(c.tile_block);
  const bool t = (c.tiles & TILE_FLUX) != 0;
  const size_t L = t ? lds_bytes_flux(c.flux_cap) : 0;
#define FLUX_ARGS \
  c.n, TILE_ARGS, cap, base, dq, qm, c.coef, c.nx, c.ny, c.flag, delta, power, flux_res
  // tiled: the blocks whose tile fits the cap, then (if any tile is larger)
  // the rest through L2
  const int cap = t ? c.flux_cap : 0;
  if (t && power == 0.0)
    hipLaunchKernelGGL((k_flux<true, false>), g, b, L, s, FLUX_ARGS);
  else if (t)
    hipLaunchKernelGGL((k_flux<true, true>), g, b, L, s, FLUX_ARGS);
  if (!t || c.tile_max > cap) {
    if (power == 0.0)
      hipLaunchKernelGGL((k_flux<false, false>), g, b, 0, s, FLUX_ARGS);
    else
      hipLaunchKernelGGL((k_flux<false, true>), g, b, 0, s, FLUX_ARGS);
  }
#undef FLUX_ARGS
}

void launch_state_update(const DevCloud &c, int rk, double *prim, const double *prim_old,
                         const double *flux_res, const FlowConsts *fcs,
                         const uint16_t *block_case, const Partials &p, double *base,
                         float *q32, bool speed, hipStream_t s) {
  // fixed 256-point blocks, whatever the points per thread: the partial
  // layout must not depend on tuning
#define SU_ARGS                                                                               \
  c.n, rk, prim, prim_old, flux_res, c.nx, c.ny, c.flag, c.perm, fcs, block_case, p.sum, p.max, \
      p.arg, base, reinterpret_cast<float4 *>(q32), speed ? 1 : 0
  if (c.su_ppt == 4)
    hipLaunchKernelGGL(k_state_update<4>, dim3(p.nblocks), dim3(64), 0, s, SU_ARGS);
  else if (c.su_ppt == 2)
    hipLaunchKernelGGL(k_state_update<2>, dim3(p.nblocks), dim3(128), 0, s, SU_ARGS);
  else
    hipLaunchKernelGGL(k_state_update<1>, dim3(p.nblocks), dim3(256), 0, s, SU_ARGS);
#undef SU_ARGS
}

void launch_reduce(const Partials &p, int n_cases, const int32_t *seg, double *out,
                   double *hist, int hist_stride, int slot, const int32_t *slot_ptr,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_reduce, dim3(n_cases), dim3(256), 0, s, seg, p.nblocks, p.sum, p.max,
                     p.arg, out, hist, hist_stride, slot, slot_ptr);
}

void launch_tick(int32_t *counter, hipStream_t s) {
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, s, counter);
}

void launch_set_int(int32_t *p, int32_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, p, v);
}

void launch_copy(const double *src, double *dst, size_t count, hipStream_t s) {
  const size_t n2 = count / 2;  // callers copy whole [4][n] double arrays
  const int blocks = (int)std::min<size_t>((n2 + 255) / 256, 8192);
  hipLaunchKernelGGL(k_copy, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, s, n2,
                     reinterpret_cast<const double2 *>(src), reinterpret_cast<double2 *>(dst));
}

void launch_forces(int n_cases, const int32_t *soff, const int32_t *m, const int32_t *l,
                   const int32_t *r, const int32_t *shape, const double *base, const double *nx,
                   const double *ny, const double *prim, int n, double pr_inf,
                   const double *qdyn, double *out3, hipStream_t s) {
  hipLaunchKernelGGL(k_forces, dim3(n_cases), dim3(256), 0, s, soff, m, l, r, shape, base, nx,
                     ny, prim, n, pr_inf, qdyn, out3);
}

void launch_kat_fluxes(int n, const double *in6, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_kat_fluxes, dim3(blocks_for(n, 128)), dim3(128), 0, s, n, in6, out);
}

void launch_kat_qtilde(int n, const double *qt4, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_kat_qtilde, dim3(blocks_for(n, 128)), dim3(128), 0, s, n, qt4, out);
}

}  // namespace lskum
