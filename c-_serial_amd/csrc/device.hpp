// device.hpp -- device helpers shared by the FP64 and mixed-precision kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace lskum {

// sub-stencil bits of a conn entry (x+, x-, y+, y-)
constexpr uint32_t XPOS = 1u, XNEG = 2u, YPOS = 4u, YNEG = 8u;

// blockIdx -> block of points such that each of the 8 XCDs (which receive
// workgroups round-robin) works on one contiguous range of points and keeps
// its neighbours' records in its own L2.
__device__ __forceinline__ int xcd_map(int v, int nb) {
  const int q = nb >> 3, r = nb & 7, xcd = v & 7, local = v >> 3;
  return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}
__device__ __forceinline__ int xcd_block() { return xcd_map(blockIdx.x, gridDim.x); }

__device__ __forceinline__ void load4(const double *p, double v[4]) {
  const double2 a = reinterpret_cast<const double2 *>(p)[0];
  const double2 b = reinterpret_cast<const double2 *>(p)[1];
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}

__device__ __forceinline__ void store4(double *p, const double v[4]) {
  reinterpret_cast<double2 *>(p)[0] = make_double2(v[0], v[1]);
  reinterpret_cast<double2 *>(p)[1] = make_double2(v[2], v[3]);
}

__device__ __forceinline__ void unpack4(const float4 a, float v[4]) {
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}

// a - b as one rounded subtraction that the compiler may not fuse with the
// product that produced a (a*b' - c in one FMA would turn the difference of
// two equal fluxes into the rounding error of one of them, and uniform flow
// would no longer give an exactly zero residual).  Needs the build's
// -ffp-contract=fast-honor-pragmas.
template <typename T>
__device__ __forceinline__ T diff_unfused(T a, T b) {
#pragma clang fp contract(off)
  return a - b;
}

// Neighbour of conn entry e: a device index (global form) or an index into
// the block's LDS tile (tiled form), with the entry's sub-stencil bits.
__device__ __forceinline__ int nbr_index(const uint32_t *conn, const uint16_t *conn16, int e,
                                         bool tiled, uint32_t &mask) {
  if (tiled) {
    const uint16_t w = conn16[e];
    mask = w >> kLocalShift;
    return w & kLocalMask;
  }
  const uint32_t w = conn[e];
  mask = w >> kMaskShift;
  return (int)(w & kIndexMask);
}

// ------------------------------------------------------------ tiles
// A block's LDS tile holds its own 256 points first (local index =
// threadIdx.x), then the other points its neighbour lists reach; tiled conn
// entries carry the local index.  The cooperative load happens before any
// thread leaves, so the barrier is reached by the whole block.
// U 16-byte gathers are in flight per thread before their LDS stores; a
// caller sizes U so that one round covers the tiles it expects (on the
// 2.5M-point cloud q_inner took 0.375 ms with 16, 0.40 with 4).
template <int STRIDE, int U = 16>
__device__ __forceinline__ void load_tile(double *lds, const int32_t *__restrict__ tile_off,
                                          const int32_t *__restrict__ tile_idx, int b,
                                          const double *__restrict__ src, int src_stride,
                                          int first_chunk, int nchunks, int dst_offset) {
  const int t0 = tile_off[b], total = (tile_off[b + 1] - t0) * nchunks;
  for (int base = threadIdx.x; base < total; base += U * blockDim.x) {
    double2 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      // every lane loads (past the end: the last chunk again), so the
      // array is fully written and stays in registers
      int t = base + u * blockDim.x;
      t = t < total ? t : total - 1;
      const int r = t / nchunks, ch = t - r * nchunks;
      v[u] = reinterpret_cast<const double2 *>(src + (size_t)src_stride *
                                                         tile_idx[t0 + r])[first_chunk + ch];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int t = base + u * blockDim.x;
      if (t < total) {
        const int r = t / nchunks, ch = t - r * nchunks;
        reinterpret_cast<double2 *>(lds + (size_t)STRIDE * r + dst_offset)[ch] = v[u];
      }
    }
  }
}

// A lane's neighbour list in its block's slice.  The forward stencil
// kernels read the lists from slices: the host lays out the lists of a
// stencil block's points entry by entry, entry e of the block's lane t at
// slice_off[block] + e * tile_block + t, so that the 64 lanes of a wave
// read 64 consecutive entries (one or two cache lines) per neighbour step
// where the CSR lists, 24 entries apart per lane, touched a cache line per
// lane or two.  Lists shorter than their block's longest are padded with
// their last entry.  Global form (device index | mask) and tiled form
// (local index | mask) have the same layout.
struct Slice {
  const uint32_t *c32;
  const uint16_t *c16;
  int stride;
  __device__ __forceinline__ uint32_t w32(int e) const { return c32[(size_t)e * stride]; }
  __device__ __forceinline__ uint32_t w16(int e) const {
    return (uint32_t)c16[(size_t)e * stride];
  }
  __device__ __forceinline__ uint32_t raw(int e, bool tiled) const {
    return tiled ? w16(e) : w32(e);
  }
};

// The slice of the calling lane of stencil block b.
__device__ __forceinline__ Slice slice_of(const int32_t *__restrict__ slice_off,
                                          const uint32_t *conn, const uint16_t *conn16, int b) {
  const size_t s = (size_t)slice_off[b] + threadIdx.x;
  return {conn + s, conn16 + s, (int)blockDim.x};
}

// The raw conn entry e (tiled: the 16-bit local form) and its parts.  The
// stencil kernels load the entry of neighbour e + 2 while neighbour e is
// processed and split it only when it is used, so no instruction waits on
// the load before then.
__device__ __forceinline__ uint32_t conn_raw(const uint32_t *conn, const uint16_t *conn16, int e,
                                             bool tiled) {
  return tiled ? (uint32_t)conn16[e] : conn[e];
}

__device__ __forceinline__ int conn_index(uint32_t w, bool tiled) {
  return tiled ? (int)(w & kLocalMask) : (int)(w & kIndexMask);
}

__device__ __forceinline__ uint32_t conn_mask(uint32_t w, bool tiled) {
  return tiled ? w >> kLocalShift : w >> kMaskShift;
}

// The next DEPTH raw conn entries of one lane's list, in registers.  The
// stencil loops fetch neighbour e + 1's record while e is processed; pop
// hands them entry e + 1 and issues the load of entry e + 1 + DEPTH
// (indices clamped to the last entry), so each conn load has DEPTH
// neighbours' work to arrive in.  The loops unroll by DEPTH, which turns
// the shift of the ring into a renaming of registers (a register copy
// would wait for its load).  With the conn arrays far larger than L2,
// one neighbour's work is not enough to cover the load on the kernels that
// run few waves per SIMD.
//
// The loads are relaxed wavefront-scope atomics, which compile to the same
// plain global loads.  A plain load would be simple enough for the compiler
// to fold the ring of loaded values back into a ring of indices and issue
// each load where its value is used, undoing the distance.
template <int DEPTH>
struct ConnAhead {
  uint32_t w[DEPTH] = {};
  __device__ __forceinline__ static uint32_t load(const Slice &l, bool tiled, int e) {
    return tiled ? (uint32_t)__hip_atomic_load(l.c16 + (size_t)e * l.stride, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WAVEFRONT)
                 : __hip_atomic_load(l.c32 + (size_t)e * l.stride, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WAVEFRONT);
  }
  // entries first .. first + DEPTH - 1; first - 1 <= last
  __device__ __forceinline__ void start(const Slice &l, bool tiled, int first, int last) {
#pragma unroll
    for (int d = 0; d < DEPTH; d++) w[d] = load(l, tiled, min(first + d, last));
  }
  __device__ __forceinline__ uint32_t pop(const Slice &l, bool tiled, int e_load, int last) {
    const uint32_t head = w[0];
#pragma unroll
    for (int d = 0; d + 1 < DEPTH; d++) w[d] = w[d + 1];
    w[DEPTH - 1] = load(l, tiled, min(e_load, last));
    return head;
  }
};
// Depths measured on a 2.5M-point cloud with 24 neighbours a point (shallower
// and deeper rings were slower).  func_delta and the mixed q_derivatives ran
// slower with any ring and load their conn entries one neighbour ahead.
constexpr int kAheadQDeriv = 3;
constexpr int kAheadQInner = 4;

// q_variables_mod.h:9-34 for one primitive state, and the denominator
// func_delta divides by (state_update_mod.h:483-485), with no products fused
// into FMAs: every kernel that forms them from a state (the q-variables and
// speed passes, the state update at any points per thread) gives the same
// bits, whatever the compiler would have fused in each.
__device__ __forceinline__ void q_of_state(double rho, double u1, double u2, double pr,
                                           double q[4]) {
#pragma clang fp contract(off)
  const double beta = 0.5 * rho / pr;
  q[0] = log(rho) + (log(beta) * 2.5) - beta * (u1 * u1 + u2 * u2);
  const double two_times_beta = 2.0 * beta;
  q[1] = two_times_beta * u1;
  q[2] = two_times_beta * u2;
  q[3] = -two_times_beta;
}

__device__ __forceinline__ double speed_of_state(double rho, double u1, double u2, double pr) {
#pragma clang fp contract(off)
  const double mod_u = sqrt(u1 * u1 + u2 * u2);
  return mod_u + 3.0 * sqrt(pr / rho);
}

// One neighbour's step, state_update_mod.h:478-488 (speed: the neighbour's
// |u| + 3 sqrt(p / rho)), its one fused product written out so that every
// kernel inlining it agrees to the last bit.
__device__ __forceinline__ double delta_step(double2 xy, double2 xyk, double speed, double cfl) {
  const double dx = xyk.x - xy.x, dy = xyk.y - xy.y;
  const double dist = sqrt(fma(dx, dx, dy * dy));
  return cfl * (dist / speed);
}

}  // namespace lskum
