// pipeline.hpp -- the pieces the persistent, pipelined stencil kernels
// share (kernels.hip, kernels_mixed.hip).  A workgroup walks a sequence of
// stencil blocks; while it computes one from LDS, the next block's tile
// entries and neighbour-list slice are in flight to registers as 16-byte
// chunks, gathered into per-thread slots: E entry slots (tile entry
// t + u tile_block, with that entry's chunks) and NC slice slots (chunk
// t + u tile_block of the slice).  A slot past the end takes the last entry
// or chunk again and writes the same bytes to the same place, so every
// thread issues the same loads and the compiler's memory-operation counts
// stay exact; the kind of every slot is fixed at compile time.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>

namespace lskum {
namespace {

struct QiBlock {
  int t0, tn, nconn;  // first tile entry, tile entries, slice chunks
  int s0;             // first slice entry
};

__device__ __forceinline__ QiBlock qi_block(const int32_t *__restrict__ tile_off,
                                            const int32_t *__restrict__ slice_off, int b) {
  QiBlock q;
  q.t0 = tile_off[b];
  q.tn = tile_off[b + 1] - q.t0;
  q.s0 = slice_off[b];
  q.nconn = (slice_off[b + 1] - q.s0) >> 3;
  return q;
}

// A 16-byte chunk in registers (of any data: chunks are only moved): a
// native vector type, which the compiler keeps in registers where an array
// of HIP's double2 struct, copied from global memory, went to scratch.
typedef double qi_chunk __attribute__((ext_vector_type(2)));

// the tile indices of this thread's entry slots (every block has tn >= 1:
// its own points)
template <int E>
__device__ __forceinline__ void pf_load_idx(int (&idx)[E], const QiBlock &q, int t, int TB,
                                            const int32_t *__restrict__ tile_idx) {
#pragma unroll
  for (int u = 0; u < E; u++) idx[u] = tile_idx[q.t0 + min(t + u * TB, q.tn - 1)];
}

// the list-slice chunks into v[at .. at + NC); a block whose points have no
// neighbours (nconn = 0) reads its first tile entry's coordinates instead
// and, in pf_store_conn, writes nothing
template <int NC, int NV>
__device__ __forceinline__ void pf_load_conn(qi_chunk (&v)[NV], int at, const QiBlock &q, int t,
                                             int TB, const uint16_t *__restrict__ conn16,
                                             const qi_chunk *fallback) {
  const qi_chunk *cn2 = q.nconn > 0 ? reinterpret_cast<const qi_chunk *>(conn16 + q.s0)
                                    : fallback;
  const int last = max(q.nconn - 1, 0);
#pragma unroll
  for (int u = 0; u < NC; u++) v[at + u] = cn2[min(t + u * TB, last)];
}

template <int NC, int NV>
__device__ __forceinline__ void pf_store_conn(qi_chunk *l2, const qi_chunk (&v)[NV], int at,
                                              const QiBlock &q, int t, int TB, int lds_conn2) {
  if (q.nconn <= 0) return;  // uniform per block
#pragma unroll
  for (int u = 0; u < NC; u++) l2[lds_conn2 + min(t + u * TB, q.nconn - 1)] = v[at + u];
}

// The pipelined kernels' layouts: E entry slots and NC list-slice chunks per
// thread, E tile_block >= tile_max and 8 NC >= slice_width (a block's slice
// is slice_width entries of 2 bytes per point).  A layout code is 10 E + NC.
#define PF_LAYOUTS(X) X(2, 3) X(2, 4) X(3, 3) X(3, 4) X(4, 3) X(4, 4)

inline int pf_layout(int tile_block, int tile_max, int slice_width) {
  const int E = (tile_max + tile_block - 1) / tile_block, NC = (slice_width + 7) / 8;
  const int code = 10 * E + std::max(NC, 3);
#define PF_HAS(e, nc) if (code == 10 * e + nc) return code;
  PF_LAYOUTS(PF_HAS)
#undef PF_HAS
  return 0;
}

// lds_conn: the double offset of the list slice, after S doubles per tile entry
inline size_t pf_lds_bytes(int tile_block, int tile_max, int slice_width, int S) {
  return sizeof(double) * (size_t)S * tile_max + sizeof(uint16_t) * (size_t)slice_width * tile_block;
}

// the workgroups that fill n_cu CUs with kernel f (a multiple of 8, at most
// nblk), or 0 where it cannot run with these LDS bytes
inline int pf_grid(const void *f, int tile_block, size_t bytes, size_t lds_limit, int nblk,
                   int n_cu) {
  if (bytes > lds_limit ||
      hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
    return 0;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, tile_block, bytes) != hipSuccess ||
      per_cu <= 0)
    return 0;
  int g = per_cu * n_cu;
  g = g >= 8 ? g / 8 * 8 : g;
  return std::min(g, nblk);
}

}  // namespace
}  // namespace lskum
