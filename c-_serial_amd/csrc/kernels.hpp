// kernels.hpp -- launchers of the q-LSKUM HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lskum {

// Per-point "hot line": one 128-byte cache line holding everything the
// derivative, inner-pass and local-time-step kernels gather from a
// neighbour.  Two copies (A, B) ping-pong the derivatives across the Jacobi
// inner passes; x, y, q and eps2 are kept in both.
enum HotSlot : int {
  H_X = 0,
  H_Y = 1,
  H_Q = 2,     // q[4]
  H_DQX = 6,   // dq/dx[4]
  H_DQY = 10,  // dq/dy[4]
  H_EPS2 = 14, // (pow(K*min_dist, 3))^2, limiters_mod.h:82-83
  H_SPEED = 15, // |u| + 3 sqrt(p/rho) of the state func_delta reads
  H_STRIDE = 16
};
// qmax[4], qmin[4] per point (64 bytes): only the flux gathers them.
constexpr int QM_STRIDE = 8;
// conn entries: neighbour index in the low 28 bits, sub-stencil mask above.
constexpr uint32_t kIndexMask = 0x0FFFFFFFu;
constexpr int kMaskShift = 28;
// tiled conn entries: local index in the block's tile (12 bits), mask above.
constexpr uint16_t kLocalMask = 0x0FFFu;
constexpr int kLocalShift = 12;
constexpr int kTileBlock = 256;  // points per block = per LDS tile

struct DevCloud {
  int n;
  const int32_t *conn_off;
  const uint32_t *conn;      // global form
  const uint16_t *conn16;    // tiled form (same order)
  const int32_t *tile_off;   // [nblocks + 1]
  const int32_t *tile_idx;   // own points first, then other neighbours
  int tile_max;              // largest tile, entries
  bool tiled;                // gather kernels use the tiles
  bool flux_tiled;           // the flux kernel uses them too
  const double *nx, *ny;
  const double *coef;  // per point: least-squares coefficients of the
                       // sub-stencils x+, x-, y+, y- (a, b each)
  const uint8_t *flag;
  const int32_t *perm;  // device index -> caller index
};

struct FlowConsts {
  double cfl, power;
  double u1_inf, u2_inf;  // q_inf[2], q_inf[3]
  double rho_inf, pr_inf;
};

struct Partials {
  double *sum, *max;
  int32_t *arg;
  int nblocks;
};

int blocks_for(int n, int block);
// LDS bytes per tiled launch of each kernel for a given tile size
size_t lds_bytes_qderiv(int tile_max);
size_t lds_bytes_qinner(int tile_max);
size_t lds_bytes_delta(int tile_max);
size_t lds_bytes_flux(int tile_max);
// raises the dynamic-LDS limits of the tiled kernels; false if they do not fit
bool prepare_tiled(int tile_max, size_t lds_limit);

void launch_q_variables(const DevCloud &c, const double *prim, double *hot_a,
                        double *hot_b, hipStream_t s);
void launch_speed(const DevCloud &c, const double *prim, double *hot, hipStream_t s);
void launch_func_delta(const DevCloud &c, const double *hot, double cfl, double *delta,
                       hipStream_t s);
void launch_q_derivatives(const DevCloud &c, double *hot, double *qm, double power,
                          hipStream_t s);
void launch_q_inner(const DevCloud &c, const double *src, double *dst, double power,
                    hipStream_t s);
void launch_flux(const DevCloud &c, const double *hot, const double *qm, const double *delta,
                 double power, double *flux_res, hipStream_t s);
// also writes q of the new state into hot_a/hot_b (when non-null) and, with
// speed, the func_delta denominator into hot_a
void launch_state_update(const DevCloud &c, int rk, double *prim, const double *prim_old,
                         const double *flux_res, const FlowConsts &fc, const Partials &p,
                         double *hot_a, double *hot_b, bool speed, hipStream_t s);
// reduces the partials into out[0] = sum, out[1] = max, out[2] = argmax
// (caller index, as a double), and also hist_slot[0] = sum when non-null
void launch_reduce(const Partials &p, double *out, double *hist_slot, hipStream_t s);
void launch_forces(int n_shape, const int32_t *m, const int32_t *l, const int32_t *r,
                   const int32_t *shape, const double *hot, const double *nx,
                   const double *ny, const double *prim, int n, double pr_inf, double qdyn,
                   double *out3, hipStream_t s);
void launch_kat_fluxes(int n, const double *in6, double *out, hipStream_t s);
void launch_kat_qtilde(int n, const double *qt4, double *out, hipStream_t s);

}  // namespace lskum
