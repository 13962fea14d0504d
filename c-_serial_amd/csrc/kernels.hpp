// kernels.hpp -- launchers of the q-LSKUM HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lskum {

// Per-point record gathered by the stencil kernels: 24 doubles (192 B),
// so one neighbour's whole state is 12 16-byte loads from 2 cache lines.
enum RecordSlot : int {
  R_X = 0,
  R_Y = 1,
  R_Q = 2,     // q[4]
  R_DQX = 6,   // dq/dx[4]
  R_DQY = 10,  // dq/dy[4]
  R_QMAX = 14, // qmax[4]
  R_QMIN = 18, // qmin[4]
  R_EPS2 = 22, // (pow(K*min_dist, 3))^2
  R_STRIDE = 24
};
// Alternate derivative buffer for the Jacobi inner passes: [dqx[4], dqy[4]].
constexpr int ALT_STRIDE = 8;
// conn entries: neighbour index in the low 28 bits, sub-stencil mask above.
constexpr uint32_t kIndexMask = 0x0FFFFFFFu;
constexpr int kMaskShift = 28;

struct DevCloud {
  int n;
  const int32_t *conn_off;
  const uint32_t *conn;
  const double *nx, *ny;
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

void launch_q_variables(const DevCloud &c, const double *prim, double *rec,
                        int block, hipStream_t s);
void launch_q_derivatives(const DevCloud &c, double *rec, double *dq_dst,
                          int dq_stride, double power, int block, hipStream_t s);
void launch_q_inner(const DevCloud &c, const double *rec, const double *dq_src,
                    int src_stride, double *dq_dst, int dst_stride, double power,
                    int block, hipStream_t s);
void launch_flux(const DevCloud &c, const double *rec, const double *delta,
                 double power, double *flux_res, int block, hipStream_t s);
void launch_func_delta(const DevCloud &c, const double *rec, const double *prim,
                       double cfl, double *delta, int block, hipStream_t s);
void launch_state_update(const DevCloud &c, int rk, double *prim,
                         const double *prim_old, const double *flux_res,
                         const FlowConsts &fc, const Partials &p, int block,
                         hipStream_t s);
// reduces the partials into out[0] = sum, out[1] = max, out[2] = argmax
// (caller index, as a double), and also out_hist[slot] = sum when non-null
void launch_reduce(const Partials &p, double *out, double *hist_slot,
                   hipStream_t s);
void launch_forces(int n_shape, const int32_t *m, const int32_t *l,
                   const int32_t *r, const int32_t *shape, const double *rec,
                   const double *nx, const double *ny, const double *prim,
                   int n, double pr_inf, double qdyn, double *out3,
                   hipStream_t s);
void launch_kat_fluxes(int n, const double *in6, double *out, hipStream_t s);
void launch_kat_qtilde(int n, const double *qt4, double *out, hipStream_t s);

}  // namespace lskum
