// context.hpp -- the device context behind the opaque lskum_ctx, and the
// host helpers its entry points share (api.hip, shape_api.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "host.hpp"
#include "kernels.hpp"
#include "lskum.h"

namespace lskum {

// One cloud's set-up on the host, in its device order.
struct CloudLayout {
  int32_t n = 0;
  std::vector<int32_t> perm, inv;           // perm[dev] = caller, inv[caller] = dev
  std::vector<int32_t> conn_off_h, conn_h;  // caller order, for sub-stencil queries
  std::vector<uint8_t> mask_h;
  std::vector<double> nx_h, ny_h;
  std::vector<int32_t> left_h, right_h, flag1_h, flag2_h;  // caller order
  std::vector<double> min_dist_h;
  // device order
  std::vector<double> x, y, eps2, nx, ny, coef;
  std::vector<float> ls;  // mixed form: [n][4] full-stencil solve coefficients
  std::vector<uint8_t> flag;
  std::vector<int32_t> off;
  std::vector<uint32_t> conn;               // local index | mask << kMaskShift
  std::vector<int32_t> fm, fl, fr, fs;      // force panels (device indices)
};

struct CaseInfo {
  int cloud = 0;
  int32_t p0 = 0;  // first device point
  int32_t n = 0;   // points (the cloud's)
  int32_t np = 0;  // points with padding
  int32_t s0 = 0, s1 = 0;  // force panel range
  double mach = 0.0, aoa = 0.0, theta = 0.0, q_inf[4] = {0, 0, 0, 0};
  double res_old = 0.0;
  int32_t max_res_point = 0;
  int32_t failed_at = 0;  // first iteration with a non-finite residual (0: none)
};

}  // namespace lskum

using namespace lskum;

struct lskum_ctx {
  int n = 0;  // device points, all cases with padding
  lskum_params prm{};
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<CloudLayout> clouds;
  std::vector<CaseInfo> cases;
  int n_cases() const { return (int)cases.size(); }
  // device state: base records, derivative buffers (dqb[cur] holds the
  // current derivatives), qmax/qmin, primitive state
  double *base = nullptr, *dqb[2] = {nullptr, nullptr}, *qm = nullptr;
  double *qsum = nullptr;  // FP64: q_derivatives' q sums, [n][8], for the inner passes
  int cur = 0;
  // mixed-precision form: FP64 coordinates, FP32 gathered data (dq32[cur]
  // holds the current derivatives); dqb/qm are then not allocated
  bool mixed = false;
  double *xy = nullptr;
  float *q32 = nullptr, *dq32[2] = {nullptr, nullptr}, *qm32 = nullptr, *eps32 = nullptr;
  float *ls32 = nullptr;  // [n][4] the full stencil's (syy, -sxy, sxx, 0)/det
  double *prim = nullptr, *prim_old = nullptr;
  double *flux_res = nullptr, *delta = nullptr, *nx = nullptr, *ny = nullptr;
  double *coef = nullptr;
  uint8_t *flag = nullptr;
  int32_t *conn_off = nullptr, *perm_d = nullptr;
  uint32_t *conn = nullptr;
  uint16_t *conn16 = nullptr;
  int32_t *slice_off = nullptr;  // the lists in slices (DevCloud)
  uint32_t *connS = nullptr;
  uint16_t *conn16S = nullptr;
  int32_t *tile_off = nullptr, *tile_idx = nullptr;
  int tile_max = 0;
  unsigned tiles = 0;
  int tile_block = kTileBlockDefault;
  int su_ppt = 1;  // points per thread of the state update
  int flux_cap = 0;  // tile entries the flux stages (0: it reads through L2)
  int slice_width = 0;
  int qi_chunks = 0, qi_grid = 0;  // pipelined inner pass (0: not used)
  int qd_chunks = 0, qd_grid = 0;  // pipelined q_derivatives (0: not used)
  int qdd_chunks = 0, qdd_grid = 0;  // the same with func_delta (0: not used)
  // cases: flow constants, 256-block -> case map, partial and force ranges
  FlowConsts *fcs = nullptr;
  uint16_t *block_case = nullptr;  // NULL for one case
  int32_t *seg = nullptr;          // NULL for one case
  int32_t *soff = nullptr;
  double *qdyn = nullptr;
  Partials part{};
  double *red = nullptr;      // [n_cases][3] sum, max, argmax of the last state update
  double *hist = nullptr;     // [n_cases][hist_cap] per-iteration sum_res_sqr
  int hist_cap = 0;
  int32_t *counter = nullptr; // iteration slot of hist within an lskum_iterate call
  // where the stage-4 reductions of the running call record their sums:
  // hist with stride hist_cap (lskum_iterate) or the library's static
  // scratch array (lskum_iterate_noalloc)
  double *hist_dst = nullptr;
  int hist_stride = 0;
  int device = 0;             // the device current at create (owns stream and buffers)
  double *red_h = nullptr;    // pinned host mirror of red / forces
  int32_t *f_m = nullptr, *f_l = nullptr, *f_r = nullptr, *f_s = nullptr;
  double *forces = nullptr;   // [n_cases][3] H, V, pitch moment
  double rho_inf = 1.0, pr_inf = 1.0 / 1.4;
  // iteration bookkeeping (q_lskum_mod.h, fpi_solver_mod.h:138-147)
  int it = 0;
  int restart = 0;
  // q and the func_delta speed hold for the current state (a stage-4 update
  // refreshes both; setting the state invalidates them)
  bool q_fresh = false, speed_fresh = false;
  // graph capture of whole iterations
  int graph_iters = 0;
  hipGraphExec_t gexec = nullptr;
  int gexec_iters = 0;
  double *gexec_hist = nullptr;
  bool capturing = false;
  // adjoint and tangent working arrays, allocated on first use
  AdjointBuffers adj{};
  bool adj_ready = false;
  // timing
  bool timing = false;
  double k_ms[6] = {0, 0, 0, 0, 0, 0};
  int k_cnt[6] = {0, 0, 0, 0, 0, 0};
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, int>> ev_pending;  // (kernel id, event index)

  DevCloud dc() const {
    DevCloud c;
    c.n = n;
    c.conn_off = conn_off;
    c.conn = conn;
    c.conn16 = conn16;
    c.slice_off = slice_off;
    c.connS = connS;
    c.conn16S = conn16S;
    c.tile_off = tile_off;
    c.tile_idx = tile_idx;
    c.tile_max = tile_max;
    c.tiles = tiles;
    c.tile_block = tile_block;
    c.su_ppt = su_ppt;
    c.flux_cap = flux_cap;
    c.slice_width = slice_width;
    c.qi_chunks = qi_chunks;
    c.qi_grid = qi_grid;
    c.qd_chunks = qd_chunks;
    c.qd_grid = qd_grid;
    c.qdd_chunks = qdd_chunks;
    c.qdd_grid = qdd_grid;
    c.xy = xy;
    c.nx = nx;
    c.ny = ny;
    c.coef = coef;
    c.flag = flag;
    c.perm = perm_d;
    return c;
  }
  MixedState ms() const {
    MixedState m;
    m.xy = xy;
    m.q32 = q32;
    m.qm32 = qm32;
    m.eps32 = eps32;
    m.ls32 = ls32;
    return m;
  }
};


namespace lskum {

int set_err(lskum_ctx *c, int code, const std::string &m);

#define HIP_TRY(ctx, call)                                                       \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return set_err(ctx, LSKUM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CTX_OK(ctx)                                        \
  do {                                                     \
    if (!(ctx)) return LSKUM_EINVAL;                       \
    if (!(ctx)->stream) return set_err(ctx, LSKUM_EINVAL, "context failed to initialise"); \
  } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t count) {
  return hipMalloc((void **)p, sizeof(T) * (count ? count : 1));
}

int check_launch(lskum_ctx *c, const char *what);
int do_snapshot(lskum_ctx *c);
int do_speed(lskum_ctx *c);
int do_func_delta(lskum_ctx *c);
int do_q_variables(lskum_ctx *c);
// keep_sums (FP64 form): also write qsum, which an inner pass that follows
// reads (the single-stage entry point cannot know, so it always writes them);
// with_delta: func_delta in the same pass (only where qdd_chunks is set)
int do_q_derivatives(lskum_ctx *c, int n_later, bool keep_sums, bool with_delta = false);
int do_q_inner(lskum_ctx *c, int n_inner);
int do_flux(lskum_ctx *c);
int do_state_update(lskum_ctx *c, int rk, bool with_hist);
int run_iteration(lskum_ctx *c);
void permute_out(const lskum_ctx *c, const double *src, double *dst, int ncomp);
void permute_in(const lskum_ctx *c, const double *src, double *dst, int ncomp);
int fetch_plain(lskum_ctx *c, const double *dev, int ncomp, std::vector<double> &out);
// least-squares coefficients of the sub-stencils of every point of a cloud
// (device order) from caller-order coordinates and normals
void compute_coef(const CloudLayout &L, const double *x, const double *y, const double *nx,
                  const double *ny, double power, std::vector<double> &coef);
// adjoint working arrays and the transpose of conn (single-case FP64, power 0)
int prepare_adjoint(lskum_ctx *c);
int recompute_stages(lskum_ctx *c);
int upload_caller(lskum_ctx *c, const double *src, double *dev);
// lskum_iterate_noalloc's static scratch array: the address of the
// device's copy, looked up once per device (init) and then only read
int noalloc_scratch_init(int device, std::string &err);
double *noalloc_scratch(int device);
int launch_iterations(lskum_ctx *ctx, int n_iter, bool allow_graph);
int finish_iterations(lskum_ctx *ctx, int n_iter, const double *sums, double *res_new,
                      double *residue);
int download_caller(lskum_ctx *c, const double *dev, double *dst);

}  // namespace lskum
