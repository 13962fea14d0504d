// api.hip -- the C ABI of lskum.h: device context, host fpi_solver loop.
//
// The host loop keeps the reference's structure (fpi_solver_mod.h:14-150,
// q_lskum_mod.h:28-37): each stage is one stream-ordered kernel launch on the
// context's stream; only the per-iteration residual sums come back, once,
// at the end of lskum_iterate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "host.hpp"
#include "kernels.hpp"
#include "lskum.h"

using namespace lskum;

struct lskum_ctx {
  int n = 0;
  lskum_params prm{};
  hipStream_t stream = nullptr;
  std::string err;
  // caller <-> device order
  std::vector<int32_t> perm, inv;  // perm[dev] = caller, inv[caller] = dev
  // host copy of the caller-order cloud (for sub-stencil queries)
  std::vector<int32_t> conn_off_h, conn_h;
  std::vector<uint8_t> mask_h;
  std::vector<double> nx_h, ny_h;
  // device state: base records, derivative buffers (dqb[cur] holds the
  // current derivatives), qmax/qmin, primitive state
  double *base = nullptr, *dqb[2] = {nullptr, nullptr}, *qm = nullptr;
  int cur = 0;
  // mixed-precision form: FP64 coordinates, FP32 gathered data (dq32[cur]
  // holds the current derivatives); dqb/qm are then not allocated
  bool mixed = false;
  double *xy = nullptr;
  float *q32 = nullptr, *dq32[2] = {nullptr, nullptr}, *qm32 = nullptr, *eps32 = nullptr;
  double *prim = nullptr, *prim_old = nullptr;
  double *flux_res = nullptr, *delta = nullptr, *nx = nullptr, *ny = nullptr;
  double *coef = nullptr;
  uint8_t *flag = nullptr;
  int32_t *conn_off = nullptr, *perm_d = nullptr;
  uint32_t *conn = nullptr;
  uint16_t *conn16 = nullptr;
  int32_t *tile_off = nullptr, *tile_idx = nullptr;
  int tile_max = 0;
  unsigned tiles = 0;
  int tile_block = kTileBlockDefault;
  Partials part{};
  double *red = nullptr;      // [3] sum, max, argmax of the last state update
  double *hist = nullptr;     // per-iteration sum_res_sqr of lskum_iterate
  int hist_cap = 0;
  double *red_h = nullptr;    // pinned host mirror of red / hist
  int32_t *f_m = nullptr, *f_l = nullptr, *f_r = nullptr, *f_s = nullptr;
  int n_shape = 0;
  double *forces = nullptr;   // [3] H, V, pitch moment
  // case
  double theta = 0.0, q_inf[4] = {0, 0, 0, 0};
  double rho_inf = 1.0, pr_inf = 1.0 / 1.4;
  // iteration bookkeeping (q_lskum_mod.h, fpi_solver_mod.h:138-147)
  int it = 0;
  double res_old = 0.0;
  int restart = 0;
  int32_t max_res_point = 0;
  // timing
  bool timing = false;
  hipEvent_t ev[2] = {nullptr, nullptr};
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
    c.tile_off = tile_off;
    c.tile_idx = tile_idx;
    c.tile_max = tile_max;
    c.tiles = tiles;
    c.tile_block = tile_block;
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
    return m;
  }
  FlowConsts fc() const {
    FlowConsts f;
    f.cfl = prm.cfl;
    f.power = prm.power;
    f.u1_inf = q_inf[1];
    f.u2_inf = q_inf[2];
    f.rho_inf = rho_inf;
    f.pr_inf = pr_inf;
    return f;
  }
};

namespace {

int set_err(lskum_ctx *c, int code, const std::string &m) {
  if (c) c->err = m;
  return code;
}

#define HIP_TRY(ctx, call)                                                       \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return set_err(ctx, LSKUM_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t count) {
  return hipMalloc((void **)p, sizeof(T) * (count ? count : 1));
}

void release(lskum_ctx *c) {
  if (!c) return;
  void *ptrs[] = {c->coef, c->base, c->dqb[0], c->dqb[1], c->qm, c->conn16, c->tile_off, c->tile_idx, c->prim, c->prim_old, c->flux_res, c->delta, c->nx,
                  c->ny, c->flag, c->conn_off, c->perm_d, c->conn, c->part.sum,
                  c->part.max, c->part.arg, c->red, c->hist, c->f_m, c->f_l, c->f_r,
                  c->f_s, c->forces, c->xy, c->q32, c->dq32[0], c->dq32[1], c->qm32,
                  c->eps32};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->red_h) (void)hipHostFree(c->red_h);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

// Kernel timing: an event pair around a launch when enabled.
struct KTimer {
  lskum_ctx *c;
  int id;
  int idx = -1;
  KTimer(lskum_ctx *ctx, int kid) : c(ctx), id(kid) {
    if (!c->timing) return;
    idx = (int)c->ev_pending.size() * 2;
    while ((int)c->ev_pool.size() < idx + 2) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) { idx = -1; return; }
      c->ev_pool.push_back(e);
    }
    (void)hipEventRecord(c->ev_pool[idx], c->stream);
  }
  ~KTimer() {
    if (idx < 0) return;
    (void)hipEventRecord(c->ev_pool[idx + 1], c->stream);
    c->ev_pending.push_back({id, idx});
  }
};

int collect_timing(lskum_ctx *c) {
  if (c->ev_pending.empty()) return LSKUM_OK;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (auto &p : c->ev_pending) {
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev_pool[p.second], c->ev_pool[p.second + 1]));
    c->k_ms[p.first] += ms;
    c->k_cnt[p.first] += 1;
  }
  c->ev_pending.clear();
  return LSKUM_OK;
}

int check_launch(lskum_ctx *c, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(c, LSKUM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return LSKUM_OK;
}

// prim_old <- prim
int do_snapshot(lskum_ctx *c) {
  HIP_TRY(c, hipMemcpyAsync(c->prim_old, c->prim, sizeof(double) * 4 * c->n,
                            hipMemcpyDeviceToDevice, c->stream));
  return LSKUM_OK;
}

// with_speed: first refresh the per-point speed from prim (not needed right
// after a stage-4 state update, which wrote it)
int do_func_delta(lskum_ctx *c, bool with_speed) {
  {
    KTimer t(c, LSKUM_K_DELTA);
    if (with_speed) launch_speed(c->dc(), c->prim, c->base, c->stream);  // both forms: FP64
    launch_func_delta(c->dc(), c->base, c->prm.cfl, c->delta, c->stream);
  }
  return check_launch(c, "func_delta");
}

int do_q_variables(lskum_ctx *c) {
  {
    KTimer t(c, LSKUM_K_QVAR);
    if (c->mixed)
      launch_q_variables_m(c->dc(), c->prim, c->q32, c->stream);
    else
      launch_q_variables(c->dc(), c->prim, c->base, c->stream);
  }
  return check_launch(c, "q_variables");
}

// derivatives into the buffer from which n_later Jacobi passes end in dqb[0]
int do_q_derivatives(lskum_ctx *c, int n_later) {
  c->cur = n_later % 2;
  {
    KTimer t(c, LSKUM_K_QDERIV);
    if (c->mixed)
      launch_q_derivatives_m(c->dc(), c->ms(), c->dq32[c->cur], c->prm.power, c->stream);
    else
      launch_q_derivatives(c->dc(), c->base, c->dqb[c->cur], c->qm, c->prm.power, c->stream);
  }
  return check_launch(c, "q_derivatives");
}

// n Jacobi passes, each reading dqb[cur] and writing the other buffer
int do_q_inner(lskum_ctx *c, int n_inner) {
  for (int p = 0; p < n_inner; p++) {
    KTimer t(c, LSKUM_K_QINNER);
    if (c->mixed)
      launch_q_inner_m(c->dc(), c->ms(), c->dq32[c->cur], c->dq32[1 - c->cur], c->prm.power,
                       c->stream);
    else
      launch_q_inner(c->dc(), c->base, c->dqb[c->cur], c->dqb[1 - c->cur], c->prm.power,
                     c->stream);
    c->cur = 1 - c->cur;
  }
  return check_launch(c, "q_inner");
}

int do_flux(lskum_ctx *c) {
  {
    KTimer t(c, LSKUM_K_FLUX);
    if (c->mixed)
      launch_flux_m(c->dc(), c->ms(), c->dq32[c->cur], c->delta, c->prm.power, c->flux_res,
                    c->stream);
    else
      launch_flux(c->dc(), c->base, c->dqb[c->cur], c->qm, c->delta, c->prm.power,
                  c->flux_res, c->stream);
  }
  return check_launch(c, "flux_residual");
}

// the state update also refreshes q (and, after the last stage, the speed
// func_delta reads), so the next stage needs no separate q_variables pass
int do_state_update(lskum_ctx *c, int rk, double *hist_slot) {
  {
    KTimer t(c, LSKUM_K_STATE);
    launch_state_update(c->dc(), rk, c->prim, c->prim_old, c->flux_res, c->fc(), c->part,
                        c->base, c->mixed ? c->q32 : nullptr, rk == 4, c->stream);
    launch_reduce(c->part, c->red, hist_slot, c->stream);
  }
  return check_launch(c, "state_update");
}

void permute_in(const lskum_ctx *c, const double *src, double *dst, int ncomp) {
  for (int r = 0; r < ncomp; r++)
    for (int d = 0; d < c->n; d++) dst[(size_t)r * c->n + d] = src[(size_t)r * c->n + c->perm[d]];
}

void permute_out(const lskum_ctx *c, const double *src, double *dst, int ncomp) {
  for (int r = 0; r < ncomp; r++)
    for (int d = 0; d < c->n; d++) dst[(size_t)r * c->n + c->perm[d]] = src[(size_t)r * c->n + d];
}

// fetch ncomp consecutive slots of a per-point record array: out[r][dev]
int fetch_record(lskum_ctx *c, const double *base, int stride, int slot, int ncomp,
                 std::vector<double> &out) {
  out.assign((size_t)ncomp * c->n, 0.0);
  std::vector<double> tmp((size_t)ncomp * c->n);
  HIP_TRY(c, hipMemcpy2DAsync(tmp.data(), sizeof(double) * ncomp, base + slot,
                              sizeof(double) * stride, sizeof(double) * ncomp, c->n,
                              hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int d = 0; d < c->n; d++)
    for (int r = 0; r < ncomp; r++) out[(size_t)r * c->n + d] = tmp[(size_t)d * ncomp + r];
  return LSKUM_OK;
}

// fetch a float array of ncomp values per point as out[r][dev] (doubles)
int fetch_float(lskum_ctx *c, const float *dev, int ncomp, std::vector<double> &out) {
  std::vector<float> tmp((size_t)ncomp * c->n);
  HIP_TRY(c, hipMemcpyAsync(tmp.data(), dev, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  out.resize(tmp.size());
  for (int d = 0; d < c->n; d++)
    for (int r = 0; r < ncomp; r++) out[(size_t)r * c->n + d] = tmp[(size_t)d * ncomp + r];
  return LSKUM_OK;
}

int fetch_plain(lskum_ctx *c, const double *dev, int ncomp, std::vector<double> &out) {
  out.resize((size_t)ncomp * c->n);
  HIP_TRY(c, hipMemcpyAsync(out.data(), dev, sizeof(double) * ncomp * c->n,
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return LSKUM_OK;
}

}  // namespace

extern "C" {

const char *lskum_last_error(const lskum_ctx *ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int lskum_create(const lskum_cloud *cloud, const lskum_params *params, lskum_ctx **out) {
  if (!out) return LSKUM_EINVAL;
  *out = nullptr;
  if (!cloud || !params) return LSKUM_EINVAL;
  lskum_ctx *c = new (std::nothrow) lskum_ctx;
  if (!c) return LSKUM_ENOMEM;
  std::string err;
  int rc = check_cloud(*cloud, &err);
  const lskum_params &p = *params;
  if (rc == LSKUM_OK && !(p.mach > 0.0 && p.cfl > 0.0 && p.vl_const >= 0.0 &&
                          p.inner_iterations >= 0 &&
                          (p.precision == LSKUM_FP64 || p.precision == LSKUM_MIXED) &&
                          (p.reorder >= LSKUM_REORDER_NONE && p.reorder <= LSKUM_REORDER_HILBERT) &&
                          (p.lds_tiles == LSKUM_TILES_AUTO ||
                           (p.lds_tiles >= 0 && p.lds_tiles <= 15)) &&
                          (p.tile_block == 0 || p.tile_block == 64 || p.tile_block == 128 ||
                           p.tile_block == 256))) {
    rc = LSKUM_EINVAL;
    err = "invalid parameters";
  }
  if (rc != LSKUM_OK) {
    // keep the message reachable: return the context in its failed state
    c->err = err;
    *out = c;
    return rc;
  }
  c->prm = p;
  c->mixed = p.precision == LSKUM_MIXED;
  c->tile_block = p.tile_block ? p.tile_block : kTileBlockDefault;
  const int32_t n = cloud->n_points;
  c->n = n;
  // set-up on the host: normals and sub-stencils in the caller's order
  compute_normals(*cloud, c->nx_h, c->ny_h);
  generate_connectivity(*cloud, c->nx_h, c->ny_h, c->mask_h);
  for (int32_t i = 0; i < n; i++) {
    double r = substencil_min_rel_det(*cloud, c->nx_h, c->ny_h, c->mask_h, i);
    if (!(r > 0.0)) {
      char m[160];
      std::snprintf(m, sizeof m,
                    "degenerate sub-stencil at point %d (flag_1 %d): empty or collinear", i,
                    cloud->flag_1[i]);
      c->err = m;
      *out = c;
      return LSKUM_ESTENCIL;
    }
  }
  c->conn_off_h.assign(cloud->conn_off, cloud->conn_off + n + 1);
  c->conn_h.assign(cloud->conn, cloud->conn + cloud->conn_off[n]);
  // device order: Hilbert (compact 256-point blocks whose neighbour sets fit
  // an LDS tile), or RCM / the caller's order grouped by point type
  std::vector<int32_t> base;
  if (p.reorder == LSKUM_REORDER_HILBERT) {
    hilbert_order(*cloud, c->perm);
  } else {
    if (p.reorder == LSKUM_REORDER_RCM) {
      rcm_order(*cloud, base);
    } else {
      base.resize(n);
      for (int32_t i = 0; i < n; i++) base[i] = i;
    }
    c->perm.clear();
    c->perm.reserve(n);
    for (int t = 0; t < 3; t++)
      for (int32_t v : base)
        if (cloud->flag_1[v] == t) c->perm.push_back(v);
  }
  c->inv.assign(n, 0);
  for (int32_t d = 0; d < n; d++) c->inv[c->perm[d]] = d;
  // host images of the device arrays
  std::vector<double> base_h((size_t)B_STRIDE * n, 0.0), nxd(n), nyd(n);
  std::vector<uint8_t> flag_h(n);
  std::vector<int32_t> off_h(n + 1, 0);
  std::vector<uint32_t> conn_d(cloud->conn_off[n]);
  for (int32_t d = 0; d < n; d++) {
    const int32_t v = c->perm[d];
    base_h[(size_t)B_STRIDE * d + B_X] = cloud->x[v];
    base_h[(size_t)B_STRIDE * d + B_Y] = cloud->y[v];
    // limiters_mod.h:82-83: epsi = pow(VL_CONST*min_dist, 3), used squared
    const double epsi = std::pow(p.vl_const * cloud->min_dist[v], 3);
    base_h[(size_t)B_STRIDE * d + B_EPS2] = epsi * epsi;
    nxd[d] = c->nx_h[v];
    nyd[d] = c->ny_h[v];
    flag_h[d] = (uint8_t)cloud->flag_1[v];
    const int32_t b = cloud->conn_off[v], e = cloud->conn_off[v + 1];
    off_h[d + 1] = off_h[d] + (e - b);
    for (int32_t t = b; t < e; t++)
      conn_d[off_h[d] + (t - b)] =
          (uint32_t)c->inv[cloud->conn[t]] | ((uint32_t)c->mask_h[t] << kMaskShift);
  }
  // least-squares coefficients of each sub-stencil (geometry only): the
  // sums of interior_fluxes_mod.h:51-61 over the members in conn order, then
  // (syy, -sxy)/det for the x-form solve and (sxx, -sxy)/det for the y-form
  std::vector<double> coef_h((size_t)8 * n, 0.0);
  for (int32_t d = 0; d < n; d++) {
    const int32_t v = c->perm[d];
    const double nxv = c->nx_h[v], nyv = c->ny_h[v], tx = nyv, ty = -nxv;
    double sxx[4] = {0, 0, 0, 0}, syy[4] = {0, 0, 0, 0}, sxy[4] = {0, 0, 0, 0};
    for (int32_t t = cloud->conn_off[v]; t < cloud->conn_off[v + 1]; t++) {
      const int32_t k = cloud->conn[t];
      const double delx = cloud->x[k] - cloud->x[v], dely = cloud->y[k] - cloud->y[v];
      const double dels = delx * tx + dely * ty, deln = delx * nxv + dely * nyv;
      double dsw = dels, dnw = deln;
      if (p.power != 0.0) {
        const double w = std::pow(std::sqrt(dels * dels + deln * deln), p.power);
        dsw = dels * w;
        dnw = deln * w;
      }
      for (int sb = 0; sb < 4; sb++) {
        if (!(c->mask_h[t] & (1u << sb))) continue;
        sxx[sb] = sxx[sb] + dels * dsw;
        syy[sb] = syy[sb] + deln * dnw;
        sxy[sb] = sxy[sb] + dels * dnw;
      }
    }
    for (int sb = 0; sb < 4; sb++) {
      const double det = sxx[sb] * syy[sb] - sxy[sb] * sxy[sb];
      const double inv = det != 0.0 ? 1.0 / det : 0.0;
      coef_h[(size_t)8 * d + 2 * sb] = (sb < 2 ? syy[sb] : sxx[sb]) * inv;
      coef_h[(size_t)8 * d + 2 * sb + 1] = -sxy[sb] * inv;
    }
  }
  // LDS tiles: per block of tile_block points, its own points then the
  // sorted other points its lists reach; conn16 holds local indices
  const int TB = c->tile_block;
  const int nblk = blocks_for(n, TB);
  std::vector<int32_t> toff(nblk + 1, 0), tidx;
  std::vector<uint16_t> conn16_h(conn_d.size());
  bool tiles_ok = true;
  {
    std::vector<int32_t> loc(n, -1), extra;
    for (int blk = 0; blk < nblk && tiles_ok; blk++) {
      const int32_t p0 = blk * TB, p1 = std::min(n, p0 + TB);
      extra.clear();
      for (int32_t d = p0; d < p1; d++)
        for (int32_t e = off_h[d]; e < off_h[d + 1]; e++) {
          const int32_t k = (int32_t)(conn_d[e] & kIndexMask);
          if (k < p0 || k >= p1) extra.push_back(k);
        }
      std::sort(extra.begin(), extra.end());
      extra.erase(std::unique(extra.begin(), extra.end()), extra.end());
      const int32_t size = (p1 - p0) + (int32_t)extra.size();
      if (size > (int32_t)kLocalMask + 1) {
        tiles_ok = false;
        break;
      }
      for (int32_t d = p0; d < p1; d++) {
        loc[d] = d - p0;
        tidx.push_back(d);
      }
      for (size_t j = 0; j < extra.size(); j++) {
        loc[extra[j]] = (p1 - p0) + (int32_t)j;
        tidx.push_back(extra[j]);
      }
      for (int32_t d = p0; d < p1; d++)
        for (int32_t e = off_h[d]; e < off_h[d + 1]; e++) {
          const uint32_t w = conn_d[e];
          conn16_h[e] = (uint16_t)(loc[w & kIndexMask] |
                                   ((w >> kMaskShift) << kLocalShift));
        }
      toff[blk + 1] = (int32_t)tidx.size();
      c->tile_max = std::max(c->tile_max, size);
      for (int32_t d = p0; d < p1; d++) loc[d] = -1;
      for (int32_t k : extra) loc[k] = -1;
    }
  }
  // the gather kernels stage neighbours in LDS; the FP64 flux runs at
  // higher occupancy reading its neighbours through L2, so by default it
  // does not (its FP32 records, half the size, do fit at occupancy 2)
  const unsigned auto_mask = LSKUM_TILE_QDERIV | LSKUM_TILE_QINNER | LSKUM_TILE_DELTA |
                             (c->mixed ? LSKUM_TILE_FLUX : 0u);
  const unsigned want =
      p.lds_tiles == LSKUM_TILES_AUTO ? auto_mask : (unsigned)p.lds_tiles;
  if (tiles_ok && want) {
    int dev = 0, lds_limit = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
            hipSuccess) {
      if (c->mixed)  // func_delta is FP64 in both forms
        c->tiles = prepare_tiled(want & TILE_DELTA, c->tile_max, (size_t)lds_limit) |
                   prepare_tiled_m(want & ~TILE_DELTA, c->tile_max, (size_t)lds_limit);
      else
        c->tiles = prepare_tiled(want, c->tile_max, (size_t)lds_limit);
    }
  }
  // shape points for the force reduction: the first shape_points points
  // (compute_force_coeffs_mod.h:25-27 takes m = j)
  int shape_points = 0;
  for (int32_t i = 0; i < n; i++)
    if (cloud->flag_2[i] > 0) shape_points++;
  std::vector<int32_t> fm, fl, fr, fs;
  for (int j = 0; j < shape_points; j++) {
    if (cloud->flag_1[j] == 1) continue;  // no left/right: cannot contribute a panel
    fm.push_back(c->inv[j]);
    fl.push_back(c->inv[cloud->left[j]]);
    fr.push_back(c->inv[cloud->right[j]]);
    fs.push_back(cloud->flag_2[j]);
  }
  c->n_shape = (int)fm.size();
  // device allocation and upload
  auto fail = [&](int code) {
    *out = c;
    return code;
  };
  hipError_t e;
#define ALLOC(ptr, count)                                                   \
  if ((e = dalloc(&(ptr), (count))) != hipSuccess) {                        \
    c->err = std::string("hipMalloc " #ptr ": ") + hipGetErrorString(e);   \
    return fail(LSKUM_ENOMEM);                                              \
  }
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
    c->err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
    return fail(LSKUM_EHIP);
  }
  c->part.nblocks = blocks_for(n, 256);
  ALLOC(c->base, (size_t)B_STRIDE * n);
  if (c->mixed) {
    ALLOC(c->xy, (size_t)2 * n);
    ALLOC(c->q32, (size_t)4 * n);
    ALLOC(c->dq32[0], (size_t)8 * n);
    ALLOC(c->dq32[1], (size_t)8 * n);
    ALLOC(c->qm32, (size_t)8 * n);
    ALLOC(c->eps32, (size_t)n);
  } else {
    ALLOC(c->dqb[0], (size_t)DQ_STRIDE * n);
    ALLOC(c->dqb[1], (size_t)DQ_STRIDE * n);
    ALLOC(c->qm, (size_t)QM_STRIDE * n);
  }
  ALLOC(c->prim, (size_t)4 * n);
  ALLOC(c->prim_old, (size_t)4 * n);
  ALLOC(c->flux_res, (size_t)4 * n);
  ALLOC(c->delta, (size_t)n);
  ALLOC(c->nx, (size_t)n);
  ALLOC(c->ny, (size_t)n);
  ALLOC(c->coef, (size_t)8 * n);
  ALLOC(c->flag, (size_t)n);
  ALLOC(c->conn_off, (size_t)n + 1);
  ALLOC(c->perm_d, (size_t)n);
  ALLOC(c->conn, conn_d.size());
  if (c->tiles) {
    ALLOC(c->conn16, conn16_h.size());
    ALLOC(c->tile_off, toff.size());
    ALLOC(c->tile_idx, tidx.size());
  }
  ALLOC(c->part.sum, (size_t)c->part.nblocks);
  ALLOC(c->part.max, (size_t)c->part.nblocks);
  ALLOC(c->part.arg, (size_t)c->part.nblocks);
  ALLOC(c->red, 3);
  ALLOC(c->forces, 3);
  ALLOC(c->f_m, fm.size());
  ALLOC(c->f_l, fm.size());
  ALLOC(c->f_r, fm.size());
  ALLOC(c->f_s, fm.size());
#undef ALLOC
  if ((e = hipHostMalloc((void **)&c->red_h, sizeof(double) * 8)) != hipSuccess) {
    c->err = std::string("hipHostMalloc: ") + hipGetErrorString(e);
    return fail(LSKUM_ENOMEM);
  }
  struct Up { void *dst; const void *src; size_t bytes; };
  std::vector<Up> ups = {{c->base, base_h.data(), base_h.size() * sizeof(double)},
                         {c->nx, nxd.data(), n * sizeof(double)},
                         {c->coef, coef_h.data(), coef_h.size() * sizeof(double)},
                         {c->ny, nyd.data(), n * sizeof(double)},
                         {c->flag, flag_h.data(), (size_t)n},
                         {c->conn_off, off_h.data(), (n + 1) * sizeof(int32_t)},
                         {c->perm_d, c->perm.data(), n * sizeof(int32_t)},
                         {c->conn, conn_d.data(), conn_d.size() * sizeof(uint32_t)},
                         {c->f_m, fm.data(), fm.size() * sizeof(int32_t)},
                         {c->f_l, fl.data(), fl.size() * sizeof(int32_t)},
                         {c->f_r, fr.data(), fr.size() * sizeof(int32_t)},
                         {c->f_s, fs.data(), fs.size() * sizeof(int32_t)}};
  std::vector<double> xy_h;
  std::vector<float> eps_h;
  if (c->mixed) {
    xy_h.resize((size_t)2 * n);
    eps_h.resize(n);
    for (int32_t d = 0; d < n; d++) {
      xy_h[2 * (size_t)d] = base_h[(size_t)B_STRIDE * d + B_X];
      xy_h[2 * (size_t)d + 1] = base_h[(size_t)B_STRIDE * d + B_Y];
      eps_h[d] = (float)base_h[(size_t)B_STRIDE * d + B_EPS2];
    }
    ups.push_back({c->xy, xy_h.data(), xy_h.size() * sizeof(double)});
    ups.push_back({c->eps32, eps_h.data(), eps_h.size() * sizeof(float)});
  }
  if (c->tiles) {
    ups.push_back({c->conn16, conn16_h.data(), conn16_h.size() * sizeof(uint16_t)});
    ups.push_back({c->tile_off, toff.data(), toff.size() * sizeof(int32_t)});
    ups.push_back({c->tile_idx, tidx.data(), tidx.size() * sizeof(int32_t)});
  }
  for (const Up &u : ups) {
    if (!u.bytes) continue;
    if ((e = hipMemcpy(u.dst, u.src, u.bytes, hipMemcpyHostToDevice)) != hipSuccess) {
      c->err = std::string("hipMemcpy: ") + hipGetErrorString(e);
      return fail(LSKUM_EHIP);
    }
  }
  auto zero = [&](void *ptr, size_t bytes) {
    return ptr ? hipMemset(ptr, 0, bytes) : hipSuccess;
  };
  if ((e = zero(c->qm, sizeof(double) * QM_STRIDE * n)) != hipSuccess ||
      (e = zero(c->dqb[0], sizeof(double) * DQ_STRIDE * n)) != hipSuccess ||
      (e = zero(c->dqb[1], sizeof(double) * DQ_STRIDE * n)) != hipSuccess ||
      (e = zero(c->q32, sizeof(float) * 4 * n)) != hipSuccess ||
      (e = zero(c->qm32, sizeof(float) * 8 * n)) != hipSuccess ||
      (e = zero(c->dq32[0], sizeof(float) * 8 * n)) != hipSuccess ||
      (e = zero(c->dq32[1], sizeof(float) * 8 * n)) != hipSuccess ||
      (e = hipMemset(c->flux_res, 0, sizeof(double) * 4 * n)) != hipSuccess ||
      (e = hipMemset(c->delta, 0, sizeof(double) * n)) != hipSuccess) {
    c->err = std::string("hipMemset: ") + hipGetErrorString(e);
    return fail(LSKUM_EHIP);
  }
  // initial_conditions_mod.h:23-37 with parameter_mod.h:13-30
  const double pi = 4 * std::atan(1.0);
  c->theta = p.aoa_deg * pi / 180;
  c->q_inf[0] = c->rho_inf;
  c->q_inf[1] = p.mach * std::cos(c->theta);
  c->q_inf[2] = p.mach * std::sin(c->theta);
  c->q_inf[3] = c->pr_inf;
  std::vector<double> prim0((size_t)4 * n);
  for (int r = 0; r < 4; r++)
    for (int32_t i = 0; i < n; i++) prim0[(size_t)r * n + i] = c->q_inf[r];
  if ((e = hipMemcpy(c->prim, prim0.data(), prim0.size() * sizeof(double),
                     hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(c->prim_old, prim0.data(), prim0.size() * sizeof(double),
                     hipMemcpyHostToDevice)) != hipSuccess) {
    c->err = std::string("hipMemcpy prim: ") + hipGetErrorString(e);
    return fail(LSKUM_EHIP);
  }
  *out = c;
  return LSKUM_OK;
}

void lskum_destroy(lskum_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  release(ctx);
  delete ctx;
}

#define CTX_OK(ctx)                                        \
  do {                                                     \
    if (!(ctx)) return LSKUM_EINVAL;                       \
    if (!(ctx)->stream) return set_err(ctx, LSKUM_EINVAL, "context failed to initialise"); \
  } while (0)

int lskum_set_state(lskum_ctx *ctx, const double *prim) {
  CTX_OK(ctx);
  if (!prim) return set_err(ctx, LSKUM_EINVAL, "null prim");
  std::vector<double> tmp((size_t)4 * ctx->n);
  permute_in(ctx, prim, tmp.data(), 4);
  HIP_TRY(ctx, hipMemcpyAsync(ctx->prim, tmp.data(), tmp.size() * sizeof(double),
                              hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return LSKUM_OK;
}

int lskum_get_state(lskum_ctx *ctx, double *prim) {
  return lskum_get_array(ctx, LSKUM_A_PRIM, prim);
}

int lskum_snapshot(lskum_ctx *ctx) {
  CTX_OK(ctx);
  return do_snapshot(ctx);
}

int lskum_func_delta(lskum_ctx *ctx) {
  CTX_OK(ctx);
  return do_func_delta(ctx, true);
}

int lskum_q_variables(lskum_ctx *ctx) {
  CTX_OK(ctx);
  return do_q_variables(ctx);
}

int lskum_q_derivatives(lskum_ctx *ctx) {
  CTX_OK(ctx);
  return do_q_derivatives(ctx, 0);
}

int lskum_q_inner(lskum_ctx *ctx, int32_t n_inner) {
  CTX_OK(ctx);
  if (n_inner < 0) return set_err(ctx, LSKUM_EINVAL, "n_inner < 0");
  return do_q_inner(ctx, n_inner);
}

int lskum_flux_residual(lskum_ctx *ctx) {
  CTX_OK(ctx);
  return do_flux(ctx);
}

int lskum_state_update(lskum_ctx *ctx, int32_t rk, double *sum_res_sqr, double *max_res,
                       int32_t *max_res_point) {
  CTX_OK(ctx);
  if (rk < 1 || rk > 4) return set_err(ctx, LSKUM_EINVAL, "rk outside 1..4");
  int rc = do_state_update(ctx, rk, nullptr);
  if (rc) return rc;
  if (!sum_res_sqr && !max_res && !max_res_point) return LSKUM_OK;
  HIP_TRY(ctx, hipMemcpyAsync(ctx->red_h, ctx->red, sizeof(double) * 3, hipMemcpyDeviceToHost,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  // state_update_mod.h:106-107,159-163: max_res starts at 0 and the point
  // only moves on a strictly larger value
  const double mx = ctx->red_h[1] > 0.0 ? ctx->red_h[1] : 0.0;
  if (ctx->red_h[1] > 0.0) ctx->max_res_point = (int32_t)ctx->red_h[2];
  if (sum_res_sqr) *sum_res_sqr = ctx->red_h[0];
  if (max_res) *max_res = mx;
  if (max_res_point) *max_res_point = ctx->max_res_point;
  if (!std::isfinite(ctx->red_h[0]))
    return set_err(ctx, LSKUM_ENONFINITE, "residual sum is not finite");
  return LSKUM_OK;
}

int lskum_forces(lskum_ctx *ctx, double *cl, double *cd, double *cm) {
  CTX_OK(ctx);
  const double qdyn = 0.5 * ctx->rho_inf * ctx->prm.mach * ctx->prm.mach;
  launch_forces(ctx->n_shape, ctx->f_m, ctx->f_l, ctx->f_r, ctx->f_s, ctx->base, ctx->nx, ctx->ny,
                ctx->prim, ctx->n, ctx->pr_inf, qdyn, ctx->forces, ctx->stream);
  int rc = check_launch(ctx, "forces");
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(ctx->red_h + 3, ctx->forces, sizeof(double) * 3,
                              hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const double H = ctx->red_h[3], V = ctx->red_h[4], P = ctx->red_h[5];
  // compute_force_coeffs_mod.h:62-67
  const double l = V * std::cos(ctx->theta) - H * std::sin(ctx->theta);
  const double d = H * std::cos(ctx->theta) + V * std::sin(ctx->theta);
  if (cl) *cl = l;
  if (cd) *cd = d;
  if (cm) *cm = P;
  if (!std::isfinite(l) || !std::isfinite(d) || !std::isfinite(P))
    return set_err(ctx, LSKUM_ENONFINITE, "force coefficients are not finite");
  return LSKUM_OK;
}

int lskum_iterate(lskum_ctx *ctx, int32_t n_iter, double *res_new, double *residue) {
  CTX_OK(ctx);
  if (n_iter < 0) return set_err(ctx, LSKUM_EINVAL, "n_iter < 0");
  if (n_iter == 0) return LSKUM_OK;
  if (ctx->hist_cap < n_iter) {
    if (ctx->hist) (void)hipFree(ctx->hist);
    ctx->hist = nullptr;
    ctx->hist_cap = 0;
    HIP_TRY(ctx, dalloc(&ctx->hist, (size_t)n_iter));
    ctx->hist_cap = n_iter;
  }
  const int inner = ctx->prm.inner_iterations;
  for (int t = 0; t < n_iter; t++) {
    // q and the speed come from the previous stage-4 update, except on the
    // first iteration of the call (the state may have been set since)
    int rc = do_snapshot(ctx);
    if (!rc) rc = do_func_delta(ctx, t == 0);
    for (int rk = 1; rk <= 4 && !rc; rk++) {
      if (t == 0 && rk == 1) rc = do_q_variables(ctx);
      if (!rc) rc = do_q_derivatives(ctx, inner);
      if (!rc) rc = do_q_inner(ctx, inner);
      if (!rc) rc = do_flux(ctx);
      if (!rc) rc = do_state_update(ctx, rk, rk == 4 ? ctx->hist + t : nullptr);
    }
    if (rc) return rc;
  }
  std::vector<double> sums(n_iter);
  HIP_TRY(ctx, hipMemcpyAsync(sums.data(), ctx->hist, sizeof(double) * n_iter,
                              hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  int rc = collect_timing(ctx);
  if (rc) return rc;
  bool finite = true;
  for (int t = 0; t < n_iter; t++) {
    // fpi_solver_mod.h:138-147
    const int it = ctx->it + 1;
    const double rn = std::sqrt(sums[t]) / ctx->n;
    double rs;
    if (it <= 2 && ctx->restart == 0) {
      ctx->res_old = rn;
      rs = 0;
    } else {
      rs = std::log10(rn / ctx->res_old);
    }
    ctx->it = it;
    if (res_new) res_new[t] = rn;
    if (residue) residue[t] = rs;
    if (!std::isfinite(sums[t])) finite = false;
  }
  if (!finite) return set_err(ctx, LSKUM_ENONFINITE, "residual became non-finite");
  return LSKUM_OK;
}

int lskum_get_iteration(lskum_ctx *ctx, int32_t *it, double *res_old) {
  CTX_OK(ctx);
  if (it) *it = ctx->it;
  if (res_old) *res_old = ctx->res_old;
  return LSKUM_OK;
}

int lskum_set_iteration(lskum_ctx *ctx, int32_t it, double res_old, int32_t restart) {
  CTX_OK(ctx);
  if (it < 0) return set_err(ctx, LSKUM_EINVAL, "it < 0");
  ctx->it = it;
  ctx->res_old = res_old;
  ctx->restart = restart ? 1 : 0;
  return LSKUM_OK;
}

int lskum_get_array(lskum_ctx *ctx, int32_t which, double *dst) {
  CTX_OK(ctx);
  if (!dst) return set_err(ctx, LSKUM_EINVAL, "null destination");
  std::vector<double> dev;
  int rc = LSKUM_OK, ncomp = 0;
  switch (which) {
    case LSKUM_A_PRIM: rc = fetch_plain(ctx, ctx->prim, 4, dev); ncomp = 4; break;
    case LSKUM_A_PRIM_OLD: rc = fetch_plain(ctx, ctx->prim_old, 4, dev); ncomp = 4; break;
    case LSKUM_A_Q:
      rc = ctx->mixed ? fetch_float(ctx, ctx->q32, 4, dev)
                      : fetch_record(ctx, ctx->base, B_STRIDE, B_Q, 4, dev);
      ncomp = 4;
      break;
    case LSKUM_A_DQ:
      rc = ctx->mixed ? fetch_float(ctx, ctx->dq32[ctx->cur], 8, dev)
                      : fetch_record(ctx, ctx->dqb[ctx->cur], DQ_STRIDE, 0, 8, dev);
      ncomp = 8;
      break;
    case LSKUM_A_QM:
      rc = ctx->mixed ? fetch_float(ctx, ctx->qm32, 8, dev)
                      : fetch_record(ctx, ctx->qm, QM_STRIDE, 0, 8, dev);
      ncomp = 8;
      break;
    case LSKUM_A_DELTA: rc = fetch_plain(ctx, ctx->delta, 1, dev); ncomp = 1; break;
    case LSKUM_A_FLUX_RES: rc = fetch_plain(ctx, ctx->flux_res, 4, dev); ncomp = 4; break;
    case LSKUM_A_NORMALS:
      std::memcpy(dst, ctx->nx_h.data(), sizeof(double) * ctx->n);
      std::memcpy(dst + ctx->n, ctx->ny_h.data(), sizeof(double) * ctx->n);
      return LSKUM_OK;
    default: return set_err(ctx, LSKUM_EINVAL, "unknown array id");
  }
  if (rc) return rc;
  permute_out(ctx, dev.data(), dst, ncomp);
  return LSKUM_OK;
}

int lskum_get_substencil(lskum_ctx *ctx, int32_t which, int32_t *counts, int32_t *members) {
  CTX_OK(ctx);
  if (which < 0 || which > 3 || !counts) return set_err(ctx, LSKUM_EINVAL, "bad arguments");
  const uint8_t bit = (uint8_t)(1u << which);
  int64_t w = 0;
  for (int32_t i = 0; i < ctx->n; i++) {
    int32_t cnt = 0;
    for (int32_t e = ctx->conn_off_h[i]; e < ctx->conn_off_h[i + 1]; e++) {
      if (!(ctx->mask_h[e] & bit)) continue;
      if (members) members[w] = ctx->conn_h[e];
      w++;
      cnt++;
    }
    counts[i] = cnt;
  }
  return LSKUM_OK;
}

void *lskum_stream(lskum_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int lskum_layout(lskum_ctx *ctx, int32_t *tiled, int32_t *tile_max, int32_t *order) {
  CTX_OK(ctx);
  if (tiled) *tiled = (int32_t)ctx->tiles;
  if (tile_max) *tile_max = ctx->tile_max;
  if (order) *order = ctx->prm.reorder;
  return LSKUM_OK;
}

int lskum_set_timing(lskum_ctx *ctx, int32_t enable) {
  CTX_OK(ctx);
  int rc = collect_timing(ctx);
  ctx->timing = enable != 0;
  for (int k = 0; k < 6; k++) {
    ctx->k_ms[k] = 0.0;
    ctx->k_cnt[k] = 0;
  }
  return rc;
}

int lskum_kernel_time(lskum_ctx *ctx, int32_t kernel, double *avg_ms, int32_t *launches) {
  CTX_OK(ctx);
  if (kernel < 0 || kernel > 5) return set_err(ctx, LSKUM_EINVAL, "bad kernel id");
  int rc = collect_timing(ctx);
  if (rc) return rc;
  if (avg_ms) *avg_ms = ctx->k_cnt[kernel] ? ctx->k_ms[kernel] / ctx->k_cnt[kernel] : 0.0;
  if (launches) *launches = ctx->k_cnt[kernel];
  return LSKUM_OK;
}

int lskum_kat_fluxes(int32_t n, const double *in6, double *out, int32_t mixed) {
  if (n <= 0 || !in6 || !out) return LSKUM_EINVAL;
  double *din = nullptr, *dout = nullptr;
  if (dalloc(&din, (size_t)6 * n) != hipSuccess) return LSKUM_ENOMEM;
  if (dalloc(&dout, (size_t)32 * n) != hipSuccess) {
    (void)hipFree(din);
    return LSKUM_ENOMEM;
  }
  int rc = LSKUM_OK;
  if (hipMemcpy(din, in6, sizeof(double) * 6 * n, hipMemcpyHostToDevice) != hipSuccess) rc = LSKUM_EHIP;
  if (!rc) {
    if (mixed)
      launch_kat_fluxes_m(n, din, dout, nullptr);
    else
      launch_kat_fluxes(n, din, dout, nullptr);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = LSKUM_EHIP;
  }
  if (!rc && hipMemcpy(out, dout, sizeof(double) * 32 * n, hipMemcpyDeviceToHost) != hipSuccess)
    rc = LSKUM_EHIP;
  (void)hipFree(din);
  (void)hipFree(dout);
  return rc;
}

int lskum_kat_qtilde(int32_t n, const double *qt4, double *prim_out) {
  if (n <= 0 || !qt4 || !prim_out) return LSKUM_EINVAL;
  double *din = nullptr, *dout = nullptr;
  if (dalloc(&din, (size_t)4 * n) != hipSuccess) return LSKUM_ENOMEM;
  if (dalloc(&dout, (size_t)4 * n) != hipSuccess) {
    (void)hipFree(din);
    return LSKUM_ENOMEM;
  }
  int rc = LSKUM_OK;
  if (hipMemcpy(din, qt4, sizeof(double) * 4 * n, hipMemcpyHostToDevice) != hipSuccess) rc = LSKUM_EHIP;
  if (!rc) {
    launch_kat_qtilde(n, din, dout, nullptr);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = LSKUM_EHIP;
  }
  if (!rc && hipMemcpy(prim_out, dout, sizeof(double) * 4 * n, hipMemcpyDeviceToHost) != hipSuccess)
    rc = LSKUM_EHIP;
  (void)hipFree(din);
  (void)hipFree(dout);
  return rc;
}

}  // extern "C"
