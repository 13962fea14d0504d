// api.hip -- the C ABI of lskum.h: device context, host fpi_solver loop.
//
// The host loop keeps the reference's structure (fpi_solver_mod.h:14-150,
// q_lskum_mod.h:28-37): each stage is one stream-ordered kernel launch on the
// context's stream; only the per-iteration residual sums come back, once,
// at the end of lskum_iterate.
//
// A context runs n_cases independent cases (one for lskum_create, many for
// lskum_create_batch).  Cases are concatenated in device memory, each padded
// to whole 256-point blocks when there are several, so that every stage is
// one launch over all cases: the stencil kernels need no notion of a case
// (a case's neighbour lists only reach its own points), the state update
// reads its case's flow constants per block, and the residual and force
// reductions run one block per case.  Padding points have no neighbours and
// stay at their initial state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "host.hpp"
#include "kernels.hpp"
#include "lskum.h"

#include "context.hpp"

using namespace lskum;

namespace lskum {

int set_err(lskum_ctx *c, int code, const std::string &m) {
  if (c) c->err = m;
  return code;
}

void drop_graph(lskum_ctx *c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
  c->gexec_iters = 0;
  c->gexec_hist = nullptr;
}

void release(lskum_ctx *c) {
  if (!c) return;
  drop_graph(c);
  void *ptrs[] = {c->coef, c->base, c->dqb[0], c->dqb[1], c->qm, c->conn16, c->tile_off,
                  c->tile_idx, c->prim, c->prim_old, c->flux_res, c->delta, c->nx, c->ny,
                  c->flag, c->conn_off, c->perm_d, c->conn, c->part.sum, c->part.max,
                  c->part.arg, c->red, c->hist, c->f_m, c->f_l, c->f_r, c->f_s, c->forces,
                  c->xy, c->q32, c->dq32[0], c->dq32[1], c->qm32, c->eps32, c->ls32, c->fcs,
                  c->block_case, c->seg, c->soff, c->qdyn, c->counter, c->qsum,
                  c->slice_off, c->connS, c->conn16S};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  void *aptrs[] = {c->adj.own_q, c->adj.own_dq, c->adj.own_qm, c->adj.edge, c->adj.edge_mx, c->adj.lq,
                   c->adj.ldq[0], c->adj.ldq[1], c->adj.lqm, c->adj.ls, c->adj.arg,
                   c->adj.tr_off, c->adj.tr_edge, c->adj.tr_owner, c->adj.tr_pos,
                   c->adj.tr_sbase, c->adj.tr_owner_s, c->adj.tbase,
                   c->adj.tdq[0],
                   c->adj.tdq[1], c->adj.tqm, c->adj.tqs, c->adj.lsum, c->adj.vec_in,
                   c->adj.vec_out};
  for (void *p : aptrs)
    if (p) (void)hipFree(p);
  if (c->red_h) (void)hipHostFree(c->red_h);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

// Kernel timing: an event pair around a launch when enabled (never inside a
// graph capture).
struct KTimer {
  lskum_ctx *c;
  int id;
  int idx = -1;
  KTimer(lskum_ctx *ctx, int kid) : c(ctx), id(kid) {
    if (!c->timing || c->capturing) return;
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
  launch_copy(c->prim, c->prim_old, (size_t)4 * c->n, c->stream);
  return check_launch(c, "snapshot");
}

int do_speed(lskum_ctx *c) {
  launch_speed(c->dc(), c->prim, c->base, c->stream);  // both forms: FP64
  c->speed_fresh = true;
  return check_launch(c, "speed");
}

int do_func_delta(lskum_ctx *c) {
  {
    KTimer t(c, LSKUM_K_DELTA);
    if (c->mixed)
      launch_func_delta_m(c->dc(), c->base, c->prm.cfl, c->delta, c->stream);
    else
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
  c->q_fresh = true;
  return check_launch(c, "q_variables");
}

// derivatives into the buffer from which n_later Jacobi passes end in buffer 0
int do_q_derivatives(lskum_ctx *c, int n_later, bool keep_sums, bool with_delta) {
  c->cur = n_later % 2;
  {
    KTimer t(c, LSKUM_K_QDERIV);
    if (c->mixed)
      launch_q_derivatives_m(c->dc(), c->ms(), c->dq32[c->cur], c->prm.power, c->stream,
                             c->base, c->prm.cfl, with_delta ? c->delta : nullptr);
    else if (with_delta)
      launch_q_derivatives_delta(c->dc(), c->base, c->dqb[c->cur], c->qm,
                                 keep_sums ? c->qsum : nullptr, c->prm.cfl, c->delta, c->stream);
    else
      launch_q_derivatives(c->dc(), c->base, c->dqb[c->cur], c->qm, keep_sums ? c->qsum : nullptr,
                           c->prm.power, c->stream);
  }
  return check_launch(c, "q_derivatives");
}

// n Jacobi passes, each reading buffer cur and writing the other one
int do_q_inner(lskum_ctx *c, int n_inner) {
  for (int p = 0; p < n_inner; p++) {
    KTimer t(c, LSKUM_K_QINNER);
    if (c->mixed)
      launch_q_inner_m(c->dc(), c->ms(), c->dq32[c->cur], c->dq32[1 - c->cur], c->prm.power,
                       c->stream);
    else
      launch_q_inner(c->dc(), c->qsum, c->dqb[c->cur], c->dqb[1 - c->cur], c->prm.power,
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

// The state update also refreshes q (and, after the last stage, the speed
// func_delta reads), so the next stage needs no separate q_variables pass.
// with_hist: record each case's sum in hist at the slot the counter holds.
int do_state_update(lskum_ctx *c, int rk, bool with_hist) {
  {
    KTimer t(c, LSKUM_K_STATE);
    launch_state_update(c->dc(), rk, c->prim, c->prim_old, c->flux_res, c->fcs, c->block_case,
                        c->part, c->base, c->mixed ? c->q32 : nullptr, rk == 4, c->stream);
    launch_reduce(c->part, c->n_cases(), c->seg, c->red, with_hist ? c->hist_dst : nullptr,
                  c->hist_stride, 0, c->counter, c->stream);
  }
  c->q_fresh = true;
  c->speed_fresh = rk == 4;
  return check_launch(c, "state_update");
}

// One fpi_solver iteration (fpi_solver_mod.h:14-150) with q and the speed
// already fresh; advances the hist slot counter.  func_delta reads the
// iteration's starting state, as stage 1's q_derivatives does, so where the
// pipelined kernel allows it the two share one pass.
int run_iteration(lskum_ctx *c) {
  const int inner = c->prm.inner_iterations;
  const bool fused_delta = c->mixed ? c->prm.power == 0.0 : c->qdd_chunks > 0;
  int rc = do_snapshot(c);
  if (!rc && !fused_delta) rc = do_func_delta(c);
  for (int rk = 1; rk <= 4 && !rc; rk++) {
    rc = do_q_derivatives(c, inner, inner > 0, rk == 1 && fused_delta);
    if (!rc) rc = do_q_inner(c, inner);
    if (!rc) rc = do_flux(c);
    if (!rc) rc = do_state_update(c, rk, rk == 4);
  }
  if (!rc) {
    launch_tick(c->counter, c->stream);
    rc = check_launch(c, "tick");
  }
  return rc;
}

// Captures gexec_iters = K whole iterations as one graph.
int capture_graph(lskum_ctx *c, int K) {
  drop_graph(c);
  HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  int rc = LSKUM_OK;
  for (int k = 0; k < K && !rc; k++) rc = run_iteration(c);
  c->capturing = false;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return set_err(c, LSKUM_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->gexec = nullptr;
    return set_err(c, LSKUM_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  c->gexec_iters = K;
  c->gexec_hist = c->hist;
  return LSKUM_OK;
}

// per-point arrays of all cases <-> the caller's layout: case after case,
// each [ncomp][n_case] in the case's caller order
void permute_out(const lskum_ctx *c, const double *src, double *dst, int ncomp) {
  size_t o = 0;
  for (const CaseInfo &cs : c->cases) {
    const CloudLayout &L = c->clouds[cs.cloud];
    for (int r = 0; r < ncomp; r++)
      for (int32_t d = 0; d < cs.n; d++)
        dst[o + (size_t)r * cs.n + L.perm[d]] = src[(size_t)r * c->n + cs.p0 + d];
    o += (size_t)ncomp * cs.n;
  }
}

// the inverse; padding points keep what dst held
void permute_in(const lskum_ctx *c, const double *src, double *dst, int ncomp) {
  size_t o = 0;
  for (const CaseInfo &cs : c->cases) {
    const CloudLayout &L = c->clouds[cs.cloud];
    for (int r = 0; r < ncomp; r++)
      for (int32_t d = 0; d < cs.n; d++)
        dst[(size_t)r * c->n + cs.p0 + d] = src[o + (size_t)r * cs.n + L.perm[d]];
    o += (size_t)ncomp * cs.n;
  }
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

bool params_ok(const lskum_params &p) {
  return p.mach > 0.0 && p.cfl > 0.0 && p.vl_const >= 0.0 && p.inner_iterations >= 0 &&
         (p.precision == LSKUM_FP64 || p.precision == LSKUM_MIXED) &&
         (p.reorder >= LSKUM_REORDER_NONE && p.reorder <= LSKUM_REORDER_HILBERT) &&
         (p.lds_tiles == LSKUM_TILES_AUTO || (p.lds_tiles >= 0 && p.lds_tiles <= 15)) &&
         (p.tile_block == 0 || p.tile_block == 64 || p.tile_block == 128 ||
          p.tile_block == 256) &&
         (p.points_per_thread == 0 || p.points_per_thread == 1 || p.points_per_thread == 2 ||
          p.points_per_thread == 4);
}

// Least-squares coefficients of each sub-stencil (geometry only): the sums
// of interior_fluxes_mod.h:51-61 over the members in conn order, then
// (syy, -sxy)/det for the x-form solve and (sxx, -sxy)/det for the y-form.
void compute_coef(const CloudLayout &L, const double *x, const double *y, const double *nxh,
                  const double *nyh, double power, std::vector<double> &coef) {
  const int32_t n = L.n;
  coef.assign((size_t)8 * n, 0.0);
  for (int32_t d = 0; d < n; d++) {
    const int32_t v = L.perm[d];
    const double nxv = nxh[v], nyv = nyh[v], tx = nyv, ty = -nxv;
    double sxx[4] = {0, 0, 0, 0}, syy[4] = {0, 0, 0, 0}, sxy[4] = {0, 0, 0, 0};
    for (int32_t t = L.conn_off_h[v]; t < L.conn_off_h[v + 1]; t++) {
      const int32_t k = L.conn_h[t];
      const double delx = x[k] - x[v], dely = y[k] - y[v];
      const double dels = delx * tx + dely * ty, deln = delx * nxv + dely * nyv;
      double dsw = dels, dnw = deln;
      if (power != 0.0) {
        const double w = std::pow(std::sqrt(dels * dels + deln * deln), power);
        dsw = dels * w;
        dnw = deln * w;
      }
      for (int sb = 0; sb < 4; sb++) {
        if (!(L.mask_h[t] & (1u << sb))) continue;
        sxx[sb] = sxx[sb] + dels * dsw;
        syy[sb] = syy[sb] + deln * dnw;
        sxy[sb] = sxy[sb] + dels * dnw;
      }
    }
    for (int sb = 0; sb < 4; sb++) {
      const double det = sxx[sb] * syy[sb] - sxy[sb] * sxy[sb];
      const double inv = det != 0.0 ? 1.0 / det : 0.0;
      coef[(size_t)8 * d + 2 * sb] = (sb < 2 ? syy[sb] : sxx[sb]) * inv;
      coef[(size_t)8 * d + 2 * sb + 1] = -sxy[sb] * inv;
    }
  }
}

// The mixed form's least-squares coefficients of the full stencil (geometry
// only): the sums of q_variables_mod.h:66-81 over the neighbours in conn
// order, in FP64, then (syy, -sxy, sxx)/det rounded to FP32 (padding: 0).
void compute_ls_coef(const CloudLayout &L, const double *x, const double *y, double power,
                     std::vector<float> &ls) {
  const int32_t n = L.n;
  ls.assign((size_t)4 * n, 0.0f);
  for (int32_t d = 0; d < n; d++) {
    const int32_t v = L.perm[d];
    double sxx = 0.0, syy = 0.0, sxy = 0.0;
    for (int32_t t = L.conn_off_h[v]; t < L.conn_off_h[v + 1]; t++) {
      const int32_t k = L.conn_h[t];
      const double delx = x[k] - x[v], dely = y[k] - y[v];
      const double w =
          power != 0.0 ? std::pow(std::sqrt(delx * delx + dely * dely), power) : 1.0;
      sxx = sxx + delx * delx * w;
      syy = syy + dely * dely * w;
      sxy = sxy + delx * dely * w;
    }
    const double det = sxx * syy - sxy * sxy;
    const double inv = det != 0.0 ? 1.0 / det : 0.0;
    ls[(size_t)4 * d] = (float)(syy * inv);
    ls[(size_t)4 * d + 1] = (float)(-sxy * inv);
    ls[(size_t)4 * d + 2] = (float)(sxx * inv);
  }
}

// Host set-up of one cloud: normals and sub-stencils in the caller's order,
// stencil validation, device order, device-order images, least-squares
// coefficients, force panels.
int build_layout(const lskum_cloud &cloud, const lskum_params &p, CloudLayout &L,
                 std::string &err) {
  const int32_t n = cloud.n_points;
  L.n = n;
  compute_normals(cloud, L.nx_h, L.ny_h);
  generate_connectivity(cloud, L.nx_h, L.ny_h, L.mask_h);
  for (int32_t i = 0; i < n; i++) {
    double r = substencil_min_rel_det(cloud, L.nx_h, L.ny_h, L.mask_h, i);
    if (!(r > 0.0)) {
      char m[160];
      std::snprintf(m, sizeof m,
                    "degenerate sub-stencil at point %d (flag_1 %d): empty or collinear", i,
                    cloud.flag_1[i]);
      err = m;
      return LSKUM_ESTENCIL;
    }
  }
  L.conn_off_h.assign(cloud.conn_off, cloud.conn_off + n + 1);
  L.conn_h.assign(cloud.conn, cloud.conn + cloud.conn_off[n]);
  // device order: Hilbert (compact 256-point blocks whose neighbour sets fit
  // an LDS tile), or RCM / the caller's order grouped by point type
  std::vector<int32_t> order;
  if (p.reorder == LSKUM_REORDER_HILBERT) {
    hilbert_order(cloud, L.perm);
  } else {
    if (p.reorder == LSKUM_REORDER_RCM) {
      rcm_order(cloud, order);
    } else {
      order.resize(n);
      for (int32_t i = 0; i < n; i++) order[i] = i;
    }
    L.perm.clear();
    L.perm.reserve(n);
    for (int t = 0; t < 3; t++)
      for (int32_t v : order)
        if (cloud.flag_1[v] == t) L.perm.push_back(v);
  }
  L.inv.assign(n, 0);
  for (int32_t d = 0; d < n; d++) L.inv[L.perm[d]] = d;
  L.x.resize(n);
  L.y.resize(n);
  L.eps2.resize(n);
  L.nx.resize(n);
  L.ny.resize(n);
  L.flag.resize(n);
  L.off.assign(n + 1, 0);
  L.conn.resize(cloud.conn_off[n]);
  for (int32_t d = 0; d < n; d++) {
    const int32_t v = L.perm[d];
    L.x[d] = cloud.x[v];
    L.y[d] = cloud.y[v];
    // limiters_mod.h:82-83: epsi = pow(VL_CONST*min_dist, 3), used squared
    const double epsi = std::pow(p.vl_const * cloud.min_dist[v], 3);
    L.eps2[d] = epsi * epsi;
    L.nx[d] = L.nx_h[v];
    L.ny[d] = L.ny_h[v];
    L.flag[d] = (uint8_t)cloud.flag_1[v];
    const int32_t b = cloud.conn_off[v], e = cloud.conn_off[v + 1];
    L.off[d + 1] = L.off[d] + (e - b);
    for (int32_t t = b; t < e; t++)
      L.conn[L.off[d] + (t - b)] =
          (uint32_t)L.inv[cloud.conn[t]] | ((uint32_t)L.mask_h[t] << kMaskShift);
  }
  L.left_h.assign(cloud.left, cloud.left + n);
  L.right_h.assign(cloud.right, cloud.right + n);
  L.flag1_h.assign(cloud.flag_1, cloud.flag_1 + n);
  L.flag2_h.assign(cloud.flag_2, cloud.flag_2 + n);
  L.min_dist_h.assign(cloud.min_dist, cloud.min_dist + n);
  compute_coef(L, cloud.x, cloud.y, L.nx_h.data(), L.ny_h.data(), p.power, L.coef);
  if (p.precision == LSKUM_MIXED) compute_ls_coef(L, cloud.x, cloud.y, p.power, L.ls);
  // force panels: the first shape_points points (compute_force_coeffs_mod.h:
  // 25-27 takes m = j)
  int shape_points = 0;
  for (int32_t i = 0; i < n; i++)
    if (cloud.flag_2[i] > 0) shape_points++;
  for (int j = 0; j < shape_points; j++) {
    if (cloud.flag_1[j] == 1) continue;  // no left/right: cannot contribute a panel
    L.fm.push_back(L.inv[j]);
    L.fl.push_back(L.inv[cloud.left[j]]);
    L.fr.push_back(L.inv[cloud.right[j]]);
    L.fs.push_back(cloud.flag_2[j]);
  }
  return LSKUM_OK;
}

// The context of n_cases cases over n_clouds clouds.
int create_impl(const lskum_cloud *clouds, int32_t n_clouds, const lskum_case *cases,
                int32_t n_cases, const lskum_params *params, lskum_ctx **out) {
  if (!out) return LSKUM_EINVAL;
  *out = nullptr;
  if (!clouds || !cases || !params || n_clouds < 1 || n_cases < 1) return LSKUM_EINVAL;
  lskum_ctx *c = new (std::nothrow) lskum_ctx;
  if (!c) return LSKUM_ENOMEM;
  auto fail = [&](int code, const std::string &m) {
    c->err = m;
    *out = c;  // keep the message reachable: return the context in its failed state
    return code;
  };
  const lskum_params &p = *params;
  if (!params_ok(p)) return fail(LSKUM_EINVAL, "invalid parameters");
  if (n_cases > 65535) return fail(LSKUM_EINVAL, "more than 65535 cases");
  for (int32_t j = 0; j < n_cases; j++)
    if (cases[j].cloud < 0 || cases[j].cloud >= n_clouds || !(cases[j].mach > 0.0))
      return fail(LSKUM_EINVAL, "case " + std::to_string(j) + ": bad cloud index or mach");
  c->prm = p;
  c->mixed = p.precision == LSKUM_MIXED;
  c->tile_block = p.tile_block ? p.tile_block : kTileBlockDefault;
  c->su_ppt = p.points_per_thread ? p.points_per_thread : 1;
  c->clouds.resize(n_clouds);
  for (int32_t ci = 0; ci < n_clouds; ci++) {
    std::string err;
    int rc = check_cloud(clouds[ci], &err);
    if (!rc) rc = build_layout(clouds[ci], p, c->clouds[ci], err);
    if (rc) return fail(rc, n_clouds > 1 ? "cloud " + std::to_string(ci) + ": " + err : err);
  }
  // concatenation: each case padded to whole 256-point blocks when batched
  const int pad = n_cases > 1 ? 256 : 1;
  const double pi = 4 * std::atan(1.0);
  int64_t total = 0, entries = 0;
  c->cases.resize(n_cases);
  for (int32_t j = 0; j < n_cases; j++) {
    CaseInfo &cs = c->cases[j];
    const CloudLayout &L = c->clouds[cases[j].cloud];
    cs.cloud = cases[j].cloud;
    cs.p0 = (int32_t)total;
    cs.n = L.n;
    cs.np = (L.n + pad - 1) / pad * pad;
    cs.mach = cases[j].mach;
    cs.aoa = cases[j].aoa_deg;
    // initial_conditions_mod.h:23-37 with parameter_mod.h:13-30
    cs.theta = cs.aoa * pi / 180;
    cs.q_inf[0] = c->rho_inf;
    cs.q_inf[1] = cs.mach * std::cos(cs.theta);
    cs.q_inf[2] = cs.mach * std::sin(cs.theta);
    cs.q_inf[3] = c->pr_inf;
    total += cs.np;
    entries += L.off[L.n];
  }
  if (total > (int64_t)kIndexMask || entries > (int64_t)INT32_MAX)
    return fail(LSKUM_EINVAL, "batch too large for 28-bit point / 31-bit entry indices");
  const int32_t n = (int32_t)total;
  c->n = n;
  std::vector<double> base_h((size_t)B_STRIDE * n, 0.0), nxd(n, 0.0), nyd(n, 1.0),
      coef_h((size_t)8 * n, 0.0), prim0((size_t)4 * n);
  std::vector<float> ls_h(c->mixed ? (size_t)4 * n : 0, 0.0f);
  std::vector<uint8_t> flag_h(n, 1);
  std::vector<int32_t> off_h(n + 1, 0), perm_h(n);
  std::vector<uint32_t> conn_d((size_t)entries);
  std::vector<int32_t> fm, fl, fr, fs;
  std::vector<int32_t> soff_h(n_cases + 1, 0);
  for (int32_t j = 0; j < n_cases; j++) {
    CaseInfo &cs = c->cases[j];
    const CloudLayout &L = c->clouds[cs.cloud];
    for (int32_t d = 0; d < cs.np; d++) {
      const int32_t g = cs.p0 + d;
      for (int r = 0; r < 4; r++) prim0[(size_t)r * n + g] = cs.q_inf[r];
      if (d >= cs.n) {  // padding: no neighbours, interior frame, its own state
        off_h[g + 1] = off_h[g];
        perm_h[g] = cs.n + (d - cs.n);
        continue;
      }
      base_h[(size_t)B_STRIDE * g + B_X] = L.x[d];
      base_h[(size_t)B_STRIDE * g + B_Y] = L.y[d];
      base_h[(size_t)B_STRIDE * g + B_EPS2] = L.eps2[d];
      nxd[g] = L.nx[d];
      nyd[g] = L.ny[d];
      flag_h[g] = L.flag[d];
      perm_h[g] = L.perm[d];
      std::copy(L.coef.begin() + (size_t)8 * d, L.coef.begin() + (size_t)8 * (d + 1),
                coef_h.begin() + (size_t)8 * g);
      if (c->mixed)
        std::copy(L.ls.begin() + (size_t)4 * d, L.ls.begin() + (size_t)4 * (d + 1),
                  ls_h.begin() + (size_t)4 * g);
      const int32_t cnt = L.off[d + 1] - L.off[d];
      off_h[g + 1] = off_h[g] + cnt;
      for (int32_t t = 0; t < cnt; t++) {
        const uint32_t w = L.conn[L.off[d] + t];
        conn_d[off_h[g] + t] = ((w & kIndexMask) + (uint32_t)cs.p0) | (w & ~kIndexMask);
      }
    }
    cs.s0 = (int32_t)fm.size();
    for (size_t k = 0; k < L.fm.size(); k++) {
      fm.push_back(L.fm[k] + cs.p0);
      fl.push_back(L.fl[k] + cs.p0);
      fr.push_back(L.fr[k] + cs.p0);
      fs.push_back(L.fs[k]);
    }
    cs.s1 = (int32_t)fm.size();
    soff_h[j + 1] = cs.s1;
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
  // The lists in slices for the forward stencil kernels (entry e of the
  // lane t of stencil block blk at slice_off[blk] + e * TB + t), each list
  // padded with its last entry to its block's longest; a point without
  // neighbours (batch padding) points at itself.
  std::vector<int32_t> slice_h(nblk + 1, 0);
  std::vector<uint32_t> connS_h;
  std::vector<uint16_t> conn16S_h;
  {
    int64_t total_s = 0;
    for (int blk = 0; blk < nblk; blk++) {
      const int32_t p0 = blk * TB, p1 = std::min(n, p0 + TB);
      int32_t width = 0;
      for (int32_t d = p0; d < p1; d++) width = std::max(width, off_h[d + 1] - off_h[d]);
      c->slice_width = std::max(c->slice_width, width);
      if (total_s > (int64_t)INT32_MAX) break;
      slice_h[blk] = (int32_t)total_s;
      total_s += (int64_t)width * TB;
    }
    if (total_s > (int64_t)INT32_MAX)
      return fail(LSKUM_EINVAL, "neighbour slices exceed 31-bit entry indices");
    slice_h[nblk] = (int32_t)total_s;
    connS_h.assign((size_t)total_s, 0u);
    conn16S_h.assign((size_t)total_s, 0);
    for (int blk = 0; blk < nblk; blk++) {
      const int32_t p0 = blk * TB, p1 = std::min(n, p0 + TB);
      const int64_t width = (int64_t)slice_h[blk + 1] - slice_h[blk];
      const int32_t w = (int32_t)(width / TB);
      for (int32_t d = p0; d < p1; d++) {
        const int32_t cnt = off_h[d + 1] - off_h[d];
        for (int32_t e = 0; e < w; e++) {
          const size_t at = (size_t)slice_h[blk] + (size_t)e * TB + (d - p0);
          if (cnt == 0) {
            connS_h[at] = (uint32_t)d;
            conn16S_h[at] = (uint16_t)(d - p0);
            continue;
          }
          const int32_t src = off_h[d] + std::min(e, cnt - 1);
          connS_h[at] = conn_d[src];
          conn16S_h[at] = tiles_ok ? conn16_h[src] : (uint16_t)0;
        }
      }
    }
  }
  // The gather kernels stage neighbours in LDS.  The flux stages part of
  // each neighbour record (FP64: dq, qmax/qmin; mixed: all but the
  // coordinates), at most flux_cap entries of a tile: as many as let a CU
  // hold the blocks the kernel's registers allow (a larger tile's further
  // entries are read through L2).  Without room for a block's own points it
  // reads every record through L2.
  int dev = 0, lds_limit = 0, lds_per_cu = 0;
  const bool dev_ok =
      hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
          hipSuccess &&
      hipDeviceGetAttribute(&lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                            dev) == hipSuccess;
  if (tiles_ok && dev_ok)
    c->flux_cap =
        c->mixed ? flux_tile_cap_m(TB, c->tile_max, (size_t)lds_limit, (size_t)lds_per_cu)
                 : flux_tile_cap(TB, c->tile_max, (size_t)lds_limit, (size_t)lds_per_cu);
  const unsigned auto_mask = LSKUM_TILE_QDERIV | LSKUM_TILE_QINNER | LSKUM_TILE_DELTA |
                             (c->flux_cap > 0 ? LSKUM_TILE_FLUX : 0u);
  unsigned want = p.lds_tiles == LSKUM_TILES_AUTO ? auto_mask : (unsigned)p.lds_tiles;
  if (c->flux_cap <= 0) want &= ~(unsigned)LSKUM_TILE_FLUX;
  if (tiles_ok && want && dev_ok) {
    if (c->mixed)
      c->tiles = prepare_tiled_m(want, c->tile_max, c->flux_cap, (size_t)lds_limit);
    else
      c->tiles = prepare_tiled(want, c->tile_max, c->flux_cap, (size_t)lds_limit);
  }
  // the pipelined inner pass and q_derivatives (with func_delta in stage
  // 1's), where the tiles are in use for them
  int n_cu = 0;
  if (p.power == 0.0 && dev_ok &&
      hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess) {
    const size_t lim = (size_t)lds_limit;
    const int tm = c->tile_max, sw = c->slice_width;
    // the mixed form's tiles (FP32 records) let more blocks share a CU
    // than the pipelined kernels' registers and list slices do: pipelined
    // mixed kernels were slower, so it keeps its own (whose stage-1
    // q_derivatives takes func_delta's minimum itself)
    if (!c->mixed) {
      if (c->tiles & LSKUM_TILE_QINNER)
        c->qi_chunks = prepare_qinner_pf(TB, tm, sw, nblk, n_cu, lim, &c->qi_grid);
      if (c->tiles & LSKUM_TILE_QDERIV) {
        c->qd_chunks = prepare_qderiv_pf(TB, tm, sw, nblk, n_cu, lim, false, &c->qd_grid);
        if (c->qd_chunks)
          c->qdd_chunks = prepare_qderiv_pf(TB, tm, sw, nblk, n_cu, lim, true, &c->qdd_grid);
      }
    }
  }
  // per-case constants and reduction ranges
  c->part.nblocks = blocks_for(n, 256);
  std::vector<FlowConsts> fcs_h(n_cases);
  std::vector<double> qdyn_h(n_cases);
  std::vector<int32_t> seg_h(n_cases + 1, 0);
  std::vector<uint16_t> bcase_h;
  if (n_cases > 1) bcase_h.resize(c->part.nblocks);
  for (int32_t j = 0; j < n_cases; j++) {
    const CaseInfo &cs = c->cases[j];
    FlowConsts &f = fcs_h[j];
    f.cfl = p.cfl;
    f.power = p.power;
    f.u1_inf = cs.q_inf[1];
    f.u2_inf = cs.q_inf[2];
    f.rho_inf = c->rho_inf;
    f.pr_inf = c->pr_inf;
    qdyn_h[j] = 0.5 * c->rho_inf * cs.mach * cs.mach;
    seg_h[j] = cs.p0 / 256;
    seg_h[j + 1] = n_cases > 1 ? (cs.p0 + cs.np) / 256 : c->part.nblocks;
    for (int32_t b = seg_h[j]; b < seg_h[j + 1] && n_cases > 1; b++) bcase_h[b] = (uint16_t)j;
  }
  // device allocation and upload
  hipError_t e;
#define ALLOC(ptr, count)                                                              \
  if ((e = dalloc(&(ptr), (count))) != hipSuccess)                                     \
    return fail(LSKUM_ENOMEM, std::string("hipMalloc " #ptr ": ") + hipGetErrorString(e));
  if ((e = hipGetDevice(&c->device)) != hipSuccess)
    return fail(LSKUM_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  {
    std::string serr;
    const int src = noalloc_scratch_init(c->device, serr);
    if (src) return fail(src, serr);
  }
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
    return fail(LSKUM_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  ALLOC(c->base, (size_t)B_STRIDE * n);
  ALLOC(c->xy, (size_t)2 * n);
  if (c->mixed) {
    ALLOC(c->q32, (size_t)4 * n);
    ALLOC(c->dq32[0], (size_t)8 * n);
    ALLOC(c->dq32[1], (size_t)8 * n);
    ALLOC(c->qm32, (size_t)8 * n);
    ALLOC(c->eps32, (size_t)n);
    ALLOC(c->ls32, (size_t)4 * n);
  } else {
    ALLOC(c->dqb[0], (size_t)DQ_STRIDE * n);
    ALLOC(c->dqb[1], (size_t)DQ_STRIDE * n);
    ALLOC(c->qm, (size_t)QM_STRIDE * n);
    ALLOC(c->qsum, (size_t)8 * n);
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
  ALLOC(c->slice_off, slice_h.size());
  ALLOC(c->connS, connS_h.size());
  ALLOC(c->conn16S, conn16S_h.size());
  if (c->tiles) {
    ALLOC(c->conn16, conn16_h.size());
    ALLOC(c->tile_off, toff.size());
    ALLOC(c->tile_idx, tidx.size());
  }
  ALLOC(c->part.sum, (size_t)c->part.nblocks);
  ALLOC(c->part.max, (size_t)c->part.nblocks);
  ALLOC(c->part.arg, (size_t)c->part.nblocks);
  ALLOC(c->red, (size_t)3 * n_cases);
  ALLOC(c->forces, (size_t)3 * n_cases);
  ALLOC(c->fcs, (size_t)n_cases);
  ALLOC(c->qdyn, (size_t)n_cases);
  ALLOC(c->soff, (size_t)n_cases + 1);
  ALLOC(c->counter, 1);
  if (n_cases > 1) {
    ALLOC(c->seg, (size_t)n_cases + 1);
    ALLOC(c->block_case, bcase_h.size());
  }
  ALLOC(c->f_m, fm.size());
  ALLOC(c->f_l, fm.size());
  ALLOC(c->f_r, fm.size());
  ALLOC(c->f_s, fm.size());
#undef ALLOC
  if ((e = hipHostMalloc((void **)&c->red_h, sizeof(double) * 6 * (size_t)n_cases)) !=
      hipSuccess)
    return fail(LSKUM_ENOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  struct Up { void *dst; const void *src; size_t bytes; };
  std::vector<Up> ups = {{c->base, base_h.data(), base_h.size() * sizeof(double)},
                         {c->nx, nxd.data(), n * sizeof(double)},
                         {c->coef, coef_h.data(), coef_h.size() * sizeof(double)},
                         {c->ny, nyd.data(), n * sizeof(double)},
                         {c->flag, flag_h.data(), (size_t)n},
                         {c->conn_off, off_h.data(), (n + 1) * sizeof(int32_t)},
                         {c->perm_d, perm_h.data(), n * sizeof(int32_t)},
                         {c->conn, conn_d.data(), conn_d.size() * sizeof(uint32_t)},
                         {c->slice_off, slice_h.data(), slice_h.size() * sizeof(int32_t)},
                         {c->connS, connS_h.data(), connS_h.size() * sizeof(uint32_t)},
                         {c->conn16S, conn16S_h.data(), conn16S_h.size() * sizeof(uint16_t)},
                         {c->f_m, fm.data(), fm.size() * sizeof(int32_t)},
                         {c->f_l, fl.data(), fl.size() * sizeof(int32_t)},
                         {c->f_r, fr.data(), fr.size() * sizeof(int32_t)},
                         {c->f_s, fs.data(), fs.size() * sizeof(int32_t)},
                         {c->fcs, fcs_h.data(), fcs_h.size() * sizeof(FlowConsts)},
                         {c->qdyn, qdyn_h.data(), qdyn_h.size() * sizeof(double)},
                         {c->soff, soff_h.data(), soff_h.size() * sizeof(int32_t)},
                         {c->prim, prim0.data(), prim0.size() * sizeof(double)},
                         {c->prim_old, prim0.data(), prim0.size() * sizeof(double)}};
  if (n_cases > 1) {
    ups.push_back({c->seg, seg_h.data(), seg_h.size() * sizeof(int32_t)});
    ups.push_back({c->block_case, bcase_h.data(), bcase_h.size() * sizeof(uint16_t)});
  }
  std::vector<double> xy_h((size_t)2 * n);
  for (int32_t d = 0; d < n; d++) {
    xy_h[2 * (size_t)d] = base_h[(size_t)B_STRIDE * d + B_X];
    xy_h[2 * (size_t)d + 1] = base_h[(size_t)B_STRIDE * d + B_Y];
  }
  ups.push_back({c->xy, xy_h.data(), xy_h.size() * sizeof(double)});
  std::vector<float> eps_h;
  if (c->mixed) {
    eps_h.resize(n);
    for (int32_t d = 0; d < n; d++) eps_h[d] = (float)base_h[(size_t)B_STRIDE * d + B_EPS2];
    ups.push_back({c->eps32, eps_h.data(), eps_h.size() * sizeof(float)});
    ups.push_back({c->ls32, ls_h.data(), ls_h.size() * sizeof(float)});
  }
  if (c->tiles) {
    ups.push_back({c->conn16, conn16_h.data(), conn16_h.size() * sizeof(uint16_t)});
    ups.push_back({c->tile_off, toff.data(), toff.size() * sizeof(int32_t)});
    ups.push_back({c->tile_idx, tidx.data(), tidx.size() * sizeof(int32_t)});
  }
  for (const Up &u : ups) {
    if (!u.bytes) continue;
    if ((e = hipMemcpy(u.dst, u.src, u.bytes, hipMemcpyHostToDevice)) != hipSuccess)
      return fail(LSKUM_EHIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
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
      (e = hipMemset(c->delta, 0, sizeof(double) * n)) != hipSuccess ||
      (e = hipMemset(c->counter, 0, sizeof(int32_t))) != hipSuccess)
    return fail(LSKUM_EHIP, std::string("hipMemset: ") + hipGetErrorString(e));
  *out = c;
  return LSKUM_OK;
}

// Adjoint working arrays and the transpose of conn (single-case FP64
// contexts with power = 0).
int prepare_adjoint(lskum_ctx *c) {
  if (c->n_cases() != 1 || c->mixed || c->prm.power != 0.0)
    return set_err(c, LSKUM_EINVAL,
                   "the adjoint and tangent need a single-case FP64 context with power 0");
  if (c->adj_ready) return LSKUM_OK;
  const CloudLayout &L = c->clouds[0];
  const int32_t n = c->n;
  const size_t E = (size_t)L.off[n];
  std::vector<int32_t> tr_off(n + 1, 0), tr_edge(E), tr_owner(E), tr_pos(E);
  for (size_t e = 0; e < E; e++) tr_off[(L.conn[e] & kIndexMask) + 1]++;
  for (int32_t j = 0; j < n; j++) tr_off[j + 1] += tr_off[j];
  std::vector<int32_t> pos(tr_off.begin(), tr_off.end() - 1);
  for (int32_t i = 0; i < n; i++)
    for (int32_t e = L.off[i]; e < L.off[i + 1]; e++) {
      const int32_t j = (int32_t)(L.conn[e] & kIndexMask);
      tr_edge[pos[j]] = e;
      tr_owner[pos[j]++] = i;
    }
  // record slots: per block of 256 points, its points' r-th records side by
  // side, as many rows as its largest transposed degree
  const int32_t nb = (n + 255) / 256;
  std::vector<int32_t> sbase(nb);
  int64_t slots = 0;
  for (int32_t B = 0; B < nb; B++) {
    int32_t dmax = 0;
    for (int32_t j = 256 * B; j < std::min(n, 256 * B + 256); j++)
      dmax = std::max(dmax, tr_off[j + 1] - tr_off[j]);
    sbase[B] = (int32_t)slots;
    slots += (int64_t)256 * dmax;
    if (slots > (int64_t)INT32_MAX)
      return set_err(c, LSKUM_EINVAL, "too many pairs for the adjoint's 31-bit record slots");
  }
  const size_t S = (size_t)slots;
  std::vector<int32_t> owner_s(S, 0);
  for (int32_t j = 0; j < n; j++)
    for (int32_t t = tr_off[j]; t < tr_off[j + 1]; t++) {
      const int32_t slot = sbase[j / 256] + 256 * (t - tr_off[j]) + j % 256;
      tr_pos[tr_edge[t]] = slot;
      owner_s[slot] = tr_owner[t];
    }
  AdjointBuffers &a = c->adj;
  hipError_t e;
#define AALLOC(ptr, count)                                                   \
  if ((e = dalloc(&(ptr), (count))) != hipSuccess)                           \
    return set_err(c, LSKUM_ENOMEM, std::string("hipMalloc " #ptr ": ") + hipGetErrorString(e));
  AALLOC(a.own_q, (size_t)4 * n);
  AALLOC(a.own_dq, (size_t)8 * n);
  AALLOC(a.own_qm, (size_t)8 * n);
  AALLOC(a.edge, EDGE_REC * S);
  AALLOC(a.edge_mx, S);
  AALLOC(a.lq, (size_t)4 * n);
  AALLOC(a.ldq[0], (size_t)8 * n);
  AALLOC(a.ldq[1], (size_t)8 * n);
  AALLOC(a.lqm, (size_t)8 * n);
  AALLOC(a.ls, (size_t)8 * n);
  AALLOC(a.arg, (size_t)8 * n);
  AALLOC(a.tr_off, (size_t)n + 1);
  AALLOC(a.tr_edge, E);
  AALLOC(a.tr_owner, E);
  AALLOC(a.tr_pos, E);
  AALLOC(a.tr_sbase, (size_t)nb);
  AALLOC(a.tr_owner_s, S);
  AALLOC(a.tbase, (size_t)B_STRIDE * n);
  AALLOC(a.tdq[0], (size_t)DQ_STRIDE * n);
  AALLOC(a.tdq[1], (size_t)DQ_STRIDE * n);
  AALLOC(a.tqm, (size_t)8 * n);
  AALLOC(a.tqs, (size_t)8 * n);
  AALLOC(a.lsum, (size_t)4 * n);
  AALLOC(a.vec_in, (size_t)4 * n);
  AALLOC(a.vec_out, (size_t)4 * n);
#undef AALLOC
  HIP_TRY(c, hipMemcpy(a.tr_off, tr_off.data(), tr_off.size() * sizeof(int32_t),
                       hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(a.tr_edge, tr_edge.data(), E * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(a.tr_owner, tr_owner.data(), E * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(a.tr_pos, tr_pos.data(), E * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(a.tr_sbase, sbase.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(a.tr_owner_s, owner_s.data(), S * sizeof(int32_t),
                       hipMemcpyHostToDevice));
  int dev = 0, lds_limit = 0;
  a.vjp_tiled =
      hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
          hipSuccess &&
      prepare_flux_vjp_tiles(c->dc(), (size_t)lds_limit);
  a.ls_tiled = lds_limit > 0 && prepare_ls_adj_tiles(c->dc(), (size_t)lds_limit);
  c->adj_ready = true;
  return LSKUM_OK;
}

// q, dq (after the inner passes) and qmax/qmin of the current state, and
// the origin of every extremum: the checkpoint recomputation
int recompute_stages(lskum_ctx *c) {
  const int inner = c->prm.inner_iterations;
  int rc = do_q_variables(c);
  if (rc) return rc;
  // q_derivatives with the extrema's origins, into the buffer from which
  // the inner passes end in buffer 0
  c->cur = inner % 2;
  launch_q_derivatives_arg(c->dc(), c->base, c->dqb[c->cur], c->qm, inner > 0 ? c->qsum : nullptr,
                           c->adj.arg, c->adj.lsum, c->stream);
  rc = check_launch(c, "q_derivatives");
  return rc ? rc : do_q_inner(c, inner);
}

int upload_caller(lskum_ctx *c, const double *src, double *dev) {
  std::vector<double> tmp((size_t)4 * c->n, 0.0);
  permute_in(c, src, tmp.data(), 4);
  HIP_TRY(c, hipMemcpyAsync(dev, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return LSKUM_OK;
}

int download_caller(lskum_ctx *c, const double *dev, double *dst) {
  std::vector<double> tmp;
  int rc = fetch_plain(c, dev, 4, tmp);
  if (rc) return rc;
  permute_out(c, tmp.data(), dst, 4);
  return LSKUM_OK;
}

}  // namespace lskum

namespace lskum {

// The scratch array of lskum_iterate_noalloc: the per-iteration residual
// sums of the running call, [n_cases][n_iter].  Each device holds its own
// copy; host code reaches a device's copy only through the address
// hipGetSymbolAddress gave for it (the array's name is a host-side stand-in).
static __device__ __align__(256) double scratch[LSKUM_NOALLOC_HIST_MAX];

constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Looks up the current device's copy once per device (lskum_create calls it
// with the context's device current).
int noalloc_scratch_init(int device, std::string &err) {
  if (device < 0 || device >= kMaxDevices) {
    err = "device index " + std::to_string(device) + " beyond the scratch table";
    return LSKUM_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return LSKUM_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return LSKUM_EHIP;
  }
  g_scratch_ptr[device] = (double *)p;
  return LSKUM_OK;
}

double *noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

// n_iter whole iterations from the current state; stage-4 sums go to
// hist_dst (stride hist_stride).  Graph replays only when allowed (the
// first replay of a new graph instantiates it).
int launch_iterations(lskum_ctx *ctx, int n_iter, bool allow_graph) {
  int rc = LSKUM_OK;
  // q and the speed of the current state: from the last stage-4 update, else
  // computed here
  if (!ctx->speed_fresh) rc = do_speed(ctx);
  if (!rc && !ctx->q_fresh) rc = do_q_variables(ctx);
  if (rc) return rc;
  launch_set_int(ctx->counter, 0, ctx->stream);
  int done = 0;
  const int K = ctx->graph_iters;
  if (allow_graph && K > 0 && n_iter >= K && !ctx->timing) {
    if (!ctx->gexec || ctx->gexec_iters != K || ctx->gexec_hist != ctx->hist)
      rc = capture_graph(ctx, K);
    for (; !rc && done + K <= n_iter; done += K) {
      hipError_t e = hipGraphLaunch(ctx->gexec, ctx->stream);
      if (e != hipSuccess) rc = set_err(ctx, LSKUM_EHIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    }
  }
  for (; !rc && done < n_iter; done++) rc = run_iteration(ctx);
  return rc;
}

// Residual bookkeeping of n_iter iterations (fpi_solver_mod.h:138-147) from
// their stage-4 sums, sums[case * n_iter + t]; res_new may be sums itself
// (each entry is read before it is overwritten).
int finish_iterations(lskum_ctx *ctx, int n_iter, const double *sums, double *res_new,
                      double *residue) {
  const int nc = ctx->n_cases();
  int bad = -1;
  for (int t = 0; t < n_iter; t++) {
    for (int j = 0; j < nc; j++)
      if (!std::isfinite(sums[(size_t)j * n_iter + t]) && !ctx->cases[j].failed_at)
        ctx->cases[j].failed_at = ctx->it + 1 + t;
    const int it = ctx->it + 1;
    for (int j = 0; j < nc; j++) {
      CaseInfo &cs = ctx->cases[j];
      const double s = sums[(size_t)j * n_iter + t];
      const double rn = std::sqrt(s) / cs.n;
      double rs;
      if (it <= 2 && ctx->restart == 0) {
        cs.res_old = rn;
        rs = 0;
      } else {
        rs = std::log10(rn / cs.res_old);
      }
      if (res_new) res_new[(size_t)j * n_iter + t] = rn;
      if (residue) residue[(size_t)j * n_iter + t] = rs;
      if (!std::isfinite(s) && bad < 0) bad = j;
    }
    ctx->it = it;
  }
  // A batch's cases share no points, stencils or reductions, so a case that
  // diverges cannot disturb the others: the batch goes on and
  // lskum_case_status names the failed cases.  A single case fails the call.
  if (bad >= 0 && nc == 1)
    return set_err(ctx, LSKUM_ENONFINITE, "residual became non-finite");
  return LSKUM_OK;
}

}  // namespace lskum

extern "C" {

const char *lskum_last_error(const lskum_ctx *ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int lskum_create(const lskum_cloud *cloud, const lskum_params *params, lskum_ctx **out) {
  if (!params) return LSKUM_EINVAL;
  lskum_case one;
  one.cloud = 0;
  one.mach = params->mach;
  one.aoa_deg = params->aoa_deg;
  return create_impl(cloud, 1, &one, 1, params, out);
}

int lskum_create_batch(const lskum_cloud *clouds, int32_t n_clouds, const lskum_case *cases,
                       int32_t n_cases, const lskum_params *params, lskum_ctx **out) {
  return create_impl(clouds, n_clouds, cases, n_cases, params, out);
}

void lskum_destroy(lskum_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  release(ctx);
  delete ctx;
}

int lskum_case_count(lskum_ctx *ctx, int32_t *n_cases, int32_t *points) {
  CTX_OK(ctx);
  if (n_cases) *n_cases = ctx->n_cases();
  if (points)
    for (int j = 0; j < ctx->n_cases(); j++) points[j] = ctx->cases[j].n;
  return LSKUM_OK;
}

int lskum_set_state(lskum_ctx *ctx, const double *prim) {
  CTX_OK(ctx);
  if (!prim) return set_err(ctx, LSKUM_EINVAL, "null prim");
  std::vector<double> tmp;
  int rc = fetch_plain(ctx, ctx->prim, 4, tmp);  // keeps the padding points' state
  if (rc) return rc;
  permute_in(ctx, prim, tmp.data(), 4);
  HIP_TRY(ctx, hipMemcpyAsync(ctx->prim, tmp.data(), tmp.size() * sizeof(double),
                              hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->q_fresh = ctx->speed_fresh = false;
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
  int rc = do_speed(ctx);
  return rc ? rc : do_func_delta(ctx);
}

int lskum_q_variables(lskum_ctx *ctx) {
  CTX_OK(ctx);
  return do_q_variables(ctx);
}

int lskum_q_derivatives(lskum_ctx *ctx) {
  CTX_OK(ctx);
  return do_q_derivatives(ctx, 0, true);
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
  int rc = do_state_update(ctx, rk, false);
  if (rc) return rc;
  if (!sum_res_sqr && !max_res && !max_res_point) return LSKUM_OK;
  const int nc = ctx->n_cases();
  HIP_TRY(ctx, hipMemcpyAsync(ctx->red_h, ctx->red, sizeof(double) * 3 * nc,
                              hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  bool finite = true;
  for (int j = 0; j < nc; j++) {
    const double *r = ctx->red_h + 3 * j;
    // state_update_mod.h:106-107,159-163: max_res starts at 0 and the point
    // only moves on a strictly larger value
    if (r[1] > 0.0) ctx->cases[j].max_res_point = (int32_t)r[2];
    if (sum_res_sqr) sum_res_sqr[j] = r[0];
    if (max_res) max_res[j] = r[1] > 0.0 ? r[1] : 0.0;
    if (max_res_point) max_res_point[j] = ctx->cases[j].max_res_point;
    if (!std::isfinite(r[0])) finite = false;
  }
  if (!finite && nc == 1)  // a batch reports NaN for a failed case instead
    return set_err(ctx, LSKUM_ENONFINITE, "residual sum is not finite");
  return LSKUM_OK;
}

int lskum_forces(lskum_ctx *ctx, double *cl, double *cd, double *cm) {
  CTX_OK(ctx);
  const int nc = ctx->n_cases();
  launch_forces(nc, ctx->soff, ctx->f_m, ctx->f_l, ctx->f_r, ctx->f_s, ctx->base, ctx->nx,
                ctx->ny, ctx->prim, ctx->n, ctx->pr_inf, ctx->qdyn, ctx->forces, ctx->stream);
  int rc = check_launch(ctx, "forces");
  if (rc) return rc;
  double *h = ctx->red_h + 3 * nc;
  HIP_TRY(ctx, hipMemcpyAsync(h, ctx->forces, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  bool finite = true;
  for (int j = 0; j < nc; j++) {
    const double H = h[3 * j], V = h[3 * j + 1], P = h[3 * j + 2], th = ctx->cases[j].theta;
    // compute_force_coeffs_mod.h:62-67
    const double l = V * std::cos(th) - H * std::sin(th);
    const double d = H * std::cos(th) + V * std::sin(th);
    if (cl) cl[j] = l;
    if (cd) cd[j] = d;
    if (cm) cm[j] = P;
    if (!std::isfinite(l) || !std::isfinite(d) || !std::isfinite(P)) finite = false;
  }
  if (!finite && nc == 1)  // a batch reports NaN for a failed case instead
    return set_err(ctx, LSKUM_ENONFINITE, "force coefficients are not finite");
  return LSKUM_OK;
}

int lskum_set_graph(lskum_ctx *ctx, int32_t iters_per_graph) {
  CTX_OK(ctx);
  if (iters_per_graph < 0 || iters_per_graph > 1000)
    return set_err(ctx, LSKUM_EINVAL, "iters_per_graph outside 0..1000");
  if (iters_per_graph != ctx->gexec_iters) drop_graph(ctx);
  ctx->graph_iters = iters_per_graph;
  return LSKUM_OK;
}

int lskum_iterate(lskum_ctx *ctx, int32_t n_iter, double *res_new, double *residue) {
  CTX_OK(ctx);
  if (n_iter < 0) return set_err(ctx, LSKUM_EINVAL, "n_iter < 0");
  if (n_iter == 0) return LSKUM_OK;
  const int nc = ctx->n_cases();
  if (ctx->hist_cap < n_iter) {
    drop_graph(ctx);  // a captured graph writes into the old buffer
    if (ctx->hist) (void)hipFree(ctx->hist);
    ctx->hist = nullptr;
    ctx->hist_cap = 0;
    HIP_TRY(ctx, dalloc(&ctx->hist, (size_t)n_iter * nc));
    ctx->hist_cap = n_iter;
  }
  ctx->hist_dst = ctx->hist;
  ctx->hist_stride = ctx->hist_cap;
  int rc = launch_iterations(ctx, n_iter, true);
  if (rc) return rc;
  std::vector<double> sums((size_t)nc * n_iter);
  HIP_TRY(ctx, hipMemcpy2DAsync(sums.data(), sizeof(double) * n_iter, ctx->hist,
                                sizeof(double) * ctx->hist_cap, sizeof(double) * n_iter, nc,
                                hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  rc = collect_timing(ctx);
  if (rc) return rc;
  return finish_iterations(ctx, n_iter, sums.data(), res_new, residue);
}

int lskum_iterate_noalloc(lskum_ctx *ctx, int32_t n_iter, double *res_new, double *residue) {
  CTX_OK(ctx);
  if (n_iter < 0) return set_err(ctx, LSKUM_EINVAL, "n_iter < 0");
  if (n_iter == 0) return LSKUM_OK;
  if (!res_new)
    return set_err(ctx, LSKUM_EINVAL, "lskum_iterate_noalloc: res_new is required");
  if (ctx->timing)
    return set_err(ctx, LSKUM_EINVAL,
                   "lskum_iterate_noalloc: kernel timing creates events; disable it first");
  const size_t need = (size_t)n_iter * ctx->n_cases();
  if (need > (size_t)LSKUM_NOALLOC_HIST_MAX)
    return set_err(ctx, LSKUM_EINVAL,
                   "lskum_iterate_noalloc: n_iter x n_cases = " + std::to_string(need) +
                       " exceeds the scratch array's " +
                       std::to_string(LSKUM_NOALLOC_HIST_MAX) + " entries");
  double *scratch = noalloc_scratch(ctx->device);
  if (!scratch)
    return set_err(ctx, LSKUM_EHIP, "lskum_iterate_noalloc: scratch array not initialised");
  ctx->hist_dst = scratch;
  ctx->hist_stride = n_iter;
  // every slot [case][t] is written by iteration t's stage-4 reduction
  // before it is read, so nothing needs clearing
  int rc = launch_iterations(ctx, n_iter, false);
  if (rc) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(res_new, scratch, sizeof(double) * need, hipMemcpyDeviceToHost,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return finish_iterations(ctx, n_iter, res_new, res_new, residue);
}

int lskum_case_status(lskum_ctx *ctx, int32_t *failed_at) {
  CTX_OK(ctx);
  if (!failed_at) return set_err(ctx, LSKUM_EINVAL, "null output");
  for (int j = 0; j < ctx->n_cases(); j++) failed_at[j] = ctx->cases[j].failed_at;
  return LSKUM_OK;
}

int lskum_get_iteration(lskum_ctx *ctx, int32_t *it, double *res_old) {
  CTX_OK(ctx);
  if (it) *it = ctx->it;
  if (res_old)
    for (int j = 0; j < ctx->n_cases(); j++) res_old[j] = ctx->cases[j].res_old;
  return LSKUM_OK;
}

int lskum_set_iteration(lskum_ctx *ctx, int32_t it, double res_old, int32_t restart) {
  CTX_OK(ctx);
  if (it < 0) return set_err(ctx, LSKUM_EINVAL, "it < 0");
  ctx->it = it;
  for (CaseInfo &cs : ctx->cases) cs.res_old = res_old;
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
    case LSKUM_A_NORMALS: {
      size_t o = 0;
      for (const CaseInfo &cs : ctx->cases) {
        const CloudLayout &L = ctx->clouds[cs.cloud];
        std::memcpy(dst + o, L.nx_h.data(), sizeof(double) * cs.n);
        std::memcpy(dst + o + cs.n, L.ny_h.data(), sizeof(double) * cs.n);
        o += 2 * (size_t)cs.n;
      }
      return LSKUM_OK;
    }
    default: return set_err(ctx, LSKUM_EINVAL, "unknown array id");
  }
  if (rc) return rc;
  permute_out(ctx, dev.data(), dst, ncomp);
  return LSKUM_OK;
}

int lskum_get_substencil(lskum_ctx *ctx, int32_t which, int32_t *counts, int32_t *members) {
  CTX_OK(ctx);
  if (which < 0 || which > 3 || !counts) return set_err(ctx, LSKUM_EINVAL, "bad arguments");
  const CloudLayout &L = ctx->clouds[ctx->cases[0].cloud];
  const uint8_t bit = (uint8_t)(1u << which);
  int64_t w = 0;
  for (int32_t i = 0; i < L.n; i++) {
    int32_t cnt = 0;
    for (int32_t e = L.conn_off_h[i]; e < L.conn_off_h[i + 1]; e++) {
      if (!(L.mask_h[e] & bit)) continue;
      if (members) members[w] = L.conn_h[e];
      w++;
      cnt++;
    }
    counts[i] = cnt;
  }
  return LSKUM_OK;
}

void *lskum_stream(lskum_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int lskum_flux_tile(lskum_ctx *ctx, int32_t *cap) {
  CTX_OK(ctx);
  if (!cap) return set_err(ctx, LSKUM_EINVAL, "null argument");
  *cap = (ctx->tiles & TILE_FLUX) ? ctx->flux_cap : 0;
  return LSKUM_OK;
}

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

int lskum_flux_vjp(lskum_ctx *ctx, const double *lam, double *grad) {
  CTX_OK(ctx);
  if (!lam || !grad) return set_err(ctx, LSKUM_EINVAL, "null argument");
  int rc = prepare_adjoint(ctx);
  if (!rc) rc = recompute_stages(ctx);
  if (!rc) rc = upload_caller(ctx, lam, ctx->adj.vec_in);
  if (rc) return rc;
  const DevCloud dc = ctx->dc();
  AdjointBuffers &a = ctx->adj;
  launch_flux_vjp(dc, a, ctx->base, ctx->dqb[ctx->cur], ctx->qm, ctx->delta, a.vec_in, nullptr,
                  ctx->stream);
  int b = 0;  // ldq[b]: adjoint of the derivatives the flux read
  for (int p = 0; p < ctx->prm.inner_iterations; p++) {
    launch_ls_vjp(dc, a, ctx->base, a.ldq[b], a.ldq[1 - b], false, nullptr, nullptr, nullptr,
                  ctx->stream);
    b = 1 - b;
  }
  launch_ls_vjp(dc, a, ctx->base, a.ldq[b], nullptr, true, nullptr, nullptr, nullptr,
                ctx->stream);
  launch_qvar_vjp(dc, a, ctx->prim, a.vec_out, false, ctx->stream);
  rc = check_launch(ctx, "flux_vjp");
  return rc ? rc : download_caller(ctx, a.vec_out, grad);
}

int lskum_flux_jvp(lskum_ctx *ctx, const double *v, double *out) {
  CTX_OK(ctx);
  if (!v || !out) return set_err(ctx, LSKUM_EINVAL, "null argument");
  int rc = prepare_adjoint(ctx);
  if (!rc) rc = recompute_stages(ctx);
  if (!rc) rc = upload_caller(ctx, v, ctx->adj.vec_in);
  if (rc) return rc;
  const DevCloud dc = ctx->dc();
  AdjointBuffers &a = ctx->adj;
  const int inner = ctx->prm.inner_iterations;
  HIP_TRY(ctx, hipMemcpyAsync(a.tbase, ctx->base, sizeof(double) * B_STRIDE * ctx->n,
                              hipMemcpyDeviceToDevice, ctx->stream));
  launch_qvar_jvp(dc, ctx->prim, a.vec_in, a.tbase, ctx->stream);
  // the least-squares stencils are linear: the forward kernels map tangents
  int t = inner % 2;
  launch_q_derivatives(dc, a.tbase, a.tdq[t], a.tqm, a.tqs, 0.0, ctx->stream);
  launch_ext_jvp(dc, a.arg, a.tbase, a.tqm, ctx->stream);
  for (int p = 0; p < inner; p++) {
    launch_q_inner(dc, a.tqs, a.tdq[t], a.tdq[1 - t], 0.0, ctx->stream);
    t = 1 - t;
  }
  launch_flux_jvp(dc, ctx->base, ctx->dqb[ctx->cur], ctx->qm, ctx->delta, a.tbase, a.tdq[t],
                  a.tqm, a.vec_out, ctx->stream);
  rc = check_launch(ctx, "flux_jvp");
  return rc ? rc : download_caller(ctx, a.vec_out, out);
}

int lskum_forces_grad(lskum_ctx *ctx, double *dcl, double *dcd, double *dcm) {
  CTX_OK(ctx);
  if (ctx->n_cases() != 1) return set_err(ctx, LSKUM_EINVAL, "single-case contexts only");
  const CloudLayout &L = ctx->clouds[0];
  const CaseInfo &cs = ctx->cases[0];
  const double qdyn = 0.5 * ctx->rho_inf * cs.mach * cs.mach;
  const double ct = std::cos(cs.theta), st = std::sin(cs.theta);
  double *outs[3] = {dcl, dcd, dcm};
  for (double *o : outs)
    if (o) std::fill(o, o + 4 * (size_t)L.n, 0.0);
  // compute_force_coeffs_mod.h:25-66: cp = -(p_m - p_inf)/qdyn per panel
  for (size_t j = 0; j < L.fm.size(); j++) {
    if (L.fs[j] != 1) continue;
    const int32_t m = L.fm[j], l = L.fl[j], r = L.fr[j];
    const double ds1 = std::sqrt((L.x[m] - L.x[l]) * (L.x[m] - L.x[l]) +
                                 (L.y[m] - L.y[l]) * (L.y[m] - L.y[l]));
    const double ds2 = std::sqrt((L.x[r] - L.x[m]) * (L.x[r] - L.x[m]) +
                                 (L.y[r] - L.y[m]) * (L.y[r] - L.y[m]));
    const double ds = 0.5 * (ds1 + ds2), dcp = -1.0 / qdyn;
    const double nx = L.nx[m], ny = L.ny[m];
    const size_t at = 3 * (size_t)L.n + L.perm[m];  // the pressure of point m
    if (dcl) dcl[at] += (ny * ct - nx * st) * ds * dcp;
    if (dcd) dcd[at] += (nx * ct + ny * st) * ds * dcp;
    if (dcm) dcm[at] += (-ny * ds * (L.x[m] - 0.25) + nx * ds * L.y[m]) * dcp;
  }
  return LSKUM_OK;
}

int lskum_autotune(const lskum_cloud *clouds, int32_t n_clouds, const lskum_case *cases,
                   int32_t n_cases, const lskum_params *params, int32_t iters,
                   lskum_tune_result *best, lskum_tune_result *trials, int32_t max_trials,
                   int32_t *n_trials) {
  if (!params || !best || iters < 1 || iters > 1000) return LSKUM_EINVAL;
  const int32_t blocks[3] = {64, 128, 256};
  const int32_t masks[2] = {LSKUM_TILE_QDERIV | LSKUM_TILE_QINNER | LSKUM_TILE_DELTA,
                            LSKUM_TILE_QDERIV | LSKUM_TILE_QINNER | LSKUM_TILE_DELTA |
                                LSKUM_TILE_FLUX};
  const int32_t ppts[2] = {2, 4};
  int32_t count = 0;
  best->ms_per_iter = -1.0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return LSKUM_EHIP;
  }
  // one candidate: built once, timed over one graph of `iters` iterations
  // after one warm-up replay of it
  auto trial = [&](int32_t tb, int32_t mask, int32_t ppt) -> int {
    lskum_params p = *params;
    p.tile_block = tb;
    p.lds_tiles = mask;
    p.points_per_thread = ppt;
    lskum_ctx *ctx = nullptr;
    int rc = create_impl(clouds, n_clouds, cases, n_cases, &p, &ctx);
    if (rc) {
      lskum_destroy(ctx);
      return rc;
    }
    lskum_tune_result r;
    r.tile_block = tb;
    r.lds_tiles = (int32_t)ctx->tiles;  // what fits: a tile too big for LDS is dropped
    r.points_per_thread = ctx->su_ppt;
    rc = lskum_set_graph(ctx, iters);
    if (!rc) rc = lskum_iterate(ctx, iters, nullptr, nullptr);
    if (!rc && hipEventRecord(e0, ctx->stream) != hipSuccess) rc = LSKUM_EHIP;
    if (!rc) rc = lskum_iterate(ctx, iters, nullptr, nullptr);
    if (!rc && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = LSKUM_EHIP;
    float ms = 0.f;
    if (!rc && (hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = LSKUM_EHIP;
    lskum_destroy(ctx);
    if (rc) return rc;
    r.ms_per_iter = ms / iters;
    if (trials && count < max_trials) trials[count] = r;
    count++;
    if (best->ms_per_iter < 0.0 || r.ms_per_iter < best->ms_per_iter) *best = r;
    return LSKUM_OK;
  };
  int rc = LSKUM_OK;
  // block shape and tiles over a grid, then the state update's points per
  // thread on the best of them (its kernel is independent of the others')
  for (int32_t tb : blocks)
    for (int32_t mask : masks)
      if (!rc) rc = trial(tb, mask, 1);
  const lskum_tune_result grid_best = *best;
  for (int32_t ppt : ppts)
    if (!rc) rc = trial(grid_best.tile_block, grid_best.lds_tiles, ppt);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (n_trials) *n_trials = count;
  return rc;
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
