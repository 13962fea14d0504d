// diagnostics.hip -- diagnostics of the current state: entropy
// (compute_entropy_mod.h:5-24), vorticity and enstrophy
// (compute_enstrophy_mod.h:5-76), stagnation pressure extremes and the
// stagnation-pressure loss J (stagnation_values_mod.h:8-80).
//
// One stencil kernel computes every per-point quantity -- the vorticity by
// the same weighted least squares over the point's neighbours as
// q_derivatives, the rest pointwise -- and reduces them per 256-point block
// (sums, and the first minimum / maximum in the caller's point order, as
// the reference's strict comparisons in point order keep them); one block
// per case then reduces the block partials in a fixed order.  Nothing is
// atomic, so the results are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "context.hpp"

namespace lskum {

namespace {

constexpr double kGamma = 1.4;
constexpr int PART = 8;  // per block / case: sum entropy, sum vort^2, sum J term,
                         // min p0, its point, max p0, its point, unused

struct MinMax {
  double mn, mx;
  int amn, amx;
};

// first minimum / maximum in caller order: a smaller value, or an equal value
// at a smaller caller index, wins
__device__ __forceinline__ void mm_merge(MinMax &a, const MinMax &b) {
  if (b.mn < a.mn || (b.mn == a.mn && b.amn < a.amn)) {
    a.mn = b.mn;
    a.amn = b.amn;
  }
  if (b.mx > a.mx || (b.mx == a.mx && b.amx < a.amx)) {
    a.mx = b.mx;
    a.amx = b.amx;
  }
}

__device__ void block_reduce_diag(double s[3], MinMax &m) {
  __shared__ double ss[3][256];
  __shared__ MinMax sm[256];
  const int t = threadIdx.x;
  for (int j = 0; j < 3; j++) ss[j][t] = s[j];
  sm[t] = m;
  __syncthreads();
  for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
    if (t < w) {
      for (int j = 0; j < 3; j++) ss[j][t] = ss[j][t] + ss[j][t + w];
      MinMax a = sm[t];
      mm_merge(a, sm[t + w]);
      sm[t] = a;
    }
    __syncthreads();
  }
  for (int j = 0; j < 3; j++) s[j] = ss[j][0];
  m = sm[0];
}

// p0_inf[case], log(pr_inf) shared by all cases
template <bool POW>
__global__ __launch_bounds__(256) void k_diag(
    int n, const int32_t *__restrict__ off, const uint32_t *__restrict__ conn,
    const double *__restrict__ base, const double *__restrict__ prim,
    const int32_t *__restrict__ perm, const uint16_t *__restrict__ block_case,
    const double *__restrict__ p0_inf, double log_pinf, double power,
    double *__restrict__ entropy, double *__restrict__ vort, double *__restrict__ part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double s[3] = {0.0, 0.0, 0.0};
  MinMax m = {INFINITY, -INFINITY, 0x7fffffff, 0x7fffffff};
  // padding points of a batch have no neighbours and take no part
  if (i < n && off[i + 1] > off[i]) {
    const double rho = prim[i], u1 = prim[(size_t)n + i], u2 = prim[(size_t)2 * n + i];
    const double pr = prim[(size_t)3 * n + i];
    // compute_entropy_mod.h:18-22
    double temp1 = pow(rho, kGamma);
    temp1 = pr / temp1;
    temp1 = log(temp1);
    const double ent = fabs(temp1 - log_pinf);
    // stagnation_values_mod.h:26-31 (and 69-72)
    const double gammaPower = kGamma / (kGamma - 1);
    const double angle = sqrt(kGamma * pr / rho);
    const double mach_t = sqrt(u1 * u1 + u2 * u2) / angle;
    const double p0 = pr * pow(1 + ((kGamma - 1) / 2) * mach_t * mach_t, gammaPower);
    const double p0i = p0_inf[block_case ? block_case[blockIdx.x] : 0];
    // compute_enstrophy_mod.h:31-70
    const double x_i = base[(size_t)B_STRIDE * i + B_X], y_i = base[(size_t)B_STRIDE * i + B_Y];
    double sxx = 0., syy = 0., sxy = 0., sxu1 = 0., sxu2 = 0., syu1 = 0., syu2 = 0.;
    for (int e = off[i]; e < off[i + 1]; e++) {
      const int k = (int)(conn[e] & kIndexMask);
      const double delx = base[(size_t)B_STRIDE * k + B_X] - x_i;
      const double dely = base[(size_t)B_STRIDE * k + B_Y] - y_i;
      const double du1 = prim[(size_t)n + k] - u1, du2 = prim[(size_t)2 * n + k] - u2;
      if (POW) {
        const double weights = pow(sqrt(delx * delx + dely * dely), power);
        sxx = sxx + delx * delx * weights;
        syy = syy + dely * dely * weights;
        sxy = sxy + delx * dely * weights;
        sxu1 = sxu1 + weights * delx * du1;
        sxu2 = sxu2 + weights * delx * du2;
        syu1 = syu1 + weights * dely * du1;
        syu2 = syu2 + weights * dely * du2;
      } else {  // weights == 1.0 exactly
        sxx = sxx + delx * delx;
        syy = syy + dely * dely;
        sxy = sxy + delx * dely;
        sxu1 = sxu1 + delx * du1;
        sxu2 = sxu2 + delx * du2;
        syu1 = syu1 + dely * du1;
        syu2 = syu2 + dely * du2;
      }
    }
    const double one_by_det = 1.0 / (sxx * syy - sxy * sxy);
    const double du2_dx = (sxu2 * syy - syu2 * sxy) * one_by_det;
    const double du1_dy = (syu1 * sxx - sxu1 * sxy) * one_by_det;
    const double w = du2_dx - du1_dy;
    entropy[i] = ent;
    vort[i] = w;
    s[0] = ent;
    s[1] = w * w;
    s[2] = (p0i - p0) * (p0i - p0);
    m.mn = m.mx = p0;
    m.amn = m.amx = perm[i];
  }
  block_reduce_diag(s, m);
  if (threadIdx.x == 0) {
    double *o = part + (size_t)PART * blockIdx.x;
    o[0] = s[0];
    o[1] = s[1];
    o[2] = s[2];
    o[3] = m.mn;
    o[4] = (double)m.amn;
    o[5] = m.mx;
    o[6] = (double)m.amx;
    o[7] = 0.0;
  }
}

// one block per case over its block partials [seg[c], seg[c+1]) in a fixed order
__global__ __launch_bounds__(256) void k_diag_reduce(const int32_t *__restrict__ seg, int nb,
                                                     const double *__restrict__ part,
                                                     double *__restrict__ out) {
  const int c = blockIdx.x;
  const int b0 = seg ? seg[c] : 0, b1 = seg ? seg[c + 1] : nb;
  double s[3] = {0.0, 0.0, 0.0};
  MinMax m = {INFINITY, -INFINITY, 0x7fffffff, 0x7fffffff};
  for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
    const double *p = part + (size_t)PART * b;
    for (int j = 0; j < 3; j++) s[j] = s[j] + p[j];
    MinMax q = {p[3], p[5], (int)p[4], (int)p[6]};
    mm_merge(m, q);
  }
  block_reduce_diag(s, m);
  if (threadIdx.x == 0) {
    double *o = out + (size_t)PART * c;
    o[0] = s[0];
    o[1] = s[1];
    o[2] = s[2];
    o[3] = m.mn;
    o[4] = (double)m.amn;
    o[5] = m.mx;
    o[6] = (double)m.amx;
    o[7] = 0.0;
  }
}

struct DiagResult {
  std::vector<double> entropy, vort;  // caller order, case after case
  std::vector<double> per_case;       // [n_cases][PART]
  std::vector<double> p0_inf;         // per case
};

// stagnation_values_mod.h:20-21 for the case's Mach number
double case_p0_inf(const lskum_ctx *c, const CaseInfo &cs) {
  const double g = kGamma, gammaPower = g / (g - 1);
  return c->pr_inf * (std::pow((1 + ((g - 1) / 2) * cs.mach * cs.mach), gammaPower));
}

int run_diag(lskum_ctx *c, bool arrays, DiagResult &r) {
  const int n = c->n, nc = c->n_cases();
  const int nb = blocks_for(n, 256);
  r.p0_inf.resize(nc);
  for (int j = 0; j < nc; j++) r.p0_inf[j] = case_p0_inf(c, c->cases[j]);
  double *d = nullptr;  // entropy [n], vort [n], part [nb][PART], out [nc][PART], p0_inf [nc]
  const size_t count = 2 * (size_t)n + (size_t)PART * nb + (size_t)PART * nc + nc;
  if (dalloc(&d, count) != hipSuccess)
    return set_err(c, LSKUM_ENOMEM, "diagnostics: device allocation failed");
  double *ent = d, *vort = d + n, *part = vort + n, *out = part + (size_t)PART * nb;
  double *p0i = out + (size_t)PART * nc;
  int rc = LSKUM_OK;
  hipError_t e = hipMemcpyAsync(p0i, r.p0_inf.data(), nc * sizeof(double), hipMemcpyHostToDevice,
                                c->stream);
  if (e == hipSuccess) {
    const double log_pinf = std::log(c->pr_inf);
    if (c->prm.power == 0.0)
      hipLaunchKernelGGL(k_diag<false>, dim3(nb), dim3(256), 0, c->stream, n, c->conn_off,
                         c->conn, c->base, c->prim, c->perm_d, c->block_case, p0i, log_pinf,
                         c->prm.power, ent, vort, part);
    else
      hipLaunchKernelGGL(k_diag<true>, dim3(nb), dim3(256), 0, c->stream, n, c->conn_off,
                         c->conn, c->base, c->prim, c->perm_d, c->block_case, p0i, log_pinf,
                         c->prm.power, ent, vort, part);
    hipLaunchKernelGGL(k_diag_reduce, dim3(nc), dim3(256), 0, c->stream, c->seg, nb, part, out);
    e = hipGetLastError();
  }
  r.per_case.resize((size_t)PART * nc);
  std::vector<double> tmp;
  if (e == hipSuccess)
    e = hipMemcpyAsync(r.per_case.data(), out, r.per_case.size() * sizeof(double),
                       hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && arrays) {
    tmp.resize(2 * (size_t)n);
    e = hipMemcpyAsync(tmp.data(), d, tmp.size() * sizeof(double), hipMemcpyDeviceToHost,
                       c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess)
    return set_err(c, LSKUM_EHIP, std::string("diagnostics: ") + hipGetErrorString(e));
  if (arrays) {
    size_t total = 0;
    for (const CaseInfo &cs : c->cases) total += cs.n;
    r.entropy.resize(total);
    r.vort.resize(total);
    permute_out(c, tmp.data(), r.entropy.data(), 1);
    permute_out(c, tmp.data() + n, r.vort.data(), 1);
  }
  return rc;
}

}  // namespace

}  // namespace lskum

extern "C" {

int lskum_compute_entropy(lskum_ctx *ctx, double *entropy, double *total_entropy) {
  CTX_OK(ctx);
  DiagResult r;
  int rc = run_diag(ctx, entropy != nullptr, r);
  if (rc) return rc;
  if (entropy) std::copy(r.entropy.begin(), r.entropy.end(), entropy);
  if (total_entropy)
    for (int j = 0; j < ctx->n_cases(); j++) total_entropy[j] = r.per_case[PART * j];
  return LSKUM_OK;
}

int lskum_compute_enstrophy(lskum_ctx *ctx, double *vorticity, double *vorticity_sqr,
                            double *total_enstrophy) {
  CTX_OK(ctx);
  DiagResult r;
  int rc = run_diag(ctx, vorticity || vorticity_sqr, r);
  if (rc) return rc;
  for (size_t k = 0; k < r.vort.size(); k++) {
    if (vorticity) vorticity[k] = r.vort[k];
    if (vorticity_sqr) vorticity_sqr[k] = r.vort[k] * r.vort[k];
  }
  if (total_enstrophy)
    for (int j = 0; j < ctx->n_cases(); j++) total_enstrophy[j] = r.per_case[PART * j + 1];
  return LSKUM_OK;
}

int lskum_stagnation_pressure(lskum_ctx *ctx, double *values, int32_t *index) {
  CTX_OK(ctx);
  DiagResult r;
  int rc = run_diag(ctx, false, r);
  if (rc) return rc;
  for (int j = 0; j < ctx->n_cases(); j++) {
    const double *p = &r.per_case[(size_t)PART * j];
    if (values) {
      values[4 * j] = p[3];
      values[4 * j + 1] = p[5];
      values[4 * j + 2] = p[3] / r.p0_inf[j];
      values[4 * j + 3] = p[5] / r.p0_inf[j];
    }
    if (index) {
      index[2 * j] = (int32_t)p[4];
      index[2 * j + 1] = (int32_t)p[6];
    }
  }
  return LSKUM_OK;
}

int lskum_objective_J(lskum_ctx *ctx, double *J) {
  CTX_OK(ctx);
  if (!J) return set_err(ctx, LSKUM_EINVAL, "null output");
  DiagResult r;
  int rc = run_diag(ctx, false, r);
  if (rc) return rc;
  for (int j = 0; j < ctx->n_cases(); j++) {
    // stagnation_values_mod.h:61 with plen = the case's points
    const double constant = 1 / (std::pow(r.p0_inf[j], 2) * ctx->cases[j].n);
    J[j] = r.per_case[(size_t)PART * j + 2] * constant;
  }
  return LSKUM_OK;
}

}  // extern "C"
