// shape.cpp -- Hicks-Henne shape parametrisation of a cloud's wall and its
// propagation to the other points by inverse-distance weights (no reference
// counterpart; the design variables of a shape-sensitivity study).
//
// Bump j with its peak at chord fraction p_j and width exponent t:
//   b_j(x) = sin(pi x^m_j)^t,   m_j = ln(1/2) / ln(p_j),
// so b_j(0) = b_j(1) = 0 and b_j(p_j) = 1 (Hicks and Henne, "Wing design by
// numerical optimization", J. Aircraft 15, 1978).  n_per_side bumps per
// surface have their peaks evenly spaced over [p_lo, p_hi]; design variables
// 0 .. n_per_side-1 act on the upper surface (+y), the rest on the lower
// surface (-y), so a positive amplitude moves the surface outward.  A wall
// point's side follows the wall order of the clouds (clockwise from the
// trailing edge: the first half lower, the second upper, the trailing and
// leading edge points on neither).
//
// Every other point p moves by the inverse-distance average of the wall
// displacements, weights |p - w|^-3, faded by (1 - d/R)^2 with d the
// distance to the nearest wall point and R the decay radius (points farther
// than R, the far field among them, stay fixed).  The map is linear in the
// amplitudes: the displacement fields per unit amplitude are returned, and a
// shape is y + sum_j alpha_j dy_j (x does not move).
#include <algorithm>
#include <cmath>
#include <vector>

#include "host.hpp"

namespace {

double bump(double x, double peak, int t) {
  if (!(x > 0.0) || !(x < 1.0)) return 0.0;
  const double m = std::log(0.5) / std::log(peak);
  return std::pow(std::sin(3.14159265358979323846 * std::pow(x, m)), t);
}

}  // namespace

extern "C" {

int lskum_hicks_henne_fields(const lskum_cloud *c, int32_t n_per_side, double p_lo, double p_hi,
                             int32_t width_exponent, double decay_radius, double *dy) {
  if (!c || !dy || c->n_points <= 0 || n_per_side < 1 || n_per_side > 64 || !(p_lo > 0.0) ||
      !(p_hi < 1.0) || !(p_lo <= p_hi) || width_exponent < 1 || !(decay_radius > 0.0))
    return LSKUM_EINVAL;
  const int32_t n = c->n_points;
  int32_t nw = 0;
  while (nw < n && c->flag_1[nw] == 0) nw++;
  if (nw < 4) return LSKUM_EINVAL;  // wall points must come first
  const int nps = n_per_side;
  // per wall point: its side (1 upper, 0 lower, -1 the trailing or leading
  // edge) and the signed bump values of that side's variables
  std::vector<int> side(nw, -1);
  std::vector<double> bw((size_t)nw * nps, 0.0);
  for (int32_t w = 1; w < nw; w++) {
    if (2 * w == nw) continue;
    side[w] = 2 * w > nw ? 1 : 0;
    for (int j = 0; j < nps; j++) {
      const double peak = nps == 1 ? 0.5 * (p_lo + p_hi) : p_lo + (p_hi - p_lo) * j / (nps - 1);
      const double b = bump(c->x[w], peak, width_exponent);
      bw[(size_t)w * nps + j] = side[w] == 1 ? b : -b;
    }
  }
  // variable of (side, j): upper j, lower nps + j
  auto var = [&](int sd, int j) { return sd == 1 ? j : nps + j; };
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t p = 0; p < n; p++) {
    if (p < nw) {
      for (int j = 0; j < 2 * nps; j++) dy[(size_t)j * n + p] = 0.0;
      if (side[p] >= 0)
        for (int j = 0; j < nps; j++) dy[(size_t)var(side[p], j) * n + p] = bw[(size_t)p * nps + j];
      continue;
    }
    double dmin2 = 1e300;
    for (int32_t w = 0; w < nw; w++) {
      const double dx = c->x[p] - c->x[w], dyw = c->y[p] - c->y[w];
      dmin2 = std::min(dmin2, dx * dx + dyw * dyw);
    }
    const double dmin = std::sqrt(dmin2);
    const double fade = dmin < decay_radius ? 1.0 - dmin / decay_radius : 0.0;
    double acc[2][64], wsum = 0.0;
    for (int j = 0; j < nps; j++) acc[0][j] = acc[1][j] = 0.0;
    if (fade > 0.0) {
      for (int32_t w = 0; w < nw; w++) {
        const double dx = c->x[p] - c->x[w], dyw = c->y[p] - c->y[w];
        const double d2 = dx * dx + dyw * dyw;
        const double wt = 1.0 / (d2 * std::sqrt(d2));
        wsum += wt;
        if (side[w] < 0) continue;
        const double *b = &bw[(size_t)w * nps];
        double *a = acc[side[w]];
        for (int j = 0; j < nps; j++) a[j] += wt * b[j];
      }
    }
    const double sc = fade > 0.0 ? fade * fade / wsum : 0.0;
    for (int sd = 0; sd < 2; sd++)
      for (int j = 0; j < nps; j++) dy[(size_t)var(sd, j) * n + p] = acc[sd][j] * sc;
  }
  return LSKUM_OK;
}

}  // extern "C"

namespace lskum {

void coef_adjoint(int32_t n, const double *x, const double *y, const double *nxa,
                  const double *nya, const int32_t *off, const uint32_t *conn,
                  const double *gc, double *ge, double *gn) {
  constexpr uint32_t kIndex = 0x0FFFFFFFu;
  constexpr int kShift = 28;
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t d = 0; d < n; d++) {
    const double nx = nxa[d], ny = nya[d], tx = ny, ty = -nx;
    double sxx[4] = {0, 0, 0, 0}, syy[4] = {0, 0, 0, 0}, sxy[4] = {0, 0, 0, 0};
    for (int32_t e = off[d]; e < off[d + 1]; e++) {
      const int32_t k = (int32_t)(conn[e] & kIndex);
      const uint32_t m = conn[e] >> kShift;
      const double delx = x[k] - x[d], dely = y[k] - y[d];
      const double dels = delx * tx + dely * ty, deln = delx * nx + dely * ny;
      for (int sb = 0; sb < 4; sb++) {
        if (!(m & (1u << sb))) continue;
        sxx[sb] += dels * dels;
        syy[sb] += deln * deln;
        sxy[sb] += dels * deln;
      }
    }
    double bxx[4], byy[4], bxy[4];
    for (int sb = 0; sb < 4; sb++) {
      bxx[sb] = byy[sb] = bxy[sb] = 0.0;
      const double det = sxx[sb] * syy[sb] - sxy[sb] * sxy[sb];
      if (det == 0.0) continue;  // coefficients 0 whatever the geometry
      const double inv = 1.0 / det, A = sb < 2 ? syy[sb] : sxx[sb];
      const double c0 = gc[(size_t)8 * d + 2 * sb], c1 = gc[(size_t)8 * d + 2 * sb + 1];
      const double invb = c0 * A - c1 * sxy[sb];
      const double detb = -invb * inv * inv;
      bxx[sb] = detb * syy[sb];
      byy[sb] = detb * sxx[sb];
      bxy[sb] = -2.0 * detb * sxy[sb] - c1 * inv;
      if (sb < 2)
        byy[sb] += c0 * inv;
      else
        bxx[sb] += c0 * inv;
    }
    for (int32_t e = off[d]; e < off[d + 1]; e++) {
      const int32_t k = (int32_t)(conn[e] & kIndex);
      const uint32_t m = conn[e] >> kShift;
      const double delx = x[k] - x[d], dely = y[k] - y[d];
      const double dels = delx * tx + dely * ty, deln = delx * nx + dely * ny;
      double bs = 0.0, bn = 0.0;
      for (int sb = 0; sb < 4; sb++) {
        if (!(m & (1u << sb))) continue;
        bs += 2.0 * bxx[sb] * dels + bxy[sb] * deln;
        bn += 2.0 * byy[sb] * deln + bxy[sb] * dels;
      }
      ge[2 * (size_t)e] += bs * ny + bn * nx;
      ge[2 * (size_t)e + 1] += -bs * nx + bn * ny;
      gn[2 * (size_t)d] += -bs * dely + bn * delx;
      gn[2 * (size_t)d + 1] += bs * delx + bn * dely;
    }
  }
}

}  // namespace lskum
