// cloud.cpp -- synthetic NACA 4-digit meshfree clouds and the legacy
// partGrid text format (main.cpp:16-58).
//
// Clouds are deterministic from the seed: a project RNG (splitmix64), exact
// kNN with ties broken by index, and -ffp-contract=off so that neighbour lists
// are bit-identical wherever the library is built for x86-64.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "host.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

lskum_cloud lskum_cloud_h::view() const {
  lskum_cloud v;
  v.n_points = n;
  v.x = x.data();
  v.y = y.data();
  v.left = left.data();
  v.right = right.data();
  v.flag_1 = flag1.data();
  v.flag_2 = flag2.data();
  v.min_dist = min_dist.data();
  v.conn_off = conn_off.data();
  v.conn = conn.data();
  return v;
}

namespace {

constexpr double kPi = 3.14159265358979323846;

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [-1, 1)
  double sym() { return (double)(next() >> 11) * (2.0 / 9007199254740992.0) - 1.0; }
};

struct Naca {
  double m, p, t;
  // half thickness, closed trailing edge (-0.1036 coefficient)
  double yt(double x) const {
    return 5.0 * t *
           (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x +
            0.2843 * x * x * x - 0.1036 * x * x * x * x);
  }
  void camber(double x, double *yc, double *dyc) const {
    if (m <= 0.0 || p <= 0.0) {
      *yc = 0.0;
      *dyc = 0.0;
    } else if (x < p) {
      *yc = m / (p * p) * (2.0 * p * x - x * x);
      *dyc = 2.0 * m / (p * p) * (p - x);
    } else {
      *yc = m / ((1.0 - p) * (1.0 - p)) * ((1.0 - 2.0 * p) + 2.0 * p * x - x * x);
      *dyc = 2.0 * m / ((1.0 - p) * (1.0 - p)) * (p - x);
    }
  }
  // Surface point at parameter th in [0, 2pi): x = (1 + cos th)/2; the lower
  // surface for th in [0, pi], the upper for (pi, 2pi): clockwise from the TE.
  void surf(double th, double *X, double *Y) const {
    double xc = 0.5 * (1.0 + std::cos(th));
    if (xc < 0.0) xc = 0.0;
    double y_t = yt(xc), yc, dyc;
    camber(xc, &yc, &dyc);
    double a = std::atan(dyc);
    if (std::sin(th) >= 0.0) {
      *X = xc + y_t * std::sin(a);
      *Y = yc - y_t * std::cos(a);
    } else {
      *X = xc - y_t * std::sin(a);
      *Y = yc + y_t * std::cos(a);
    }
  }
};

// Surface parametrised by arc-length fraction s in [0, 1) from the TE,
// clockwise, through a fine tabulation of the th-parametrisation.
struct ArcSurface {
  const Naca &af;
  std::vector<double> th, cum;
  double total = 0.0;
  explicit ArcSurface(const Naca &a, int nfine = 1 << 16) : af(a) {
    th.resize(nfine + 1);
    cum.resize(nfine + 1);
    double px, py;
    af.surf(0.0, &px, &py);
    cum[0] = 0.0;
    th[0] = 0.0;
    for (int i = 1; i <= nfine; i++) {
      th[i] = 2.0 * kPi * i / nfine;
      double x, y;
      af.surf(i == nfine ? 0.0 : th[i], &x, &y);
      cum[i] = cum[i - 1] + std::hypot(x - px, y - py);
      px = x;
      py = y;
    }
    total = cum[nfine];
  }
  // the arc-length fraction of the point at parameter t in [0, 2pi)
  double fraction(double t) const {
    const double f = t / (2.0 * kPi) * (double)(th.size() - 1);
    size_t lo = (size_t)f;
    if (lo >= th.size() - 1) lo = th.size() - 2;
    const double w = f - (double)lo;
    return (cum[lo] + w * (cum[lo + 1] - cum[lo])) / total;
  }
  void at(double s, double *X, double *Y) const {
    s -= std::floor(s);
    double target = s * total;
    size_t hi = std::upper_bound(cum.begin(), cum.end(), target) - cum.begin();
    if (hi == 0) hi = 1;
    if (hi >= cum.size()) hi = cum.size() - 1;
    size_t lo = hi - 1;
    double w = (target - cum[lo]) / (cum[hi] - cum[lo]);
    af.surf(th[lo] + w * (th[hi] - th[lo]), X, Y);
  }
};

// Exact kNN over 2-D points: a static k-d tree.
struct KdTree {
  const std::vector<double> &x, &y;
  std::vector<int32_t> idx;
  struct Node { int32_t lo, hi, left, right; double split; int axis; };
  std::vector<Node> nodes;
  KdTree(const std::vector<double> &xx, const std::vector<double> &yy) : x(xx), y(yy) {
    idx.resize(x.size());
    for (size_t i = 0; i < idx.size(); i++) idx[i] = (int32_t)i;
    nodes.reserve(2 * x.size() / 8 + 16);
    build(0, (int32_t)idx.size());
  }
  int32_t build(int32_t lo, int32_t hi) {
    int32_t id = (int32_t)nodes.size();
    nodes.push_back({lo, hi, -1, -1, 0.0, 0});
    if (hi - lo <= 8) return id;
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (int32_t i = lo; i < hi; i++) {
      xmin = std::min(xmin, x[idx[i]]); xmax = std::max(xmax, x[idx[i]]);
      ymin = std::min(ymin, y[idx[i]]); ymax = std::max(ymax, y[idx[i]]);
    }
    int axis = (xmax - xmin) >= (ymax - ymin) ? 0 : 1;
    const std::vector<double> &c = axis == 0 ? x : y;
    int32_t mid = (lo + hi) / 2;
    std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi,
                     [&](int32_t a, int32_t b) { return c[a] != c[b] ? c[a] < c[b] : a < b; });
    double split = c[idx[mid]];
    int32_t l = build(lo, mid);
    int32_t r = build(mid, hi);
    nodes[id].left = l;
    nodes[id].right = r;
    nodes[id].split = split;
    nodes[id].axis = axis;
    return id;
  }
  // k nearest of point q (excluding q) with squared distance <= r2;
  // result sorted by (d2, index).
  template <class Pred>
  void query(int32_t q, int k, double r2, std::vector<std::pair<double, int32_t>> &best,
             const Pred &ok) const {
    best.clear();
    double qx = x[q], qy = y[q];
    auto worse = [](const std::pair<double, int32_t> &a, const std::pair<double, int32_t> &b) {
      return a.first != b.first ? a.first < b.first : a.second < b.second;
    };
    // heap with the worst candidate on top
    search(0, q, qx, qy, k, r2, best, worse, ok);
    std::sort(best.begin(), best.end(), worse);
  }
  template <class Cmp, class Pred>
  void search(int32_t id, int32_t q, double qx, double qy, int k, double r2,
              std::vector<std::pair<double, int32_t>> &best, Cmp cmp, const Pred &ok) const {
    const Node &nd = nodes[id];
    if (nd.left < 0) {
      for (int32_t i = nd.lo; i < nd.hi; i++) {
        int32_t p = idx[i];
        if (p == q) continue;
        double dx = x[p] - qx, dy = y[p] - qy;
        double d2 = dx * dx + dy * dy;
        if (d2 > r2) continue;
        std::pair<double, int32_t> cand(d2, p);
        if ((int)best.size() == k && !cmp(cand, best.front())) continue;
        if (!ok(q, p)) continue;
        if ((int)best.size() < k) {
          best.push_back(cand);
          std::push_heap(best.begin(), best.end(), cmp);
        } else if (cmp(cand, best.front())) {
          std::pop_heap(best.begin(), best.end(), cmp);
          best.back() = cand;
          std::push_heap(best.begin(), best.end(), cmp);
        }
      }
      return;
    }
    double diff = (nd.axis == 0 ? qx : qy) - nd.split;
    int32_t first = diff < 0 ? nd.left : nd.right;
    int32_t second = diff < 0 ? nd.right : nd.left;
    search(first, q, qx, qy, k, r2, best, cmp, ok);
    double bound = (int)best.size() < k ? r2 : std::min(r2, best.front().first);
    // points at exactly the split plane distance may tie; keep them (<=)
    if (diff * diff <= bound) search(second, q, qx, qy, k, r2, best, cmp, ok);
  }
};

// Visibility against the body outline (the closed wall polygon): a
// neighbour pair is rejected when its segment crosses the body, or when both
// points are wall points on opposite surfaces (their chord runs through the
// body; near a sharp trailing edge it would couple the two sides of the
// airfoil).  Edges are bucketed on a uniform grid over the body's box.
struct Body {
  const std::vector<double> &x, &y;
  int nw;
  double x0, y0, x1, y1, cw, ch;
  int gx, gy;
  std::vector<std::vector<int32_t>> cells;  // edge e joins wall e and e+1
  Body(const std::vector<double> &xx, const std::vector<double> &yy, int n_wall)
      : x(xx), y(yy), nw(n_wall) {
    x0 = y0 = 1e300;
    x1 = y1 = -1e300;
    for (int i = 0; i < nw; i++) {
      x0 = std::min(x0, x[i]); x1 = std::max(x1, x[i]);
      y0 = std::min(y0, y[i]); y1 = std::max(y1, y[i]);
    }
    double pad = 1e-9 + 1e-6 * (x1 - x0);
    x0 -= pad; y0 -= pad; x1 += pad; y1 += pad;
    gx = std::max(1, std::min(4096, nw));
    gy = std::max(1, (int)std::ceil(gx * (y1 - y0) / (x1 - x0)));
    cw = (x1 - x0) / gx;
    ch = (y1 - y0) / gy;
    cells.resize((size_t)gx * gy);
    for (int e = 0; e < nw; e++) {
      int a = e, b = (e + 1) % nw;
      int cx0, cy0, cx1, cy1;
      cell_range(std::min(x[a], x[b]), std::min(y[a], y[b]), std::max(x[a], x[b]),
                 std::max(y[a], y[b]), cx0, cy0, cx1, cy1);
      for (int cy = cy0; cy <= cy1; cy++)
        for (int cx = cx0; cx <= cx1; cx++) cells[(size_t)cy * gx + cx].push_back(e);
    }
  }
  void cell_range(double ax, double ay, double bx, double by, int &cx0, int &cy0, int &cx1,
                  int &cy1) const {
    cx0 = std::max(0, std::min(gx - 1, (int)std::floor((ax - x0) / cw)));
    cx1 = std::max(0, std::min(gx - 1, (int)std::floor((bx - x0) / cw)));
    cy0 = std::max(0, std::min(gy - 1, (int)std::floor((ay - y0) / ch)));
    cy1 = std::max(0, std::min(gy - 1, (int)std::floor((by - y0) / ch)));
  }
  static double orient(double ax, double ay, double bx, double by, double cx, double cy) {
    return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
  }
  // -1 lower surface, +1 upper surface, 0 trailing/leading edge point
  int side(int i) const {
    if (i == 0 || 2 * i == nw) return 0;
    return 2 * i < nw ? -1 : 1;
  }
  bool visible(int32_t p, int32_t q) const {
    const bool pw = p < nw, qw = q < nw;
    if (pw && qw) return side(p) * side(q) >= 0;
    const double ax = x[p], ay = y[p], bx = x[q], by = y[q];
    if (std::max(ax, bx) < x0 || std::min(ax, bx) > x1 || std::max(ay, by) < y0 ||
        std::min(ay, by) > y1)
      return true;
    int cx0, cy0, cx1, cy1;
    cell_range(std::min(ax, bx), std::min(ay, by), std::max(ax, bx), std::max(ay, by), cx0, cy0,
               cx1, cy1);
    for (int cy = cy0; cy <= cy1; cy++)
      for (int cx = cx0; cx <= cx1; cx++)
        for (int32_t e : cells[(size_t)cy * gx + cx]) {
          const int32_t a = e, b = (e + 1) % nw;
          if (a == p || a == q || b == p || b == q) continue;  // touches an end point
          const double o1 = orient(ax, ay, bx, by, x[a], y[a]);
          const double o2 = orient(ax, ay, bx, by, x[b], y[b]);
          const double o3 = orient(x[a], y[a], x[b], y[b], ax, ay);
          const double o4 = orient(x[a], y[a], x[b], y[b], bx, by);
          if (((o1 <= 0 && o2 >= 0) || (o1 >= 0 && o2 <= 0)) &&
              ((o3 <= 0 && o4 >= 0) || (o3 >= 0 && o4 <= 0)))
            return false;
        }
    return true;
  }
};

// sub-stencil balance of the neighbour selection
constexpr int kMinSub = 3;
constexpr double kMinRelDet = 0.05;

double derive_stretch(int L, double h_first, double depth) {
  // solve h_first * (q^L - 1)/(q - 1) = depth for q > 1
  double target = depth / h_first;
  if (target <= L) return 1.0 + 1e-9;
  double lo = 1.0 + 1e-12, hi = 2.0;
  for (int it = 0; it < 200; it++) {
    double mid = 0.5 * (lo + hi);
    double s = (std::pow(mid, L) - 1.0) / (mid - 1.0);
    if (s < target) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

}  // namespace

extern "C" {

int lskum_cloud_generate(const lskum_gen_params *p, lskum_cloud_h **out) {
  if (!p || !out) return LSKUM_EINVAL;
  *out = nullptr;
  if (p->n_wall < 8 || p->n_radial < 2 || p->n_circ < 8 || p->k < 3 ||
      p->k > 255 || p->r_far <= 1.0 || p->jitter < 0.0 || p->jitter >= 0.5 ||
      p->radius_factor <= 0.0 || p->naca_t <= 0 || p->naca_t >= 40 ||
      !(p->wall_blend >= 0.0 && p->wall_blend <= 1.0))
    return LSKUM_EINVAL;
  const int64_t n64 = (int64_t)p->n_wall + (int64_t)p->n_radial * p->n_circ;
  if (n64 >= (1 << 28)) return LSKUM_EINVAL;
  lskum_cloud_h *c = new (std::nothrow) lskum_cloud_h;
  if (!c) return LSKUM_ENOMEM;
  const int32_t nw = p->n_wall, L = p->n_radial, M = p->n_circ;
  const int32_t n = (int32_t)n64;
  Naca af{p->naca_m / 100.0, p->naca_p / 10.0, p->naca_t / 100.0};
  ArcSurface arc(af);
  const double R = p->r_far;
  c->n = n;
  c->x.assign(n, 0.0);
  c->y.assign(n, 0.0);
  c->left.assign(n, -1);
  c->right.assign(n, -1);
  c->flag1.assign(n, 1);
  c->flag2.assign(n, 0);
  c->min_dist.assign(n, 0.0);
  std::vector<double> h(n, 0.0);

  // radial distribution g_r = (q^r - 1)/(q^L - 1), r = 0..L
  const double h_circ = arc.total / M;
  const double depth = R - 0.5;
  const double q = p->stretch > 0.0 ? p->stretch : derive_stretch(L, h_circ, depth);
  std::vector<double> g(L + 1);
  for (int r = 0; r <= L; r++)
    g[r] = (q == 1.0) ? (double)r / L : (std::pow(q, r) - 1.0) / (std::pow(q, L) - 1.0);
  g[L] = 1.0;

  // The far-field circle is parametrised from the direction of the camber
  // line at the trailing edge (0 for a symmetric section), so that the rings
  // leave the trailing-edge wedge along its bisector rather than along +x,
  // which for an aft-cambered section runs along its lower surface.
  double yc_te, beta;
  af.camber(1.0, &yc_te, &beta);
  beta = std::atan(beta);
  auto ring_pos = [&](double s, double gg, double *X, double *Y) {
    double sx, sy;
    arc.at(s, &sx, &sy);
    const double th = -2.0 * kPi * s + beta;
    double fx = 0.5 + R * std::cos(th);
    double fy = R * std::sin(th);
    *X = (1.0 - gg) * sx + gg * fx;
    *Y = (1.0 - gg) * sy + gg * fy;
  };

  // wall: cosine chord spacing, clockwise from the trailing edge, or its
  // arc-length positions blended toward uniform arc length
  const double wb = p->wall_blend;
  for (int32_t i = 0; i < nw; i++) {
    if (wb > 0.0) {
      const double s = (1.0 - wb) * arc.fraction(2.0 * kPi * i / nw) + wb * ((double)i / nw);
      arc.at(s, &c->x[i], &c->y[i]);
    } else {
      af.surf(2.0 * kPi * i / nw, &c->x[i], &c->y[i]);
    }
    c->flag1[i] = 0;
    c->flag2[i] = 1;
    c->left[i] = (i + nw - 1) % nw;
    c->right[i] = (i + 1) % nw;
  }
  for (int32_t i = 0; i < nw; i++) {
    int32_t l = c->left[i], r = c->right[i];
    double a = std::hypot(c->x[i] - c->x[l], c->y[i] - c->y[l]);
    double b = std::hypot(c->x[r] - c->x[i], c->y[r] - c->y[i]);
    h[i] = std::max(std::max(a, b), g[1] * depth);
  }
  // rings 1..L (L = far field), jittered in (s, g) except the far field
  SplitMix64 rng{p->seed};
  for (int32_t r = 1; r <= L; r++) {
    for (int32_t j = 0; j < M; j++) {
      int32_t id = nw + (r - 1) * M + j;
      double s = (j + 0.5) / M, gg = g[r];
      // local spacing from the unjittered map
      double x0, y0, x1, y1, xa, ya, xb, yb;
      ring_pos(s, gg, &x0, &y0);
      ring_pos(s + 1.0 / M, gg, &x1, &y1);
      ring_pos(s - 1.0 / M, gg, &xa, &ya);
      double circ = std::max(std::hypot(x1 - x0, y1 - y0), std::hypot(x0 - xa, y0 - ya));
      ring_pos(s, g[r - 1], &xa, &ya);
      double rad = std::hypot(x0 - xa, y0 - ya);
      if (r < L) {
        ring_pos(s, g[r + 1], &xb, &yb);
        rad = std::max(rad, std::hypot(xb - x0, yb - y0));
      }
      h[id] = std::max(circ, rad);
      if (r < L) {
        double us = rng.sym(), ug = rng.sym();
        double dg = std::min(g[r] - g[r - 1], g[r + 1] - g[r]);
        s += p->jitter * us / M;
        gg += p->jitter * ug * dg;
        ring_pos(s, gg, &c->x[id], &c->y[id]);
        c->flag1[id] = 1;
      } else {
        c->x[id] = x0;
        c->y[id] = y0;
        c->flag1[id] = 2;
        c->left[id] = nw + (r - 1) * M + (j + M - 1) % M;
        c->right[id] = nw + (r - 1) * M + (j + 1) % M;
      }
    }
  }
  // Ring points inside the body.  The rings blend each surface point toward
  // the far-field circle, which at the trailing edge lies along +x; where a
  // cambered section's lower surface descends toward the trailing edge, that
  // direction enters the thin wedge, and near-wall ring points land inside
  // the body.  Such a point moves to the foot of its nearest wall edge plus
  // its ring's wall distance along the edge's outward normal.  (A symmetric
  // section's surfaces approach the trailing edge from outside that
  // direction, so none of its points move.)
  {
    double area2 = 0.0;  // twice the signed area of the wall polygon
    for (int32_t e = 0; e < nw; e++) {
      const int32_t b = (e + 1) % nw;
      area2 += c->x[e] * c->y[b] - c->x[b] * c->y[e];
    }
    const double sgn = area2 > 0.0 ? 1.0 : -1.0;  // outward = sgn * (dy, -dx)
    double bx0 = 1e300, bx1 = -1e300, by0 = 1e300, by1 = -1e300;
    for (int32_t e = 0; e < nw; e++) {
      bx0 = std::min(bx0, c->x[e]); bx1 = std::max(bx1, c->x[e]);
      by0 = std::min(by0, c->y[e]); by1 = std::max(by1, c->y[e]);
    }
    auto inside = [&](double px, double py) {
      bool in = false;
      for (int32_t e = 0; e < nw; e++) {
        const int32_t b = (e + 1) % nw;
        const double ya = c->y[e], yb = c->y[b];
        if ((ya > py) != (yb > py) &&
            px < c->x[e] + (py - ya) * (c->x[b] - c->x[e]) / (yb - ya))
          in = !in;
      }
      return in;
    };
    for (int32_t r = 1; r < L; r++) {
#pragma omp parallel for schedule(dynamic, 64)
      for (int32_t j = 0; j < M; j++) {
        const int32_t id = nw + (r - 1) * M + j;
        const double px = c->x[id], py = c->y[id];
        // the ring's distance from the wall at this arc position
        const double s = (j + 0.5) / M;
        double sx, sy, fx_far, fy_far;
        ring_pos(s, 0.0, &sx, &sy);
        ring_pos(s, 1.0, &fx_far, &fy_far);
        const double dist = g[r] * std::hypot(fx_far - sx, fy_far - sy);
        if (px < bx0 || px > bx1 || py < by0 || py > by1 || !inside(px, py)) continue;
        int32_t best = 0;
        double bd = 1e300, fx = 0.0, fy = 0.0;
        for (int32_t e = 0; e < nw; e++) {
          const int32_t b = (e + 1) % nw;
          const double ex = c->x[b] - c->x[e], ey = c->y[b] - c->y[e];
          const double ee = ex * ex + ey * ey;
          double t = ee > 0.0 ? ((px - c->x[e]) * ex + (py - c->y[e]) * ey) / ee : 0.0;
          t = std::max(0.0, std::min(1.0, t));
          const double qx = c->x[e] + t * ex, qy = c->y[e] + t * ey;
          const double d = (px - qx) * (px - qx) + (py - qy) * (py - qy);
          if (d < bd) { bd = d; best = e; fx = qx; fy = qy; }
        }
        const int32_t b = (best + 1) % nw;
        const double ex = c->x[b] - c->x[best], ey = c->y[b] - c->y[best];
        const double el = std::hypot(ex, ey);
        if (!(el > 0.0)) continue;
        const double nxp = fx + dist * sgn * ey / el, nyp = fy - dist * sgn * ex / el;
        if (!inside(nxp, nyp)) {
          c->x[id] = nxp;
          c->y[id] = nyp;
        }
      }
    }
  }
  // neighbours: k nearest visible points within radius_factor * h, ordered
  // by (distance, index), then balanced so that every sub-stencil the flux
  // needs (x+/x- along the point's tangent; y+/y- for interior points) has
  // at least kMinSub members and a least-squares matrix that is not nearly
  // singular: nearest qualifying candidates within twice the radius replace
  // the farthest members no sub-stencil depends on.
  std::vector<double> nxv, nyv;
  {
    lskum_cloud v{};
    v.n_points = n;
    v.x = c->x.data();
    v.y = c->y.data();
    v.left = c->left.data();
    v.right = c->right.data();
    v.flag_1 = c->flag1.data();
    lskum::compute_normals(v, nxv, nyv);
  }
  KdTree tree(c->x, c->y);
  Body body(c->x, c->y, nw);
  auto visible = [&](int32_t a, int32_t b) { return body.visible(a, b); };
  std::vector<int32_t> cnt(n, 0);
  std::vector<int32_t> nb((size_t)n * p->k, -1);
  const int k = p->k;
#pragma omp parallel
  {
    std::vector<std::pair<double, int32_t>> cand;
    cand.reserve(4 * k + 1);
    std::vector<char> used;
#pragma omp for schedule(dynamic, 4096)
    for (int32_t i = 0; i < n; i++) {
      const double rr = p->radius_factor * h[i];
      tree.query(i, 4 * k, 4.0 * rr * rr, cand, visible);
      int nc = (int)cand.size();
      used.assign(nc, 0);
      // sub-stencil membership of each candidate in i's (t, n) frame
      const double tx = nyv[i], ty = -nxv[i], nxi = nxv[i], nyi = nyv[i];
      const int f = c->flag1[i];
      auto bits = [&](int a) {
        const int32_t q = cand[a].second;
        const double dx = c->x[q] - c->x[i], dy = c->y[q] - c->y[i];
        const double ds = dx * tx + dy * ty, dn = dx * nxi + dy * nyi;
        unsigned b = 0;
        if (ds <= 0.0) b |= 1u;
        if (ds >= 0.0) b |= 2u;
        if (f == 1) {
          if (dn <= 0.0) b |= 4u;
          if (dn >= 0.0) b |= 8u;
        } else {
          b |= (f == 0) ? 8u : 4u;
        }
        return b;
      };
      auto rel_det = [&](unsigned sb) {
        double sxx = 0, syy = 0, sxy = 0;
        int m = 0;
        for (int a = 0; a < nc; a++) {
          if (!used[a] || !(bits(a) & sb)) continue;
          const int32_t q = cand[a].second;
          const double dx = c->x[q] - c->x[i], dy = c->y[q] - c->y[i];
          const double ds = dx * tx + dy * ty, dn = dx * nxi + dy * nyi;
          sxx += ds * ds;
          syy += dn * dn;
          sxy += ds * dn;
          m++;
        }
        if (m < kMinSub || !(sxx > 0.0 && syy > 0.0)) return -1.0;
        return (sxx * syy - sxy * sxy) / (sxx * syy);
      };
      int taken = 0;
      for (int a = 0; a < nc && taken < k; a++)
        if (cand[a].first <= rr * rr) {
          used[a] = 1;
          taken++;
        }
      const unsigned need = (f == 1) ? 15u : (f == 0 ? 11u : 7u);
      for (unsigned sb = 1; sb <= 8; sb <<= 1) {
        if (!(need & sb)) continue;
        for (int guard = 0; guard < nc && rel_det(sb) < kMinRelDet; guard++) {
          int add = -1;
          for (int a = 0; a < nc; a++)
            if (!used[a] && (bits(a) & sb)) { add = a; break; }
          if (add < 0) break;
          used[add] = 1;
          if (++taken <= k) continue;
          // drop the farthest member whose removal keeps every needed
          // sub-stencil at kMinSub members
          for (int a = nc - 1; a >= 0; a--) {
            if (!used[a] || a == add) continue;
            used[a] = 0;
            bool ok = true;
            for (unsigned t = 1; t <= 8 && ok; t <<= 1) {
              if (!(need & t) || !(bits(a) & t)) continue;
              int m = 0;
              for (int b2 = 0; b2 < nc; b2++) m += used[b2] && (bits(b2) & t);
              ok = m >= kMinSub;
            }
            if (ok) { taken--; break; }
            used[a] = 1;
          }
        }
      }
      // Fallback where the nearest candidates are nearly collinear, as for
      // first-ring points at a thin trailing edge, whose rings keep the
      // wall's tangential clustering but sit a full radial spacing apart:
      // widen the candidate set and add, one at a time, the candidate that
      // raises the sub-stencil's determinant most, dropping the farthest
      // member whose removal leaves every needed sub-stencil with kMinSub
      // members and no lower determinant than the target (or than it had).
      for (unsigned sb = 1; sb <= 8; sb <<= 1) {
        if (!(need & sb)) continue;
        for (int widen = 1; widen <= 2 && rel_det(sb) < kMinRelDet; widen++) {
          std::vector<int32_t> ids;
          for (int a = 0; a < nc; a++)
            if (used[a]) ids.push_back(cand[a].second);
          const double scale = (double)(2 << widen);  // 4x, then 8x the radius
          tree.query(i, 4 * k << (2 * widen), scale * scale * rr * rr, cand, visible);
          nc = (int)cand.size();
          used.assign(nc, 0);
          for (int a = 0; a < nc; a++)
            used[a] = std::find(ids.begin(), ids.end(), cand[a].second) != ids.end();
          for (int step = 0; step < 8 && rel_det(sb) < kMinRelDet; step++) {
            const double cur = rel_det(sb);
            int best = -1;
            double best_det = cur;
            for (int a = 0; a < nc; a++) {
              if (used[a] || !(bits(a) & sb)) continue;
              used[a] = 1;
              const double d = rel_det(sb);
              used[a] = 0;
              if (d > best_det) { best_det = d; best = a; }
            }
            if (best < 0) break;
            used[best] = 1;
            if (++taken <= k) continue;
            double before[4];
            for (int t = 0; t < 4; t++) before[t] = (need >> t & 1u) ? rel_det(1u << t) : 0.0;
            bool dropped = false;
            for (int a = nc - 1; a >= 0 && !dropped; a--) {
              if (!used[a] || a == best) continue;
              used[a] = 0;
              bool ok = true;
              for (int t = 0; t < 4 && ok; t++) {
                if (!(need >> t & 1u)) continue;
                ok = rel_det(1u << t) >= std::min(before[t], kMinRelDet);
              }
              if (ok) {
                taken--;
                dropped = true;
              } else {
                used[a] = 1;
              }
            }
            if (!dropped) {  // nothing can make room: keep the list as it was
              used[best] = 0;
              taken--;
              break;
            }
          }
        }
      }
      int t = 0;
      for (int a = 0; a < nc; a++)
        if (used[a] && t < k) nb[(size_t)i * k + t++] = cand[a].second;
      cnt[i] = t;
      c->min_dist[i] = t ? std::sqrt(cand[0].first) : 0.0;
      for (int a = 0; a < nc; a++)
        if (used[a]) { c->min_dist[i] = std::sqrt(cand[a].first); break; }
    }
  }
  c->conn_off.assign(n + 1, 0);
  for (int32_t i = 0; i < n; i++) c->conn_off[i + 1] = c->conn_off[i] + cnt[i];
  c->conn.resize(c->conn_off[n]);
  for (int32_t i = 0; i < n; i++)
    for (int32_t t = 0; t < cnt[i]; t++) c->conn[c->conn_off[i] + t] = nb[(size_t)i * k + t];
  *out = c;
  return LSKUM_OK;
}

int lskum_cloud_read_partgrid(const char *path, lskum_cloud_h **out) {
  if (!path || !out) return LSKUM_EINVAL;
  *out = nullptr;
  FILE *f = std::fopen(path, "r");
  if (!f) return LSKUM_EIO;
  lskum_cloud_h *c = new (std::nothrow) lskum_cloud_h;
  if (!c) { std::fclose(f); return LSKUM_ENOMEM; }
  int n = 0;
  int rc = LSKUM_OK;
  if (std::fscanf(f, "%d", &n) != 1 || n <= 0 || n >= (1 << 28)) rc = LSKUM_EIO;
  if (rc == LSKUM_OK) {
    c->n = n;
    c->x.resize(n); c->y.resize(n); c->min_dist.resize(n);
    c->left.resize(n); c->right.resize(n); c->flag1.resize(n); c->flag2.resize(n);
    c->conn_off.assign(n + 1, 0);
    for (int i = 0; i < n && rc == LSKUM_OK; i++) {
      int l, r, f1, f2, nb;
      if (std::fscanf(f, "%lf %lf %d %d %d %d %lf %d", &c->x[i], &c->y[i], &l, &r,
                      &f1, &f2, &c->min_dist[i], &nb) != 8 || nb < 0 || nb > 255) {
        rc = LSKUM_EIO;
        break;
      }
      // 1-based in the file (0 = none) -> 0-based (-1 = none)
      c->left[i] = l - 1;
      c->right[i] = r - 1;
      c->flag1[i] = f1;
      c->flag2[i] = f2;
      for (int t = 0; t < nb; t++) {
        int v;
        if (std::fscanf(f, "%d", &v) != 1) { rc = LSKUM_EIO; break; }
        c->conn.push_back(v - 1);
      }
      c->conn_off[i + 1] = (int32_t)c->conn.size();
    }
  }
  std::fclose(f);
  if (rc != LSKUM_OK) { delete c; return rc; }
  lskum_cloud v = c->view();
  std::string err;
  if (lskum::check_cloud(v, &err) != LSKUM_OK) { delete c; return LSKUM_EIO; }
  *out = c;
  return LSKUM_OK;
}

int lskum_cloud_write_partgrid(const lskum_cloud_h *c, const char *path) {
  if (!c || !path) return LSKUM_EINVAL;
  FILE *f = std::fopen(path, "w");
  if (!f) return LSKUM_EIO;
  std::fprintf(f, "%d\n", c->n);
  for (int32_t i = 0; i < c->n; i++) {
    std::fprintf(f, "%.17g %.17g %d %d %d %d %.17g %d", c->x[i], c->y[i],
                 c->left[i] + 1, c->right[i] + 1, c->flag1[i], c->flag2[i],
                 c->min_dist[i], c->conn_off[i + 1] - c->conn_off[i]);
    for (int32_t e = c->conn_off[i]; e < c->conn_off[i + 1]; e++)
      std::fprintf(f, " %d", c->conn[e] + 1);
    std::fputc('\n', f);
  }
  int bad = std::ferror(f);
  if (std::fclose(f) != 0 || bad) return LSKUM_EIO;
  return LSKUM_OK;
}

int lskum_cloud_from_arrays(const lskum_cloud *src, lskum_cloud_h **out) {
  if (!src || !out) return LSKUM_EINVAL;
  *out = nullptr;
  std::string err;
  if (lskum::check_cloud(*src, &err) != LSKUM_OK) return LSKUM_EINVAL;
  lskum_cloud_h *c = new (std::nothrow) lskum_cloud_h;
  if (!c) return LSKUM_ENOMEM;
  int32_t n = src->n_points;
  c->n = n;
  c->x.assign(src->x, src->x + n);
  c->y.assign(src->y, src->y + n);
  c->min_dist.assign(src->min_dist, src->min_dist + n);
  c->left.assign(src->left, src->left + n);
  c->right.assign(src->right, src->right + n);
  c->flag1.assign(src->flag_1, src->flag_1 + n);
  c->flag2.assign(src->flag_2, src->flag_2 + n);
  c->conn_off.assign(src->conn_off, src->conn_off + n + 1);
  c->conn.assign(src->conn, src->conn + src->conn_off[n]);
  *out = c;
  return LSKUM_OK;
}

int lskum_cloud_view(const lskum_cloud_h *c, lskum_cloud *view) {
  if (!c || !view) return LSKUM_EINVAL;
  *view = c->view();
  return LSKUM_OK;
}

int lskum_cloud_stats(const lskum_cloud_h *c, int32_t *min_nbhs,
                      int32_t *max_nbhs, double *min_rel_det) {
  if (!c) return LSKUM_EINVAL;
  lskum_cloud v = c->view();
  int32_t lo = 1 << 30, hi = 0;
  for (int32_t i = 0; i < c->n; i++) {
    int32_t d = c->conn_off[i + 1] - c->conn_off[i];
    lo = std::min(lo, d);
    hi = std::max(hi, d);
  }
  if (min_nbhs) *min_nbhs = lo;
  if (max_nbhs) *max_nbhs = hi;
  if (min_rel_det) {
    std::vector<double> nx, ny;
    std::vector<uint8_t> mask;
    lskum::compute_normals(v, nx, ny);
    lskum::generate_connectivity(v, nx, ny, mask);
    for (int t = 0; t < 3; t++) min_rel_det[t] = -1.0;
    for (int32_t i = 0; i < c->n; i++) {
      double r = lskum::substencil_min_rel_det(v, nx, ny, mask, i);
      int t = c->flag1[i];
      if (min_rel_det[t] < 0.0 || r < min_rel_det[t]) min_rel_det[t] = r;
    }
  }
  return LSKUM_OK;
}

int lskum_cloud_connectivity(const lskum_cloud_h *c, double *nx, double *ny, uint8_t *mask) {
  if (!c || !nx || !ny || !mask) return LSKUM_EINVAL;
  lskum_cloud v = c->view();
  std::vector<double> a, b;
  std::vector<uint8_t> m;
  lskum::compute_normals(v, a, b);
  lskum::generate_connectivity(v, a, b, m);
  std::copy(a.begin(), a.end(), nx);
  std::copy(b.begin(), b.end(), ny);
  std::copy(m.begin(), m.end(), mask);
  return LSKUM_OK;
}

int lskum_cloud_rcm(const lskum_cloud_h *c, int32_t *perm) {
  if (!c || !perm) return LSKUM_EINVAL;
  std::vector<int32_t> p;
  lskum::rcm_order(c->view(), p);
  std::copy(p.begin(), p.end(), perm);
  return LSKUM_OK;
}

void lskum_cloud_free(lskum_cloud_h *c) { delete c; }

}  // extern "C"
