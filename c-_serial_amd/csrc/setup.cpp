// setup.cpp -- one-time host set-up of a cloud: validation, normals,
// sub-stencil classification and the device point orders (RCM, Hilbert).
//
// Built with -ffp-contract=off: the sub-stencil tests compare dels/deln with
// zero exactly as the reference does, so the lists are bit-identical to the
// reference's on any x86-64 host.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <queue>

#include "host.hpp"

namespace lskum {

int check_cloud(const lskum_cloud &c, std::string *err) {
  char buf[256];
  auto fail = [&](const char *m, long a) {
    std::snprintf(buf, sizeof buf, "%s (%ld)", m, a);
    if (err) *err = buf;
    return LSKUM_EINVAL;
  };
  if (c.n_points <= 0) return fail("cloud has no points", c.n_points);
  if (!c.x || !c.y || !c.left || !c.right || !c.flag_1 || !c.flag_2 ||
      !c.min_dist || !c.conn_off || !c.conn)
    return fail("cloud has a null array", 0);
  if (c.n_points >= (1 << 28)) return fail("too many points", c.n_points);
  const int32_t n = c.n_points;
  if (c.conn_off[0] != 0) return fail("conn_off[0] must be 0", c.conn_off[0]);
  for (int32_t i = 0; i < n; i++) {
    if (c.conn_off[i + 1] < c.conn_off[i])
      return fail("conn_off not monotone at point", i);
    if (c.conn_off[i + 1] - c.conn_off[i] > 255)
      return fail("more than 255 neighbours at point", i);
    int f = c.flag_1[i];
    if (f < 0 || f > 2) return fail("flag_1 outside 0..2 at point", i);
    if (f != 1) {
      if (c.left[i] < 0 || c.left[i] >= n || c.right[i] < 0 || c.right[i] >= n)
        return fail("boundary point without valid left/right", i);
    }
    if (!(c.min_dist[i] >= 0.0)) return fail("min_dist not >= 0 at point", i);
  }
  for (int64_t e = 0; e < c.conn_off[n]; e++)
    if (c.conn[e] < 0 || c.conn[e] >= n)
      return fail("neighbour index out of range at entry", (long)e);
  return LSKUM_OK;
}

void compute_normals(const lskum_cloud &c, std::vector<double> &nx_out,
                     std::vector<double> &ny_out) {
  const int32_t n = c.n_points;
  nx_out.assign(n, 0.0);
  ny_out.assign(n, 1.0);
  for (int32_t m = 0; m < n; m++) {
    if (c.flag_1[m] == 0 || c.flag_1[m] == 2) {
      int32_t l = c.left[m], r = c.right[m];
      double nx1 = c.y[m] - c.y[l], nx2 = c.y[r] - c.y[m];
      double ny1 = c.x[m] - c.x[l], ny2 = c.x[r] - c.x[m];
      double nx = 0.5 * (nx1 + nx2);
      double ny = 0.5 * (ny1 + ny2);
      double det = std::sqrt(nx * nx + ny * ny);
      nx_out[m] = -nx / det;
      ny_out[m] = ny / det;
    } else {
      nx_out[m] = 0;
      ny_out[m] = 1;
    }
  }
}

void generate_connectivity(const lskum_cloud &c, const std::vector<double> &nxv,
                           const std::vector<double> &nyv,
                           std::vector<uint8_t> &mask) {
  const int32_t n = c.n_points;
  mask.assign(c.conn_off[n], 0);
  for (int32_t i = 0; i < n; i++) {
    const double xi = c.x[i], yi = c.y[i], nx = nxv[i], ny = nyv[i];
    const double tx = ny, ty = -nx;
    const int f = c.flag_1[i];
    for (int32_t e = c.conn_off[i]; e < c.conn_off[i + 1]; e++) {
      const int32_t k = c.conn[e];
      const double delx = c.x[k] - xi, dely = c.y[k] - yi;
      const double dels = delx * tx + dely * ty;
      const double deln = delx * nx + dely * ny;
      uint8_t m = 0;
      if (dels <= 0.0) m |= XPOS;
      if (dels >= 0.0) m |= XNEG;
      if (f == 0) {
        m |= YNEG;
      } else if (f == 1) {
        if (deln <= 0.0) m |= YPOS;
        if (deln >= 0.0) m |= YNEG;
      } else {
        m |= YPOS;
      }
      mask[e] = m;
    }
  }
}

double substencil_min_rel_det(const lskum_cloud &c,
                              const std::vector<double> &nxv,
                              const std::vector<double> &nyv,
                              const std::vector<uint8_t> &mask, int32_t i) {
  const double nx = nxv[i], ny = nyv[i], tx = ny, ty = -nx;
  const int f = c.flag_1[i];
  uint8_t need = (f == 0) ? (XPOS | XNEG | YNEG)
                 : (f == 1) ? (XPOS | XNEG | YPOS | YNEG)
                            : (XPOS | XNEG | YPOS);
  double worst = 1.0;
  for (int s = 0; s < 4; s++) {
    uint8_t bit = (uint8_t)(1u << s);
    if (!(need & bit)) continue;
    double sxx = 0, syy = 0, sxy = 0;
    for (int32_t e = c.conn_off[i]; e < c.conn_off[i + 1]; e++) {
      if (!(mask[e] & bit)) continue;
      int32_t k = c.conn[e];
      double delx = c.x[k] - c.x[i], dely = c.y[k] - c.y[i];
      double dels = delx * tx + dely * ty, deln = delx * nx + dely * ny;
      sxx += dels * dels;
      syy += deln * deln;
      sxy += dels * deln;
    }
    double rel = (sxx > 0 && syy > 0) ? (sxx * syy - sxy * sxy) / (sxx * syy) : 0.0;
    worst = std::min(worst, rel);
  }
  return worst;
}

void rcm_order(const lskum_cloud &c, std::vector<int32_t> &perm) {
  const int32_t n = c.n_points;
  // symmetrised adjacency, sorted and de-duplicated
  std::vector<int64_t> deg(n + 1, 0);
  for (int32_t i = 0; i < n; i++)
    for (int32_t e = c.conn_off[i]; e < c.conn_off[i + 1]; e++) {
      int32_t k = c.conn[e];
      if (k == i) continue;
      deg[i + 1]++;
      deg[k + 1]++;
    }
  for (int32_t i = 0; i < n; i++) deg[i + 1] += deg[i];
  std::vector<int32_t> adj(deg[n]);
  std::vector<int64_t> fill(deg.begin(), deg.end() - 1);
  for (int32_t i = 0; i < n; i++)
    for (int32_t e = c.conn_off[i]; e < c.conn_off[i + 1]; e++) {
      int32_t k = c.conn[e];
      if (k == i) continue;
      adj[fill[i]++] = k;
      adj[fill[k]++] = i;
    }
  std::vector<int32_t> d(n);
  std::vector<int64_t> off(n + 1, 0);
  {
    std::vector<int32_t> compact;
    compact.reserve(adj.size());
    for (int32_t i = 0; i < n; i++) {
      auto b = adj.begin() + deg[i], en = adj.begin() + deg[i + 1];
      std::sort(b, en);
      auto u = std::unique(b, en);
      off[i] = (int64_t)compact.size();
      compact.insert(compact.end(), b, u);
      d[i] = (int32_t)(u - b);
    }
    off[n] = (int64_t)compact.size();
    adj.swap(compact);
  }
  std::vector<char> seen(n, 0);
  std::vector<int32_t> order;
  order.reserve(n);
  std::vector<int32_t> level(n, -1);
  // BFS from s returning the last node of the deepest level (lowest degree)
  auto bfs_far = [&](int32_t s, int32_t &depth) {
    std::vector<int32_t> q{s};
    level[s] = 0;
    size_t h = 0;
    int32_t far = s;
    while (h < q.size()) {
      int32_t v = q[h++];
      if (level[v] > level[far] || (level[v] == level[far] && d[v] < d[far])) far = v;
      for (int64_t e = off[v]; e < off[v + 1]; e++) {
        int32_t w = adj[e];
        if (level[w] < 0) {
          level[w] = level[v] + 1;
          q.push_back(w);
        }
      }
    }
    depth = level[far];
    for (int32_t v : q) level[v] = -1;
    return far;
  };
  for (int32_t start = 0; start < n; start++) {
    if (seen[start]) continue;
    // pseudo-peripheral node: repeat BFS while the eccentricity grows
    int32_t depth = 0, s = start;
    int32_t far = bfs_far(s, depth);
    for (int rep = 0; rep < 8; rep++) {
      int32_t d2 = 0;
      int32_t far2 = bfs_far(far, d2);
      if (d2 <= depth) break;
      s = far;
      far = far2;
      depth = d2;
    }
    s = far;
    // Cuthill-McKee BFS, neighbours by increasing degree then index
    size_t h = order.size();
    order.push_back(s);
    seen[s] = 1;
    std::vector<int32_t> nb;
    while (h < order.size()) {
      int32_t v = order[h++];
      nb.clear();
      for (int64_t e = off[v]; e < off[v + 1]; e++)
        if (!seen[adj[e]]) nb.push_back(adj[e]);
      std::sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) {
        return d[a] != d[b] ? d[a] < d[b] : a < b;
      });
      for (int32_t w : nb) {
        seen[w] = 1;
        order.push_back(w);
      }
    }
  }
  perm.assign(order.rbegin(), order.rend());
}

void hilbert_order(const lskum_cloud &c, std::vector<int32_t> &perm) {
  const int32_t n = c.n_points;
  double x0 = c.x[0], x1 = c.x[0], y0 = c.y[0], y1 = c.y[0];
  for (int32_t i = 1; i < n; i++) {
    x0 = std::min(x0, c.x[i]); x1 = std::max(x1, c.x[i]);
    y0 = std::min(y0, c.y[i]); y1 = std::max(y1, c.y[i]);
  }
  const int bits = 21;
  const uint32_t side = 1u << bits;
  const double sx = (x1 > x0) ? (side - 1) / (x1 - x0) : 0.0;
  const double sy = (y1 > y0) ? (side - 1) / (y1 - y0) : 0.0;
  std::vector<std::pair<uint64_t, int32_t>> key(n);
  for (int32_t i = 0; i < n; i++) {
    uint32_t xi = (uint32_t)((c.x[i] - x0) * sx), yi = (uint32_t)((c.y[i] - y0) * sy);
    uint64_t d = 0;
    for (uint32_t s = side >> 1; s > 0; s >>= 1) {  // xy -> Hilbert distance
      const uint32_t rx = (xi & s) ? 1u : 0u, ry = (yi & s) ? 1u : 0u;
      d += (uint64_t)s * s * ((3u * rx) ^ ry);
      if (ry == 0) {  // rotate the quadrant; only bits below s matter later
        if (rx == 1) {
          xi = side - 1 - xi;
          yi = side - 1 - yi;
        }
        std::swap(xi, yi);
      }
    }
    key[i] = {d, i};
  }
  std::sort(key.begin(), key.end());
  perm.resize(n);
  for (int32_t i = 0; i < n; i++) perm[i] = key[i].second;
}

}  // namespace lskum
