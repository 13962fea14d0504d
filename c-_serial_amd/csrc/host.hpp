// host.hpp -- host-side set-up shared by the cloud tools and the context.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lskum.h"

struct lskum_cloud_h {
  int32_t n = 0;
  std::vector<double> x, y, min_dist;
  std::vector<int32_t> left, right, flag1, flag2, conn_off, conn;
  lskum_cloud view() const;
};

namespace lskum {

// Sub-stencil membership bits of one neighbour, as generate_connectivity
// assigns them (generate_connectivity_mod.h:37-71, :122-143, :188-209).
enum : uint32_t { XPOS = 1u, XNEG = 2u, YPOS = 4u, YNEG = 8u };

// Validated copy of a caller's cloud.
int check_cloud(const lskum_cloud &c, std::string *err);

// point_normals_mod.h:17-55
void compute_normals(const lskum_cloud &c, std::vector<double> &nx,
                     std::vector<double> &ny);

// generate_connectivity_mod.h:221-249: one mask per conn entry.
void generate_connectivity(const lskum_cloud &c, const std::vector<double> &nx,
                           const std::vector<double> &ny,
                           std::vector<uint8_t> &mask);

// Smallest relative determinant det/(sxx*syy) of the sub-stencil least-squares
// matrices (the det of interior_fluxes_mod.h:104) over the point's
// sub-stencils; 0 for an empty or collinear sub-stencil.
double substencil_min_rel_det(const lskum_cloud &c,
                              const std::vector<double> &nx,
                              const std::vector<double> &ny,
                              const std::vector<uint8_t> &mask, int32_t i);

// Reverse Cuthill-McKee order of the symmetrised neighbour graph.
// perm[new] = old.
void rcm_order(const lskum_cloud &c, std::vector<int32_t> &perm);

// Hilbert-curve order of the point coordinates (21 bits per axis over the
// bounding box; ties by index): consecutive points are close in space, so a
// block of consecutive points reaches a small set of neighbours.
void hilbert_order(const lskum_cloud &c, std::vector<int32_t> &perm);

// The reverse of the sub-stencil coefficients for the shape gradient: per
// point d, from the adjoints gc[8d..8d+7] of its coefficients (a, b) =
// (syy, -sxy)/det (x+, x-) or (sxx, -sxy)/det (y+, y-), the adjoints of its
// pairs' offsets (added to ge[2e], ge[2e+1]) and of its normal (added to
// gn[2d], gn[2d+1]).  conn entries are index | mask << 28.  Points are
// independent, so the loop runs on every host thread.
void coef_adjoint(int32_t n, const double *x, const double *y, const double *nx,
                  const double *ny, const int32_t *off, const uint32_t *conn,
                  const double *gc, double *ge, double *gn);

}  // namespace lskum
