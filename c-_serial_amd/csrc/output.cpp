// output.cpp -- solution and surface-pressure files of a state, the
// reference's two writers with their defects fixed.
//
// print_primal_output (post_processing_mod.h:7-27) writes a header "n it
// res_old" and then, per point, x y flag_1 rho u1 u2 p Mach entropy -- but
// without a newline or a separator between points, so the file cannot be
// read back, at the stream's default 6 digits, and with Mach computed as
// sqrt(|u|^2 / c) where |u| / c is meant.  Here: one point per line,
// 17 significant digits, Mach = |u| / c with c = sqrt(gamma p / rho).
//
// compute_cl_cd_cm writes its Cp file (compute_force_coeffs_mod.h:20-21,54)
// with `fout << flag_2, x, cp;`: the comma operator sends only flag_2, with
// no separator.  Here: per shape point, in the force loop's order (the
// first shape_points points), "flag_2 x cp" on a line, cp = -(p - p_inf) /
// (rho_inf M^2 / 2).
#include <cmath>
#include <cstdio>

#include "host.hpp"

extern "C" {

int lskum_solution_write(const char *path, const lskum_cloud *c, int32_t it, double res_old,
                         const double *prim, const double *entropy) {
  if (!path || !c || !prim || c->n_points <= 0 || it < 0) return LSKUM_EINVAL;
  FILE *f = std::fopen(path, "w");
  if (!f) return LSKUM_EIO;
  const int32_t n = c->n_points;
  const double gamma = 1.4;
  bool ok = std::fprintf(f, "%d %d %.17g\n", n, it, res_old) > 0;
  for (int32_t i = 0; i < n && ok; i++) {
    const double rho = prim[i], u1 = prim[n + i], u2 = prim[2 * (size_t)n + i];
    const double pr = prim[3 * (size_t)n + i];
    const double sos = std::sqrt(gamma * pr / rho);
    const double mach = std::sqrt(u1 * u1 + u2 * u2) / sos;
    ok = std::fprintf(f, "%.17g %.17g %d %.17g %.17g %.17g %.17g %.17g %.17g\n", c->x[i], c->y[i],
                      c->flag_1[i], rho, u1, u2, pr, mach, entropy ? entropy[i] : 0.0) > 0;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? LSKUM_OK : LSKUM_EIO;
}

int lskum_cp_write(const char *path, const lskum_cloud *c, const double *prim, double mach) {
  if (!path || !c || !prim || c->n_points <= 0 || !(mach > 0.0)) return LSKUM_EINVAL;
  FILE *f = std::fopen(path, "w");
  if (!f) return LSKUM_EIO;
  const int32_t n = c->n_points;
  const double rho_inf = 1.0, pr_inf = 1.0 / 1.4;  // parameter_mod.h:7-8
  const double temp = 0.5 * rho_inf * mach * mach;
  // the force loop runs over the first shape_points points (m = j,
  // compute_force_coeffs_mod.h:25-27): as many as have flag_2 > 0
  int32_t shape_points = 0;
  for (int32_t i = 0; i < n; i++)
    if (c->flag_2[i] > 0) shape_points++;
  bool ok = true;
  for (int32_t m = 0; m < shape_points && ok; m++) {
    double cp = prim[3 * (size_t)n + m] - pr_inf;
    cp = -cp / temp;
    ok = std::fprintf(f, "%d %.17g %.17g\n", c->flag_2[m], c->x[m], cp) > 0;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? LSKUM_OK : LSKUM_EIO;
}

}  // extern "C"
