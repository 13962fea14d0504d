// restart.cpp -- the reference's restart.dat layout (initial_conditions_mod.h:
// 7-21 reads it): a header "n itr res_old", then per point nine numbers, of
// which the reader skips the first five and takes prim[1..4] (rho, u1, u2,
// p).  The reference has no writer for this layout (print_primal_output,
// post_processing_mod.h:7-27, writes x, y, flag_1 and then prim, which that
// reader would misalign); this one puts x, y, flag_1, flag_2 and min_dist in
// the skipped fields, one point per line, doubles at 17 significant digits
// so that a state survives the round trip bit for bit.
#include <cstdio>

#include "host.hpp"

extern "C" {

int lskum_restart_write(const char *path, const lskum_cloud *c, int32_t it, double res_old,
                        const double *prim) {
  if (!path || !c || !prim || c->n_points <= 0 || it < 0) return LSKUM_EINVAL;
  FILE *f = std::fopen(path, "w");
  if (!f) return LSKUM_EIO;
  const int32_t n = c->n_points;
  bool ok = std::fprintf(f, "%d %d %.17g\n", n, it, res_old) > 0;
  for (int32_t i = 0; i < n && ok; i++)
    ok = std::fprintf(f, "%.17g %.17g %d %d %.17g %.17g %.17g %.17g %.17g\n", c->x[i], c->y[i],
                      c->flag_1[i], c->flag_2[i], c->min_dist[i], prim[i], prim[n + i],
                      prim[2 * (size_t)n + i], prim[3 * (size_t)n + i]) > 0;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? LSKUM_OK : LSKUM_EIO;
}

int lskum_restart_read(const char *path, int32_t n, int32_t *it, double *res_old,
                       double *prim) {
  if (!path || n <= 0 || !it || !res_old || !prim) return LSKUM_EINVAL;
  FILE *f = std::fopen(path, "r");
  if (!f) return LSKUM_EIO;
  double nop;
  int rc = LSKUM_OK;
  if (std::fscanf(f, "%lf %d %lf", &nop, it, res_old) != 3)
    rc = LSKUM_EIO;
  else if (nop != (double)n || *it < 0)
    rc = LSKUM_EINVAL;  // a different cloud, or no iteration count
  for (int32_t i = 0; i < n && rc == LSKUM_OK; i++) {
    double skip[5];
    if (std::fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf", &skip[0], &skip[1], &skip[2],
                    &skip[3], &skip[4], &prim[i], &prim[n + i], &prim[2 * (size_t)n + i],
                    &prim[3 * (size_t)n + i]) != 9)
      rc = LSKUM_EIO;
  }
  std::fclose(f);
  return rc;
}

}  // extern "C"
