// kernels.hpp -- launchers of the q-LSKUM HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lskum {

// Per-point records, each one 64-byte segment so that every kernel writes
// whole segments and a neighbour gather touches one segment per record:
//  base  [x, y, q0..q3, eps2, speed]  (q rewritten every stage)
//  dq    [dq/dx 0..3, dq/dy 0..3]     (two buffers: the Jacobi inner
//                                      passes ping-pong between them)
//  qm    [qmax 0..3, qmin 0..3]       (only the flux gathers it)
enum BaseSlot : int {
  B_X = 0,
  B_Y = 1,
  B_Q = 2,      // q[4]
  B_EPS2 = 6,   // (pow(K*min_dist, 3))^2, limiters_mod.h:82-83
  B_SPEED = 7,  // |u| + 3 sqrt(p/rho) of the state func_delta reads
  B_STRIDE = 8
};
constexpr int DQ_STRIDE = 8;
constexpr int QM_STRIDE = 8;
// the flux reverse's record of one (point, neighbour) pair: the adjoints of
// the neighbour's q[4], of its increment d[4] and of its extremum [4]
constexpr int EDGE_REC = 12;
// conn entries: neighbour index in the low 28 bits, sub-stencil mask above.
constexpr uint32_t kIndexMask = 0x0FFFFFFFu;
constexpr int kMaskShift = 28;
// tiled conn entries: local index in the block's tile (12 bits), mask above.
constexpr uint16_t kLocalMask = 0x0FFFu;
constexpr int kLocalShift = 12;
constexpr int kTileBlockDefault = 256;  // points per block = per LDS tile

struct DevCloud {
  int n;
  const int32_t *conn_off;
  const uint32_t *conn;      // global form
  const uint16_t *conn16;    // tiled form (same order)
  // the same lists in slices, which the forward stencil kernels read: the
  // lists of a stencil block's points entry by entry, entry e of lane t of
  // block b at slice_off[b] + e * tile_block + t (lists padded with their
  // last entry to the block's longest)
  const int32_t *slice_off;  // [stencil blocks + 1]
  const uint32_t *connS;     // global form
  const uint16_t *conn16S;   // tiled form
  const int32_t *tile_off;   // [nblocks + 1]
  const int32_t *tile_idx;   // own points first, then other neighbours
  int tile_max;              // largest tile, entries
  unsigned tiles;            // LSKUM_TILE_* bits of the kernels using them
  int tile_block;            // points per block of the stencil kernels
  int su_ppt;                // points per thread of the state update: 1, 2, 4
  int flux_cap;              // entries of a tile the flux stages (<= tile_max)
  int slice_width;           // longest list of any point (entries of a slice row)
  int qi_chunks;             // pipelined inner pass: its layout code (0: not used)
  int qi_grid;               // its workgroups
  int qd_chunks;             // pipelined q_derivatives: its layout code (0: not used)
  int qd_grid;
  int qdd_chunks;            // the same with func_delta (0: not used)
  int qdd_grid;
  const double *xy;   // [n][2] coordinates (also in the base records)
  const double *nx, *ny;
  const double *coef;  // per point: least-squares coefficients of the
                       // sub-stencils x+, x-, y+, y- (a, b each)
  const uint8_t *flag;
  const int32_t *perm;  // device index -> caller index
};

struct FlowConsts {
  double cfl, power;
  double u1_inf, u2_inf;  // q_inf[2], q_inf[3]
  double rho_inf, pr_inf;
};

struct Partials {
  double *sum, *max;
  int32_t *arg;
  int nblocks;
};

int blocks_for(int n, int block);
// LDS bytes per tiled launch of each kernel for a given tile size
constexpr unsigned TILE_QDERIV = 1, TILE_QINNER = 2, TILE_DELTA = 4, TILE_FLUX = 8;
size_t lds_bytes_qderiv(int tile_max);
size_t lds_bytes_qinner(int tile_max);
size_t lds_bytes_delta(int tile_max);
size_t lds_bytes_flux(int tile_max);
// The number of tile entries the FP64 flux (mixed: flux_tile_cap_m) stages
// per block so that a CU (lds_per_cu bytes of LDS) still holds as many of
// its blocks as its registers allow; 0 when not even the block's own points
// fit.  A block whose tile is larger reads the entries beyond the cap
// through L2.
int flux_tile_cap(int block, int tile_max, size_t lds_limit, size_t lds_per_cu);
// raises the dynamic-LDS limits of the kernels in `tiles`; returns the
// subset whose tiles fit in lds_limit bytes
unsigned prepare_tiled(unsigned tiles, int tile_max, int flux_cap, size_t lds_limit);
// The pipelined inner pass (tiled, power 0) for a layout whose tiles hold at
// most tile_max entries and whose lists at most slice_width: the layout
// code of its per-thread gather slots (0 when no built layout covers the
// tile and lists, or they do not fit lds_limit), and, through *grid, the
// workgroups that fill n_cu CUs (a multiple of 8, at most nblk).
int prepare_qinner_pf(int tile_block, int tile_max, int slice_width, int nblk, int n_cu,
                      size_t lds_limit, int *grid);
// The same for the pipelined q_derivatives (tiled, power 0), with_delta:
// the variant that also computes func_delta.
int prepare_qderiv_pf(int tile_block, int tile_max, int slice_width, int nblk, int n_cu,
                      size_t lds_limit, bool with_delta, int *grid);

void launch_q_variables(const DevCloud &c, const double *prim, double *base, hipStream_t s);
void launch_speed(const DevCloud &c, const double *prim, double *base, hipStream_t s);
void launch_func_delta(const DevCloud &c, const double *base, double cfl, double *delta,
                       hipStream_t s);
// q_derivatives also writes qs [n][8]: sum w dx (q_k - q_i) and sum w dy
// (q_k - q_i), which the inner passes reuse (qs may be null when none run)
void launch_q_derivatives(const DevCloud &c, const double *base, double *dq, double *qm,
                          double *qs, double power, hipStream_t s);
// launch_q_derivatives (power 0) and launch_func_delta from the same base
// records in one pipelined pass; only where c.qdd_chunks is set
void launch_q_derivatives_delta(const DevCloud &c, const double *base, double *dq, double *qm,
                                double *qs, double cfl, double *delta, hipStream_t s);
void launch_q_inner(const DevCloud &c, const double *qs, const double *dq_src, double *dq_dst,
                    double power, hipStream_t s);
// q_derivatives (power 0) that also writes arg [n][8], per qmax/qmin
// component the device index of the point it came from, and lsum [n][4],
// the geometry sums (sxx, syy, sxy) of the least-squares solve
void launch_q_derivatives_arg(const DevCloud &c, const double *base, double *dq, double *qm,
                              double *qs, int32_t *arg, double *lsum, hipStream_t s);
void launch_flux(const DevCloud &c, const double *base, const double *dq, const double *qm,
                 const double *delta, double power, double *flux_res, hipStream_t s);
// also writes q of the new state, as FP32 into q32 when non-null (mixed
// form), else into base when non-null, and, with speed, the func_delta
// denominator into base
// fcs[case] holds each case's flow constants; block_case[b] the case of
// 256-point block b (NULL: one case)
void launch_state_update(const DevCloud &c, int rk, double *prim, const double *prim_old,
                         const double *flux_res, const FlowConsts *fcs,
                         const uint16_t *block_case, const Partials &p, double *base,
                         float *q32, bool speed, hipStream_t s);
// per case c (partials seg[c]..seg[c+1]-1, or all of them when seg is NULL):
// out[3c] = sum, out[3c+1] = max, out[3c+2] = argmax (caller index, as a
// double), and, when hist is non-null, hist[c*hist_stride + slot] = sum with
// slot read from slot_ptr when that is non-null
void launch_reduce(const Partials &p, int n_cases, const int32_t *seg, double *out,
                   double *hist, int hist_stride, int slot, const int32_t *slot_ptr,
                   hipStream_t s);
void launch_tick(int32_t *counter, hipStream_t s);      // counter[0] += 1
void launch_set_int(int32_t *p, int32_t v, hipStream_t s);
// count doubles (even) from src to dst by a kernel (graph-safe)
void launch_copy(const double *src, double *dst, size_t count, hipStream_t s);
// per case c: H, V, pitch moment over shape entries soff[c]..soff[c+1]-1
void launch_forces(int n_cases, const int32_t *soff, const int32_t *m, const int32_t *l,
                   const int32_t *r, const int32_t *shape, const double *base, const double *nx,
                   const double *ny, const double *prim, int n, double pr_inf,
                   const double *qdyn, double *out3, hipStream_t s);
void launch_kat_fluxes(int n, const double *in6, double *out, hipStream_t s);
void launch_kat_qtilde(int n, const double *qt4, double *out, hipStream_t s);

// ---------------------------------------------- mixed-precision form
// Per-point arrays (device order) the mixed kernels gather: coordinates in
// FP64, everything else in FP32.  dq buffers are [n][8] floats (dq/dx 0..3,
// dq/dy 0..3), qm32 is [n][8] (qmax 0..3, qmin 0..3).
struct MixedState {
  const double *xy;  // [n][2]
  float *q32;        // [n][4]
  float *qm32;       // [n][8]
  const float *eps32;  // [n], (pow(K*min_dist,3))^2 rounded to FP32
  const float *ls32;   // [n][4], the full stencil's (syy, -sxy, sxx)/det rounded
};
size_t lds_bytes_qderiv_m(int tile_max);
size_t lds_bytes_qinner_m(int tile_max);
size_t lds_bytes_flux_m(int tile_max);
int flux_tile_cap_m(int block, int tile_max, size_t lds_limit, size_t lds_per_cu);
unsigned prepare_tiled_m(unsigned tiles, int tile_max, int flux_cap, size_t lds_limit);
void launch_q_variables_m(const DevCloud &c, const double *prim, float *q32, hipStream_t s);
// func_delta with each neighbour's step in FP32 (records and the minimum's
// storage as launch_func_delta's)
void launch_func_delta_m(const DevCloud &c, const double *base, double cfl, double *delta,
                         hipStream_t s);
// with delta set (power 0 only): func_delta's minimum in the same pass,
// the same bits as launch_func_delta_m
void launch_q_derivatives_m(const DevCloud &c, const MixedState &m, float *dq, double power,
                            hipStream_t s, const double *base = nullptr, double cfl = 0.0,
                            double *delta = nullptr);
void launch_q_inner_m(const DevCloud &c, const MixedState &m, const float *src, float *dst,
                      double power, hipStream_t s);
void launch_flux_m(const DevCloud &c, const MixedState &m, const float *dq, const double *delta,
                   double power, double *flux_res, hipStream_t s);
void launch_kat_fluxes_m(int n, const double *in6, double *out, hipStream_t s);

// ---------------------------------------------- adjoint and tangent (FP64)
// Working arrays of the reverse and tangent passes (device order): own_*
// and edge hold the flux pairs' shares (edge: [n_edges][16]); lq [n][4],
// ldq [n][8] (two, for the inner passes), lqm [n][8], ls [n][8] are
// adjoints of q, dq, qmax/qmin and of the least-squares sums; arg [n][8]
// the point each qmax/qmin came from; tr_* the transpose of conn (for each
// point, the edges that name it, owners ascending); tr_pos[e] the slot of
// edge e's record in `edge`: point j's r-th transposed edge is kept at
// tr_sbase[j / 256] + 256 r + j % 256 (the 256 points of a block in
// adjacent slots, so that the gather's lanes read adjacent records), with
// tr_owner_s the owner of each slot; tbase, tdq, tqm the
// tangents (tbase a copy of the base records with tangent q); vec_in /
// vec_out [4][n] staging.
struct AdjointBuffers {
  // edge: the flux reverse's per-pair records, EDGE_REC doubles each;
  // edge_mx: per pair, bit r set where component r's extremum is qmax
  double *own_q, *own_dq, *own_qm, *edge, *lq, *ldq[2], *lqm, *ls;
  uint8_t *edge_mx;
  int32_t *arg, *tr_off, *tr_edge, *tr_owner, *tr_pos, *tr_sbase, *tr_owner_s;
  double *tbase, *tdq[2], *tqm, *vec_in, *vec_out;
  double *tqs;     // tangent of q_derivatives' q sums
  double *lsum;    // per point sxx, syy, sxy of the least-squares solve
  bool vjp_tiled;  // the flux reverse stages neighbour records in LDS
  bool ls_tiled;   // so does the least-squares reverse with geometry terms
};
// LDS bytes of the flux reverse's tile, and whether it can use the cloud's
// tiles (they exist and fit in lds_limit; sets the kernels' LDS attribute)
size_t lds_bytes_flux_vjp(int tile_max);
bool prepare_flux_vjp_tiles(const DevCloud &c, size_t lds_limit);
bool prepare_ls_adj_tiles(const DevCloud &c, size_t lds_limit);
// Geometry adjoints the reverse kernels accumulate (+=), device order:
// edge [n_edges][2] d/d(delx, dely) of each (point, neighbour) pair, nrm
// [n][2] d/d(nx, ny), coef [n][8] d/d of the sub-stencil coefficients,
// delta [n] d/d of the local time step.
struct GeoAdjoint {
  double *edge, *nrm, *coef, *delta;
};
// flux reverse: lam [4][n] (adjoint of flux_res) -> lq, ldq[0], lqm, and
// with geo non-null the geometry adjoints (+=)
void launch_flux_vjp(const DevCloud &c, const AdjointBuffers &a, const double *base,
                     const double *dq, const double *qm, const double *delta, const double *lam,
                     const GeoAdjoint *geo, hipStream_t s);
// least-squares pass reverse from the adjoint lo of its output: adds to lq;
// an inner pass (extrema false) writes the adjoint of its input
// derivatives to ldq_in; q_derivatives (extrema true) routes lqm.  With
// geo_edge non-null also the pairs' geometry adjoints (+=), from the pass's
// output derivatives `out` and, for an inner pass, its input dq_in.
void launch_ls_vjp(const DevCloud &c, const AdjointBuffers &a, const double *base,
                   const double *lo, double *ldq_in, bool extrema, const double *out,
                   const double *dq_in, double *geo_edge, hipStream_t s);
// grad = J^T lq, or grad += J^T lq
void launch_qvar_vjp(const DevCloud &c, const AdjointBuffers &a, const double *prim,
                     double *grad, bool accumulate, hipStream_t s);
void launch_qvar_jvp(const DevCloud &c, const double *prim, const double *v, double *tbase,
                     hipStream_t s);
void launch_ext_jvp(const DevCloud &c, const int32_t *arg, const double *tbase, double *tqm,
                    hipStream_t s);
void launch_flux_jvp(const DevCloud &c, const double *base, const double *dq, const double *qm,
                     const double *delta, const double *tbase, const double *tdq,
                     const double *tqm, double *out, hipStream_t s);

// ---------------------------------------------- whole iterations (trajectory.hip)
// Stage update reverse at every point (FP64, one case): from pbar, the
// adjoint of the state the stage wrote, to pbar = the adjoint of the
// stage's input state (overwritten), pold_bar += that of the iteration's
// start state, rbar = that of the flux residual, nrm_bar += that of the
// normals.
void launch_su_vjp(const DevCloud &c, int rk, const double *prim, const double *prim_old,
                   const double *fres, const FlowConsts *fc, double *pbar, double *pold_bar,
                   double *rbar, double *nrm_bar, hipStream_t s);
// func_delta reverse from dbar [n]: the pair that set each point's delta
// gets its geometry adjoint (geo_edge +=) and the speed of its neighbour
// (base: the speed of the state prim) passes it on to pbar (+=).
void launch_delta_vjp(const DevCloud &c, const AdjointBuffers &a, const double *base,
                      const double *prim, double cfl, const double *dbar, int32_t *amin_e,
                      double *amin_v, double *geo_edge, double *pbar, hipStream_t s);
// y += x over count doubles
void launch_axpy(size_t count, const double *x, double *y, hipStream_t s);

// Forward-mode (value, tangent) run of whole iterations, every array device
// order in pairs {value, derivative along the seeded direction}.
struct Tan {
  double v, d;
};
struct TanState {
  Tan *x, *y, *nx, *ny, *coef;  // [n], [n][8]
  Tan *prim, *prim_old, *fres;  // [4][n]
  Tan *q, *speed, *delta;       // [n][4], [n], [n]
  Tan *dq[2], *qm;              // [n][8]
};
// normals of the wall and outer points (lr: [n][2] device indices of their
// left and right neighbours) and sub-stencil coefficients from x, y
void launch_tan_geometry(const DevCloud &c, const int32_t *lr, const TanState &t,
                         hipStream_t s);
// one fpi_solver iteration (eps2 read from base, power 0)
void launch_tan_iteration(const DevCloud &c, const double *base, const FlowConsts *fc,
                          double cfl, int inner, const TanState &t, hipStream_t s);

}  // namespace lskum
