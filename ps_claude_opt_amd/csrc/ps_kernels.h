// Device-side problem description and launch wrappers of the fit kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#define PS_OUT_STRIDE 10   // doubles per pair result
#define PS_STATIC_SCRATCH_BYTES ((size_t)256 << 20)  // device memory of ps_fit_items_static
#define PS_FACET_PAD 16    // zero facets after the last one in the basis table and the LDS normals

// Device pointers and scalars of one prepared problem (passed by value).
struct PsDevProblem {
  // observations, data points only
  const double *tim, *ex, *ey, *ez, *sx, *sy, *sz;  // viewer (Earth) and Sun unit vectors
  const double *alph;                                // solar phase angle per point
  const double *bright, *sig, *s2w;                  // brightness, curve mean, sigr2*weight
  const int* curve_off;                              // [ncurves+1]
  int ncurves, ndata;
  // 16-point tiles of the curves (tiles never straddle curves)
  const int* tile_off;  // [ncurves+1] first tile of each curve
  const int* tile_cv;   // [ntiles] curve of each tile
  const int* tile_info; // [ntiles][4] first point, points in the tile (1..16), curve, curve length
                        // (negative for an absolute curve, in_rel = 0)
  int ntiles;
  // facets and basis
  int nfac, lmax, mmax, ncoef, ma, mfit;
  int qs;   // columns of a Jacobian row (16 * nb)
  int nb;   // 16-column blocks of a row
  int nbs;  // blocks holding shape columns
  const double *nx, *ny, *nz, *darea;
  const double *pleg, *fc, *fs;  // curv tables
  const double* yl;              // [F + PS_FACET_PAD][16][nbs]: shape column q = 16b + r (parameter q+2) at
                                 // [f][r][b]; the padding facets are zero
  const double* cg_first;        // [ncoef]
  const int* cmap;               // [mfit] compact index -> parameter index (0-based)
  int ia1;                       // parameter 1 (the size scale) is free: compact column 0 is parameter 1,
                                 // compact column q >= 1 parameter q + 1 (else compact q is parameter q + 2)
  // model and LM controls
  double par[3], lcl, phi0;
  double alamda_start, alamda_incr, alamda_incrr;
  int n_iter_max, n_iter_min;
  double iter_diff_max;
  double f_start, f_step;
  double beta_pole[10], lambda_pole[10];
};

// One persistent launch fits a list of work items; item i is pair
// (item_n[i], item_m[i]) of problem item_p[i].  Every problem of a launch
// shares the facet geometry and shape basis (nfac, lmax, mmax, the facet
// tables, the free-parameter map); P0 carries those shared tables (it is
// problem 0) and sizes the LDS.
struct PsFitArgs {
  PsDevProblem P0;
  const PsDevProblem* probs;  // [n_problems] on the device
  const int* item_p;    // [n_items] problem index
  const int* item_n;    // [n_items] frequency index, 1-based
  const int* item_m;    // [n_items] pole index, 1-based
  const double* item_cw;  // [n_items][2] sigr2 of the convexity rows (conwR^2), conwR of Iter2's dev
  int n_items;
  int* counter;         // work-queue head, zeroed before launch
  double* scratch;      // per-workgroup scratch slabs (ps_slab layout, sized for the largest problem)
  long long scratch_stride;
  double* out;          // [n_items][PS_OUT_STRIDE]
  const double* item_a0;  // [n_items][ma] initial shape and phase-function parameters, or null (the
                          // reference's initialisation); pole, rotation rate and cls always come from (n, m)
  double* out_a;        // [n_items][ma] final parameters, or null
  int n_iter_fixed;     // > 0: exactly this many LM iterations (the reference's stopCondition > 1 mode)
                        // instead of each problem's own stop condition
  int claim_tiles;      // 16-point facet-pass tiles a wave claims at a time (1, 4, 8, 16)
  double conw_r2_rows;  // ps_eval only: sigr2 of the convexity rows
  int batch;            // 1: items of several problems (read from probs), 0: every item is on P0
};

// 16-column blocks of a row: the raw Jacobian row holds parameters 2..C+8
// (C + 7 columns); the normal-equation rows hold mfit Jacobian columns, the
// residual and two columns that carry the per-curve sums.
__host__ __device__ inline int ps_row_blocks(int ncoef, int mfit) {
  const int a = (ncoef + 7 + 15) / 16, b = (mfit + 3 + 15) / 16;
  return a > b ? a : b;
}

// Point chains of the normal-equation product: the waves of a workgroup
// form ps_tile_chains(nb) groups; group g accumulates the tiles T = g (mod
// groups) and the groups' partial matrices are added at the end.
__host__ __device__ inline int ps_tile_chains(int nb) { return nb <= 4 ? 4 : 2; }

// Doubles per point in the per-evaluation point record (one 128-byte line).
#define PS_REC 16

// Curves of the renormalisation correction, rounded up to even (two
// correction terms per curve; MFMA steps take four).
__host__ __device__ inline int ps_corr_curves(int ncurves) { return (ncurves + 1) & ~1; }
// Doubles per (chain, curve) of the per-curve sums: rows z and y of the
// chain's matrix, [qs] each.
__host__ __device__ inline int ps_curve_sums_width(int qs) { return 2 * qs; }

// One workgroup's scratch slab in HBM (offsets in doubles, 256-byte aligned).
struct PsSlab {
  size_t jbuf;   // [ndata+3][qs] raw Jacobian rows (data points, then the 3 convexity rows)
  size_t ybuf;   // [ndata+3] model brightness
  size_t tsy;    // [ntiles] per-tile brightness sums
  size_t crv;    // [ncurves] per relative curve: 1 / brightness sum
  size_t csum;   // [chains][ncurves][ps_curve_sums_width] per-curve sums of each point chain
  size_t ttab;   // [2 ps_corr_curves][qs] correction vectors a_k, b_k of each curve
  size_t part;   // [chains][16nb][16nb] partial augmented normal matrices
  size_t best;   // [K][K] accepted normal matrix
  size_t prec;   // [ndata][PS_REC] per-point records of the evaluated parameters (rotation angle,
                 // pole-frame Earth and Sun directions, phase-function terms), in facet-pass order
  size_t perm;   // [ndata] ints: the points in the order of the facet-pass tiles
  size_t pkey;   // [ndata] bytes: each point's view-direction cell
  size_t stride;
};

__host__ __device__ inline PsSlab ps_slab(const PsDevProblem& P) {
  auto al = [](size_t n) { return (n + 31) & ~(size_t)31; };
  PsSlab s;
  size_t o = 0;
  const size_t rows = (size_t)P.ndata + 3;
  const size_t K = (size_t)P.mfit + 1;
  s.jbuf = o;
  o += al(rows * P.qs);
  s.ybuf = o;
  o += al(rows);
  s.tsy = o;
  o += al((size_t)P.ntiles);
  s.crv = o;
  o += al((size_t)P.ncurves);
  s.csum = o;
  o += al((size_t)ps_tile_chains(P.nb) * P.ncurves * ps_curve_sums_width(P.qs));
  s.ttab = o;
  o += al((size_t)2 * ps_corr_curves(P.ncurves) * P.qs);
  s.part = o;
  o += al((size_t)ps_tile_chains(P.nb) * 256 * P.nb * P.nb);
  s.best = o;
  o += al(K * K);
  s.prec = o;
  o += al((size_t)PS_REC * P.ndata);
  s.perm = o;
  o += al(((size_t)P.ndata + 1) / 2);
  s.pkey = o;
  o += al(((size_t)P.ndata + 7) / 8);
  s.stride = o;
  return s;
}

// Launch API of one compiled variant of the fit kernels (namespace = the
// number of wavefronts per workgroup: psw1, psw2, psw4, psw8).
#define PS_DECLARE_FIT_VARIANT(NS)                                                                  \
  namespace NS {                                                                                    \
  size_t ps_fit_lds_bytes(const PsDevProblem& P);                                                   \
  hipError_t ps_set_fit_attributes(int nb, int nbs, size_t lds);                                    \
  int ps_fit_blocks_per_cu(int nb, int nbs, size_t lds); /* resident workgroups per CU */           \
  hipError_t ps_launch_fit(const PsFitArgs& A, int grid, hipStream_t s);                            \
  hipError_t ps_launch_eval(const PsFitArgs& A, hipStream_t s, const double* a, double* work,       \
                            double* area, double* jrows);                                           \
  hipError_t ps_launch_gaussj(int mf, const double* alpha, const double* beta, double lambda,       \
                              double* x, int* err, hipStream_t s);                                  \
  hipError_t ps_phase_cycles(unsigned long long* out, bool reset);                                  \
  hipError_t ps_launch_math(int n, const double* x, double* out, double* out2, hipStream_t s);      \
  }

PS_DECLARE_FIT_VARIANT(psw1)
PS_DECLARE_FIT_VARIANT(psw2)
PS_DECLARE_FIT_VARIANT(psw4)
PS_DECLARE_FIT_VARIANT(psw8)
