// Device-side problem description and launch wrappers of the fit kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#define PS_FIT_WAVES 4     // wavefronts per workgroup (one pair per workgroup)
#define PS_OUT_STRIDE 10   // doubles per pair result

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
  int ntiles;
  // facets and basis
  int nfac, lmax, mmax, ncoef, ma, mfit;
  int qs;   // columns of a Jacobian row (16 * nb)
  int nb;   // 16-column blocks of a row
  int nbs;  // blocks holding shape columns
  const double *nx, *ny, *nz, *darea;
  const double *pleg, *fc, *fs;  // curv tables
  const double* yl;              // [F][16][nbs]: shape column q = 16b + r (parameter q+2) at [f][r][b]
  int ylds;                      // 1: the fit kernel copies yl into LDS
  const double* cg_first;        // [ncoef]
  const int* cmap;               // [mfit] compact index -> parameter index (0-based)
  // model and LM controls
  double par[3], lcl, phi0;
  double alamda_start, alamda_incr, alamda_incrr;
  int n_iter_max, n_iter_min;
  double iter_diff_max;
  double f_start, f_step;
  double beta_pole[10], lambda_pole[10];
};

struct PsFitArgs {
  PsDevProblem P;
  double conw_r2_rows;  // sigr2 of the three convexity rows (conwR^2)
  double conw_r_dev;    // conw_r used by Iter2's dev
  const int* item_n;    // [n_items] frequency index, 1-based
  const int* item_m;    // [n_items] pole index, 1-based
  int n_items;
  int* counter;         // work-queue head, zeroed before launch
  double* scratch;      // per-workgroup scratch slabs (ps_slab layout)
  long long scratch_stride;
  double* out;          // [n_items][PS_OUT_STRIDE]
};

// Point chains of the normal-equation product: the waves of a workgroup
// form ps_tile_chains(nb) groups; group g accumulates the tiles T = g (mod
// groups) and the groups' partial matrices are added at the end.
__host__ __device__ inline int ps_tile_chains(int nb) { return nb <= 4 ? 4 : 2; }

// One workgroup's scratch slab in HBM (offsets in doubles, 256-byte aligned).
struct PsSlab {
  size_t jbuf;   // [ndata+3][qs] raw Jacobian rows (data points, then the 3 convexity rows)
  size_t ybuf;   // [ndata+3] model brightness
  size_t tsum;   // [ntiles][qs+1] per-tile column sums (column qs: brightness)
  size_t dsum;   // [ncurves][qs+1] per-curve column sums (column qs: 1 / brightness sum)
  size_t part;   // [chains][16nb][16nb] partial augmented normal matrices
  size_t best;   // [K][K] accepted normal matrix
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
  s.tsum = o;
  o += al((size_t)P.ntiles * (P.qs + 1));
  s.dsum = o;
  o += al((size_t)P.ncurves * (P.qs + 1));
  s.part = o;
  o += al((size_t)ps_tile_chains(P.nb) * 256 * P.nb * P.nb);
  s.best = o;
  o += al(K * K);
  s.stride = o;
  return s;
}

size_t ps_fit_lds_bytes(const PsDevProblem& P, bool ylds);
hipError_t ps_set_fit_attributes(int nb, int nbs, bool ylds, size_t lds);
int ps_fit_blocks_per_cu(int nb, int nbs, bool ylds, size_t lds);  // resident workgroups per CU
hipError_t ps_launch_fit(const PsFitArgs& A, int grid, hipStream_t s);
hipError_t ps_launch_eval(const PsFitArgs& A, hipStream_t s, const double* a, double* work, double* area,
                          double* jrows);
hipError_t ps_launch_gaussj(int mf, const double* alpha, const double* beta, double lambda, double* x, int* err,
                            hipStream_t s);
hipError_t ps_phase_cycles(unsigned long long* out, bool reset);
hipError_t ps_launch_math(int n, const double* x, double* out, double* out2, hipStream_t s);
