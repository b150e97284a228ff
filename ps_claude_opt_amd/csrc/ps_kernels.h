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
  double* scratch;      // per-workgroup scratch
  long long scratch_stride;
  int jrows;            // Jacobian rows held per workgroup
  double* out;          // [n_items][PS_OUT_STRIDE]
};

size_t ps_fit_lds_bytes(const PsDevProblem& P, bool ylds);
size_t ps_fit_scratch_stride(const PsDevProblem& P, int jrows);
hipError_t ps_set_fit_attributes(int nb, int nbs, bool ylds, size_t lds);
hipError_t ps_launch_fit(const PsFitArgs& A, int grid, hipStream_t s);
hipError_t ps_launch_eval(const PsFitArgs& A, hipStream_t s, const double* a, double* work, double* area,
                          double* jrows);
hipError_t ps_launch_gaussj(int mf, const double* alpha, const double* beta, double lambda, double* x, int* err,
                            hipStream_t s);
hipError_t ps_phase_cycles(unsigned long long* out, bool reset);
hipError_t ps_launch_math(int n, const double* x, double* out, double* out2, hipStream_t s);
