// C ABI of the MI355X period scan (include/ps_mi355x.h).
//
// Host orchestration replacing the reference's start_CUDA.cu: table upload
// (CUDAPrepare + the symbol copies of CUDAPrecalc/CUDAStart), the
// convexity-weight round (CUDAPrecalc, start_CUDA.cu:400-789), the scan with
// its best-pole reduction and output lines (CUDAStart, start_CUDA.cu:
// 792-1245).  One persistent kernel launch fits a whole batch of pairs; the
// host never polls per iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ps_mi355x.h"
#include "ps_kernels.h"
#include "ps_setup.h"

namespace {
constexpr int kMaxIter = 60;  // constants.h:12
}

struct ps_input {
  ps::InputFile f;
};

struct ps_ctx {
  int device = 0;
  int n_cu = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  // problem
  bool prepared = false;
  bool have_file = false;
  ps::InputFile file;
  int ndata = 0, ncurves = 0, nfac = 0, ncoef = 0, ma = 0, mfit = 0, n_max = 0, n_iter_max = 0;
  int max_lp = 0;
  double f_start = 0, f_end = 0, f_step = 0, escl = 0, conw = 0;
  PsDevProblem dp{};
  std::vector<void*> dbufs;
  size_t lds = 0;
  long long stride = 0;
  int grid_cap = 0;
  int blocks_per_cu = 0;
  // batch buffers
  double* scratch = nullptr;
  size_t scratch_elems = 0;
  int* counter = nullptr;
  int* d_items_n = nullptr;
  int* d_items_m = nullptr;
  double* d_out = nullptr;
  size_t items_cap = 0;
  // statistics of the last fit call
  double last_ms = 0;
  int last_launches = 0;
  long long last_pairs = 0, last_vis = 0, last_iters = 0, last_evals = 0;
};

namespace {

int fail(ps_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hipfail(ps_ctx* c, hipError_t e, const char* what) {
  return fail(c, PS_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define PS_HIP(ctx, call)                              \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hipfail(ctx, e_, #call); \
  } while (0)

template <typename T>
int upload(ps_ctx* c, const T* host, size_t n, const T** dev) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return hipfail(c, e, "hipMalloc");
  c->dbufs.push_back(p);
  if (n) {
    e = hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hipfail(c, e, "hipMemcpy");
  }
  *dev = static_cast<const T*>(p);
  return PS_OK;
}

void free_problem(ps_ctx* c) {
  for (void* p : c->dbufs) hipFree(p);
  c->dbufs.clear();
  c->prepared = false;
}

int ensure_batch(ps_ctx* c, int npairs, int grid) {
  if ((size_t)npairs > c->items_cap) {
    if (c->d_items_n) hipFree(c->d_items_n);
    if (c->d_items_m) hipFree(c->d_items_m);
    if (c->d_out) hipFree(c->d_out);
    c->d_items_n = c->d_items_m = nullptr;
    c->d_out = nullptr;
    size_t cap = std::max<size_t>(npairs, 1024);
    PS_HIP(c, hipMalloc(&c->d_items_n, cap * sizeof(int)));
    PS_HIP(c, hipMalloc(&c->d_items_m, cap * sizeof(int)));
    PS_HIP(c, hipMalloc(&c->d_out, cap * PS_OUT_STRIDE * sizeof(double)));
    c->items_cap = cap;
  }
  const size_t need = (size_t)grid * c->stride;
  if (need > c->scratch_elems) {
    if (c->scratch) hipFree(c->scratch);
    c->scratch = nullptr;
    PS_HIP(c, hipMalloc(&c->scratch, need * sizeof(double)));
    c->scratch_elems = need;
  }
  if (!c->counter) PS_HIP(c, hipMalloc(&c->counter, sizeof(int)));
  return PS_OK;
}

PsFitArgs make_args(ps_ctx* c, double conw_r, double conw_r_dev, int n_items) {
  PsFitArgs A;
  A.P = c->dp;
  A.conw_r2_rows = conw_r * conw_r;
  A.conw_r_dev = conw_r_dev;
  A.item_n = c->d_items_n;
  A.item_m = c->d_items_m;
  A.n_items = n_items;
  A.counter = c->counter;
  A.scratch = c->scratch;
  A.scratch_stride = c->stride;
  A.out = c->d_out;
  return A;
}

// earliest reported pole with minimal dev, NaN-safe (start_CUDA.cu:1200-1210)
int best_pole(const ps_pair_result* r) {
  int best = -1;
  for (int p = 0; p < PS_N_POLES; ++p)
    if (r[p].reported == 1 && (best < 0 || (!std::isnan(r[p].dev) && !(r[p].dev >= r[best].dev)))) best = p;
  return best;
}

}  // namespace

extern "C" {

int ps_input_read(const char* path, ps_input** out, char* err, int errlen) {
  if (!path || !out) return PS_E_ARG;
  *out = nullptr;
  ps_input* in = new ps_input();
  std::string e;
  const int rc = ps::read_input(path, in->f, e);
  if (rc) {
    if (err && errlen > 0) std::snprintf(err, errlen, "%s", e.c_str());
    delete in;
    return rc;
  }
  *out = in;
  return PS_OK;
}

int ps_input_view(const ps_input* in, ps_problem* view, double* conw, int* n_max) {
  if (!in || !view) return PS_E_ARG;
  *view = in->f.view();
  view->stop_condition = in->f.stop_condition;
  if (conw) *conw = in->f.conw;
  if (n_max) *n_max = (int)((in->f.f_start - in->f.f_end) / in->f.f_step) + 1;
  return PS_OK;
}

void ps_input_free(ps_input* in) { delete in; }

int ps_select_best_pole(const ps_pair_result* r) { return r ? best_pole(r) : -1; }

int ps_open(int device, ps_ctx** out) {
  if (!out) return PS_E_ARG;
  *out = nullptr;
  ps_ctx* c = new ps_ctx();
  c->device = device;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) {
    delete c;
    return PS_E_HIP;
  }
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return PS_E_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete c;
    return PS_E_HIP;
  }
  c->n_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return PS_E_HIP;
  }
  *out = c;
  return PS_OK;
}

const char* ps_last_error(const ps_ctx* c) { return c ? c->err.c_str() : "null context"; }

int ps_close(ps_ctx* c) {
  if (!c) return PS_E_ARG;
  hipSetDevice(c->device);
  free_problem(c);
  if (c->scratch) hipFree(c->scratch);
  if (c->counter) hipFree(c->counter);
  if (c->d_items_n) hipFree(c->d_items_n);
  if (c->d_items_m) hipFree(c->d_items_m);
  if (c->d_out) hipFree(c->d_out);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return PS_OK;
}

int ps_prepare(ps_ctx* c, const ps_problem* p) {
  if (!c || !p) return PS_E_ARG;
  PS_HIP(c, hipSetDevice(c->device));
  free_problem(c);
  const int N = p->ndata, L = p->n_curves, F = p->nfac, C = p->ncoef;
  if (N <= 0 || L <= 0 || F <= 0 || F % 4 || C <= 1 || p->lmax < 0 || p->mmax < 0 || p->mmax > p->lmax)
    return fail(c, PS_E_ARG, "bad problem sizes");
  int count = 0;
  for (int i = 0; i < L; ++i) {
    if (p->l_points[i] <= 0) return fail(c, PS_E_ARG, "empty lightcurve");
    if (!p->in_rel[i]) return fail(c, PS_E_UNSUPPORTED, "calibrated (absolute) lightcurves are not supported");
    count += p->l_points[i];
  }
  if (count != N) return fail(c, PS_E_ARG, "l_points do not sum to ndata");
  const int ma = C + 8;
  // free parameters must be one contiguous run starting at parameter 2
  int mfit = 0, lastma = 0;
  for (int i = 0; i < ma; ++i)
    if (p->ia[i]) {
      ++mfit;
      lastma = i + 1;
    }
  if (p->ia[0]) return fail(c, PS_E_UNSUPPORTED, "the size-scale parameter must be fixed (relative data)");
  for (int i = 1; i < lastma; ++i)
    if (!p->ia[i]) return fail(c, PS_E_UNSUPPORTED, "free parameters must be contiguous from parameter 2");
  if (mfit < 1) return fail(c, PS_E_ARG, "no free parameters");
  const int K = mfit + 1;
  const int nb = std::max((C + 7 + 15) / 16, (K + 15) / 16);
  if (nb > 6 || (C - 1 + 15) / 16 < nb - 1) return fail(c, PS_E_UNSUPPORTED, "more than 96 Jacobian columns (lmax too large)");
  // host-derived tables
  const int nbs = (C - 1 + 15) / 16;
  std::vector<double> alph(N), s2w(N), yl((size_t)F * 16 * nbs, 0.0);
  for (int i = 0; i < N; ++i) {
    const double d = std::fma(p->ee[2 * N + i], p->ee0[2 * N + i],
                              std::fma(p->ee[N + i], p->ee0[N + i], p->ee[i] * p->ee0[i]));
    alph[i] = std::acos(std::fmin(1.0, std::fmax(-1.0, d)));
    s2w[i] = p->sigr2[i] * p->weight[i];
  }
  for (int f = 0; f < F; ++f)
    for (int q = 0; q <= C - 2; ++q) yl[((size_t)f * 16 + (q & 15)) * nbs + (q >> 4)] = p->d_sphere[(size_t)(q + 1) * F + f];
  std::vector<int> off(L + 1, 0), toff(L + 1, 0), tcv, cmap;
  int maxlp = 3;
  for (int i = 0; i < L; ++i) {
    off[i + 1] = off[i] + p->l_points[i];
    maxlp = std::max(maxlp, p->l_points[i]);
    const int nt = (p->l_points[i] + 15) / 16;
    toff[i + 1] = toff[i] + nt;
    for (int t = 0; t < nt; ++t) tcv.push_back(i);
  }
  for (int i = 0; i < ma; ++i)
    if (p->ia[i]) cmap.push_back(i);
  PsDevProblem& d = c->dp;
  d = PsDevProblem{};
  int rc;
#define UP(field, ptr, n) \
  if ((rc = upload(c, ptr, n, &d.field)) != PS_OK) return rc
  UP(tim, p->tim, N);
  UP(ex, p->ee, N);
  UP(ey, p->ee + N, N);
  UP(ez, p->ee + 2 * N, N);
  UP(sx, p->ee0, N);
  UP(sy, p->ee0 + N, N);
  UP(sz, p->ee0 + 2 * N, N);
  UP(alph, alph.data(), N);
  UP(bright, p->brightness, N);
  UP(sig, p->sig, N);
  UP(s2w, s2w.data(), N);
  UP(curve_off, off.data(), L + 1);
  UP(tile_off, toff.data(), L + 1);
  UP(tile_cv, tcv.data(), tcv.size());
  UP(nx, p->normal, F);
  UP(ny, p->normal + F, F);
  UP(nz, p->normal + 2 * F, F);
  UP(darea, p->d_area, F);
  UP(pleg, p->pleg, (size_t)(p->mmax + 1) * (p->lmax + 1) * F);
  UP(fc, p->fc, (size_t)(p->mmax + 1) * F);
  UP(fs, p->fs, (size_t)(p->mmax + 1) * F);
  UP(yl, yl.data(), yl.size());
  UP(cg_first, p->cg_first, C);
  UP(cmap, cmap.data(), cmap.size());
#undef UP
  d.ncurves = L;
  d.ndata = N;
  d.ntiles = (int)tcv.size();
  d.nfac = F;
  d.lmax = p->lmax;
  d.mmax = p->mmax;
  d.ncoef = C;
  d.ma = ma;
  d.mfit = mfit;
  d.nb = nb;
  d.qs = 16 * nb;
  d.nbs = nbs;
  for (int i = 0; i < 3; ++i) d.par[i] = p->par[i];
  d.lcl = std::log(p->cl);
  d.phi0 = p->phi0;
  d.alamda_start = p->alamda_start;
  d.alamda_incr = p->alamda_incr;
  d.alamda_incrr = p->alamda_incrr;
  // start_CUDA.cu:855-869
  int n_iter_min = p->n_iter_min;
  d.n_iter_max = 0;
  d.iter_diff_max = -1;
  if (p->stop_condition > 1) {
    d.n_iter_max = (int)p->stop_condition;
    d.iter_diff_max = 0;
    n_iter_min = 0;
  }
  if (p->stop_condition < 1) {
    d.n_iter_max = kMaxIter;
    d.iter_diff_max = p->stop_condition;
  }
  d.n_iter_min = n_iter_min;
  d.f_start = p->f_start;
  d.f_step = p->f_step;
  for (int i = 0; i < PS_N_POLES; ++i) {
    d.beta_pole[i] = p->beta_pole[i];
    d.lambda_pole[i] = p->lambda_pole[i];
  }
  c->ndata = N;
  c->ncurves = L;
  c->nfac = F;
  c->ncoef = C;
  c->ma = ma;
  c->mfit = mfit;
  c->f_start = p->f_start;
  c->f_end = p->f_end;
  c->f_step = p->f_step;
  c->n_max = (int)((p->f_start - p->f_end) / p->f_step) + 1;
  c->n_iter_max = d.n_iter_max;
  c->escl = p->escl;
  c->max_lp = maxlp;
  c->stride = (long long)ps_slab(d).stride;
  // The shape basis is read from L2 by default, which leaves room for two
  // workgroups per CU; PS_BASIS_LDS=1 stages it in LDS instead (one
  // workgroup per CU) when it fits.
  const size_t lds_cap = 160 * 1024;
  const char* bl = std::getenv("PS_BASIS_LDS");
  d.ylds = (bl && bl[0] == '1' && ps_fit_lds_bytes(d, true) <= lds_cap) ? 1 : 0;
  c->lds = ps_fit_lds_bytes(d, d.ylds != 0);
  if (c->lds > lds_cap) return fail(c, PS_E_UNSUPPORTED, "problem needs more LDS than a workgroup can have");
  PS_HIP(c, ps_set_fit_attributes(nb, d.nbs, d.ylds != 0, c->lds));
  const int per_cu = ps_fit_blocks_per_cu(nb, d.nbs, d.ylds != 0, c->lds);
  if (per_cu < 1) return fail(c, PS_E_HIP, "fit kernel cannot be resident (occupancy query failed)");
  c->grid_cap = c->n_cu * per_cu;
  c->blocks_per_cu = per_cu;
  c->prepared = true;
  return PS_OK;
}

int ps_prepare_file(ps_ctx* c, const char* path) {
  if (!c || !path) return PS_E_ARG;
  c->file = ps::InputFile();
  std::string e;
  int rc = ps::read_input(path, c->file, e);
  if (rc) return fail(c, rc, e);
  c->have_file = true;
  ps_problem p = c->file.view();
  p.stop_condition = c->file.stop_condition;
  rc = ps_prepare(c, &p);
  if (rc) return rc;
  c->conw = c->file.conw;
  return PS_OK;
}

int ps_info(const ps_ctx* c, int* ints, double* dbls) {
  if (!c || !c->prepared) return PS_E_STATE;
  if (ints) {
    const int v[8] = {c->ndata, c->ncurves, c->nfac, c->ncoef, c->ma, c->mfit, c->n_max, c->n_iter_max};
    std::memcpy(ints, v, sizeof v);
  }
  if (dbls) {
    const double v[4] = {c->f_start, c->f_step, c->escl, c->have_file ? c->conw : 0.0};
    std::memcpy(dbls, v, sizeof v);
  }
  return PS_OK;
}

int ps_fit_pairs(ps_ctx* c, int npairs, const int* n_list, const int* m_list, double conw_r, double conw_r_dev,
                 ps_pair_result* out) {
  if (!c || npairs < 0 || (npairs && (!n_list || !m_list || !out))) return PS_E_ARG;
  if (!c->prepared) return fail(c, PS_E_STATE, "ps_fit_pairs before ps_prepare");
  for (int i = 0; i < npairs; ++i)
    if (n_list[i] < 1 || m_list[i] < 1 || m_list[i] > PS_N_POLES) return fail(c, PS_E_ARG, "pair index out of range");
  c->last_ms = 0;
  c->last_launches = 0;
  c->last_pairs = c->last_vis = c->last_iters = c->last_evals = 0;
  if (npairs == 0) return PS_OK;
  PS_HIP(c, hipSetDevice(c->device));
  const int grid = std::min(npairs, c->grid_cap);
  int rc = ensure_batch(c, npairs, grid);
  if (rc) return rc;
  PS_HIP(c, hipMemcpyAsync(c->d_items_n, n_list, npairs * sizeof(int), hipMemcpyHostToDevice, c->stream));
  PS_HIP(c, hipMemcpyAsync(c->d_items_m, m_list, npairs * sizeof(int), hipMemcpyHostToDevice, c->stream));
  const int zero = 0;
  PS_HIP(c, hipMemcpyAsync(c->counter, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
  PsFitArgs A = make_args(c, conw_r, conw_r_dev, npairs);
  PS_HIP(c, hipEventRecord(c->ev0, c->stream));
  PS_HIP(c, ps_launch_fit(A, grid, c->stream));
  PS_HIP(c, hipEventRecord(c->ev1, c->stream));
  std::vector<double> res((size_t)npairs * PS_OUT_STRIDE);
  PS_HIP(c, hipMemcpyAsync(res.data(), c->d_out, res.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PS_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  PS_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  c->last_launches = 1;
  if (std::getenv("PS_PHASE_TIMING")) {
    unsigned long long cyc[16];
    if (ps_phase_cycles(cyc, true) == hipSuccess) {
      std::fprintf(stderr, "phase cycles (thread 0 of each workgroup, summed):");
      for (int i = 0; i < 9; ++i) std::fprintf(stderr, " %d:%.4g", i, (double)cyc[i]);
      std::fprintf(stderr, "\n");
    }
  }
  for (int i = 0; i < npairs; ++i) {
    const double* o = &res[(size_t)i * PS_OUT_STRIDE];
    ps_pair_result& r = out[i];
    r.dev = o[0];
    r.dark = o[1];
    r.period = o[2];
    r.lambda_deg = o[3];
    r.beta_deg = o[4];
    r.chisq = o[5];
    r.niter = (int)o[6];
    r.reported = (int)o[7];
    r.n_vis = (long long)o[8];
    c->last_vis += r.n_vis;
    c->last_iters += r.niter;
    c->last_evals += r.niter > 0 ? r.niter + 1 : 0;
  }
  c->last_pairs = npairs;
  return PS_OK;
}

int ps_last_stats(const ps_ctx* c, double* kernel_ms, int* launches, long long* pairs, long long* vis_pairs,
                  long long* iterations, long long* evaluations) {
  if (!c) return PS_E_ARG;
  if (kernel_ms) *kernel_ms = c->last_ms;
  if (launches) *launches = c->last_launches;
  if (pairs) *pairs = c->last_pairs;
  if (vis_pairs) *vis_pairs = c->last_vis;
  if (iterations) *iterations = c->last_iters;
  if (evaluations) *evaluations = c->last_evals;
  return PS_OK;
}

int ps_precalc(ps_ctx* c, double* conw_r, int* new_conw) {
  if (!c || !conw_r || !new_conw) return PS_E_ARG;
  const int tests = 10;  // max_test_periods, start_CUDA.cu:433
  std::vector<int> nl, ml;
  for (int n = 1; n <= tests; ++n)
    for (int m = 1; m <= PS_N_POLES; ++m) {
      nl.push_back(n);
      ml.push_back(m);
    }
  std::vector<ps_pair_result> r(nl.size());
  // the device conw_r of the reference's precalc is the bit pattern of a
  // host pointer (start_CUDA.cu:476): a subnormal whose square is 0
  int rc = ps_fit_pairs(c, (int)nl.size(), nl.data(), ml.data(), *conw_r, 0.0, r.data());
  if (rc) return rc;
  double sum_dark = 0.0;
  for (int n = 0; n < tests; ++n) {
    const int b = best_pole(&r[n * PS_N_POLES]);
    if (b >= 0) sum_dark = sum_dark + r[n * PS_N_POLES + b].dark;
  }
  const double ave = sum_dark / tests;
  if (ave < 1.0) *new_conw = 1;
  if (ave >= 1.0) *conw_r = *conw_r * 2;
  return PS_OK;
}

int ps_conw_search(ps_ctx* c, double* conw_r) {
  if (!c || !conw_r) return PS_E_ARG;
  if (!c->have_file) return fail(c, PS_E_STATE, "ps_conw_search needs ps_prepare_file");
  double cw = c->conw / c->escl / c->escl;
  int new_conw = 0;
  while (new_conw != 1 && (cw * c->escl * c->escl) < 10.0) {
    int rc = ps_precalc(c, &cw, &new_conw);
    if (rc) return rc;
  }
  *conw_r = cw;
  return PS_OK;
}

int ps_start(ps_ctx* c, int n_start, int n_count, double conw_r, ps_line* out) {
  if (!c || n_start < 1 || n_count < 0 || (n_count && !out)) return PS_E_ARG;
  if (!c->prepared) return fail(c, PS_E_STATE, "ps_start before ps_prepare");
  std::vector<int> nl, ml;
  for (int n = n_start; n < n_start + n_count; ++n)
    for (int m = 1; m <= PS_N_POLES; ++m) {
      nl.push_back(n);
      ml.push_back(m);
    }
  std::vector<ps_pair_result> r(nl.size());
  int rc = ps_fit_pairs(c, (int)nl.size(), nl.data(), ml.data(), conw_r, conw_r, r.data());
  if (rc) return rc;
  for (int i = 0; i < n_count; ++i) {
    ps_line& L = out[i];
    const int b = best_pole(&r[(size_t)i * PS_N_POLES]);
    if (b < 0) {
      L = ps_line{NAN, NAN, NAN, NAN, NAN, NAN};
      continue;
    }
    const ps_pair_result& x = r[(size_t)i * PS_N_POLES + b];
    L.period_h = 24 * x.period;
    L.rms = x.dev;
    L.chi2 = x.dev * x.dev * c->ndata;
    L.dark_or_conw = (n_start + i == 1) ? conw_r * c->escl * c->escl : x.dark;
    L.lambda_deg = x.lambda_deg;
    L.beta_deg = x.beta_deg;
  }
  return PS_OK;
}

int ps_unprepare(ps_ctx* c) {
  if (!c) return PS_E_ARG;
  hipSetDevice(c->device);
  free_problem(c);
  return PS_OK;
}

int ps_eval(ps_ctx* c, const double* a, double conw_r2_rows, double* alpha, double* beta, double* chisq,
            double* area, double* jrows) {
  if (!c || !a) return PS_E_ARG;
  if (!c->prepared) return fail(c, PS_E_STATE, "ps_eval before ps_prepare");
  PS_HIP(c, hipSetDevice(c->device));
  int rc = ensure_batch(c, 1, 1);
  if (rc) return rc;
  const int K = c->mfit + 1, ma = c->ma, mf = c->mfit;
  double *da = nullptr, *dw = nullptr, *dar = nullptr, *dj = nullptr;
  PS_HIP(c, hipMalloc(&da, ma * sizeof(double)));
  PS_HIP(c, hipMalloc(&dw, (size_t)K * K * sizeof(double)));
  PS_HIP(c, hipMalloc(&dar, c->nfac * sizeof(double)));
  PS_HIP(c, hipMemcpy(da, a, ma * sizeof(double), hipMemcpyHostToDevice));
  PsFitArgs A = make_args(c, 0.0, 0.0, 1);
  A.conw_r2_rows = conw_r2_rows;
  std::vector<double> w((size_t)K * K), ar(c->nfac);
  if (alpha || beta || chisq || area) {
    PS_HIP(c, ps_launch_eval(A, c->stream, da, dw, dar, nullptr));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    PS_HIP(c, hipMemcpy(w.data(), dw, w.size() * sizeof(double), hipMemcpyDeviceToHost));
    PS_HIP(c, hipMemcpy(ar.data(), dar, ar.size() * sizeof(double), hipMemcpyDeviceToHost));
    if (alpha)
      for (int i = 0; i < mf; ++i)
        for (int j = 0; j < mf; ++j) alpha[i * mf + j] = w[(size_t)i * K + j];
    if (beta)
      for (int j = 0; j < mf; ++j) beta[j] = w[(size_t)mf * K + j];
    if (chisq) *chisq = w[(size_t)mf * K + mf];
    if (area) std::memcpy(area, ar.data(), ar.size() * sizeof(double));
  }
  if (jrows) {
    PS_HIP(c, hipMalloc(&dj, (size_t)(c->ndata + 3) * (ma + 1) * sizeof(double)));
    PS_HIP(c, ps_launch_eval(A, c->stream, da, dw, dar, dj));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    PS_HIP(c, hipMemcpy(jrows, dj, (size_t)(c->ndata + 3) * (ma + 1) * sizeof(double), hipMemcpyDeviceToHost));
    hipFree(dj);
  }
  hipFree(da);
  hipFree(dw);
  hipFree(dar);
  return PS_OK;
}

int ps_gaussj(ps_ctx* c, int mfit, const double* alpha, const double* beta, double lambda, double* x, int* err) {
  if (!c || mfit <= 0 || !alpha || !beta || !x || !err) return PS_E_ARG;
  if ((size_t)(mfit + 1) * (mfit + 1) * 8 + 4096 > 64 * 1024) return fail(c, PS_E_ARG, "system too large");
  PS_HIP(c, hipSetDevice(c->device));
  double *da = nullptr, *db = nullptr, *dx = nullptr;
  int* de = nullptr;
  PS_HIP(c, hipMalloc(&da, (size_t)mfit * mfit * sizeof(double)));
  PS_HIP(c, hipMalloc(&db, mfit * sizeof(double)));
  PS_HIP(c, hipMalloc(&dx, mfit * sizeof(double)));
  PS_HIP(c, hipMalloc(&de, sizeof(int)));
  PS_HIP(c, hipMemcpy(da, alpha, (size_t)mfit * mfit * sizeof(double), hipMemcpyHostToDevice));
  PS_HIP(c, hipMemcpy(db, beta, mfit * sizeof(double), hipMemcpyHostToDevice));
  PS_HIP(c, ps_launch_gaussj(mfit, da, db, lambda, dx, de, c->stream));
  PS_HIP(c, hipStreamSynchronize(c->stream));
  PS_HIP(c, hipMemcpy(x, dx, mfit * sizeof(double), hipMemcpyDeviceToHost));
  PS_HIP(c, hipMemcpy(err, de, sizeof(int), hipMemcpyDeviceToHost));
  hipFree(da);
  hipFree(db);
  hipFree(dx);
  hipFree(de);
  return PS_OK;
}

int ps_math_probe(ps_ctx* c, int n, const double* x, double* out, double* out2) {
  if (!c || n <= 0 || !x || !out || !out2) return PS_E_ARG;
  PS_HIP(c, hipSetDevice(c->device));
  double *dx = nullptr, *d1 = nullptr, *d2 = nullptr;
  PS_HIP(c, hipMalloc(&dx, n * sizeof(double)));
  PS_HIP(c, hipMalloc(&d1, 3 * (size_t)n * sizeof(double)));
  PS_HIP(c, hipMalloc(&d2, n * sizeof(double)));
  PS_HIP(c, hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice));
  PS_HIP(c, ps_launch_math(n, dx, d1, d2, c->stream));
  PS_HIP(c, hipStreamSynchronize(c->stream));
  PS_HIP(c, hipMemcpy(out, d1, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  PS_HIP(c, hipMemcpy(out2, d2, n * sizeof(double), hipMemcpyDeviceToHost));
  hipFree(dx);
  hipFree(d1);
  hipFree(d2);
  return PS_OK;
}

}  // extern "C"
