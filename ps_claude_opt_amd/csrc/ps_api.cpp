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
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ps_mi355x.h"
#include "ps_kernels.h"
#include "ps_setup.h"

namespace {
constexpr int kMaxIter = 60;  // constants.h:12
}

#include "ps_ctx.h"

// Device memory of ps_fit_items_static, the entry point for callers that
// cannot allocate: the work-queue counter (its first 256 bytes) and the
// workgroups' scratch slabs.  Every GPU the library runs on holds one copy
// for the life of the process.  In host code `scratch` names a host-side
// stand-in, not device memory: only the address that hipGetSymbolAddress
// returns for the current device (kept per device at ps_open) may be used.
static __device__ __align__(256) double scratch[PS_STATIC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
double* g_static_scratch[kMaxDevices] = {};  // scratch's address on each device (host table)
std::mutex g_static_scratch_mu;               // guards the table (contexts may open concurrently)
}

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

// Device buffers of one call (the test entry points), freed on every
// return path, error returns included.
struct CallBufs {
  std::vector<void*> p;
  CallBufs() = default;
  CallBufs(const CallBufs&) = delete;
  CallBufs& operator=(const CallBufs&) = delete;
  ~CallBufs() {
    for (void* q : p) hipFree(q);
  }
  template <typename T>
  hipError_t alloc(T** out, size_t n) {
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) {
      p.push_back(q);
      *out = static_cast<T*>(q);
    }
    return e;
  }
};

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

// Work-queue reset inside a captured scan step (a kernel, not a memset node).
__global__ void ps_reset_counter_kernel(int* counter) {
  if (threadIdx.x == 0) *counter = 0;
}

// The compiled variants of the fit kernels, by wavefronts per workgroup.
struct FitVariant {
  int waves;
  size_t (*lds_bytes)(const PsDevProblem&);
  hipError_t (*set_attributes)(int, int, size_t);
  int (*blocks_per_cu)(int, int, size_t);
  hipError_t (*launch)(const PsFitArgs&, int, hipStream_t);
  hipError_t (*phase_cycles)(unsigned long long*, bool);
  hipError_t (*gaussj)(int, const double*, const double*, double, double*, int*, hipStream_t);
};
const FitVariant kVariants[4] = {
    {1, psw1::ps_fit_lds_bytes, psw1::ps_set_fit_attributes, psw1::ps_fit_blocks_per_cu, psw1::ps_launch_fit,
     psw1::ps_phase_cycles, psw1::ps_launch_gaussj},
    {2, psw2::ps_fit_lds_bytes, psw2::ps_set_fit_attributes, psw2::ps_fit_blocks_per_cu, psw2::ps_launch_fit,
     psw2::ps_phase_cycles, psw2::ps_launch_gaussj},
    {4, psw4::ps_fit_lds_bytes, psw4::ps_set_fit_attributes, psw4::ps_fit_blocks_per_cu, psw4::ps_launch_fit,
     psw4::ps_phase_cycles, psw4::ps_launch_gaussj},
    {8, psw8::ps_fit_lds_bytes, psw8::ps_set_fit_attributes, psw8::ps_fit_blocks_per_cu, psw8::ps_launch_fit,
     psw8::ps_phase_cycles, psw8::ps_launch_gaussj},
};
int variant_index(int waves) { return waves == 1 ? 0 : waves == 2 ? 1 : waves == 8 ? 3 : 2; }
// The variant the next launch uses: the configured one if it can be
// resident, else the 4-wave one.
int active_variant(const ps_ctx* c) {
  const int v = variant_index(c->waves);
  return c->per_cu[v] > 0 ? v : 2;
}

void drop_graph(ps_ctx* c) {
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
}

// The scan step on the context's stream: reset of the work counter, then
// the persistent fit launch -- eagerly, or as a captured HIP graph that is
// replayed while the launch arguments, grid and variant stay the same.
hipError_t launch_step(ps_ctx* c, const PsFitArgs& A, int grid, int v) {
  const FitVariant& V = kVariants[v];
  if (!c->use_graphs) {
    const int zero = 0;
    hipError_t e = hipMemcpyAsync(c->counter, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return e;
    return V.launch(A, grid, c->stream);
  }
  if (!c->gexec || grid != c->ggrid || V.waves != c->gwaves || std::memcmp(&A, &c->gargs, sizeof(A)) != 0) {
    drop_graph(c);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ps_reset_counter_kernel, dim3(1), dim3(64), 0, c->stream, c->counter);
    hipError_t el = hipGetLastError();
    if (el == hipSuccess) el = V.launch(A, grid, c->stream);
    e = hipStreamEndCapture(c->stream, &g);
    if (el != hipSuccess || e != hipSuccess) {
      if (g) hipGraphDestroy(g);
      return el != hipSuccess ? el : e;
    }
    e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
      c->gexec = nullptr;
      return e;
    }
    std::memcpy(&c->gargs, &A, sizeof(A));
    c->ggrid = grid;
    c->gwaves = V.waves;
    ++c->graph_captures;
  }
  ++c->graph_replays;
  return hipGraphLaunch(c->gexec, c->stream);
}

void free_problem(ps_ctx* c) {
  drop_graph(c);
  for (void* p : c->dbufs) hipFree(p);
  c->dbufs.clear();
  c->d_probs = nullptr;
  c->dps.clear();
  c->hp.clear();
  c->prepared = false;
}

int ensure_batch(ps_ctx* c, int n_items, int grid) {
  if ((size_t)n_items > c->items_cap) {
    for (void* q : {(void*)c->d_items_p, (void*)c->d_items_n, (void*)c->d_items_m, (void*)c->d_items_cw,
                    (void*)c->d_out})
      if (q) hipFree(q);
    c->d_items_p = c->d_items_n = c->d_items_m = nullptr;
    c->d_items_cw = c->d_out = nullptr;
    c->items_cap = 0;
    size_t cap = std::max<size_t>(n_items, 1024);
    PS_HIP(c, hipMalloc(&c->d_items_p, cap * sizeof(int)));
    PS_HIP(c, hipMalloc(&c->d_items_n, cap * sizeof(int)));
    PS_HIP(c, hipMalloc(&c->d_items_m, cap * sizeof(int)));
    PS_HIP(c, hipMalloc(&c->d_items_cw, cap * 2 * sizeof(double)));
    PS_HIP(c, hipMalloc(&c->d_out, cap * PS_OUT_STRIDE * sizeof(double)));
    c->items_cap = cap;
  }
  const size_t need = (size_t)grid * c->stride;
  if (need > c->scratch_elems) {
    if (c->scratch) hipFree(c->scratch);
    c->scratch = nullptr;
    c->scratch_elems = 0;
    PS_HIP(c, hipMalloc(&c->scratch, need * sizeof(double)));
    c->scratch_elems = need;
  }
  if (!c->counter) PS_HIP(c, hipMalloc(&c->counter, sizeof(int)));
  return PS_OK;
}

PsFitArgs make_args(ps_ctx* c, int n_items) {
  PsFitArgs A;
  std::memset(&A, 0, sizeof(A));  // padding too: graph replays compare argument blocks bytewise
  A.P0 = c->dps[0];
  A.probs = c->d_probs;
  A.item_p = c->d_items_p;
  A.item_n = c->d_items_n;
  A.item_m = c->d_items_m;
  A.item_cw = c->d_items_cw;
  A.n_items = n_items;
  A.counter = c->counter;
  A.scratch = c->scratch;
  A.scratch_stride = c->stride;
  A.out = c->d_out;
  A.batch = c->dps.size() > 1 ? 1 : 0;
  A.claim_tiles = c->claim_tiles;
  return A;
}

// earliest reported pole with minimal dev, NaN-safe (start_CUDA.cu:1200-1210)
int best_pole(const ps_pair_result* r) {
  int best = -1;
  for (int p = 0; p < PS_N_POLES; ++p)
    if (r[p].reported == 1 && (best < 0 || (!std::isnan(r[p].dev) && !(r[p].dev >= r[best].dev)))) best = p;
  return best;
}

// Validation of one problem; writes mfit.
int check_problem(ps_ctx* c, const ps_problem* p, int* mfit_out) {
  const int N = p->ndata, L = p->n_curves, F = p->nfac, C = p->ncoef;
  if (N <= 0 || L <= 0 || F <= 0 || F % 4 || C <= 1 || p->lmax < 0 || p->mmax < 0 || p->mmax > p->lmax)
    return fail(c, PS_E_ARG, "bad problem sizes");
  int count = 0;
  for (int i = 0; i < L; ++i) {
    if (p->l_points[i] <= 0) return fail(c, PS_E_ARG, "empty lightcurve");
    count += p->l_points[i];
  }
  if (count != N) return fail(c, PS_E_ARG, "l_points do not sum to ndata");
  const int ma = C + 8;
  // free parameters must be one contiguous run starting at parameter 2, or
  // at parameter 1 (the size scale, free with absolute data)
  int mfit = 0, lastma = 0;
  for (int i = 0; i < ma; ++i)
    if (p->ia[i]) {
      ++mfit;
      lastma = i + 1;
    }
  for (int i = 1; i < lastma; ++i)
    if (!p->ia[i])
      return fail(c, PS_E_UNSUPPORTED, "free parameters must be contiguous from parameter 2 (or 1)");
  if (mfit < 1) return fail(c, PS_E_ARG, "no free parameters");
  const int nb = ps_row_blocks(C, mfit);
  if (nb > 6 || (C - 1 + 15) / 16 < nb - 1)
    return fail(c, PS_E_UNSUPPORTED, "more than 96 Jacobian columns (lmax too large)");
  *mfit_out = mfit;
  return PS_OK;
}

bool same_doubles(const double* a, const double* b, size_t n) { return std::memcmp(a, b, n * sizeof(double)) == 0; }

// Problems of one launch must share the shape basis: sizes, free-parameter
// map and the facet tables, bit for bit.
int check_shared(ps_ctx* c, const ps_problem* a, const ps_problem* b) {
  const int F = a->nfac, C = a->ncoef;
  if (b->nfac != F || b->ncoef != C || b->lmax != a->lmax || b->mmax != a->mmax)
    return fail(c, PS_E_UNSUPPORTED, "problems of a batch must share nfac, lmax and mmax");
  if (std::memcmp(a->ia, b->ia, (size_t)(C + 8) * sizeof(int)) != 0)
    return fail(c, PS_E_UNSUPPORTED, "problems of a batch must share the free-parameter mask");
  const size_t npl = (size_t)(a->mmax + 1) * (a->lmax + 1) * F, nfc = (size_t)(a->mmax + 1) * F;
  if (!same_doubles(a->normal, b->normal, 3 * (size_t)F) || !same_doubles(a->d_area, b->d_area, F) ||
      !same_doubles(a->pleg, b->pleg, npl) || !same_doubles(a->fc, b->fc, nfc) || !same_doubles(a->fs, b->fs, nfc) ||
      !same_doubles(a->d_sphere, b->d_sphere, (size_t)C * F))
    return fail(c, PS_E_UNSUPPORTED, "problems of a batch must share the facet tables");
  return PS_OK;
}

// Uploads n problems (the shared tables once, from problem 0) and sizes the
// launch.  conw[i] / have_file[i] describe period_search_in headers.
int prepare_many(ps_ctx* c, int n, const ps_problem* const* ps_, const double* conw, const bool* have_file) {
  PS_HIP(c, hipSetDevice(c->device));
  free_problem(c);
  if (n < 1) return fail(c, PS_E_ARG, "no problems");
  int mfit = 0;
  for (int i = 0; i < n; ++i) {
    int mf = 0;
    int rc = check_problem(c, ps_[i], &mf);
    if (rc) return rc;
    if (i == 0) mfit = mf;
    if (i > 0 && (rc = check_shared(c, ps_[0], ps_[i])) != PS_OK) return rc;
  }
  const ps_problem* p0 = ps_[0];
  const int F = p0->nfac, C = p0->ncoef, ma = C + 8;
  const int nb = ps_row_blocks(C, mfit);
  const int nbs = (C - 1 + 15) / 16;
  PsDevProblem shared{};
  int rc;
#define UP(dst, ptr, cnt) \
  if ((rc = upload(c, ptr, cnt, &dst)) != PS_OK) return rc
  {
    // PS_FACET_PAD zero facets after the last one: the facet loop's
    // prefetches run up to three 4-facet steps past the end
    std::vector<double> yl((size_t)(F + PS_FACET_PAD) * 16 * nbs, 0.0);
    for (int f = 0; f < F; ++f)
      for (int q = 0; q <= C - 2; ++q)
        yl[((size_t)f * 16 + (q & 15)) * nbs + (q >> 4)] = p0->d_sphere[(size_t)(q + 1) * F + f];
    std::vector<int> cmap;
    for (int i = 0; i < ma; ++i)
      if (p0->ia[i]) cmap.push_back(i);
    UP(shared.nx, p0->normal, F);
    UP(shared.ny, p0->normal + F, F);
    UP(shared.nz, p0->normal + 2 * F, F);
    UP(shared.darea, p0->d_area, F);
    UP(shared.pleg, p0->pleg, (size_t)(p0->mmax + 1) * (p0->lmax + 1) * F);
    UP(shared.fc, p0->fc, (size_t)(p0->mmax + 1) * F);
    UP(shared.fs, p0->fs, (size_t)(p0->mmax + 1) * F);
    UP(shared.yl, yl.data(), yl.size());
    UP(shared.cmap, cmap.data(), cmap.size());
  }
  shared.nfac = F;
  shared.lmax = p0->lmax;
  shared.mmax = p0->mmax;
  shared.ncoef = C;
  shared.ma = ma;
  shared.mfit = mfit;
  shared.nb = nb;
  shared.qs = 16 * nb;
  shared.nbs = nbs;
  shared.ia1 = p0->ia[0] ? 1 : 0;
  c->stride = 0;
  for (int pi = 0; pi < n; ++pi) {
    const ps_problem* p = ps_[pi];
    const int N = p->ndata, L = p->n_curves;
    PsDevProblem d = shared;
    HostProb h;
    std::vector<double> alph(N), s2w(N);
    for (int i = 0; i < N; ++i) {
      const double dd = std::fma(p->ee[2 * N + i], p->ee0[2 * N + i],
                                 std::fma(p->ee[N + i], p->ee0[N + i], p->ee[i] * p->ee0[i]));
      alph[i] = std::acos(std::fmin(1.0, std::fmax(-1.0, dd)));
      s2w[i] = p->sigr2[i] * p->weight[i];
    }
    std::vector<int> off(L + 1, 0), toff(L + 1, 0), tcv, tinfo;
    for (int i = 0; i < L; ++i) {
      off[i + 1] = off[i] + p->l_points[i];
      const int nt = (p->l_points[i] + 15) / 16;
      toff[i + 1] = toff[i] + nt;
      for (int t = 0; t < nt; ++t) {
        tcv.push_back(i);
        const int n_in = std::min(16, p->l_points[i] - 16 * t);
        // the curve length carries the curve's kind: negative for an
        // absolute (calibrated) curve
        tinfo.insert(tinfo.end(), {off[i] + 16 * t, n_in, i, p->in_rel[i] ? p->l_points[i] : -p->l_points[i]});
      }
    }
    UP(d.tim, p->tim, N);
    UP(d.ex, p->ee, N);
    UP(d.ey, p->ee + N, N);
    UP(d.ez, p->ee + 2 * N, N);
    UP(d.sx, p->ee0, N);
    UP(d.sy, p->ee0 + N, N);
    UP(d.sz, p->ee0 + 2 * N, N);
    UP(d.alph, alph.data(), N);
    UP(d.bright, p->brightness, N);
    UP(d.sig, p->sig, N);
    UP(d.s2w, s2w.data(), N);
    UP(d.curve_off, off.data(), L + 1);
    UP(d.tile_off, toff.data(), L + 1);
    UP(d.tile_cv, tcv.data(), tcv.size());
    UP(d.tile_info, tinfo.data(), tinfo.size());
    UP(d.cg_first, p->cg_first, C);
    d.ncurves = L;
    d.ndata = N;
    d.ntiles = (int)tcv.size();
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
    h.ndata = N;
    h.ncurves = L;
    h.f_start = p->f_start;
    h.f_end = p->f_end;
    h.f_step = p->f_step;
    h.n_max = (int)((p->f_start - p->f_end) / p->f_step) + 1;
    h.n_iter_max = d.n_iter_max;
    h.escl = p->escl;
    h.conw = conw ? conw[pi] : 0.0;
    h.have_file = have_file ? have_file[pi] : false;
    h.l_points.assign(p->l_points, p->l_points + L);
    h.tim.assign(p->tim, p->tim + N);
    h.bright.assign(p->brightness, p->brightness + N);
    h.ee.assign(p->ee, p->ee + 3 * (size_t)N);
    h.ee0.assign(p->ee0, p->ee0 + 3 * (size_t)N);
    for (int i = 0; i < PS_N_POLES; ++i) {
      h.beta_pole[i] = p->beta_pole[i];
      h.lambda_pole[i] = p->lambda_pole[i];
    }
    h.cg_first.assign(p->cg_first, p->cg_first + C);
    for (int i = 0; i < 3; ++i) h.par[i] = p->par[i];
    h.lcl = d.lcl;
    h.phi0 = p->phi0;
    c->dps.push_back(d);
    c->hp.push_back(h);
    c->stride = std::max(c->stride, (long long)ps_slab(d).stride);
  }
#undef UP
  {
    void* q = nullptr;
    PS_HIP(c, hipMalloc(&q, sizeof(PsDevProblem) * n));
    c->dbufs.push_back(q);
    PS_HIP(c, hipMemcpy(q, c->dps.data(), sizeof(PsDevProblem) * n, hipMemcpyHostToDevice));
    c->d_probs = static_cast<PsDevProblem*>(q);
  }
  c->nfac = F;
  c->ncoef = C;
  c->ma = ma;
  c->mfit = mfit;
  c->ia.assign(p0->ia, p0->ia + ma);
  const size_t lds_cap = 160 * 1024;
  for (int v = 0; v < 4; ++v) {
    const FitVariant& V = kVariants[v];
    c->lds[v] = V.lds_bytes(c->dps[0]);
    c->per_cu[v] = 0;
    if (c->lds[v] <= lds_cap && V.set_attributes(nb, nbs, c->lds[v]) == hipSuccess)
      c->per_cu[v] = V.blocks_per_cu(nb, nbs, c->lds[v]);
  }
  if (c->lds[2] > lds_cap) return fail(c, PS_E_UNSUPPORTED, "problem needs more LDS than a workgroup can have");
  if (c->per_cu[2] < 1) return fail(c, PS_E_HIP, "fit kernel cannot be resident (occupancy query failed)");
  c->prepared = true;
  return PS_OK;
}

// Fits a list of work items in one persistent launch.  Items are queued
// largest problem first (the work of a pair grows with the problem's point
// count), results come back in the caller's order.
int fit_items(ps_ctx* c, int n_items, const int* prob, const int* n_list, const int* m_list, const double* cw_rows,
              const double* cw_dev, ps_pair_result* out, const double* a0 = nullptr, double* a_final = nullptr,
              int n_iter_fixed = 0) {
  c->last_ms = 0;
  c->last_launches = 0;
  c->last_pairs = c->last_vis = c->last_iters = c->last_evals = 0;
  if (n_items == 0) return PS_OK;
  const int np = (int)c->dps.size();
  for (int i = 0; i < n_items; ++i) {
    const int pi = prob ? prob[i] : 0;
    if (pi < 0 || pi >= np) return fail(c, PS_E_ARG, "problem index out of range");
    if (n_list[i] < 1 || m_list[i] < 1 || m_list[i] > PS_N_POLES) return fail(c, PS_E_ARG, "pair index out of range");
  }
  PS_HIP(c, hipSetDevice(c->device));
  std::vector<int> order(n_items);
  for (int i = 0; i < n_items; ++i) order[i] = i;
  if (prob && np > 1)
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return c->hp[prob[a]].ndata > c->hp[prob[b]].ndata; });
  std::vector<int> ip(n_items), in(n_items), im(n_items);
  std::vector<double> icw(2 * (size_t)n_items);
  for (int k = 0; k < n_items; ++k) {
    const int i = order[k];
    ip[k] = prob ? prob[i] : 0;
    in[k] = n_list[i];
    im[k] = m_list[i];
    icw[2 * k] = cw_rows[i];
    icw[2 * k + 1] = cw_dev[i];
  }
  const int v = active_variant(c);
  const int per_cu = c->per_cu[v];
  const int wpc = c->wg_per_cu > 0 ? std::min(c->wg_per_cu, per_cu) : per_cu;
  const int grid = std::min(n_items, c->n_cu * wpc);
  int rc = ensure_batch(c, n_items, grid);
  if (rc) return rc;
  PS_HIP(c, hipMemcpyAsync(c->d_items_p, ip.data(), n_items * sizeof(int), hipMemcpyHostToDevice, c->stream));
  PS_HIP(c, hipMemcpyAsync(c->d_items_n, in.data(), n_items * sizeof(int), hipMemcpyHostToDevice, c->stream));
  PS_HIP(c, hipMemcpyAsync(c->d_items_m, im.data(), n_items * sizeof(int), hipMemcpyHostToDevice, c->stream));
  PS_HIP(c, hipMemcpyAsync(c->d_items_cw, icw.data(), icw.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  const int ma = c->ma;
  if ((a0 || a_final) && (size_t)n_items > c->a_cap) {
    if (c->d_a0) hipFree(c->d_a0);
    if (c->d_afinal) hipFree(c->d_afinal);
    c->d_a0 = c->d_afinal = nullptr;
    c->a_cap = 0;
    PS_HIP(c, hipMalloc(&c->d_a0, (size_t)n_items * ma * sizeof(double)));
    PS_HIP(c, hipMalloc(&c->d_afinal, (size_t)n_items * ma * sizeof(double)));
    c->a_cap = n_items;
  }
  std::vector<double> a0p;
  if (a0) {
    a0p.resize((size_t)n_items * ma);
    for (int k = 0; k < n_items; ++k)
      std::memcpy(&a0p[(size_t)k * ma], a0 + (size_t)order[k] * ma, ma * sizeof(double));
    PS_HIP(c, hipMemcpyAsync(c->d_a0, a0p.data(), a0p.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  PsFitArgs A = make_args(c, n_items);
  A.item_a0 = a0 ? c->d_a0 : nullptr;
  A.out_a = a_final ? c->d_afinal : nullptr;
  A.n_iter_fixed = n_iter_fixed;
  PS_HIP(c, hipEventRecord(c->ev0, c->stream));
  PS_HIP(c, launch_step(c, A, grid, v));
  PS_HIP(c, hipEventRecord(c->ev1, c->stream));
  std::vector<double> res((size_t)n_items * PS_OUT_STRIDE);
  PS_HIP(c, hipMemcpyAsync(res.data(), c->d_out, res.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> afp;
  if (a_final) {
    afp.resize((size_t)n_items * ma);
    PS_HIP(c, hipMemcpyAsync(afp.data(), c->d_afinal, afp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  PS_HIP(c, hipStreamSynchronize(c->stream));
  if (a_final)
    for (int k = 0; k < n_items; ++k)
      std::memcpy(a_final + (size_t)order[k] * ma, &afp[(size_t)k * ma], ma * sizeof(double));
  float ms = 0;
  PS_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  c->last_launches = 1;
  if (std::getenv("PS_PHASE_TIMING")) {
    unsigned long long cyc[16];
    if (kVariants[v].phase_cycles(cyc, true) == hipSuccess) {
      std::fprintf(stderr, "phase cycles (thread 0 of each workgroup, summed):");
      for (int i = 0; i < 10; ++i) std::fprintf(stderr, " %d:%.4g", i, (double)cyc[i]);
      std::fprintf(stderr, "\n");
    }
  }
  for (int k = 0; k < n_items; ++k) {
    const double* o = &res[(size_t)k * PS_OUT_STRIDE];
    ps_pair_result& r = out[order[k]];
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
  c->last_pairs = n_items;
  return PS_OK;
}

// Output lines of a scan from its pair results (start_CUDA.cu:1200-1229).
void make_lines(const HostProb& h, int n_start, int n_count, double conw_r, const ps_pair_result* r, ps_line* out) {
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
    L.chi2 = x.dev * x.dev * h.ndata;
    L.dark_or_conw = (n_start + i == 1) ? conw_r * h.escl * h.escl : x.dark;
    L.lambda_deg = x.lambda_deg;
    L.beta_deg = x.beta_deg;
  }
}

// Conw decision of one round from its 10 x 10 results (start_CUDA.cu:
// 700-789): mean dark-facet percentage of the best poles.
void conw_decide(const ps_pair_result* r, double* conw_r, int* new_conw) {
  const int tests = 10;
  double sum_dark = 0.0;
  for (int n = 0; n < tests; ++n) {
    const int b = best_pole(&r[n * PS_N_POLES]);
    if (b >= 0) sum_dark = sum_dark + r[n * PS_N_POLES + b].dark;
  }
  const double ave = sum_dark / tests;
  if (ave < 1.0) *new_conw = 1;
  if (ave >= 1.0) *conw_r = *conw_r * 2;
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
  // the static scratch's address on this device, once per device
  if (device < kMaxDevices) {
    std::lock_guard<std::mutex> lock(g_static_scratch_mu);
    if (!g_static_scratch[device]) {
      void* sp = nullptr;
      if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) != hipSuccess || !sp) {
        delete c;
        return PS_E_HIP;
      }
      g_static_scratch[device] = static_cast<double*>(sp);
    }
    c->static_scratch = g_static_scratch[device];
  }
  if (const char* w = std::getenv("PS_WAVES_PER_PAIR")) {
    const int n = std::atoi(w);
    c->waves = (n == 1 || n == 2 || n == 8) ? n : 4;
  }
  if (const char* g = std::getenv("PS_WG_PER_CU")) c->wg_per_cu = std::max(0, std::atoi(g));
  if (const char* t = std::getenv("PS_TILE_POINTS")) {
    const int n = std::atoi(t);
    c->claim_tiles = (n == 64 || n == 128 || n == 256) ? n / 16 : 1;
  }
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
  for (void* q : {(void*)c->d_items_p, (void*)c->d_items_n, (void*)c->d_items_m, (void*)c->d_items_cw, (void*)c->d_out,
                  (void*)c->d_a0, (void*)c->d_afinal})
    if (q) hipFree(q);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return PS_OK;
}

int ps_prepare(ps_ctx* c, const ps_problem* p) {
  if (!c || !p) return PS_E_ARG;
  return prepare_many(c, 1, &p, nullptr, nullptr);
}

int ps_prepare_batch(ps_ctx* c, int n, const ps_problem* problems) {
  if (!c || n < 1 || !problems) return PS_E_ARG;
  std::vector<const ps_problem*> v(n);
  for (int i = 0; i < n; ++i) v[i] = &problems[i];
  return prepare_many(c, n, v.data(), nullptr, nullptr);
}

int ps_prepare_files(ps_ctx* c, int n, const char* const* paths) {
  if (!c || n < 1 || !paths) return PS_E_ARG;
  std::vector<ps::InputFile> files(n);
  std::vector<ps_problem> views(n);
  std::vector<const ps_problem*> v(n);
  std::vector<double> conw(n);
  std::unique_ptr<bool[]> hf(new bool[n]);
  for (int i = 0; i < n; ++i) {
    if (!paths[i]) return PS_E_ARG;
    std::string e;
    const int rc = ps::read_input(paths[i], files[i], e);
    if (rc) return fail(c, rc, std::string(paths[i]) + ": " + e);
    views[i] = files[i].view();
    views[i].stop_condition = files[i].stop_condition;
    v[i] = &views[i];
    conw[i] = files[i].conw;
    hf[i] = true;
  }
  return prepare_many(c, n, v.data(), conw.data(), hf.get());
}

int ps_prepare_file(ps_ctx* c, const char* path) {
  if (!c || !path) return PS_E_ARG;
  return ps_prepare_files(c, 1, &path);
}

int ps_n_problems(const ps_ctx* c) { return (c && c->prepared) ? (int)c->hp.size() : 0; }

int ps_problem_info(const ps_ctx* c, int i, int* ints, double* dbls) {
  if (!c || !c->prepared) return PS_E_STATE;
  if (i < 0 || i >= (int)c->hp.size()) return PS_E_ARG;
  const HostProb& h = c->hp[i];
  if (ints) {
    const int v[8] = {h.ndata, h.ncurves, c->nfac, c->ncoef, c->ma, c->mfit, h.n_max, h.n_iter_max};
    std::memcpy(ints, v, sizeof v);
  }
  if (dbls) {
    const double v[4] = {h.f_start, h.f_step, h.escl, h.have_file ? h.conw : 0.0};
    std::memcpy(dbls, v, sizeof v);
  }
  return PS_OK;
}

int ps_info(const ps_ctx* c, int* ints, double* dbls) { return ps_problem_info(c, 0, ints, dbls); }

int ps_fit_pairs(ps_ctx* c, int npairs, const int* n_list, const int* m_list, double conw_r, double conw_r_dev,
                 ps_pair_result* out) {
  if (!c || npairs < 0 || (npairs && (!n_list || !m_list || !out))) return PS_E_ARG;
  if (!c->prepared) return fail(c, PS_E_STATE, "ps_fit_pairs before ps_prepare");
  std::vector<double> cr(npairs, conw_r * conw_r), cd(npairs, conw_r_dev);
  return fit_items(c, npairs, nullptr, n_list, m_list, cr.data(), cd.data(), out);
}

int ps_fit_items(ps_ctx* c, int n_items, const int* prob, const int* n_list, const int* m_list, const double* conw_r,
                 const double* conw_r_dev, ps_pair_result* out) {
  if (!c || n_items < 0 || (n_items && (!prob || !n_list || !m_list || !conw_r || !conw_r_dev || !out)))
    return PS_E_ARG;
  if (!c->prepared) return fail(c, PS_E_STATE, "ps_fit_items before ps_prepare");
  std::vector<double> cr(n_items);
  for (int i = 0; i < n_items; ++i) cr[i] = conw_r[i] * conw_r[i];
  return fit_items(c, n_items, prob, n_list, m_list, cr.data(), conw_r_dev, out);
}

int ps_fit_items_from(ps_ctx* c, int n_items, const int* prob, const int* n_list, const int* m_list,
                      const double* conw_r, const double* conw_r_dev, const double* a0, int n_iter,
                      ps_pair_result* out, double* a_final) {
  if (!c || n_items < 0 || n_iter < 0 ||
      (n_items && (!prob || !n_list || !m_list || !conw_r || !conw_r_dev || !out)))
    return PS_E_ARG;
  if (!c->prepared) return fail(c, PS_E_STATE, "ps_fit_items_from before ps_prepare");
  std::vector<double> cr(n_items);
  for (int i = 0; i < n_items; ++i) cr[i] = conw_r[i] * conw_r[i];
  return fit_items(c, n_items, prob, n_list, m_list, cr.data(), conw_r_dev, out, a0, a_final, n_iter);
}

int ps_initial_params(const ps_ctx* c, int n_items, const int* prob, const int* n_list, const int* m_list,
                      double* a0) {
  if (!c || n_items < 0 || (n_items && (!prob || !n_list || !m_list || !a0))) return PS_E_ARG;
  if (!c->prepared) return PS_E_STATE;
  constexpr double kPi = 3.14159265358979323846;
  const int C = c->ncoef, ma = c->ma, np = (int)c->hp.size();
  for (int k = 0; k < n_items; ++k) {
    const int p = prob[k], n = n_list[k], m = m_list[k];
    if (p < 0 || p >= np || n < 1 || m < 1 || m > PS_N_POLES) return PS_E_ARG;
    const HostProb& h = c->hp[p];
    double* a = a0 + (size_t)k * ma;
    // Start.cu:4346-4418, in the kernel's operation order
    for (int i = 0; i < C; ++i) a[i] = h.cg_first[i];
    const double period = 1.0 / (h.f_start - (double)(n - 1) * h.f_step);
    const double d2r = kPi / 180.0;
    a[C] = d2r * 90 - d2r * h.beta_pole[m - 1];
    a[C + 1] = d2r * h.lambda_pole[m - 1];
    a[C + 2] = (24.0 * 2.0 * kPi) / period;
    a[C + 3] = h.par[0];
    a[C + 4] = h.par[1];
    a[C + 5] = h.par[2];
    a[C + 6] = h.lcl;
    a[C + 7] = 1;
  }
  return PS_OK;
}

int ps_free_params(const ps_ctx* c, int* ia) {
  if (!c || !ia) return PS_E_ARG;
  if (!c->prepared) return PS_E_STATE;
  std::memcpy(ia, c->ia.data(), c->ia.size() * sizeof(int));
  return PS_OK;
}

int ps_warm_n_features(void) { return PS_WARM_FEATURES; }

int ps_warm_features(const ps_ctx* c, int n_items, const int* prob, const int* n_list, const int* m_list,
                     float* out) {
  if (!c || n_items < 0 || (n_items && (!prob || !n_list || !m_list || !out))) return PS_E_ARG;
  if (!c->prepared) return PS_E_STATE;
  const int np = (int)c->hp.size();
  constexpr double kPi = 3.14159265358979323846;
  // per problem: the lightcurves that enter the features (the first 8 curves
  // spanning at most 2 days with at least 16 points) and the aspect summary
  struct Sel {
    std::vector<int> first, count;
    double me[3] = {0, 0, 0}, ms[3] = {0, 0, 0}, mphase = 0;
  };
  std::vector<Sel> sel(np);
  for (int p = 0; p < np; ++p) {
    const HostProb& h = c->hp[p];
    const int N = h.ndata;
    int off = 0, used = 0;
    for (size_t ci = 0; ci < h.l_points.size() && (int)sel[p].first.size() < PS_WARM_CURVES; ++ci) {
      const int lp = h.l_points[ci];
      if (lp >= 16 && h.tim[off + lp - 1] - h.tim[off] <= 2.0) {
        sel[p].first.push_back(off);
        sel[p].count.push_back(lp);
        for (int j = off; j < off + lp; ++j) {
          for (int d = 0; d < 3; ++d) {
            sel[p].me[d] += h.ee[(size_t)d * N + j];
            sel[p].ms[d] += h.ee0[(size_t)d * N + j];
          }
          const double cs = h.ee[j] * h.ee0[j] + h.ee[N + j] * h.ee0[N + j] + h.ee[2 * N + j] * h.ee0[2 * N + j];
          sel[p].mphase += std::acos(std::fmax(-1.0, std::fmin(1.0, cs)));
        }
        used += lp;
      }
      off += lp;
    }
    if (used > 0) {
      for (int d = 0; d < 3; ++d) {
        sel[p].me[d] /= used;
        sel[p].ms[d] /= used;
      }
      sel[p].mphase /= used;
    }
  }
  for (int k = 0; k < n_items; ++k) {
    const int p = prob[k];
    if (p < 0 || p >= np || n_list[k] < 1 || m_list[k] < 1 || m_list[k] > PS_N_POLES) return PS_E_ARG;
    const HostProb& h = c->hp[p];
    const Sel& S = sel[p];
    float* f = out + (size_t)k * PS_WARM_FEATURES;
    for (int i = 0; i < PS_WARM_FEATURES; ++i) f[i] = 0.f;
    const double period_h = 1.0 / (h.f_start - (double)(n_list[k] - 1) * h.f_step);
    const double turns_per_day = 24.0 / period_h;
    for (size_t ci = 0; ci < S.first.size(); ++ci) {
      double sum[PS_WARM_BINS] = {0}, mean = 0;
      int cnt[PS_WARM_BINS] = {0};
      const int j0 = S.first[ci], lp = S.count[ci];
      for (int j = j0; j < j0 + lp; ++j) mean += h.bright[j];
      mean /= lp;
      for (int j = j0; j < j0 + lp; ++j) {
        double ph = h.tim[j] * turns_per_day;
        ph -= std::floor(ph);
        int b = (int)(ph * PS_WARM_BINS);
        b = b < 0 ? 0 : (b >= PS_WARM_BINS ? PS_WARM_BINS - 1 : b);
        sum[b] += h.bright[j] / mean - 1.0;
        cnt[b] += 1;
      }
      for (int b = 0; b < PS_WARM_BINS; ++b)
        f[ci * PS_WARM_BINS + b] = cnt[b] ? (float)(sum[b] / cnt[b]) : 0.f;
    }
    const int base = PS_WARM_CURVES * PS_WARM_BINS;
    const double be = h.beta_pole[m_list[k] - 1] * kPi / 180.0, la = h.lambda_pole[m_list[k] - 1] * kPi / 180.0;
    f[base + 0] = (float)(std::cos(be) * std::cos(la));
    f[base + 1] = (float)(std::cos(be) * std::sin(la));
    f[base + 2] = (float)std::sin(be);
    for (int d = 0; d < 3; ++d) {
      f[base + 3 + d] = (float)S.me[d];
      f[base + 6 + d] = (float)S.ms[d];
    }
    f[base + 9] = (float)S.mphase;
    f[base + 10] = (float)std::log(period_h / 10.0);
  }
  return PS_OK;
}

int ps_fit_items_static(ps_ctx* c, int n_items, const int* d_prob, const int* d_n, const int* d_m,
                        const double* d_cw, double* d_out) {
  if (!c) return PS_E_ARG;
  if (!c->prepared) return fail(c, PS_E_STATE, "ps_fit_items_static before ps_prepare");
  if (n_items <= 0 || !d_n || !d_m || !d_cw || !d_out || (c->dps.size() > 1 && !d_prob))
    return fail(c, PS_E_ARG, "ps_fit_items_static: bad item list");
  if (!c->static_scratch) return fail(c, PS_E_HIP, "no static scratch on this device");
  // the counter in the first 256 bytes, then as many slabs as fit
  constexpr size_t head = 256 / sizeof(double), total = PS_STATIC_SCRATCH_BYTES / sizeof(double);
  const long long max_slabs = (long long)((total - head) / (size_t)c->stride);
  if (max_slabs < 1)
    return fail(c, PS_E_ARG, "ps_fit_items_static: one scratch slab of this problem exceeds the static scratch");
  PS_HIP(c, hipSetDevice(c->device));
  const int v = active_variant(c);
  const int per_cu = c->per_cu[v];
  const int wpc = c->wg_per_cu > 0 ? std::min(c->wg_per_cu, per_cu) : per_cu;
  const int grid = (int)std::min<long long>({(long long)n_items, (long long)c->n_cu * wpc, max_slabs});
  // the allocating path's arguments with the caller's buffers and the
  // static array in place of the context's
  PsFitArgs A = make_args(c, n_items);
  A.item_p = d_prob;
  A.item_n = d_n;
  A.item_m = d_m;
  A.item_cw = d_cw;
  A.counter = reinterpret_cast<int*>(c->static_scratch);
  A.scratch = c->static_scratch + head;
  A.out = d_out;
  // the counter is cleared by a kernel (no memset node if the caller
  // captures the stream), then the persistent launch
  PS_HIP(c, hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(ps_reset_counter_kernel, dim3(1), dim3(64), 0, c->stream, A.counter);
  PS_HIP(c, hipGetLastError());
  PS_HIP(c, kVariants[v].launch(A, grid, c->stream));
  PS_HIP(c, hipEventRecord(c->ev1, c->stream));
  PS_HIP(c, hipStreamSynchronize(c->stream));
  float ms = 0;
  PS_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_ms = ms;
  c->last_launches = 1;
  c->last_pairs = n_items;
  c->last_vis = c->last_iters = c->last_evals = 0;
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
  conw_decide(r.data(), conw_r, new_conw);
  return PS_OK;
}

int ps_conw_search(ps_ctx* c, double* conw_r) {
  if (!c || !conw_r) return PS_E_ARG;
  if (!c->prepared || !c->hp[0].have_file) return fail(c, PS_E_STATE, "ps_conw_search needs ps_prepare_file");
  const HostProb& h = c->hp[0];
  double cw = h.conw / h.escl / h.escl;
  int new_conw = 0;
  while (new_conw != 1 && (cw * h.escl * h.escl) < 10.0) {
    int rc = ps_precalc(c, &cw, &new_conw);
    if (rc) return rc;
  }
  *conw_r = cw;
  return PS_OK;
}

int ps_conw_search_batch(ps_ctx* c, double* conw_r) {
  if (!c || !conw_r) return PS_E_ARG;
  if (!c->prepared) return fail(c, PS_E_STATE, "ps_conw_search_batch before ps_prepare_files");
  const int np = (int)c->hp.size();
  for (int i = 0; i < np; ++i)
    if (!c->hp[i].have_file) return fail(c, PS_E_STATE, "ps_conw_search_batch needs ps_prepare_files");
  // every problem runs the loop of period_search_BOINC.cpp:647-771; the
  // rounds of all undecided problems share one launch
  std::vector<double> cw(np);
  std::vector<int> done(np, 0);
  for (int i = 0; i < np; ++i) cw[i] = c->hp[i].conw / c->hp[i].escl / c->hp[i].escl;
  for (;;) {
    std::vector<int> act;
    for (int i = 0; i < np; ++i)
      if (!done[i] && (cw[i] * c->hp[i].escl * c->hp[i].escl) < 10.0) act.push_back(i);
    if (act.empty()) break;
    std::vector<int> ip, nl, ml;
    std::vector<double> cr, cd;
    for (int i : act)
      for (int n = 1; n <= 10; ++n)
        for (int m = 1; m <= PS_N_POLES; ++m) {
          ip.push_back(i);
          nl.push_back(n);
          ml.push_back(m);
          cr.push_back(cw[i] * cw[i]);
          cd.push_back(0.0);
        }
    std::vector<ps_pair_result> r(ip.size());
    int rc = fit_items(c, (int)ip.size(), ip.data(), nl.data(), ml.data(), cr.data(), cd.data(), r.data());
    if (rc) return rc;
    for (size_t k = 0; k < act.size(); ++k) conw_decide(&r[k * 100], &cw[act[k]], &done[act[k]]);
  }
  for (int i = 0; i < np; ++i) conw_r[i] = cw[i];
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
  make_lines(c->hp[0], n_start, n_count, conw_r, r.data(), out);
  return PS_OK;
}

int ps_start_batch(ps_ctx* c, const int* n_start, const int* n_count, const double* conw_r, ps_line* out) {
  if (!c || !n_start || !n_count || !conw_r) return PS_E_ARG;
  if (!c->prepared) return fail(c, PS_E_STATE, "ps_start_batch before ps_prepare");
  const int np = (int)c->hp.size();
  std::vector<int> ip, nl, ml;
  std::vector<double> cr, cd;
  size_t lines = 0;
  for (int i = 0; i < np; ++i) {
    if (n_start[i] < 1 || n_count[i] < 0) return PS_E_ARG;
    lines += n_count[i];
    for (int n = n_start[i]; n < n_start[i] + n_count[i]; ++n)
      for (int m = 1; m <= PS_N_POLES; ++m) {
        ip.push_back(i);
        nl.push_back(n);
        ml.push_back(m);
        cr.push_back(conw_r[i] * conw_r[i]);
        cd.push_back(conw_r[i]);
      }
  }
  if (lines && !out) return PS_E_ARG;
  if ((long long)ip.size() > 0x7fffffffLL) return fail(c, PS_E_ARG, "too many pairs in one call");
  std::vector<ps_pair_result> r(ip.size());
  int rc = fit_items(c, (int)ip.size(), ip.data(), nl.data(), ml.data(), cr.data(), cd.data(), r.data());
  if (rc) return rc;
  size_t li = 0, ri = 0;
  for (int i = 0; i < np; ++i) {
    make_lines(c->hp[i], n_start[i], n_count[i], conw_r[i], &r[ri], out + li);
    li += n_count[i];
    ri += (size_t)n_count[i] * PS_N_POLES;
  }
  return PS_OK;
}

int ps_set_launch(ps_ctx* c, int waves_per_pair, int wg_per_cu) {
  if (!c || (waves_per_pair != 1 && waves_per_pair != 2 && waves_per_pair != 4 && waves_per_pair != 8) ||
      wg_per_cu < 0)
    return PS_E_ARG;
  c->waves = waves_per_pair;
  c->wg_per_cu = wg_per_cu;
  return PS_OK;
}

int ps_set_tile_points(ps_ctx* c, int points) {
  if (!c || (points != 16 && points != 64 && points != 128 && points != 256)) return PS_E_ARG;
  c->claim_tiles = points / 16;
  return PS_OK;
}

int ps_get_tile_points(const ps_ctx* c, int* points) {
  if (!c || !points) return PS_E_ARG;
  *points = 16 * c->claim_tiles;
  return PS_OK;
}

int ps_set_graphs(ps_ctx* c, int on) {
  if (!c) return PS_E_ARG;
  c->use_graphs = on != 0;
  if (!c->use_graphs) drop_graph(c);
  return PS_OK;
}

int ps_graph_stats(const ps_ctx* c, int* captures, int* replays) {
  if (!c) return PS_E_ARG;
  if (captures) *captures = c->graph_captures;
  if (replays) *replays = c->graph_replays;
  return PS_OK;
}

int ps_get_launch(const ps_ctx* c, int* waves_per_pair, int* wg_per_cu) {
  if (!c || !c->prepared) return PS_E_STATE;
  const int v = active_variant(c);
  const int per_cu = c->per_cu[v];
  if (waves_per_pair) *waves_per_pair = kVariants[v].waves;
  if (wg_per_cu) *wg_per_cu = c->wg_per_cu > 0 ? std::min(c->wg_per_cu, per_cu) : per_cu;
  return PS_OK;
}

int ps_launch_options(const ps_ctx* c, int* waves, int* max_wg_per_cu, int cap) {
  if (!c || !c->prepared) return PS_E_STATE;
  int n = 0;
  for (int v = 0; v < 4; ++v)
    if (c->per_cu[v] > 0) {
      if (n < cap) {
        if (waves) waves[n] = kVariants[v].waves;
        if (max_wg_per_cu) max_wg_per_cu[n] = c->per_cu[v];
      }
      ++n;
    }
  return n;
}

int ps_autotune(ps_ctx* c, int n_items, const int* prob, const int* n_list, const int* m_list, const double* conw_r,
                int* waves_per_pair, int* wg_per_cu, double* best_ms) {
  if (!c || n_items <= 0 || !prob || !n_list || !m_list || !conw_r) return PS_E_ARG;
  if (!c->prepared) return fail(c, PS_E_STATE, "ps_autotune before ps_prepare");
  std::vector<ps_pair_result> r(n_items);
  int bw = 4, bg = 0, bt = 1;
  double bms = 0.0;
  bool have = false;
  // waves per pair x workgroups per CU at one tile per claim, then the
  // claim size (16 .. 256 points) at the best of those
  for (int v = 0; v < 4; ++v) {
    const int per_cu = c->per_cu[v];
    if (per_cu < 1) continue;
    for (int g = 1; g <= per_cu; ++g) {
      c->waves = kVariants[v].waves;
      c->wg_per_cu = g;
      c->claim_tiles = 1;
      int rc = ps_fit_items(c, n_items, prob, n_list, m_list, conw_r, conw_r, r.data());
      if (rc) return rc;
      if (!have || c->last_ms < bms) {
        have = true;
        bms = c->last_ms;
        bw = c->waves;
        bg = g;
      }
    }
  }
  c->waves = bw;
  c->wg_per_cu = bg;
  for (int t : {4, 8, 16}) {
    c->claim_tiles = t;
    int rc = ps_fit_items(c, n_items, prob, n_list, m_list, conw_r, conw_r, r.data());
    if (rc) return rc;
    if (c->last_ms < bms) {
      bms = c->last_ms;
      bt = t;
    }
  }
  c->claim_tiles = bt;
  if (waves_per_pair) *waves_per_pair = bw;
  if (wg_per_cu) *wg_per_cu = bg;
  if (best_ms) *best_ms = bms;
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
  CallBufs bufs;
  PS_HIP(c, bufs.alloc(&da, ma));
  PS_HIP(c, bufs.alloc(&dw, (size_t)K * K));
  PS_HIP(c, bufs.alloc(&dar, c->nfac));
  PS_HIP(c, hipMemcpy(da, a, ma * sizeof(double), hipMemcpyHostToDevice));
  PsFitArgs A = make_args(c, 1);
  A.conw_r2_rows = conw_r2_rows;
  std::vector<double> w((size_t)K * K), ar(c->nfac);
  if (alpha || beta || chisq || area) {
    PS_HIP(c, psw4::ps_launch_eval(A, c->stream, da, dw, dar, nullptr));
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
    const size_t nj = (size_t)(c->hp[0].ndata + 3) * (ma + 1);
    PS_HIP(c, bufs.alloc(&dj, nj));
    PS_HIP(c, psw4::ps_launch_eval(A, c->stream, da, dw, dar, dj));
    PS_HIP(c, hipStreamSynchronize(c->stream));
    PS_HIP(c, hipMemcpy(jrows, dj, nj * sizeof(double), hipMemcpyDeviceToHost));
  }
  return PS_OK;
}

int ps_gaussj(ps_ctx* c, int mfit, const double* alpha, const double* beta, double lambda, double* x, int* err) {
  if (!c || mfit <= 0 || !alpha || !beta || !x || !err) return PS_E_ARG;
  if ((size_t)(mfit + 1) * (mfit + 1) * 8 + 4096 > 64 * 1024) return fail(c, PS_E_ARG, "system too large");
  PS_HIP(c, hipSetDevice(c->device));
  double *da = nullptr, *db = nullptr, *dx = nullptr;
  int* de = nullptr;
  CallBufs bufs;
  PS_HIP(c, bufs.alloc(&da, (size_t)mfit * mfit));
  PS_HIP(c, bufs.alloc(&db, mfit));
  PS_HIP(c, bufs.alloc(&dx, mfit));
  PS_HIP(c, bufs.alloc(&de, 1));
  PS_HIP(c, hipMemcpy(da, alpha, (size_t)mfit * mfit * sizeof(double), hipMemcpyHostToDevice));
  PS_HIP(c, hipMemcpy(db, beta, mfit * sizeof(double), hipMemcpyHostToDevice));
  // the Gauss-Jordan of the configured wave count (ps_set_launch)
  PS_HIP(c, kVariants[variant_index(c->waves)].gaussj(mfit, da, db, lambda, dx, de, c->stream));
  PS_HIP(c, hipStreamSynchronize(c->stream));
  PS_HIP(c, hipMemcpy(x, dx, mfit * sizeof(double), hipMemcpyDeviceToHost));
  PS_HIP(c, hipMemcpy(err, de, sizeof(int), hipMemcpyDeviceToHost));
  return PS_OK;
}

int ps_math_probe(ps_ctx* c, int n, const double* x, double* out, double* out2) {
  if (!c || n <= 0 || !x || !out || !out2) return PS_E_ARG;
  PS_HIP(c, hipSetDevice(c->device));
  double *dx = nullptr, *d1 = nullptr, *d2 = nullptr;
  CallBufs bufs;
  PS_HIP(c, bufs.alloc(&dx, n));
  PS_HIP(c, bufs.alloc(&d1, 3 * (size_t)n));
  PS_HIP(c, bufs.alloc(&d2, n));
  PS_HIP(c, hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice));
  PS_HIP(c, psw4::ps_launch_math(n, dx, d1, d2, c->stream));
  PS_HIP(c, hipStreamSynchronize(c->stream));
  PS_HIP(c, hipMemcpy(out, d1, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  PS_HIP(c, hipMemcpy(out2, d2, n * sizeof(double), hipMemcpyDeviceToHost));
  return PS_OK;
}

}  // extern "C"
