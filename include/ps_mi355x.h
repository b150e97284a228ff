/* ps_mi355x -- C ABI of the MI355X period scan.
 *
 * The drop-in boundary for the reference's GPU layer.  The reference exposes
 * four C++ functions (start_CUDA.h:1-18) that read most of their inputs from
 * process globals; this ABI passes everything explicitly, returns int status
 * codes (never exits the process) and writes into caller-owned buffers.
 *
 *   reference                                    this ABI
 *   CUDAPrepare   (start_CUDA.cu:207-352)  ->    ps_prepare
 *   CUDAPrecalc   (start_CUDA.cu:400-789)  ->    ps_precalc (one conw round)
 *   CUDAStart     (start_CUDA.cu:792-1245) ->    ps_start
 *   CUDAUnprepare (start_CUDA.cu:356-373)  ->    ps_unprepare
 *
 * The application-level pieces the reference keeps in main()
 * (period_search_BOINC.cpp:343-866: input parsing, geometry tables, the
 * convexity-weight loop) are exported too, so a caller can either hand over
 * its own tables (ps_problem) or let the library read period_search_in.
 *
 * All arithmetic is IEEE FP64.  Functions are synchronous; the library owns
 * its HIP stream and device memory.  Status codes: PS_OK or a negative
 * PS_E_* value; ps_last_error() describes the last failure on a context.
 */
#ifndef PS_MI355X_H
#define PS_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

#define PS_OK 0
#define PS_E_ARG (-1)         /* bad argument / size limit */
#define PS_E_INPUT (-2)       /* malformed period_search_in */
#define PS_E_HIP (-3)         /* HIP runtime failure */
#define PS_E_UNSUPPORTED (-4) /* problem outside what the kernels support */
#define PS_E_STATE (-5)       /* call order (e.g. ps_start before ps_prepare) */

#define PS_N_POLES 10

typedef struct ps_ctx ps_ctx;

/* Everything the scan reads.  Arrays are 0-based; N = ndata (data points
 * only, the three convexity rows are added by the library), L = n_curves,
 * F = nfac, C = ncoef, ma = C + 8.  The struct mirrors the globals and
 * arguments of CUDAPrepare/CUDAPrecalc/CUDAStart (globals.h:6-22,
 * start_CUDA.h:1-18). */
typedef struct {
  /* observations */
  int ndata, n_curves;
  const int* l_points;     /* [L] */
  const int* in_rel;       /* [L], 1 = relative curve */
  const double* tim;       /* [N] days minus the epoch */
  const double* ee;        /* [3*N] asteroid->Earth unit vectors, component-major */
  const double* ee0;       /* [3*N] asteroid->Sun unit vectors, component-major */
  const double* brightness;/* [N] */
  const double* sig;       /* [N] curve mean brightness */
  const double* sigr2;     /* [N] 1/sig^2 */
  const double* weight;    /* [N] */
  /* shape basis */
  int nfac, lmax, mmax, ncoef;
  const double* normal;    /* [3*F] unit facet normals, component-major */
  const double* d_area;    /* [F] */
  const double* pleg;      /* [(mmax+1)*(lmax+1)*F] P_l^m at facet normals */
  const double* fc;        /* [(mmax+1)*F] cos(m phi) */
  const double* fs;        /* [(mmax+1)*F] sin(m phi) */
  const double* d_sphere;  /* [C*F] basis row k-1 = d_sphere[k] */
  const double* cg_first;  /* [C] initial ellipsoid coefficients */
  const int* ia;           /* [ma] free (1) / fixed (0), ia[i] is parameter i+1 */
  /* model and LM controls */
  double par[3];           /* phase function A, D, k */
  double cl;               /* initial Lambert weight (not its log) */
  double phi0;             /* radians */
  double alamda_start, alamda_incr, alamda_incrr;
  double stop_condition;
  int n_iter_min;
  /* scan */
  double f_start, f_end, f_step;   /* 1/hours */
  double escl;                     /* initial size scale, for the first line */
  double beta_pole[PS_N_POLES], lambda_pole[PS_N_POLES]; /* degrees */
} ps_problem;

/* One output line of period_search_out (start_CUDA.cu:1222-1226). */
typedef struct {
  double period_h, rms, chi2, dark_or_conw, lambda_deg, beta_deg;
} ps_line;

/* Per-pair result of one LM fit (Start.cu:5056-5125), plus the number of
 * visible (point, facet) pairs the fit evaluated (its algorithmic work). */
typedef struct {
  int reported, niter;
  double dev, dark, period, lambda_deg, beta_deg, chisq;
  long long n_vis;
} ps_pair_result;

/* ---- host-only helpers (no device needed) ---- */

/* The parsed period_search_in with its geometry tables, as the reference's
 * main() builds them (period_search_BOINC.cpp:343-866, trifac.c, areanorm.c,
 * sphfunc.c, ellfit.c).  ps_input_view fills *view with pointers into the
 * object (valid until ps_input_free) plus conw (header value) and n_max
 * (number of trial frequencies). */
typedef struct ps_input ps_input;
int ps_input_read(const char* period_search_in, ps_input** out, char* err, int errlen);
int ps_input_view(const ps_input* in, ps_problem* view, double* conw, int* n_max);
void ps_input_free(ps_input* in);

/* Best pole of one frequency (start_CUDA.cu:1200-1210): the earliest
 * reported pole whose dev is not NaN and not larger than the incumbent's
 * (a NaN incumbent is replaced).  r points to PS_N_POLES results; returns
 * the pole index 0..9 or -1 if none is reported. */
int ps_select_best_pole(const ps_pair_result* r);

/* Library / context */
int ps_open(int device, ps_ctx** out);
const char* ps_last_error(const ps_ctx* ctx);
int ps_close(ps_ctx* ctx);

/* CUDAPrepare + the table uploads of CUDAPrecalc/CUDAStart: copies the
 * problem (deep copy) to the device.  Replaces any previous problem. */
int ps_prepare(ps_ctx* ctx, const ps_problem* prob);

/* Parse period_search_in and build the geometry tables exactly as the
 * reference's main() does (period_search_BOINC.cpp:343-866), then
 * ps_prepare() the result. */
int ps_prepare_file(ps_ctx* ctx, const char* period_search_in);

/* Query derived sizes after ps_prepare*: ints[0..7] = ndata, n_curves, nfac,
 * ncoef, ma, mfit, n_max, n_iter_max; dbls[0..3] = f_start, f_step, escl,
 * conw (conw only after ps_prepare_file, else 0). */
int ps_info(const ps_ctx* ctx, int* ints, double* dbls);

/* One round of the convexity-weight search (CUDAPrecalc): fits the first 10
 * trial frequencies x 10 poles with the three convexity rows weighted by
 * *conw_r, and sets *new_conw = 1 if the mean dark-facet percentage of the
 * best poles is below 1, else doubles *conw_r. */
int ps_precalc(ps_ctx* ctx, double* conw_r, int* new_conw);

/* The whole convexity-weight loop of main() (period_search_BOINC.cpp:
 * 647-771): starts from conw/escl^2 (conw from the input file).  Needs
 * ps_prepare_file.  Writes the final conwR. */
int ps_conw_search(ps_ctx* ctx, double* conw_r);

/* CUDAStart: scans frequencies n_start .. n_start+n_count-1 (1-based) with
 * the final conw_r; writes one line per frequency into out[n_count].  The
 * first line of the whole scan (n == 1) carries conw_r*escl^2 in the dark
 * column, as the reference does. */
int ps_start(ps_ctx* ctx, int n_start, int n_count, double conw_r, ps_line* out);

/* Fits an explicit list of (frequency index n, pole m), both 1-based.
 * conw_r weights the convexity rows; conw_r_dev is the value the dev
 * computation subtracts with (conw_r for the scan, 0 for the conw search,
 * start_CUDA.cu:476).  Results into out[npairs]. */
int ps_fit_pairs(ps_ctx* ctx, int npairs, const int* n_list, const int* m_list,
                 double conw_r, double conw_r_dev, ps_pair_result* out);

/* Statistics of the last ps_fit_pairs / ps_start / ps_precalc call:
 * device time of the fit kernel in milliseconds (HIP events on the library's
 * stream), fit-kernel launches, pairs fitted, and the summed visible
 * (point, facet) pairs, LM iterations and normal-equation evaluations of
 * those pairs.  Any pointer may be NULL. */
int ps_last_stats(const ps_ctx* ctx, double* kernel_ms, int* launches, long long* pairs,
                  long long* vis_pairs, long long* iterations, long long* evaluations);

/* CUDAUnprepare: frees the device copy of the problem. */
int ps_unprepare(ps_ctx* ctx);

/* ---- batches of asteroids (no counterpart in the reference, whose main()
 * scans one asteroid per process; for callers that scan many asteroids
 * with one persistent launch) ----
 *
 * A batch holds n problems that share the shape basis: the same nfac, lmax,
 * mmax, free-parameter mask and facet tables (every period_search_in with
 * the same nrows / lmax / mmax / fixed-parameter flags produces them), each
 * with its own observations, scan range and controls.  Work items of all
 * problems share one persistent launch, queued largest problem first; a
 * pair's result does not depend on what else is in the launch.  The
 * single-problem calls above address problem 0. */
int ps_prepare_batch(ps_ctx* ctx, int n, const ps_problem* problems);
int ps_prepare_files(ps_ctx* ctx, int n, const char* const* period_search_in_paths);
int ps_n_problems(const ps_ctx* ctx);
/* ps_info of problem i. */
int ps_problem_info(const ps_ctx* ctx, int i, int* ints, double* dbls);
/* Pair (n_list[k], m_list[k]) of problem prob[k], convexity rows weighted
 * by conw_r[k], dev computed with conw_r_dev[k] (see ps_fit_pairs). */
int ps_fit_items(ps_ctx* ctx, int n_items, const int* prob, const int* n_list, const int* m_list,
                 const double* conw_r, const double* conw_r_dev, ps_pair_result* out);
/* The convexity-weight loop of every problem (needs ps_prepare_files); the
 * rounds of the problems still searching share one launch. */
int ps_conw_search_batch(ps_ctx* ctx, double* conw_r /* [n] */);
/* Scans frequencies n_start[i] .. n_start[i]+n_count[i]-1 of every problem
 * i with conw_r[i]; the lines of problem i follow those of problems < i. */
int ps_start_batch(ps_ctx* ctx, const int* n_start, const int* n_count, const double* conw_r, ps_line* out);

/* ---- allocation-free fits (no counterpart in the reference; for callers
 * that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc) ----
 *
 * ps_fit_items on caller-owned DEVICE buffers, allocating nothing: item k is
 * pair (d_n[k], d_m[k]) (1-based) of problem d_prob[k] (d_prob may be NULL
 * when one problem is prepared), with d_cw[2k] the SQUARE of the convexity
 * rows' weight (conw_r[k]^2 for ps_fit_items' conw_r, the factor the rows'
 * squared residuals carry) and d_cw[2k+1] the conwR of Iter2's dev
 * (ps_fit_items' conw_r_dev[k], not squared).  Results go to
 * d_out[k*10 ..]: dev, dark, period [days], lambda, beta, chisq, LM
 * iterations, reported (1), visible (point, facet) pairs, unused.
 *
 * The work counter and the workgroups' scratch slabs live in a file-scope
 * device array of 256 MiB (each GPU the library runs on holds one for the
 * life of the process): the launch uses at most as many workgroups as slabs
 * fit, and a problem whose single slab exceeds it is rejected with
 * PS_E_ARG.  Every call on a GPU shares that GPU's array and leaves its
 * contents behind (the counter is cleared by a kernel at each call): calls
 * that overlap on one GPU race.  Synchronous on the context's stream;
 * results are bit-identical to ps_fit_items. */
int ps_fit_items_static(ps_ctx* ctx, int n_items, const int* d_prob, const int* d_n, const int* d_m,
                        const double* d_cw, double* d_out);

/* ---- learned warm start (no counterpart in the reference; for callers
 * that start each LM fit from a predicted shape instead of the ellipsoid of
 * ellfit.c) ----
 *
 * Features of work item (problem p, frequency n, pole m), PS_WARM_FEATURES
 * floats:
 *   [0, 512)   up to 8 lightcurves (the first curves of the problem spanning
 *              at most 2 days with at least 16 points) folded at the trial
 *              period (phase = frac(t[days] * 24 / P[h])) into 64 bins: mean
 *              of brightness / curve mean - 1 per bin, 0 for empty bins;
 *   512..514   unit vector of the initial pole m (ecliptic);
 *   515..520   mean asteroid->Earth and asteroid->Sun unit vectors of those
 *              curves' points;
 *   521        their mean solar phase angle (rad);
 *   522        log(P[h] / 10);  523..527 zero. */
#define PS_WARM_CURVES 8
#define PS_WARM_BINS 64
#define PS_WARM_FEATURES 528
/* The reference's initial parameter vector of each item (Start.cu:
 * 4346-4418), a0[k*ma ...], and the free-parameter mask (ia[i] = parameter
 * i+1 is fitted) shared by the problems of the context. */
int ps_initial_params(const ps_ctx* ctx, int n_items, const int* prob, const int* n_list, const int* m_list,
                      double* a0);
int ps_free_params(const ps_ctx* ctx, int* ia /* [ma] */);
int ps_warm_n_features(void);
int ps_warm_features(const ps_ctx* ctx, int n_items, const int* prob, const int* n_list, const int* m_list,
                     float* out /* [n_items][PS_WARM_FEATURES] */);
/* Learned frequency screen (extension; the reference fits every trial
 * frequency, start_CUDA.cu:792-1245).  Fold features of the trial
 * frequencies n_first .. n_first + n_count - 1 (1-based) of problem `prob`,
 * PS_SCREEN_FEATURES floats per frequency: every curve of at least 8 points,
 * made relative (b / mean - 1), is folded at the trial frequency through
 * its harmonic powers P_k = |sum y exp(2 pi i k nu t)|^2 / sum y^2, k = 1..4,
 * summarised over the short (spanning at most 2 days) and the long curves,
 * with the log trial period and the problem's curve and point counts.  The
 * features do not depend on the pole.  Host code (up to 16 threads). */
#define PS_SCREEN_HARMONICS 4
#define PS_SCREEN_FEATURES 32
int ps_screen_n_features(void);
int ps_screen_features(const ps_ctx* ctx, int prob, int n_first, int n_count,
                       float* out /* [n_count][PS_SCREEN_FEATURES] */);
/* ps_fit_items starting from caller-given parameters: a0[k*ma + i] for the
 * shape coefficients (i < ncoef) and the phase-function / Lambert
 * parameters (i = ncoef+3 .. ncoef+6); the pole, rotation rate and the
 * Lommel-Seeliger weight are set from (n, m) as in the reference
 * (Start.cu:4346-4418).  a0 may be NULL (the reference's initialisation).
 * n_iter > 0 runs exactly n_iter LM iterations (the reference's
 * stopCondition > 1 mode) instead of each problem's own stop condition.
 * a_final (may be NULL) receives every item's final parameter vector. */
int ps_fit_items_from(ps_ctx* ctx, int n_items, const int* prob, const int* n_list, const int* m_list,
                      const double* conw_r, const double* conw_r_dev, const double* a0, int n_iter,
                      ps_pair_result* out, double* a_final);

/* The warm-start network: n_layers hidden layers of n_hidden GELU units
 * and a linear output of n_out, trained with BF16 operands and FP32
 * accumulation (MFMA) and Adam, evaluated in FP32.  Parameters are FP32 in
 * the order W0 [n_in][n_hidden], b0, W1, b1, ..., W_out [n_hidden][n_out],
 * b_out (row-major).  Inputs and targets are used as given (standardise
 * them first). */
typedef struct ps_mlp ps_mlp;
int ps_mlp_create(ps_ctx* ctx, int n_in, int n_hidden, int n_layers, int n_out, unsigned long long seed,
                  ps_mlp** out);
long long ps_mlp_n_params(const ps_mlp* net);
int ps_mlp_get_params(const ps_mlp* net, float* params);
int ps_mlp_set_params(ps_mlp* net, const float* params);
/* y[rows][n_out] = net(x[rows][n_in]); bf16 = 1 evaluates with the training
 * precision (BF16 operands), 0 in FP32. */
int ps_mlp_forward(ps_mlp* net, int rows, const float* x, float* y, int bf16);
/* Mean squared error over rows x n_out and its gradient with respect to
 * every parameter (training precision), grads in ps_mlp_get_params order. */
int ps_mlp_loss_grad(ps_mlp* net, int rows, const float* x, const float* t, float* loss, float* grads);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over `epochs` passes of seeded shuffled
 * minibatches; loss[e] = mean training loss of epoch e (may be NULL). */
int ps_mlp_train(ps_mlp* net, int n, const float* x, const float* t, int epochs, int batch, float lr,
                 unsigned long long seed, float* loss);
int ps_mlp_free(ps_mlp* net);

/* ---- launch configuration (no counterpart in the reference, whose kernel
 * launch shapes are compile-time constants; for callers tuning throughput
 * on their own workloads) ----
 *
 * One LM fit runs in one workgroup of waves_per_pair wavefronts (1, 2, 4
 * or 8); wg_per_cu resident workgroups share a CU (0 = as many as fit).
 * Every configuration gives bit-identical results: it changes only which
 * wavefront computes what, never the order of an operation chain.  A wave
 * count whose workgroup cannot be resident for the prepared problem falls
 * back to 4.  Defaults: 4 and 0 (the environment variables
 * PS_WAVES_PER_PAIR and PS_WG_PER_CU override them at ps_open). */
int ps_set_launch(ps_ctx* ctx, int waves_per_pair, int wg_per_cu);
/* The configuration the next launch will use (after ps_prepare*). */
int ps_get_launch(const ps_ctx* ctx, int* waves_per_pair, int* wg_per_cu);
/* Points per tile claim of the facet pass: a wave takes 16, 64, 128 or
 * 256 of the evaluation's view-sorted points at a time (16-point MFMA
 * tiles, one work-counter claim per group).  Results do not depend on it.
 * Default 16 (PS_TILE_POINTS overrides it at ps_open). */
int ps_set_tile_points(ps_ctx* ctx, int points);
int ps_get_tile_points(const ps_ctx* ctx, int* points);
/* The wave counts the prepared problem can run with and each one's
 * maximum resident workgroups per CU (up to cap entries); returns how many
 * there are. */
int ps_launch_options(const ps_ctx* ctx, int* waves, int* max_wg_per_cu, int cap);
/* HIP graphs (no counterpart in the reference, which launches each kernel
 * from the host; for callers whose scan steps are short enough that launch
 * overhead shows): with on != 0, every fit launch of this context (the
 * work-counter reset and the persistent LM kernel) is captured once into a
 * HIP graph and replayed while its arguments, grid and variant stay the
 * same; item and result copies stay outside the graph.  Results are
 * bit-identical with graphs on and off.  Default off. */
int ps_set_graphs(ps_ctx* ctx, int on);
/* Graph captures and replays so far on this context (either may be NULL). */
int ps_graph_stats(const ps_ctx* ctx, int* captures, int* replays);
/* Times the configurations on the given work items (as ps_fit_items with
 * conw_r_dev = conw_r): every wave count x workgroups per CU at 16 points
 * per claim, then 64, 128 and 256 points per claim at the best of those.
 * Keeps the fastest for later calls on this context (ps_get_launch,
 * ps_get_tile_points) and reports it with its fit-kernel time in
 * milliseconds. */
int ps_autotune(ps_ctx* ctx, int n_items, const int* prob, const int* n_list, const int* m_list,
                const double* conw_r, int* waves_per_pair, int* wg_per_cu, double* best_ms);

/* ---- kernel-level entry points used by the parity tests ---- */

/* One normal-equation evaluation (the reference's mrqcof, Start.cu:
 * 4464-4498 + 2262-2828 + 4791-4807) at parameters a[ma] (a[i] = parameter
 * i+1): alpha[mfit*mfit] (symmetric), beta[mfit], *chisq, area[nfac] (facet
 * areas of a), jrows[(ndata+3)*(ma+1)] raw Jacobian rows (column 0 = model
 * brightness, column c = d/da_c; column 1 is not computed).  Any output
 * pointer may be NULL. */
int ps_eval(ps_ctx* ctx, const double* a, double conw_r2_rows, double* alpha,
            double* beta, double* chisq, double* area, double* jrows);

/* The damped Gauss-Jordan solve of mrqmin step 1 (Start.cu:2838-3021) on
 * the device: x = (alpha with diagonal*(1+lambda))^-1 beta, by the fit
 * kernels' solver of the wave count ps_set_launch configured (4 by
 * default).  Returns the reference error code (0, 1, 2) in *err. */
int ps_gaussj(ps_ctx* ctx, int mfit, const double* alpha, const double* beta,
              double lambda, double* x, int* err);

/* Transcendental functions of the kernels evaluated on the device:
 * out[3*i .. 3*i+2] = exp(x[i]), exp2(x[i]), sin(x[i]); out2[i] = cos(x[i]). */
int ps_math_probe(ps_ctx* ctx, int n, const double* x, double* out, double* out2);

#ifdef __cplusplus
}
#endif
#endif
