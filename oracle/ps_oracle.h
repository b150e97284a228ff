/* CPU FP64 oracle of the PeriodSearch period scan -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or, for bench.py, as the timed
 * CPU baseline).  The product path (ps_claude_opt_amd) never calls it.
 *
 * The C API below is what the Python tests bind with ctypes.
 */
#ifndef PS_ORACLE_H
#define PS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque parsed + set-up problem (input file, geometry tables). */
typedef struct pso_problem pso_problem;

/* Work-decomposition parameters of the GPU kernel (wavefronts per pair,
 * points per tile).  The oracle's summation order does not depend on them;
 * they are kept so callers can state the configuration they compare. */
typedef struct {
  int n_waves;
  int tile;
} pso_order;

/* Per-pair result: the reference's FinishPole outputs plus counters. */
typedef struct {
  int reported;      /* isReported */
  int niter;         /* LM iterations run */
  double dev;        /* dev_best */
  double dark;       /* dark facet %, NaN -> 1 */
  double period;     /* days */
  double lambda_deg; /* pole longitude */
  double beta_deg;   /* pole latitude */
  double chisq;      /* final Chisq */
  long long n_evals; /* mrqcof evaluations */
  long long n_vis;   /* visible (point, facet) pairs summed over evaluations */
} pso_pair_result;

pso_problem* pso_load(const char* path, char* err, int errlen);
void pso_free(pso_problem* p);

/* ints: [0]=ndata(data points) [1]=l_curves(data) [2]=nfac [3]=ncoef [4]=ma
 *       [5]=mfit [6]=lastma [7]=lastone [8]=n_max [9]=lmax [10]=mmax
 *       [11]=nrows [12]=n_iter_max [13]=n_iter_min [14]=onlyrel
 * dbls: [0]=f_start [1]=f_end [2]=f_step [3]=escl [4]=conw [5]=jd0
 *       [6]=phi0(rad) [7]=alamda_start [8]=alamda_incr [9]=alamda_incrr
 *       [10]=iter_diff_max [11]=cl [12]=jdMin [13]=jdMax */
void pso_info(const pso_problem* p, int* ints, double* dbls);

/* Named tables (row-major, 0-based copies):
 *  "normal" [3*F], "d_area" [F], "d_sphere" [C*F] (k=1..C rows),
 *  "cg_first" [C], "tim" [N], "ee" [3*N], "ee0" [3*N], "alpha" [N]
 *  (solar phase per point), "brightness" [N], "sig" [N], "sigr2" [N],
 *  "weight" [N], "ia" [ma] (as doubles), "l_points" [L], "in_rel" [L],
 *  "pleg" [(mmax+1)(lmax+1)F], "par" [5] = phase A, D, k, log(cl), 1.
 * Returns number of elements, or -1. Pass out=NULL to query the size. */
int pso_table(const pso_problem* p, const char* name, double* out, int cap);

/* One mrqcof evaluation at parameter vector a[0..ma-1] (a[i] == param i+1).
 * conw_r2_rows = sigr2 of the three convexity rows (conwR^2).
 * Outputs alpha[mfit*mfit] (symmetric), beta[mfit], *chisq, and optionally
 * area[F] (curv result) and jrows[(N+3)*(ma+1)] (raw, un-renormalised
 * Jacobian rows; column 0 = ymod, column c = dY/da_c). */
int pso_mrqcof(const pso_problem* p, const double* a, double conw_r2_rows,
               pso_order order, double* alpha, double* beta, double* chisq,
               double* area, double* jrows);

/* Damped Gauss-Jordan of the reference's mrqmin step 1 on an mfit x mfit
 * system: solves (alpha with diag*(1+lambda)) x = beta in place into x.
 * Returns the reference error code (0 ok, 1 ipiv>1, 2 zero pivot). */
int pso_gaussj(int mfit, const double* alpha, const double* beta,
               double lambda, double* x);

/* Custom transcendental functions shared in spirit with the GPU kernel. */
double pso_exp(double x);
double pso_exp2(double x);
void pso_sincos(double x, double* s, double* c);

/* Full LM fit of one (frequency index n (1-based), pole m (1-based)) pair.
 * conw_r2_rows as above; conw_r_dev = conw_r used in Iter2 (0 in precalc).
 * If traj != NULL it receives per-iteration (Alamda, Chisq, Ochisq, dev_new)
 * for up to traj_cap iterations. */
int pso_fit_pair(const pso_problem* p, int n, int m, double conw_r2_rows,
                 double conw_r_dev, pso_order order, pso_pair_result* res,
                 double* traj, int traj_cap);

/* The whole application: conw search + scan, output written as text to
 * out_path (period_search_out format).  threads <= 0 -> hardware count.
 * freq_limit > 0 scans only the first freq_limit frequencies.
 * On return *conw_r_final holds the chosen conwR. */
int pso_run(pso_problem* p, const char* out_path, int threads, pso_order order,
            int freq_limit, double* conw_r_final);

/* Conw search only (the reference's CUDAPrecalc loop).  Returns rounds run. */
int pso_conw_search(pso_problem* p, int threads, pso_order order,
                    double* conw_r_final);

/* Fit an explicit list of pairs (n_list[i], m_list[i]) with the given conwR
 * and device conw_r; results into res[]. threads <= 0 -> hardware count.
 * Returns wall seconds. */
double pso_fit_pairs(const pso_problem* p, int npairs, const int* n_list,
                     const int* m_list, double conw_r, double conw_r_dev,
                     pso_order order, int threads, pso_pair_result* res);

/* pso_fit_pairs starting every pair from a0[k*ma ...] (shape and phase-
 * function / Lambert parameters; pole, rotation rate and cls from (n, m)),
 * a0 may be NULL; n_iter > 0 runs exactly n_iter iterations; the final
 * parameter vectors into a_out (may be NULL). */
double pso_fit_pairs_from(const pso_problem* p, int npairs, const int* n_list, const int* m_list,
                          double conw_r, double conw_r_dev, int threads, const double* a0, int n_iter,
                          pso_pair_result* res, double* a_out);

/* pso_fit_pairs_from with every parameter taken from a0[k*ma ...] (the
 * pole and rotation rate too), so a fit can start at a known truth. */
double pso_fit_params(const pso_problem* p, int npairs, const double* a0, double conw_r, double conw_r_dev,
                      int threads, int n_iter, pso_pair_result* res, double* a_out);

/* Iter2's rms and FinishPole's dark-facet percentage at parameters a with
 * the given Chisq (the code path the fit itself uses): dev, chck = |sum A n|
 * and dark (any pointer may be NULL). */
void pso_finish(const pso_problem* p, const double* a, double chisq, double conw_r_dev, double* dev,
                double* chck, double* dark);

/* The convexity-weight decision of one search round on n_periods x 10 pair
 * results: updates *conw_r (doubled unless accepted) and sets *new_conw = 1
 * when the mean best-pole dark percentage is below 1. */
void pso_conw_decide(const pso_pair_result* res, int n_periods, double* conw_r, int* new_conw);

/* Test hook: override the LM iteration controls (n_iter_max, n_iter_min,
 * iter_diff_max) as if read from stopCondition / n_iter_min. */
void pso_set_stop(pso_problem* p, double stop_condition, int n_iter_min);

#ifdef __cplusplus
}
#endif
#endif
