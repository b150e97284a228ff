/* CPU FP64 oracle of the PeriodSearch period scan -- TEST INFRASTRUCTURE ONLY.
 *
 * A restatement, in plain sequential C++, of the algorithm of the reference's
 * live GPU path (IanSteveC/ps_claude_opt), used as the checker for the
 * MI355X kernels and as the timed CPU baseline.  Nothing in the product path
 * links or calls this file.
 *
 * Reference files followed (all paths relative to the reference root):
 *   input parsing + preprocessing ........ period_search_BOINC.cpp:343-585
 *   conw search / final setup ............ period_search_BOINC.cpp:647-866
 *   triangulation ........................ trifac.c:11-83
 *   facet areas and normals .............. areanorm.c:11-61
 *   Laplace-series basis ................. sphfunc.c:12-88
 *   initial ellipsoid fit ................ ellfit.c:12-88, ludcmp.c, lubksb.c
 *   pair initialisation .................. Start.cu:4346-4418
 *   loop predicate ....................... Start.cu:4421-4462
 *   curv (facet areas) ................... Start.cu:421-528
 *   blmatrix ............................. Start.cu:399-417
 *   point geometry ....................... Start.cu:2095-2195, 687-692
 *   brightness + Jacobian rows ........... Start.cu:2262-2535 (Appendix C)
 *   convexity pseudo-curve rows .......... Start.cu:2539-2612
 *   relative renormalisation + J^T W J ... Start.cu:2622-2828
 *   symmetrisation ....................... Start.cu:590-629
 *   damped Gauss-Jordan + atry ........... Start.cu:2838-3021
 *   accept / reject ...................... Start.cu:300-392, 4521-4539
 *   Iter2 (dev, iter_diff) ............... Start.cu:5129-5248
 *   FinishPole ........................... Start.cu:5056-5125
 *   best pole + output line .............. start_CUDA.cu:1200-1229
 *   conw decision ........................ start_CUDA.cu:400-789
 *
 * Summation order.  The reference's own order is an artefact of its 32-lane
 * warps (lane-strided partial sums + butterflies); the upstream CPU build
 * sums sequentially.  This oracle fixes ONE order, the one the MI355X kernel
 * uses, so CPU and GPU results can be compared bit for bit:
 *   - facet sums of a point: 4 partial sums over facets f = s (mod 4),
 *     combined as (P0 + P1) + (P2 + P3);
 *   - shape-Jacobian gather: one fused multiply-add chain over facets in
 *     ascending order (invisible facets add an exact zero);
 *   - points are grouped in tiles of 16 consecutive points of one curve
 *     (aligned at the curve start; the last tile of a curve is padded);
 *     tiles are numbered over all curves in file order, the three
 *     convexity rows form one more tile after the last data tile;
 *   - curve brightness sums: per tile four partial sums over the tile's
 *     points p = s (mod 4) in order (padding adds +0.0), combined as
 *     (G0 + G1) + (G2 + G3); the curve sum adds the tile sums sequentially;
 *   - J^T W J, J^T W r and chi^2: one augmented product (W Z)^T Z,
 *     Z = [cJ|r|z|y], split into G chains (G = 4 for rows of at most 64
 *     columns, else 2): chain g holds the tiles T = g (mod G), every element
 *     a fused multiply-add chain (weighted side times plain side) over its
 *     tiles' points in order; the chains are combined as
 *     (C0 + C1) + (C2 + C3) (G = 4) or C0 + C1 (G = 2);
 *   - per relative curve, the chain's rows z and y over its tiles of the
 *     curve (the same chains, cleared after each curve), the chains
 *     combined as the matrices (a chain without a tile of the curve gives
 *     0.0);
 *   - the renormalisation correction: one fused multiply-add chain per
 *     element of the combined matrix over the curves in order, two terms
 *     per curve (see mrqcof).
 * Transcendentals (exp, exp2, sin/cos of the LM parameters) use this file's
 * own Cody-Waite + Taylor implementations, which the kernel restates
 * identically; host-side setup uses the C library as the reference does.
 * FMA contraction is off (-ffp-contract=off); every fused multiply-add is an
 * explicit fma() call.
 */
#include "ps_oracle.h"

#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr double PI = 3.14159265358979323846;
constexpr double DEG2RAD = PI / 180.0;
constexpr double RAD2DEG = 180.0 / PI;
constexpr double TINY = 1e-8;          /* constants.h:19 */
constexpr int N_POLES = 10;            /* constants.h:20 */
constexpr int MAX_N_ITER = 60;         /* constants.h:12 */

/* ------------------------------------------------------------------ */
/* transcendental functions (Cody-Waite reduction + Taylor polynomials) */
/* ------------------------------------------------------------------ */

/* 1/n! for n = 2..13, correctly rounded */
const double EXP_C[14] = {
    1.0, 1.0, 0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
    0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
    0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22,
    0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33};
constexpr double INV_LN2 = 0x1.71547652b82fep+0;
constexpr double LN2_HI = 0x1.62e42fee00000p-1;
constexpr double LN2_LO = 0x1.a39ef35793c76p-33;
constexpr double LN2 = 0x1.62e42fefa39efp-1;

/* exp(r) for |r| <= ~0.35: Horner over the Taylor series to r^13 */
inline double exp_poly(double r) {
  double p = EXP_C[13];
  for (int n = 12; n >= 0; --n) p = std::fma(p, r, EXP_C[n]);
  return p;
}

double o_exp(double x) {
  if (std::isnan(x)) return x;
  if (x > 709.8) return INFINITY;
  if (x < -745.2) return 0.0;
  double k = std::rint(x * INV_LN2);
  double r = std::fma(-k, LN2_HI, x);
  r = std::fma(-k, LN2_LO, r);
  return std::ldexp(exp_poly(r), (int)k);
}

double o_exp2(double y) {
  if (std::isnan(y)) return y;
  if (y > 1024.0) return INFINITY;
  if (y < -1075.0) return 0.0;
  double k = std::rint(y);
  double r = y - k;
  return std::ldexp(exp_poly(r * LN2), (int)k);
}

/* (-1)^j / (2j+1)!  and (-1)^j / (2j)! */
const double SIN_C[10] = {1.0, -0x1.5555555555555p-3, 0x1.1111111111111p-7,
                          -0x1.a01a01a01a01ap-13, 0x1.71de3a556c734p-19,
                          -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33,
                          -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4ap-49,
                          -0x1.2f49b46814157p-57};
const double COS_C[10] = {1.0, -0x1.0000000000000p-1, 0x1.5555555555555p-5,
                          -0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-16,
                          -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29,
                          -0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-45,
                          -0x1.6827863b97d97p-53};
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
constexpr double PIO2_1 = 0x1.921fb54400000p+0;
constexpr double PIO2_2 = 0x1.0b4611a600000p-34;
constexpr double PIO2_3 = 0x1.3198a2e000000p-69;

void o_sincos(double x, double* s, double* c) {
  if (!std::isfinite(x)) { *s = NAN; *c = NAN; return; }
  double k = std::rint(x * TWO_OVER_PI);
  double r = std::fma(-k, PIO2_1, x);
  r = std::fma(-k, PIO2_2, r);
  r = std::fma(-k, PIO2_3, r);
  double z = r * r;
  double ps = SIN_C[9];
  for (int j = 8; j >= 1; --j) ps = std::fma(ps, z, SIN_C[j]);
  double sn = std::fma(r * z, ps, r);
  double pc = COS_C[9];
  for (int j = 8; j >= 1; --j) pc = std::fma(pc, z, COS_C[j]);
  double cs = std::fma(z, pc, 1.0);
  long long q = (long long)k & 3;
  switch (q) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
  }
}

inline double dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
  return std::fma(a2, b2, std::fma(a1, b1, a0 * b0));
}

/* ------------------------------------------------------------------ */
/* problem                                                              */
/* ------------------------------------------------------------------ */
struct Problem {
  /* header */
  double P_start = 0, period_step_coef = 0, P_end = 0;
  int ia_prd = 1;
  double jd00 = 0, phi0 = 0, conw = 0;
  int lmax = 0, mmax = 0, nrows = 0;
  double par[4] = {0, 0, 0, 0};
  int ia_par[4] = {0, 0, 0, 0};
  double cl = 0;
  int ia_cl = 0;
  double stop_condition = 0;
  int n_iter_min = 0;
  double alamda_incr = 0, alamda_incrr = 0, alamda_start = 0;
  /* data (data curves only; the convexity rows are appended per run) */
  int l_curves = 0;
  std::vector<int> l_points, in_rel;          /* 0-based curve index */
  std::vector<double> tim, bright, sig, sigr2, weight; /* 0-based point index */
  std::vector<double> ee[3], ee0[3];
  std::vector<double> alpha;                  /* solar phase per point */
  int ndata = 0, onlyrel = 1, max_lp = 0;
  double jdMin = 0, jdMax = 0, jd0 = 0, escl = 0;
  /* geometry */
  int nfac = 0, ncoef = 0;
  std::vector<double> nor[3], d_area;         /* [F] */
  std::vector<double> dsph;                   /* [C][F], row k-1 = d_sphere[k] */
  std::vector<double> pleg;                   /* [(mmax+1)][(lmax+1)][F] */
  std::vector<double> fc, fs;                 /* [(mmax+1)][F] */
  std::vector<double> cg_first;               /* [C] */
  /* fit controls */
  std::vector<int> ia;                        /* 1..ma */
  int ma = 0, mfit = 0, lastma = 0, lastone = 0, ncoef0 = 0;
  int n_iter_max = 0;
  double iter_diff_max = 0;
  double f_start = 0, f_end = 0, f_step = 0;
  int n_max = 0;
  double lcl = 0;                             /* log(cl) */
  double beta_pole[N_POLES + 1], lambda_pole[N_POLES + 1];
};

/* ---------- geometry setup (host side in the reference) ---------- */

/* trifac.c:11-83: vertex triplets of the standard octant triangulation */
void trifac(int nrows, std::vector<int>& ifp /* [(F+1)*4] 1-based */) {
  std::vector<std::vector<int>> nod(2 * nrows + 1, std::vector<int>(4 * nrows + 1, 0));
  int nnod = 1;
  nod[0][0] = nnod;
  for (int i = 1; i <= nrows; i++)
    for (int j = 0; j <= 4 * i - 1; j++) {
      nnod++;
      nod[i][j] = nnod;
      if (j == 0) nod[i][4 * i] = nnod;
    }
  for (int i = nrows - 1; i >= 1; i--)
    for (int j = 0; j <= 4 * i - 1; j++) {
      nnod++;
      nod[2 * nrows - i][j] = nnod;
      if (j == 0) nod[2 * nrows - i][4 * i] = nnod;
    }
  nod[2 * nrows][0] = nnod + 1;
  int ntri = 0;
  auto put = [&](int a, int b, int c) {
    ntri++;
    ifp[ntri * 4 + 1] = a;
    ifp[ntri * 4 + 2] = b;
    ifp[ntri * 4 + 3] = c;
  };
  for (int j1 = 1; j1 <= nrows; j1++)
    for (int j3 = 1; j3 <= 4; j3++) {
      int j0 = (j3 - 1) * j1;
      put(nod[j1 - 1][j0 - (j3 - 1)], nod[j1][j0], nod[j1][j0 + 1]);
      for (int j2 = j0 + 1; j2 <= j0 + j1 - 1; j2++) {
        put(nod[j1][j2], nod[j1 - 1][j2 - (j3 - 1)], nod[j1 - 1][j2 - 1 - (j3 - 1)]);
        put(nod[j1 - 1][j2 - (j3 - 1)], nod[j1][j2], nod[j1][j2 + 1]);
      }
    }
  for (int j1 = nrows + 1; j1 <= 2 * nrows; j1++)
    for (int j3 = 1; j3 <= 4; j3++) {
      int j0 = (j3 - 1) * (2 * nrows - j1);
      put(nod[j1][j0], nod[j1 - 1][j0 + 1 + (j3 - 1)], nod[j1 - 1][j0 + (j3 - 1)]);
      for (int j2 = j0 + 1; j2 <= j0 + (2 * nrows - j1); j2++) {
        put(nod[j1][j2], nod[j1 - 1][j2 + (j3 - 1)], nod[j1][j2 - 1]);
        put(nod[j1][j2], nod[j1 - 1][j2 + 1 + (j3 - 1)], nod[j1 - 1][j2 + (j3 - 1)]);
      }
    }
}

/* NR ludcmp/lubksb (ludcmp.c, lubksb.c), 1-based, row-major [n+1][n+1] */
int ludcmp(std::vector<double>& a, int n, std::vector<int>& indx) {
  const double tiny = 1.0e-20;
  int imax = -999;
  std::vector<double> v(n + 1);
  auto A = [&](int i, int j) -> double& { return a[i * (n + 1) + j]; };
  for (int i = 1; i <= n; i++) {
    double big = 0.0;
    for (int j = 1; j <= n; j++) {
      double temp = std::fabs(A(i, j));
      if (temp > big) big = temp;
    }
    if (big == 0.0) return 4;
    v[i] = 1.0 / big;
  }
  for (int j = 1; j <= n; j++) {
    for (int i = 1; i < j; i++) {
      double sum = A(i, j);
      for (int k = 1; k < i; k++) sum -= A(i, k) * A(k, j);
      A(i, j) = sum;
    }
    double big = 0.0;
    for (int i = j; i <= n; i++) {
      double sum = A(i, j);
      for (int k = 1; k < j; k++) sum -= A(i, k) * A(k, j);
      A(i, j) = sum;
      double dum = v[i] * std::fabs(sum);
      if (dum >= big) { big = dum; imax = i; }
    }
    if (j != imax) {
      for (int k = 1; k <= n; k++) std::swap(A(imax, k), A(j, k));
      v[imax] = v[j];
    }
    indx[j] = imax;
    if (A(j, j) == 0.0) A(j, j) = tiny;
    if (j != n) {
      double dum = 1.0 / A(j, j);
      for (int i = j + 1; i <= n; i++) A(i, j) *= dum;
    }
  }
  return 0;
}

void lubksb(const std::vector<double>& a, int n, const std::vector<int>& indx, std::vector<double>& b) {
  int ii = 0;
  auto A = [&](int i, int j) { return a[i * (n + 1) + j]; };
  for (int i = 1; i <= n; i++) {
    int ip = indx[i];
    double sum = b[ip];
    b[ip] = b[i];
    if (ii)
      for (int j = ii; j <= i - 1; j++) sum -= A(i, j) * b[j];
    else if (sum)
      ii = i;
    b[i] = sum;
  }
  for (int i = n; i >= 1; i--) {
    double sum = b[i];
    for (int j = i + 1; j <= n; j++) sum -= A(i, j) * b[j];
    b[i] = sum / A(i, i);
  }
}

/* period_search_BOINC.cpp:781-857 + trifac/areanorm/sphfunc/ellfit */
int setup_geometry(Problem& P, double a, double b, double c) {
  const int nrows = P.nrows;
  P.ncoef = 0;
  for (int m = 0; m <= P.mmax; m++)
    for (int l = m; l <= P.lmax; l++) {
      P.ncoef++;
      if (m != 0) P.ncoef++;
    }
  /* vertex directions */
  int nvmax = 4 * nrows * nrows + 8;
  std::vector<double> t(nvmax + 1), f(nvmax + 1);
  double dth = PI / (2 * nrows);
  int k = 1;
  t[1] = 0; f[1] = 0;
  for (int i = 1; i <= nrows; i++) {
    double dph = PI / (2 * i);
    for (int j = 0; j <= 4 * i - 1; j++) { k++; t[k] = i * dth; f[k] = j * dph; }
  }
  for (int i = nrows - 1; i >= 1; i--) {
    double dph = PI / (2 * i);
    for (int j = 0; j <= 4 * i - 1; j++) { k++; t[k] = PI - i * dth; f[k] = j * dph; }
  }
  int ndir = k + 1;
  t[ndir] = PI; f[ndir] = 0;
  const int nf = 8 * nrows * nrows;
  P.nfac = nf;
  std::vector<int> ifp((nf + 1) * 4, 0);
  trifac(nrows, ifp);
  /* areanorm.c:11-61 */
  std::vector<double> x(ndir + 1), y(ndir + 1), z(ndir + 1), at(nf + 1), af(nf + 1);
  for (int i = 1; i <= ndir; i++) {
    double st = std::sin(t[i]);
    x[i] = st * std::cos(f[i]);
    y[i] = st * std::sin(f[i]);
    z[i] = std::cos(t[i]);
  }
  for (int d = 0; d < 3; d++) P.nor[d].assign(nf, 0.0);
  P.d_area.assign(nf, 0.0);
  for (int i = 1; i <= nf; i++) {
    double vx[4], vy[4], vz[4];
    int i1 = ifp[i * 4 + 1];
    for (int j = 2; j <= 3; j++) {
      int ij = ifp[i * 4 + j];
      vx[j] = x[ij] - x[i1];
      vy[j] = y[ij] - y[i1];
      vz[j] = z[ij] - z[i1];
    }
    double c1 = vy[2] * vz[3] - vy[3] * vz[2];
    double c2 = vz[2] * vx[3] - vz[3] * vx[2];
    double c3 = vx[2] * vy[3] - vx[3] * vy[2];
    double clen2 = c1 * c1 + c2 * c2 + c3 * c3;
    double clen = std::sqrt(clen2);
    P.nor[0][i - 1] = c1 / clen;
    P.nor[1][i - 1] = c2 / clen;
    P.nor[2][i - 1] = c3 / clen;
    at[i] = std::acos(P.nor[2][i - 1]);
    af[i] = std::atan2(P.nor[1][i - 1], P.nor[0][i - 1]);
    P.d_area[i - 1] = 0.5 * clen;
  }
  /* sphfunc.c:12-88 */
  const int L = P.lmax, M = P.mmax;
  std::vector<double> ts((nf + 1) * (L + 1)), tc((nf + 1) * (L + 1));
  P.fc.assign((M + 1) * nf, 0.0);
  P.fs.assign((M + 1) * nf, 0.0);
  for (int i = 1; i <= nf; i++) {
    ts[i * (L + 1) + 0] = 1;
    tc[i * (L + 1) + 0] = 1;
    ts[i * (L + 1) + 1] = std::sin(at[i]);
    tc[i * (L + 1) + 1] = std::cos(at[i]);
    for (int j = 2; j <= L; j++) {
      ts[i * (L + 1) + j] = ts[i * (L + 1) + 1] * ts[i * (L + 1) + j - 1];
      tc[i * (L + 1) + j] = tc[i * (L + 1) + 1] * tc[i * (L + 1) + j - 1];
    }
    P.fs[0 * nf + i - 1] = 0;
    P.fc[0 * nf + i - 1] = 1;
    for (int j = 1; j <= M; j++) {
      P.fs[j * nf + i - 1] = std::sin(j * af[i]);
      P.fc[j * nf + i - 1] = std::cos(j * af[i]);
    }
  }
  const int LM1 = L + 1;
  std::vector<double> aleg(LM1 * LM1 * LM1, 0.0); /* aleg[n][l][m] */
  auto AL = [&](int n, int l, int m) -> double& { return aleg[(n * LM1 + l) * LM1 + m]; };
  AL(0, 0, 0) = 1;
  if (L >= 1) AL(1, 1, 0) = 1;
  for (int l = 1; l <= L; l++) AL(0, l, l) = AL(0, l - 1, l - 1) * (2 * l - 1);
  for (int m = 0; m <= M; m++)
    for (int l = m + 1; l <= L; l++) {
      int ibot;
      if ((2 * ((l - m) / 2)) == (l - m)) {
        AL(0, l, m) = -(l + m - 1) * AL(0, l - 2, m) / (1 * (l - m));
        ibot = 2;
      } else
        ibot = 1;
      if (l != 1)
        for (int n = ibot; n <= l - m; n = n + 2)
          AL(n, l, m) = ((2 * l - 1) * AL(n - 1, l - 1, m) - (l + m - 1) * AL(n, l - 2, m)) / (1 * (l - m));
    }
  P.pleg.assign((M + 1) * (L + 1) * nf, 0.0);
  P.dsph.assign(P.ncoef * nf, 0.0);
  for (int i = 1; i <= nf; i++) {
    int kk = 0;
    for (int m = 0; m <= M; m++)
      for (int l = m; l <= L; l++) {
        double pl = 0;
        int ibot = ((2 * ((l - m) / 2)) == (l - m)) ? 0 : 1;
        for (int n = ibot; n <= l - m; n = n + 2)
          pl = pl + AL(n, l, m) * tc[i * (L + 1) + n] * ts[i * (L + 1) + m];
        P.pleg[(m * (L + 1) + l) * nf + i - 1] = pl;
        kk++;
        P.dsph[(kk - 1) * nf + i - 1] = P.fc[m * nf + i - 1] * pl;
        if (m != 0) {
          kk++;
          P.dsph[(kk - 1) * nf + i - 1] = P.fs[m * nf + i - 1] * pl;
        }
      }
  }
  /* ellfit.c:12-88 */
  const int nc = P.ncoef;
  std::vector<double> er(nf + 1), fmat((nf + 1) * (nc + 1), 0.0), fitmat((nc + 1) * (nc + 1), 0.0), fitvec(nc + 1, 0.0);
  for (int i = 1; i <= nf; i++) {
    double st = std::sin(at[i]);
    double q1 = a * st * std::cos(af[i]);
    double q2 = b * st * std::sin(af[i]);
    double q3 = c * std::cos(at[i]);
    double sum = q1 * q1 + q2 * q2 + q3 * q3;
    er[i] = 2 * (std::log(a * b * c) - std::log(sum));
  }
  for (int i = 1; i <= nf; i++) {
    int n = 0;
    for (int m = 0; m <= M; m++)
      for (int l = m; l <= L; l++) {
        n++;
        double pl = P.pleg[(m * (L + 1) + l) * nf + i - 1];
        if (m != 0) {
          fmat[i * (nc + 1) + n] = pl * std::cos(m * af[i]);
          n++;
          fmat[i * (nc + 1) + n] = pl * std::sin(m * af[i]);
        } else
          fmat[i * (nc + 1) + n] = pl;
      }
  }
  for (int i = 1; i <= nc; i++) {
    for (int j = 1; j <= nc; j++) {
      double s = 0;
      for (int kk = 1; kk <= nf; kk++) s = s + fmat[kk * (nc + 1) + i] * fmat[kk * (nc + 1) + j];
      fitmat[i * (nc + 1) + j] = s;
    }
    double s = 0;
    for (int j = 1; j <= nf; j++) s = s + fmat[j * (nc + 1) + i] * er[j];
    fitvec[i] = s;
  }
  std::vector<int> indx(nc + 1);
  if (ludcmp(fitmat, nc, indx)) return 4;
  lubksb(fitmat, nc, indx, fitvec);
  P.cg_first.assign(nc, 0.0);
  for (int i = 1; i <= nc; i++) P.cg_first[i - 1] = fitvec[i];
  return 0;
}

void set_stop(Problem& P, double stop_condition, int n_iter_min) {
  P.stop_condition = stop_condition;
  P.n_iter_min = n_iter_min;
  /* start_CUDA.cu:855-869 */
  P.n_iter_max = 0;
  P.iter_diff_max = -1;
  if (stop_condition > 1) {
    P.n_iter_max = (int)stop_condition;
    P.iter_diff_max = 0;
    P.n_iter_min = 0;
  }
  if (stop_condition < 1) {
    P.n_iter_max = MAX_N_ITER;
    P.iter_diff_max = stop_condition;
  }
}

/* period_search_BOINC.cpp:343-585 (parse + preprocessing) and :773-862 */
int parse_and_setup(Problem& P, const char* path, std::string& err) {
  FILE* in = std::fopen(path, "r");
  if (!in) { err = "cannot open input"; return 1; }
  char line[1024];
  auto eol = [&]() { if (!std::fgets(line, sizeof line, in)) line[0] = 0; };
  int ok = 1;
  ok &= std::fscanf(in, "%lf %lf %lf %d", &P.P_start, &P.period_step_coef, &P.P_end, &P.ia_prd) == 4; eol();
  ok &= std::fscanf(in, "%lf", &P.jd00) == 1; eol();
  ok &= std::fscanf(in, "%lf", &P.phi0) == 1; eol();
  ok &= std::fscanf(in, "%lf", &P.conw) == 1; eol();
  ok &= std::fscanf(in, "%d %d", &P.lmax, &P.mmax) == 2; eol();
  ok &= std::fscanf(in, "%d", &P.nrows) == 1; eol();
  for (int i = 1; i <= 3; i++) { ok &= std::fscanf(in, "%lf %d", &P.par[i], &P.ia_par[i]) == 2; eol(); }
  ok &= std::fscanf(in, "%lf %d", &P.cl, &P.ia_cl) == 2; eol();
  ok &= std::fscanf(in, "%lf", &P.stop_condition) == 1; eol();
  ok &= std::fscanf(in, "%d", &P.n_iter_min) == 1; eol();
  ok &= std::fscanf(in, "%lf", &P.alamda_incr) == 1; eol();
  P.alamda_incrr = (P.alamda_incr != 0.0) ? 1.0 / P.alamda_incr : DBL_MAX;
  ok &= std::fscanf(in, "%lf", &P.alamda_start) == 1; eol();
  ok &= std::fscanf(in, "%d", &P.l_curves) == 1;
  if (!ok || P.l_curves <= 0) { std::fclose(in); err = "bad header"; return 2; }
  if (P.lmax < 0 || P.mmax < 0 || P.mmax > P.lmax || P.lmax > 15 || P.nrows < 1) {
    std::fclose(in); err = "bad lmax/mmax/nrows"; return 2;
  }
  P.l_points.assign(P.l_curves, 0);
  P.in_rel.assign(P.l_curves, 0);
  std::vector<double> al(P.l_curves);
  double al0 = PI, al0Abs = PI;
  int ial0 = -1, ial0_abs = -1;
  P.jdMin = 1e20; P.jdMax = -1e40;
  P.onlyrel = 1; P.max_lp = 0;
  const double a0 = 1.05, b0 = 1.00, c0 = 0.95;
  for (int i = 0; i < P.l_curves; i++) {
    double ave = 0;
    int lp, itemp;
    if (std::fscanf(in, "%d %d", &lp, &itemp) != 2 || lp <= 0) { std::fclose(in); err = "bad curve header"; return 2; }
    eol();
    P.l_points[i] = lp;
    P.in_rel[i] = 1 - itemp;
    if (P.in_rel[i] == 0) P.onlyrel = 0;
    if (lp > P.max_lp) P.max_lp = lp;
    for (int j = 1; j <= lp; j++) {
      double t, br, e0[4], e[4];
      if (std::fscanf(in, "%lf %lf", &t, &br) != 2 ||
          std::fscanf(in, "%lf %lf %lf", &e0[1], &e0[2], &e0[3]) != 3 ||
          std::fscanf(in, "%lf %lf %lf", &e[1], &e[2], &e[3]) != 3) {
        std::fclose(in); err = "bad data line"; return 2;
      }
      P.ndata++;
      P.tim.push_back(t);
      P.bright.push_back(br);
      if (t < P.jdMin) P.jdMin = t;
      if (t > P.jdMax) P.jdMax = t;
      double e0Len = std::sqrt(e0[1] * e0[1] + e0[2] * e0[2] + e0[3] * e0[3]);
      double elen = std::sqrt(e[1] * e[1] + e[2] * e[2] + e[3] * e[3]);
      ave += br;
      for (int kk = 1; kk <= 3; kk++) {
        P.ee[kk - 1].push_back(e[kk] / elen);
        P.ee0[kk - 1].push_back(e0[kk] / e0Len);
      }
      if (j == 1) {
        double cosAlpha = (e[1] * e0[1] + e[2] * e0[2] + e[3] * e0[3]) / (elen * e0Len);
        al[i] = std::acos(cosAlpha);
        if (al[i] < al0) { al0 = al[i]; ial0 = P.ndata; }
        if ((al[i] < al0Abs) && (P.in_rel[i] == 0)) { al0Abs = al[i]; ial0_abs = P.ndata; }
      }
    }
    ave /= lp;
    for (int j = 1; j <= lp; j++) {
      P.sig.push_back(ave);
      P.sigr2.push_back(1.0 / (ave * ave));
    }
  }
  std::vector<double> weightLc(P.l_curves + 1, -1.0);
  for (;;) {
    int lcNumber;
    double w;
    if (std::fscanf(in, "%d", &lcNumber) != 1) break;
    if (std::fscanf(in, "%lf", &w) != 1) break;
    if (lcNumber >= 1 && lcNumber <= P.l_curves) weightLc[lcNumber] = w;
  }
  std::fclose(in);
  P.jd0 = P.jd00;
  if (P.jd0 <= 0) P.jd0 = (int)P.jdMin;
  for (int i = 0; i < P.ndata; i++) P.tim[i] = P.tim[i] - P.jd0;
  P.phi0 = P.phi0 * DEG2RAD;
  P.weight.clear();
  for (int i = 1; i <= P.l_curves; i++)
    for (int j = 1; j <= P.l_points[i - 1]; j++) P.weight.push_back(weightLc[i] == -1 ? 1.0 : weightLc[i]);
  if (P.onlyrel == 0) { al0 = al0Abs; ial0 = ial0_abs; }
  if (ial0 < 1) { err = "no phase reference point"; return 2; }
  double rfit = std::sqrt(2 * P.sig[ial0 - 1] / (0.5 * PI * (1 + std::cos(al0))));
  P.escl = rfit / std::sqrt((a0 * b0 + b0 * c0 + a0 * c0) / 3);
  if (P.onlyrel == 0) P.escl *= 0.8;
  /* solar phase per point (pair-independent part of Start.cu:2113-2114) */
  P.alpha.resize(P.ndata);
  for (int i = 0; i < P.ndata; i++) {
    double d = dot3(P.ee[0][i], P.ee[1][i], P.ee[2][i], P.ee0[0][i], P.ee0[1][i], P.ee0[2][i]);
    P.alpha[i] = std::acos(std::fmin(1.0, std::fmax(-1.0, d)));
  }
  int rc = setup_geometry(P, a0 * P.escl, b0 * P.escl, c0 * P.escl);
  if (rc) { err = "singular ellipsoid fit"; return rc; }
  /* ia[] (period_search_BOINC.cpp:843-857, start_CUDA.cu:850-853) */
  const int C = P.ncoef;
  P.ma = C + 5 + 3;
  P.ia.assign(P.ma + 1, 0);
  P.ia[C + 1] = 1;
  P.ia[C + 2] = 1;
  P.ia[C + 3] = P.ia_prd;
  for (int i = 1; i <= C; i++) P.ia[i] = 1;
  if (P.onlyrel == 1) P.ia[1] = 0;
  for (int i = 1; i <= 3; i++) P.ia[C + 3 + i] = P.ia_par[i];
  P.ia[C + 3 + 3 + 1] = P.ia_cl;
  P.ia[C + 3 + 3 + 2] = 0;
  P.mfit = 0; P.lastma = 0;
  for (int m = 1; m <= P.ma; m++)
    if (P.ia[m]) { P.mfit++; P.lastma = m; }
  P.lastone = 1;
  for (int m = 2; m <= P.lastma; m++) {
    if (!P.ia[m]) break;
    P.lastone = m;
  }
  P.ncoef0 = P.ma - 2 - 3;
  P.f_start = 1 / P.P_start;
  P.f_end = 1 / P.P_end;
  P.f_step = 0.5 / (P.jdMax - P.jdMin) / 24 * P.period_step_coef;
  P.n_max = (int)((P.f_start - P.f_end) / P.f_step) + 1;
  P.lcl = std::log(P.cl);
  set_stop(P, P.stop_condition, P.n_iter_min);
  const double lp[N_POLES + 1] = {0, 0, 90, 180, 270, 60, 180, 300, 60, 180, 300};
  const double bp[N_POLES + 1] = {0, 0, 0, 0, 0, 60, 60, 60, -60, -60, -60};
  for (int i = 0; i <= N_POLES; i++) { P.lambda_pole[i] = lp[i]; P.beta_pole[i] = bp[i]; }
  return 0;
}

/* ------------------------------------------------------------------ */
/* per-pair evaluation                                                  */
/* ------------------------------------------------------------------ */

/* Start.cu:421-528: Area_f = Darea_f * exp(sum_{m,l} P_l^m (c cos + s sin)) */
void curv(const Problem& P, const double* a /* a[0]=param 1 */, double* area) {
  const int nf = P.nfac, L = P.lmax, M = P.mmax;
  for (int f = 0; f < nf; f++) {
    double g = 0;
    int n = 0;
    for (int l = 0; l <= L; l++) {
      double fsum = a[n++];      /* m = 0: cos(0*phi) == 1 */
      g = std::fma(P.pleg[(0 * (L + 1) + l) * nf + f], fsum, g);
    }
    for (int m = 1; m <= M; m++) {
      double fcm = P.fc[m * nf + f], fsm = P.fs[m * nf + f];
      for (int l = m; l <= L; l++) {
        double fsum = a[n] * fcm;
        fsum = std::fma(a[n + 1], fsm, fsum);
        n += 2;
        g = std::fma(P.pleg[(m * (L + 1) + l) * nf + f], fsum, g);
      }
    }
    area[f] = P.d_area[f] * o_exp(g);
  }
}

struct PairInv {
  double omega, A, invD, kph, invD2, cl, cls;
  double sb, cb, sl, clam; /* sin/cos of beta' and lambda */
};

PairInv pair_invariants(const Problem& P, const double* a) {
  const int nc = P.ncoef0; /* a[nc] (1-based) == omega */
  PairInv v;
  v.omega = a[nc - 1];
  v.A = a[nc];
  v.invD = 1.0 / a[nc + 1];
  v.kph = a[nc + 2];
  v.invD2 = v.invD * v.invD;
  v.cl = o_exp(a[P.ma - 2]);
  v.cls = a[P.ma - 1];
  o_sincos(a[P.ncoef], &v.sb, &v.cb);       /* beta' = param C+1 */
  o_sincos(a[P.ncoef + 1], &v.sl, &v.clam); /* lambda = param C+2 */
  return v;
}

struct PointGeo {
  double g[3], g0[3];             /* viewer / Sun in the body frame */
  double db[3], db0[3];           /* d/d beta' */
  double dl[3], dl0[3];           /* d/d lambda */
  double dw[3], dw0[3];           /* d/d omega (z component 0) */
  double scale, ff, d2, alph;
};

/* Body-frame geometry of one point.  g = Rz(phi) P(beta') Rz(lambda) e with
 * Rz(lambda) e = u, P(beta') u = v; same rotation as Start.cu:2095-2195. */
void body_vectors(const PairInv& v, double cf, double sf, double t,
                  double ex, double ey, double ez, double* g, double* db, double* dl, double* dw) {
  double ux = std::fma(v.sl, ey, v.clam * ex);
  double uy = std::fma(-v.sl, ex, v.clam * ey);
  double uz = ez;
  double vx = std::fma(-v.sb, uz, v.cb * ux);
  double vy = uy;
  double vz = std::fma(v.sb, ux, v.cb * uz);
  g[0] = std::fma(sf, vy, cf * vx);
  g[1] = std::fma(-sf, vx, cf * vy);
  g[2] = vz;
  db[0] = -(cf * vz);
  db[1] = sf * vz;
  db[2] = vx;
  double w = v.cb * uy;
  dl[0] = std::fma(cf, w, -(sf * ux));
  dl[1] = std::fma(-sf, w, -(cf * ux));
  dl[2] = v.sb * uy;
  dw[0] = t * g[1];
  dw[1] = -(t * g[0]);
  dw[2] = 0.0;
}

void point_geometry(const Problem& P, const PairInv& v, int pt, PointGeo& G) {
  const double t = P.tim[pt];
  double f = std::fma(v.omega, t, P.phi0);
  const double INV2PI = 1.0 / (2.0 * PI), TWOPI = 2.0 * PI;
  f = std::fma(-TWOPI, std::round(f * INV2PI), f);
  double sf, cf;
  o_sincos(f, &sf, &cf);
  const double alph = P.alpha[pt];
  G.alph = alph;
  G.ff = o_exp2(-1.44269504088896 * (alph * v.invD));
  G.scale = std::fma(v.kph, alph, std::fma(v.A, G.ff, 1.0));
  G.d2 = ((v.A * G.ff) * alph) * v.invD2;
  body_vectors(v, cf, sf, t, P.ee[0][pt], P.ee[1][pt], P.ee[2][pt], G.g, G.db, G.dl, G.dw);
  body_vectors(v, cf, sf, t, P.ee0[0][pt], P.ee0[1][pt], P.ee0[2][pt], G.g0, G.db0, G.dl0, G.dw0);
}

/* Facet sums of one data point and its raw Jacobian row (Appendix C).
 * row[0] = ymod, row[c] = dY/da_c for c = 1..ma. */
long long point_row(const Problem& P, const PairInv& v, const double* area, int pt, double* row) {
  PointGeo G;
  point_geometry(P, v, pt, G);
  const int nf = P.nfac, C = P.ncoef;
  double U[4][3] = {}, U0[4][3] = {}, t4[4] = {0, 0, 0, 0}, t5[4] = {0, 0, 0, 0};
  std::vector<double> gat(C + 1, 0.0);
  long long nvis = 0;
  for (int f = 0; f < nf; f++) {
    const double n0 = P.nor[0][f], n1 = P.nor[1][f], n2 = P.nor[2][f];
    double mu = dot3(G.g[0], G.g[1], G.g[2], n0, n1, n2);
    double mu0 = dot3(G.g0[0], G.g0[1], G.g0[2], n0, n1, n2);
    if (mu > TINY && mu0 > TINY) {
      const int s = f & 3;
      const double ar = area[f];
      /* Lommel-Seeliger + Lambert: S = mu mu0 (cls rd + cl), rd = 1 / (mu
         + mu0); dS/dmu = mu0 (c2 mu0 + cl) and dS/dmu0 = mu (c2 mu + cl)
         with c2 = cls rd^2 (Start.cu:947-974, regrouped);
         the brightness sum ar S = cls t5 + cl t4 is formed after the loop */
      double dnom = mu + mu0;
      double rd = 1.0 / dnom;
      double am = ar * (mu * mu0);
      double w = am * std::fma(v.cls, rd, v.cl);
      double c2 = v.cls * (rd * rd);
      double ad = ar * (mu0 * std::fma(c2, mu0, v.cl));
      double ad0 = ar * (mu * std::fma(c2, mu, v.cl));
      U[s][0] = std::fma(ad, n0, U[s][0]);
      U[s][1] = std::fma(ad, n1, U[s][1]);
      U[s][2] = std::fma(ad, n2, U[s][2]);
      U0[s][0] = std::fma(ad0, n0, U0[s][0]);
      U0[s][1] = std::fma(ad0, n1, U0[s][1]);
      U0[s][2] = std::fma(ad0, n2, U0[s][2]);
      t4[s] = t4[s] + am;
      t5[s] = std::fma(am, rd, t5[s]);
      /* shape gather; column 1 (Y_1 == 1) is the size-scale column, used
         only by absolute curves */
      for (int c = 1; c <= C; c++) gat[c] = std::fma(w, P.dsph[(c - 1) * nf + f], gat[c]);
      nvis++;
    }
  }
  auto comb = [](const double* x) { return (x[0] + x[1]) + (x[2] + x[3]); };
  double T4 = comb(t4), T5 = comb(t5);
  double B = std::fma(v.cls, T5, v.cl * T4);
  double Uc[3], U0c[3];
  for (int d = 0; d < 3; d++) {
    double a4[4] = {U[0][d], U[1][d], U[2][d], U[3][d]};
    double b4[4] = {U0[0][d], U0[1][d], U0[2][d], U0[3][d]};
    Uc[d] = comb(a4);
    U0c[d] = comb(b4);
  }
  auto tsum = [&](const double* dv, const double* dv0) {
    double r = dv[0] * Uc[0];
    r = std::fma(dv[1], Uc[1], r);
    r = std::fma(dv[2], Uc[2], r);
    r = std::fma(dv0[0], U0c[0], r);
    r = std::fma(dv0[1], U0c[1], r);
    r = std::fma(dv0[2], U0c[2], r);
    return r;
  };
  double T1 = tsum(G.db, G.db0), T2 = tsum(G.dl, G.dl0), T3 = tsum(G.dw, G.dw0);
  const double sc = G.scale;
  row[0] = B * sc;
  /* column 1, the size scale: every area carries the factor exp(a_1), so
     d y / d a_1 = y */
  row[1] = row[0];
  for (int c = 2; c <= C; c++) row[c] = sc * gat[c];
  row[C + 1] = sc * T1;
  row[C + 2] = sc * T2;
  row[C + 3] = sc * T3;
  row[C + 4] = B * G.ff;
  row[C + 5] = B * G.d2;
  row[C + 6] = B * G.alph;
  row[C + 7] = (sc * T4) * v.cl;
  row[C + 8] = sc * T5;
  return nvis;
}

/* Start.cu:2539-2612: three convexity rows y_j = sum Area n_j,
 * J_j[c] = sum (Area n_j) Y_c (c <= C), 0 beyond. */
void convexity_rows(const Problem& P, const double* area, double* rows /* 3*(ma+1) */) {
  const int nf = P.nfac, C = P.ncoef, ma = P.ma;
  for (int j = 0; j < 3; j++) {
    double* row = rows + j * (ma + 1);
    for (int c = 0; c <= ma; c++) row[c] = 0.0;
    double ym = 0;
    for (int f = 0; f < nf; f++) ym = ym + area[f] * P.nor[j][f];
    row[0] = ym;
    row[1] = ym; /* d y_j / d a_1 = y_j */
    for (int c = 2; c <= C; c++) {
      double acc = 0;
      for (int f = 0; f < nf; f++) acc = std::fma(area[f] * P.nor[j][f], P.dsph[(c - 1) * nf + f], acc);
      row[c] = acc;
    }
  }
}

struct EvalOut {
  long long nvis = 0;
};

/* 16-column blocks of a row as the device lays it out: the raw row holds
 * parameters 2..C+8, the normal-equation rows mfit Jacobian columns, the
 * residual and the two per-curve sum columns. */
int row_blocks(const Problem& P) {
  const int a = (P.ncoef + 7 + 15) / 16, b = (P.mfit + 3 + 15) / 16;
  return a > b ? a : b;
}

/* Number of point chains of the normal-equation product (one per group of
 * waves that share the column blocks). */
int tile_chains(int nb) { return nb <= 4 ? 4 : 2; }

/* One mrqcof evaluation (Start.cu:4464-4498, 2262-2828, 4791-4807):
 * alpha (mfit x mfit, symmetric), beta (mfit), chisq; area[] = curv(a).
 * The convexity rows use sigr2 = conw_r2_rows and weight 1.
 *
 * Relative curves (Start.cu:2656-2715) are renormalised algebraically.  With
 * L = sum y over the curve, c = sig lp / L and the renormalised row
 * X = c (J - (y / L) d), d = sum J, the chains accumulate (W Z)^T Z over the
 * rows Z = [c J | r | z | y] (r = obs - c y; z and y are the last two
 * columns of the device's 16 NB-column rows; weighted side
 * [w c J | w r | 1 | w y], the last two only on relative curves).  Row z
 * then holds D = c d, row y holds U = c sum w y J, Y = sum w y^2 and
 * R = sum w y r, written out and cleared at the end of each curve, and
 *   X^T W X = [c J]^T W [c J] - (a b^T + b a^T),  a = U - (Y / 2L) D,
 *   X^T W r = [c J]^T W r - R b,                  b = D / L,
 * applied as one fused multiply-add chain per element over the curves.  So
 * the rows are read once, with no per-curve pass before the product.
 * Absolute curves (in_rel = 0) enter raw: X = J, r = obs - y; the column of
 * parameter 1 (the size scale) is d y / d a_1 = y, zero for relative curves.
 * The reference's GPU path computes no absolute data curves
 * (Start.cu:4685-4716); this follows its legacy forms (Start.cu:3363-3692). */
EvalOut mrqcof(const Problem& P, const double* a, double conw_r2_rows, pso_order ord,
               double* alpha, double* beta, double* chisq, double* area, double* jrows_out) {
  EvalOut out;
  const int ma = P.ma, mf = P.mfit, K = mf + 1; /* augmented width */
  const int K2 = 16 * row_blocks(P);            /* the device's row width: sum columns z, y last */
  const int ZC = K2 - 2, YC = K2 - 1;
  const int TL = 16; /* tile of 16 points: fixes the chain of every point */
  const int NC = P.l_curves;
  (void)ord;
  curv(P, a, area);
  PairInv v = pair_invariants(P, a);
  /* compact column map: compact q (0..mf-1) -> parameter index */
  std::vector<int> cmap;
  for (int c = 1; c <= ma; c++)
    if (P.ia[c]) cmap.push_back(c);
  const int G = tile_chains(row_blocks(P));
  /* raw rows of every point, then the convexity rows */
  std::vector<double> rows((size_t)(P.ndata + 3) * (ma + 1));
  for (int p = 0; p < P.ndata; p++) out.nvis += point_row(P, v, area, p, &rows[(size_t)p * (ma + 1)]);
  convexity_rows(P, area, &rows[(size_t)P.ndata * (ma + 1)]);
  if (jrows_out) std::memcpy(jrows_out, rows.data(), sizeof(double) * rows.size());
  /* brightness sum of each relative curve: per 16-point tile four partial
     sums over points p = s (mod 4), (G0+G1)+(G2+G3), then the tile sums in
     order; kept as its reciprocal */
  std::vector<int> c0(NC + 1, 0);
  for (int ci = 0; ci < NC; ci++) c0[ci + 1] = c0[ci] + P.l_points[ci];
  std::vector<double> rave(NC, 0.0);
  for (int ci = 0; ci < NC; ci++) {
    if (!P.in_rel[ci]) continue;
    const int lp = P.l_points[ci];
    double s = 0.0;
    for (int t0 = 0; t0 < lp; t0 += 16) {
      double g4[4] = {0.0, 0.0, 0.0, 0.0};
      for (int p = 0; p < 16; p++) {
        const double val = (t0 + p < lp) ? rows[(size_t)(c0[ci] + t0 + p) * (ma + 1)] : 0.0;
        g4[p & 3] = g4[p & 3] + val;
      }
      s = s + ((g4[0] + g4[1]) + (g4[2] + g4[3]));
    }
    rave[ci] = 1.0 / s;
  }
  /* G chains: lower triangles of (W Z)^T Z, K2 x K2; per (chain, relative
     curve) the rows z and y at the end of the chain's tiles of the curve */
  std::vector<double> acc((size_t)G * K2 * K2, 0.0);
  std::vector<double> rz((size_t)G * NC * K2, 0.0), ry((size_t)G * NC * K2, 0.0);
  std::vector<char> seen((size_t)G * NC, 0);
  std::vector<double> X((size_t)TL * K2, 0.0), XW((size_t)TL * K2, 0.0);
  /* the columns that carry values (the others stay zero and no element
     read later depends on them) */
  std::vector<int> act;
  for (int q = 0; q <= mf; q++) act.push_back(q);
  act.push_back(ZC);
  act.push_back(YC);
  int tile = 0;
  auto syrk_tile = [&](int np) {
    double* A = acc.data() + (size_t)(tile % G) * K2 * K2;
    for (int p = 0; p < np; p++)
      for (size_t ii = 0; ii < act.size(); ii++) {
        const int i = act[ii];
        const double wi = XW[p * K2 + i];
        for (size_t jj = 0; jj <= ii; jj++) {
          const int j = act[jj];
          A[i * K2 + j] = std::fma(wi, X[p * K2 + j], A[i * K2 + j]);
        }
      }
    tile++;
  };
  for (int ci = 0; ci < NC; ci++) {
    const int lp = P.l_points[ci];
    const bool rel = P.in_rel[ci] != 0;
    for (int t0 = 0; t0 < lp; t0 += TL) {
      const int np = lp - t0 < TL ? lp - t0 : TL;
      const int g = tile % G;
      for (int p = 0; p < np; p++) {
        const int gp = c0[ci] + t0 + p;
        const double* r = &rows[(size_t)gp * (ma + 1)];
        const double y = r[0];
        const double coef = rel ? (P.sig[gp] * (double)lp) * rave[ci] : 1.0;
        const double s2 = P.sigr2[gp] * P.weight[gp];
        for (int q = 0; q < mf; q++) {
          const int c = cmap[q];
          const double xr = (c == 1) ? (rel ? 0.0 : y) : r[c];
          X[p * K2 + q] = coef * xr;
          XW[p * K2 + q] = s2 * X[p * K2 + q];
        }
        const double res = P.bright[gp] - coef * y;
        X[p * K2 + mf] = res;
        XW[p * K2 + mf] = s2 * res;
        X[p * K2 + ZC] = 0.0;
        XW[p * K2 + ZC] = rel ? 1.0 : 0.0;
        X[p * K2 + YC] = y;
        XW[p * K2 + YC] = rel ? s2 * y : 0.0;
      }
      syrk_tile(np);
      /* the chain's last tile of this relative curve (its next tile is G
         tiles on): rows z, y out */
      if (rel) {
        if (t0 + G * TL >= lp) {
          double* A = acc.data() + (size_t)g * K2 * K2;
          double* oz = &rz[((size_t)g * NC + ci) * K2];
          double* oy = &ry[((size_t)g * NC + ci) * K2];
          for (int j = 0; j <= ZC; j++) oz[j] = A[ZC * K2 + j];
          for (int j = 0; j <= YC; j++) oy[j] = A[YC * K2 + j];
          for (int j = 0; j <= ZC; j++) A[ZC * K2 + j] = 0.0;
          for (int j = 0; j <= YC; j++) A[YC * K2 + j] = 0.0;
          seen[(size_t)g * NC + ci] = 1;
        }
      }
    }
  }
  /* convexity pseudo-curve: 3 rows, brightness 0, sigr2 = conwR^2, weight 1 */
  {
    const double* cr = &rows[(size_t)P.ndata * (ma + 1)];
    for (int p = 0; p < 3; p++) {
      const double* r = &cr[p * (ma + 1)];
      const double s2 = conw_r2_rows * 1.0;
      for (int q = 0; q < mf; q++) {
        X[p * K2 + q] = (cmap[q] == 1) ? r[0] : r[cmap[q]];
        XW[p * K2 + q] = s2 * X[p * K2 + q];
      }
      X[p * K2 + mf] = 0.0 - r[0];
      XW[p * K2 + mf] = s2 * X[p * K2 + mf];
      X[p * K2 + ZC] = XW[p * K2 + ZC] = 0.0;
      X[p * K2 + YC] = XW[p * K2 + YC] = 0.0;
    }
    syrk_tile(3);
  }
  /* the chains' partial matrices, (C0+C1)+(C2+C3) (or C0+C1) */
  std::vector<double> tot((size_t)K * K);
  const size_t KK = (size_t)K2 * K2;
  for (int i = 0; i < K; i++)
    for (int j = 0; j <= i; j++) {
      const double* a0 = acc.data() + (size_t)i * K2 + j;
      tot[(size_t)i * K + j] = (G == 4)   ? (a0[0] + a0[KK]) + (a0[2 * KK] + a0[3 * KK])
                               : (G == 2) ? a0[0] + a0[KK]
                                          : a0[0];
    }
  /* per relative curve: the chains' rows combined the same way (a chain
     without a tile of the curve gives 0.0); correction vectors
     a = U - (Y / 2L) D (a[mf] = R), b = D / L (b[mf] = 0), zero for
     absolute curves and for the padding curve that makes the count even */
  const int NCP = (NC + 1) & ~1;
  std::vector<double> T((size_t)2 * NCP * K, 0.0);
  auto comb = [&](const std::vector<double>& src, int ci, int j) {
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    for (int g = 0; g < G; g++)
      if (seen[(size_t)g * NC + ci]) x[g] = src[((size_t)g * NC + ci) * K2 + j];
    return (G == 4) ? (x[0] + x[1]) + (x[2] + x[3]) : x[0] + x[1];
  };
  for (int ci = 0; ci < NC; ci++) {
    if (!P.in_rel[ci]) continue;
    const double rv = rave[ci];
    const double Y = comb(ry, ci, YC);
    double* ta = &T[(size_t)(2 * ci) * K];
    double* tb = &T[(size_t)(2 * ci + 1) * K];
    for (int q = 0; q < mf; q++) {
      const double D = comb(rz, ci, q), U = comb(ry, ci, q);
      ta[q] = std::fma(-(0.5 * (Y * rv)), D, U);
      tb[q] = rv * D;
    }
    ta[mf] = comb(ry, ci, mf);
    tb[mf] = 0.0;
  }
  for (int i = 0; i < K; i++)
    for (int j = 0; j <= i; j++) {
      double t = tot[(size_t)i * K + j];
      for (int k2 = 0; k2 < 2 * NCP; k2++) t = std::fma(-T[(size_t)k2 * K + i], T[(size_t)(k2 ^ 1) * K + j], t);
      tot[(size_t)i * K + j] = t;
    }
  for (int i = 0; i < mf; i++)
    for (int j = 0; j <= i; j++) {
      alpha[i * mf + j] = tot[i * K + j];
      alpha[j * mf + i] = tot[i * K + j];
    }
  for (int j = 0; j < mf; j++) beta[j] = tot[mf * K + j];
  *chisq = tot[mf * K + mf];
  return out;
}

/* Start.cu:2838-3021: damped full-pivot Gauss-Jordan (NR gaussj pivot rule,
 * last occurrence of the maximum in row-major order wins).  cov is mf x mf
 * (0-based), x holds beta on entry and the solution on exit.
 * Returns 0, 1 (ipiv > 1) or 2 (zero pivot). */
int gaussj(int mf, std::vector<double>& cov, std::vector<double>& x) {
  std::vector<int> ipiv(mf, 0);
  for (int i = 0; i < mf; i++) {
    double big = 0.0;
    int irow = -1, icol = -1;
    bool err1 = false;
    for (int j = 0; j < mf; j++) {
      if (ipiv[j] == 1) continue;
      for (int k = 0; k < mf; k++) {
        if (ipiv[k] == 0) {
          double t = std::fabs(cov[j * mf + k]);
          if (t >= big) { big = t; irow = j; icol = k; }
        } else if (ipiv[k] > 1)
          err1 = true;
      }
    }
    if (err1) return 1;
    if (icol < 0) return 2; /* nothing selectable (all NaN) */
    ipiv[icol] += 1;
    if (irow != icol) {
      for (int l = 0; l < mf; l++) std::swap(cov[irow * mf + l], cov[icol * mf + l]);
      std::swap(x[irow], x[icol]);
    }
    const double piv = cov[icol * mf + icol];
    if (piv == 0.0) return 2;
    const double pivinv = 1.0 / piv;
    cov[icol * mf + icol] = 1.0;
    x[icol] *= pivinv;
    for (int l = 0; l < mf; l++) cov[icol * mf + l] *= pivinv;
    for (int ll = 0; ll < mf; ll++) {
      if (ll == icol) continue;
      const double dum = cov[ll * mf + icol];
      for (int c = 0; c < mf; c++) {
        const double base = (c == icol) ? 0.0 : cov[ll * mf + c];
        cov[ll * mf + c] = std::fma(-cov[icol * mf + c], dum, base);
      }
      x[ll] = std::fma(-x[icol], dum, x[ll]);
    }
  }
  return 0;
}

/* Iter2's rms (Start.cu:5129-5248): with the areas of the accepted
 * parameters a, v = sum_f Area_f n_f, rchisq = Chisq - |v|^2 conw_r^2 (the
 * convexity rows' share of Chisq), dev = sqrt(rchisq / ndata) over the data
 * points only (the reference divides by its ndata - 3), chck = |v|. */
void iter2_dev(const Problem& P, const double* a, double chisq, double conw_r_dev, double* area, double* dev,
               double* chck) {
  const int nf = P.nfac;
  curv(P, a, area);
  double vv[3];
  for (int d = 0; d < 3; d++) {
    double s = 0;
    for (int f = 0; f < nf; f++) s = std::fma(area[f], P.nor[d][f], s);
    vv[d] = s;
  }
  double aa = std::fma(vv[2], vv[2], std::fma(vv[1], vv[1], vv[0] * vv[0]));
  const double conwr2 = conw_r_dev * conw_r_dev;
  double rchisq = std::fma(-aa, conwr2, chisq);
  *dev = std::sqrt(rchisq / (double)(P.ndata));
  *chck = std::sqrt(aa);
}

/* FinishPole's dark-facet percentage (Start.cu:5056-5125):
 * 100 |sum A n| / sum A, NaN -> 1. */
double dark_percent(const Problem& P, const double* area, double chck) {
  double tot = 0;
  for (int f = 0; f < P.nfac; f++) tot = tot + area[f];
  tot = 1.0 / tot;
  double dd = (chck * 100.0) * tot;
  if (std::isnan(dd)) dd = 1.0;
  return dd;
}

/* ------------------------------------------------------------------ */
/* LM fit of one pair: Start.cu:4346-5248 as one sequential state machine */
/* ------------------------------------------------------------------ */
struct FitOut {
  pso_pair_result r;
};

void fit_pair(const Problem& P, int n, int m, double conw_r2_rows, double conw_r_dev, pso_order ord,
              pso_pair_result* res, double* traj, int traj_cap, const double* a0 = nullptr,
              double* a_out = nullptr, int n_iter_fixed = 0, bool a0_full = false) {
  const int ma = P.ma, mf = P.mfit, C = P.ncoef, nf = P.nfac;
  std::vector<double> cgg(ma), atry(ma), alpha(mf * mf), beta(mf), covar(mf * mf), da(mf), area(nf);
  /* Start.cu:4346-4418 */
  const double period = 1.0 / (P.f_start - (double)(n - 1) * P.f_step);
  for (int i = 0; i < C; i++) cgg[i] = P.cg_first[i];
  cgg[C] = DEG2RAD * 90 - DEG2RAD * P.beta_pole[m];
  cgg[C + 1] = DEG2RAD * P.lambda_pole[m];
  cgg[C + 2] = (24.0 * 2.0 * PI) / period;
  for (int i = 1; i <= 3; i++) cgg[C + 2 + i] = P.par[i];
  cgg[C + 6] = P.lcl;
  cgg[C + 7] = 1;
  if (a0 && a0_full) {
    /* caller-given start for every parameter (pole and rotation rate too) */
    for (int i = 0; i < ma; i++) cgg[i] = a0[i];
  } else if (a0) {
    /* caller-given start: shape and phase-function / Lambert parameters;
       pole, rotation rate and cls as above */
    for (int i = 0; i < C; i++) cgg[i] = a0[i];
    for (int i = C + 3; i <= C + 6; i++) cgg[i] = a0[i];
  }
  double Alamda = -1, iter_diff = 1e40, dev_old = 1e30, dev_new = 0, chck = 0;
  double Chisq = 0, Ochisq = 0;
  int Niter = 0;
  long long nevals = 0, nvis = 0;
  for (;;) {
    /* Iter1Begin */
    /* n_iter_fixed > 0: exactly that many iterations (stopCondition > 1 mode) */
    const int it_max = n_iter_fixed > 0 ? n_iter_fixed : P.n_iter_max;
    const double diff_max = n_iter_fixed > 0 ? 0.0 : P.iter_diff_max;
    const int it_min = n_iter_fixed > 0 ? 0 : P.n_iter_min;
    bool isNiter = ((Niter < it_max) && (iter_diff > diff_max)) || (Niter < it_min);
    if (!isNiter) break;
    bool isAlambda = false;
    if (Alamda < 0) { isAlambda = true; Alamda = P.alamda_start; }
    if (isAlambda) {
      double cs;
      nvis += mrqcof(P, cgg.data(), conw_r2_rows, ord, alpha.data(), beta.data(), &cs, area.data(), nullptr).nvis;
      nevals++;
      Ochisq = cs;
    }
    /* mrqmin step 1 */
    if (isAlambda) atry = cgg;
    {
      std::vector<double> cov(alpha);
      const double ccc = 1 + Alamda;
      for (int j = 0; j < mf; j++) cov[j * mf + j] *= ccc;
      std::vector<double> x(beta);
      int err = gaussj(mf, cov, x);
      for (int j = 0; j < mf; j++) da[j] = x[j];
      if (err == 0 || err == 2) {
        int cnt = 0;
        for (int l = 1; l <= ma; l++)
          if (P.ia[l]) { atry[l - 1] = cgg[l - 1] + da[cnt]; cnt++; }
      }
    }
    /* mrqcof at atry */
    {
      double cs;
      nvis += mrqcof(P, atry.data(), conw_r2_rows, ord, covar.data(), da.data(), &cs, area.data(), nullptr).nvis;
      nevals++;
      Chisq = cs;
    }
    /* mrqmin step 2 (Start.cu:300-392) */
    if (Chisq < Ochisq) {
      beta = da;
      alpha = covar;
      Alamda = Alamda * P.alamda_incrr;
      cgg = atry;
    } else {
      Alamda = P.alamda_incr * Alamda;
      Chisq = Ochisq;
    }
    Niter++;
    /* Iter2 (Start.cu:5129-5248) */
    if (Niter == 1 || Chisq < Ochisq) {
      Ochisq = Chisq;
      iter2_dev(P, cgg.data(), Chisq, conw_r_dev, area.data(), &dev_new, &chck);
      double diff = dev_old - dev_new;
      if (diff > 1e-10) { iter_diff = diff; dev_old = dev_new; }
    }
    if (traj && Niter <= traj_cap) {
      double* tr = traj + (size_t)(Niter - 1) * 4;
      tr[0] = Alamda; tr[1] = Chisq; tr[2] = Ochisq; tr[3] = dev_new;
    }
  }
  /* FinishPole (Start.cu:5056-5125) with dev_best = 1e40 fresh per pair */
  pso_pair_result r;
  std::memset(&r, 0, sizeof r);
  r.reported = 1;
  r.niter = Niter;
  r.dev = 1e40;
  r.chisq = Chisq;
  r.n_evals = nevals;
  r.n_vis = nvis;
  if (!(dev_new >= 1e40)) {
    r.dev = dev_new;
    r.dark = dark_percent(P, area.data(), chck);
    r.period = (2 * PI) / cgg[C + 2];
    double la = RAD2DEG * cgg[C + 1];
    r.lambda_deg = la + (la < 0 ? 360 : 0);
    r.beta_deg = 90 - RAD2DEG * cgg[C];
  }
  *res = r;
  if (a_out)
    for (int i = 0; i < ma; i++) a_out[i] = cgg[i];
}

/* start_CUDA.cu:1200-1210: earliest reported pole with minimal dev, NaN-safe */
int best_pole(const pso_pair_result* r /* N_POLES */) {
  int best = -1;
  for (int p = 0; p < N_POLES; p++)
    if (r[p].reported == 1 && (best < 0 || (!std::isnan(r[p].dev) && !(r[p].dev >= r[best].dev)))) best = p;
  return best;
}

double fit_many(const Problem& P, int npairs, const int* nl, const int* ml, double conw_r, double conw_r_dev,
                pso_order ord, int threads, pso_pair_result* res, const double* a0 = nullptr,
                double* a_out = nullptr, int n_iter_fixed = 0, bool a0_full = false) {
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  if (threads > npairs) threads = npairs > 0 ? npairs : 1;
  std::atomic<int> next(0);
  const double cr2 = conw_r * conw_r;
  auto t0 = std::chrono::steady_clock::now();
  auto work = [&]() {
    for (;;) {
      int i = next.fetch_add(1);
      if (i >= npairs) break;
      fit_pair(P, nl ? nl[i] : 1, ml ? ml[i] : 1, cr2, conw_r_dev, ord, &res[i], nullptr, 0,
               a0 ? a0 + (size_t)i * P.ma : nullptr, a_out ? a_out + (size_t)i * P.ma : nullptr, n_iter_fixed,
               a0_full);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < threads; t++) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

/* The decision that ends one round of the convexity-weight search
 * (start_CUDA.cu:747-789): the mean over the n_periods test periods of the
 * best pole's dark-facet percentage (a period with no reported pole adds
 * nothing, the mean still divides by n_periods); below 1 accepts the
 * weight, otherwise it doubles. */
void conw_decide(const pso_pair_result* res, int n_periods, double* conw_r, int* new_conw) {
  double sum_dark = 0;
  for (int n = 0; n < n_periods; n++) {
    int b = best_pole(&res[n * N_POLES]);
    if (b >= 0) sum_dark = sum_dark + res[n * N_POLES + b].dark;
  }
  double ave_dark = sum_dark / n_periods;
  if (ave_dark < 1.0) *new_conw = 1;
  if (ave_dark >= 1.0) *conw_r = *conw_r * 2;
}

/* start_CUDA.cu:400-789 + period_search_BOINC.cpp:647-771 */
int conw_search(Problem& P, int threads, pso_order ord, double* conw_out) {
  double conwR = P.conw / P.escl / P.escl;
  int newConw = 0, rounds = 0;
  const int max_test_periods = 10;
  while ((newConw != 1) && ((conwR * P.escl * P.escl) < 10.0)) {
    std::vector<int> nl, ml;
    for (int n = 1; n <= max_test_periods; n++)
      for (int m = 1; m <= N_POLES; m++) { nl.push_back(n); ml.push_back(m); }
    std::vector<pso_pair_result> res(nl.size());
    /* device conw_r during precalc is the bit pattern of a host pointer
       (start_CUDA.cu:476): a subnormal whose square is exactly 0 */
    fit_many(P, (int)nl.size(), nl.data(), ml.data(), conwR, 0.0, ord, threads, res.data());
    rounds++;
    conw_decide(res.data(), max_test_periods, &conwR, &newConw);
  }
  *conw_out = conwR;
  return rounds;
}

int run_all(Problem& P, const char* out_path, int threads, pso_order ord, int freq_limit, double* conw_final) {
  double conwR;
  conw_search(P, threads, ord, &conwR);
  if (conw_final) *conw_final = conwR;
  int nmax = P.n_max;
  if (freq_limit > 0 && freq_limit < nmax) nmax = freq_limit;
  std::vector<int> nl, ml;
  for (int n = 1; n <= nmax; n++)
    for (int m = 1; m <= N_POLES; m++) { nl.push_back(n); ml.push_back(m); }
  std::vector<pso_pair_result> res(nl.size());
  fit_many(P, (int)nl.size(), nl.data(), ml.data(), conwR, conwR, ord, threads, res.data());
  FILE* out = std::fopen(out_path, "w");
  if (!out) return 1;
  for (int n = 0; n < nmax; n++) {
    int b = best_pole(&res[n * N_POLES]);
    if (b < 0) continue;
    const pso_pair_result& r = res[n * N_POLES + b];
    double dark = (n == 0) ? conwR * P.escl * P.escl : r.dark;
    std::fprintf(out, "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n", 24 * r.period, r.dev, r.dev * r.dev * (P.ndata),
                 dark, std::round(r.lambda_deg), std::round(r.beta_deg));
  }
  std::fclose(out);
  return 0;
}

}  // namespace

/* ------------------------------------------------------------------ */
/* C API                                                                */
/* ------------------------------------------------------------------ */
struct pso_problem {
  Problem P;
};

extern "C" {

pso_problem* pso_load(const char* path, char* err, int errlen) {
  pso_problem* h = new pso_problem();
  std::string e;
  if (parse_and_setup(h->P, path, e)) {
    if (err && errlen > 0) std::snprintf(err, errlen, "%s", e.c_str());
    delete h;
    return nullptr;
  }
  return h;
}

void pso_free(pso_problem* p) { delete p; }

void pso_info(const pso_problem* h, int* I, double* D) {
  const Problem& P = h->P;
  int ii[15] = {P.ndata, P.l_curves, P.nfac, P.ncoef, P.ma, P.mfit, P.lastma, P.lastone, P.n_max,
                P.lmax, P.mmax, P.nrows, P.n_iter_max, P.n_iter_min, P.onlyrel};
  double dd[14] = {P.f_start, P.f_end, P.f_step, P.escl, P.conw, P.jd0, P.phi0, P.alamda_start,
                   P.alamda_incr, P.alamda_incrr, P.iter_diff_max, P.cl, P.jdMin, P.jdMax};
  if (I) std::memcpy(I, ii, sizeof ii);
  if (D) std::memcpy(D, dd, sizeof dd);
}

int pso_table(const pso_problem* h, const char* name, double* out, int cap) {
  const Problem& P = h->P;
  std::vector<double> v;
  std::string n(name);
  const int F = P.nfac, N = P.ndata;
  if (n == "normal") { for (int d = 0; d < 3; d++) v.insert(v.end(), P.nor[d].begin(), P.nor[d].end()); }
  else if (n == "d_area") v = P.d_area;
  else if (n == "d_sphere") v = P.dsph;
  else if (n == "cg_first") v = P.cg_first;
  else if (n == "tim") v = P.tim;
  else if (n == "ee") { for (int d = 0; d < 3; d++) v.insert(v.end(), P.ee[d].begin(), P.ee[d].end()); }
  else if (n == "ee0") { for (int d = 0; d < 3; d++) v.insert(v.end(), P.ee0[d].begin(), P.ee0[d].end()); }
  else if (n == "alpha") v = P.alpha;
  else if (n == "brightness") v = P.bright;
  else if (n == "sig") v = P.sig;
  else if (n == "sigr2") v = P.sigr2;
  else if (n == "weight") v = P.weight;
  else if (n == "ia") { for (int i = 1; i <= P.ma; i++) v.push_back(P.ia[i]); }
  else if (n == "l_points") { for (int x : P.l_points) v.push_back(x); }
  else if (n == "in_rel") { for (int x : P.in_rel) v.push_back(x); }
  else if (n == "pleg") v = P.pleg;
  else if (n == "par") v = {P.par[1], P.par[2], P.par[3], P.lcl, 1.0};
  else return -1;
  (void)F; (void)N;
  if (out) {
    if ((int)v.size() > cap) return -1;
    std::memcpy(out, v.data(), sizeof(double) * v.size());
  }
  return (int)v.size();
}

int pso_mrqcof(const pso_problem* h, const double* a, double conw_r2_rows, pso_order order, double* alpha,
               double* beta, double* chisq, double* area, double* jrows) {
  const Problem& P = h->P;
  std::vector<double> ar(P.nfac);
  mrqcof(P, a, conw_r2_rows, order, alpha, beta, chisq, ar.data(), jrows);
  if (area) std::memcpy(area, ar.data(), sizeof(double) * P.nfac);
  return 0;
}

int pso_gaussj(int mfit, const double* alpha, const double* beta, double lambda, double* x) {
  std::vector<double> cov(alpha, alpha + (size_t)mfit * mfit), xv(beta, beta + mfit);
  const double ccc = 1 + lambda;
  for (int j = 0; j < mfit; j++) cov[j * mfit + j] *= ccc;
  int err = gaussj(mfit, cov, xv);
  std::memcpy(x, xv.data(), sizeof(double) * mfit);
  return err;
}

double pso_exp(double x) { return o_exp(x); }
double pso_exp2(double x) { return o_exp2(x); }
void pso_sincos(double x, double* s, double* c) { o_sincos(x, s, c); }

int pso_fit_pair(const pso_problem* h, int n, int m, double conw_r2_rows, double conw_r_dev, pso_order order,
                 pso_pair_result* res, double* traj, int traj_cap) {
  if (m < 1 || m > N_POLES || n < 1) return 1;
  fit_pair(h->P, n, m, conw_r2_rows, conw_r_dev, order, res, traj, traj_cap);
  return 0;
}

int pso_run(pso_problem* h, const char* out_path, int threads, pso_order order, int freq_limit, double* conw_r_final) {
  return run_all(h->P, out_path, threads, order, freq_limit, conw_r_final);
}

int pso_conw_search(pso_problem* h, int threads, pso_order order, double* conw_r_final) {
  return conw_search(h->P, threads, order, conw_r_final);
}

double pso_fit_pairs(const pso_problem* h, int npairs, const int* n_list, const int* m_list, double conw_r,
                     double conw_r_dev, pso_order order, int threads, pso_pair_result* res) {
  return fit_many(h->P, npairs, n_list, m_list, conw_r, conw_r_dev, order, threads, res);
}

double pso_fit_pairs_from(const pso_problem* h, int npairs, const int* n_list, const int* m_list, double conw_r,
                          double conw_r_dev, int threads, const double* a0, int n_iter, pso_pair_result* res,
                          double* a_out) {
  return fit_many(h->P, npairs, n_list, m_list, conw_r, conw_r_dev, pso_order{4, 16}, threads, res, a0, a_out, n_iter);
}

void pso_set_stop(pso_problem* h, double stop_condition, int n_iter_min) { set_stop(h->P, stop_condition, n_iter_min); }

double pso_fit_params(const pso_problem* h, int npairs, const double* a0, double conw_r, double conw_r_dev,
                      int threads, int n_iter, pso_pair_result* res, double* a_out) {
  return fit_many(h->P, npairs, nullptr, nullptr, conw_r, conw_r_dev, pso_order{4, 16}, threads, res, a0, a_out,
                  n_iter, true);
}

void pso_finish(const pso_problem* h, const double* a, double chisq, double conw_r_dev, double* dev,
                double* chck, double* dark) {
  const Problem& P = h->P;
  std::vector<double> area(P.nfac);
  double d = 0, c = 0;
  iter2_dev(P, a, chisq, conw_r_dev, area.data(), &d, &c);
  if (dev) *dev = d;
  if (chck) *chck = c;
  if (dark) *dark = dark_percent(P, area.data(), c);
}

void pso_conw_decide(const pso_pair_result* res, int n_periods, double* conw_r, int* new_conw) {
  conw_decide(res, n_periods, conw_r, new_conw);
}

}  // extern "C"

#ifdef PSO_MAIN
/* CLI: ps_oracle <period_search_in> <period_search_out> [threads] [freq_limit] */
int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s period_search_in period_search_out [threads] [freq_limit]\n", argv[0]);
    return 2;
  }
  char err[256] = {0};
  pso_problem* h = pso_load(argv[1], err, sizeof err);
  if (!h) { std::fprintf(stderr, "error: %s\n", err); return 2; }
  int threads = argc > 3 ? std::atoi(argv[3]) : 0;
  int flim = argc > 4 ? std::atoi(argv[4]) : 0;
  pso_order o = {4, 16};
  double cw = 0;
  auto t0 = std::chrono::steady_clock::now();
  int rc = pso_run(h, argv[2], threads, o, flim, &cw);
  auto t1 = std::chrono::steady_clock::now();
  std::fprintf(stderr, "oracle: conwR %.6g, %.3f s\n", cw, std::chrono::duration<double>(t1 - t0).count());
  pso_free(h);
  return rc;
}
#endif
