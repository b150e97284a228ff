// Fold features of the learned frequency screen (include/ps_mi355x.h,
// ps_screen_features).  No counterpart in the reference: its scan fits every
// trial frequency (start_CUDA.cu:792-1245).
//
// For trial frequency n (1-based) the rotation rate is nu_n = 24 (f_start -
// (n - 1) f_step) turns per day (the scan's own frequency grid,
// Start.cu:4346-4418).  Every curve with at least kMinPoints points is made
// relative, y_j = b_j / mean - 1, and folded at nu_n through the harmonic
// sums S_k = sum_j y_j exp(2 pi i k nu_n t_j), k = 1..PS_SCREEN_HARMONICS.
// P_k = |S_k|^2 / sum_j y_j^2 is about 1 when the curve's scatter is
// incoherent at nu_n and grows towards the curve's point count when the
// curve repeats with the trial period (a rotational lightcurve is
// double-peaked, so most of its power sits in k = 2).  The features
// summarise P_k over the short ("dense", spanning at most 2 days) and the
// long ("sparse") curves of the problem; they do not depend on the pole, so
// one feature row serves the 10 poles of a frequency.  Row layout (group g =
// 0 dense, 1 sparse; zeros where a group has no curve):
//   [12 g + k - 1]      log of the mean of P_k over the group's curves
//   [12 g + 4 + k - 1]  log of the largest P_k / (points of the curve)
//   [12 g + 8 + k - 1]  log of sum |S_k|^2 / sum sum y^2 over the group
//   [24] log(period / 10 h), [25] fraction of dense curves longer than the
//   trial period, [26], [27] log of the mean points per dense and per sparse
//   curve, [28], [29] log(1 + dense curves), log(1 + sparse curves), then 0.
//
// exp(2 pi i k nu_n t_j) is advanced from frequency to frequency by one
// complex rotation per point (the grid is uniform in nu); the phases start
// from the fractional turn of nu t in double precision, so 20,000 steps
// accumulate errors of order 1e-12 in the sums.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/ps_mi355x.h"
#include "ps_ctx.h"

namespace {

constexpr int kMinPoints = 8;
constexpr double kDenseSpanDays = 2.0;
constexpr double kEps = 1e-6;
constexpr int H = PS_SCREEN_HARMONICS;

struct CurveFold {
  int first = 0, count = 0;
  bool dense = false;
  double var = 0.0;  // sum_j y_j^2
  double span = 0.0; // days
};

}  // namespace

int ps_screen_n_features(void) { return PS_SCREEN_FEATURES; }

int ps_screen_features(const ps_ctx* c, int prob, int n_first, int n_count, float* out) {
  if (!c || n_count < 0 || (n_count && !out)) return PS_E_ARG;
  if (!c->prepared) return PS_E_STATE;
  if (prob < 0 || prob >= (int)c->hp.size()) return PS_E_ARG;
  const HostProb& h = c->hp[prob];
  if (n_count == 0) return PS_OK;
  if (n_first < 1 || n_first + n_count - 1 > h.n_max) return PS_E_ARG;
  static_assert(PS_SCREEN_FEATURES >= 6 * H + 6 && H == 4, "screen feature layout");
  constexpr double kTwoPi = 6.28318530717958647692;

  // the curves that enter the features
  std::vector<CurveFold> cv;
  std::vector<double> y(h.tim.size(), 0.0);
  int off = 0;
  for (int lp : h.l_points) {
    if (lp >= kMinPoints) {
      double mean = 0.0;
      for (int j = off; j < off + lp; ++j) mean += h.bright[j];
      mean /= lp;
      CurveFold f;
      f.first = off;
      f.count = lp;
      if (mean > 0.0) {
        for (int j = off; j < off + lp; ++j) {
          y[j] = h.bright[j] / mean - 1.0;
          f.var += y[j] * y[j];
        }
      }
      f.span = h.tim[off + lp - 1] - h.tim[off];
      f.dense = f.span <= kDenseSpanDays;
      if (f.var > 0.0) cv.push_back(f);
    }
    off += lp;
  }
  const int nc = (int)cv.size();
  // pw[(n * nc + c) * H + k]: P_k of curve c at frequency n_first + n
  std::vector<double> pw((size_t)n_count * nc * H, 0.0);
  std::atomic<int> next{0};
  auto worker = [&]() {
    // exp(2 pi i nu t_j) at the current frequency (zr, zi) and its step
    // rotation (rr, ri), in plain doubles (std::complex products call the
    // C99 Annex G helper when fast-math is off)
    std::vector<double> zr, zi, rr, ri;
    for (;;) {
      const int ci = next.fetch_add(1);
      if (ci >= nc) break;
      const CurveFold& f = cv[ci];
      zr.resize(f.count);
      zi.resize(f.count);
      rr.resize(f.count);
      ri.resize(f.count);
      const double nu0 = 24.0 * (h.f_start - (double)(n_first - 1) * h.f_step);
      const double dnu = -24.0 * h.f_step;
      for (int j = 0; j < f.count; ++j) {
        const double t = h.tim[f.first + j];
        double ph = nu0 * t;
        ph -= std::floor(ph);
        double sp = dnu * t;
        sp -= std::floor(sp);
        zr[j] = std::cos(kTwoPi * ph);
        zi[j] = std::sin(kTwoPi * ph);
        rr[j] = std::cos(kTwoPi * sp);
        ri[j] = std::sin(kTwoPi * sp);
      }
      const double* yy = y.data() + f.first;
      for (int n = 0; n < n_count; ++n) {
        double sr[H] = {0.0}, si[H] = {0.0};
        for (int j = 0; j < f.count; ++j) {
          const double a = zr[j], b = zi[j];
          const double a2 = a * a - b * b, b2 = 2.0 * a * b;   // z^2
          const double a3 = a2 * a - b2 * b, b3 = a2 * b + b2 * a;  // z^3
          const double a4 = a2 * a2 - b2 * b2, b4 = 2.0 * a2 * b2;  // z^4
          const double w = yy[j];
          sr[0] += w * a;
          si[0] += w * b;
          sr[1] += w * a2;
          si[1] += w * b2;
          sr[2] += w * a3;
          si[2] += w * b3;
          sr[3] += w * a4;
          si[3] += w * b4;
          zr[j] = a * rr[j] - b * ri[j];
          zi[j] = a * ri[j] + b * rr[j];
        }
        double* p = pw.data() + ((size_t)n * nc + ci) * H;
        for (int k = 0; k < H; ++k) p[k] = (sr[k] * sr[k] + si[k] * si[k]) / f.var;
      }
    }
  };
  const int nthreads = std::max(1, std::min<int>({nc, 16, (int)std::max(1u, std::thread::hardware_concurrency())}));
  std::vector<std::thread> pool;
  for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();

  // per-problem constants
  int ndense = 0, nsparse = 0;
  double pts[2] = {0.0, 0.0};
  for (const CurveFold& f : cv) {
    (f.dense ? ndense : nsparse) += 1;
    pts[f.dense ? 0 : 1] += f.count;
  }
  for (int n = 0; n < n_count; ++n) {
    float* o = out + (size_t)n * PS_SCREEN_FEATURES;
    for (int i = 0; i < PS_SCREEN_FEATURES; ++i) o[i] = 0.f;
    const double period_h = 1.0 / (h.f_start - (double)(n_first + n - 1) * h.f_step);
    double mean[2][H] = {}, mx[2][H] = {}, pool_p[2][H] = {}, var[2] = {0.0, 0.0};
    int overlap = 0;
    for (int ci = 0; ci < nc; ++ci) {
      const CurveFold& f = cv[ci];
      const int g = f.dense ? 0 : 1;
      const double* p = pw.data() + ((size_t)n * nc + ci) * H;
      for (int k = 0; k < H; ++k) {
        mean[g][k] += p[k];
        mx[g][k] = std::max(mx[g][k], p[k] / f.count);
        pool_p[g][k] += p[k] * f.var;  // |S_k|^2
      }
      var[g] += f.var;
      if (f.dense && 24.0 * f.span > period_h) ++overlap;
    }
    const int cnt[2] = {ndense, nsparse};
    for (int g = 0; g < 2; ++g) {
      if (cnt[g] == 0) continue;
      for (int k = 0; k < H; ++k) {
        o[3 * H * g + k] = (float)std::log(mean[g][k] / cnt[g] + kEps);      // mean power
        o[3 * H * g + H + k] = (float)std::log(mx[g][k] + kEps);             // best curve's coherence
        o[3 * H * g + 2 * H + k] = (float)std::log(pool_p[g][k] / var[g] + kEps);  // pooled power
      }
    }
    const int b = 6 * H;
    o[b + 0] = (float)std::log(period_h / 10.0);
    o[b + 1] = ndense ? (float)overlap / ndense : 0.f;
    o[b + 2] = ndense ? (float)std::log(pts[0] / ndense) : 0.f;
    o[b + 3] = nsparse ? (float)std::log(pts[1] / nsparse) : 0.f;
    o[b + 4] = (float)std::log(1.0 + ndense);
    o[b + 5] = (float)std::log(1.0 + nsparse);
  }
  return PS_OK;
}
