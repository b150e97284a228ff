// Host-side setup of the period scan (see ps_setup.h).
//
// Follows period_search_BOINC.cpp:343-866 for the input grammar and the
// derived quantities, trifac.c:11-83 for the triangulation, areanorm.c:11-61
// for normals/areas, sphfunc.c:12-88 for the Laplace basis and ellfit.c:12-88
// (with NR LU decomposition, ludcmp.c / lubksb.c) for the initial shape.
#include "ps_setup.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace ps {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kDeg = kPi / 180.0;

struct Reader {
  FILE* f;
  char buf[1024];
  void skip_line() {
    if (!std::fgets(buf, sizeof buf, f)) buf[0] = 0;
  }
};

// Vertex triplets of the octant triangulation, 1-based facet and vertex
// numbers (trifac.c:11-83).
std::vector<int> triangulate(int nrows) {
  const int nf = 8 * nrows * nrows;
  std::vector<int> tri(3 * (nf + 1), 0);
  const int W = 4 * nrows + 1;
  std::vector<int> node((2 * nrows + 1) * W, 0);
  auto N = [&](int i, int j) -> int& { return node[i * W + j]; };
  int v = 1;
  N(0, 0) = v;
  for (int i = 1; i <= nrows; ++i)
    for (int j = 0; j < 4 * i; ++j) {
      N(i, j) = ++v;
      if (j == 0) N(i, 4 * i) = v;
    }
  for (int i = nrows - 1; i >= 1; --i)
    for (int j = 0; j < 4 * i; ++j) {
      N(2 * nrows - i, j) = ++v;
      if (j == 0) N(2 * nrows - i, 4 * i) = v;
    }
  N(2 * nrows, 0) = v + 1;
  int t = 0;
  auto emit = [&](int a, int b, int c) {
    ++t;
    tri[3 * t + 0] = a;
    tri[3 * t + 1] = b;
    tri[3 * t + 2] = c;
  };
  for (int r = 1; r <= nrows; ++r)
    for (int q = 0; q < 4; ++q) {
      const int j0 = q * r;
      emit(N(r - 1, j0 - q), N(r, j0), N(r, j0 + 1));
      for (int j = j0 + 1; j <= j0 + r - 1; ++j) {
        emit(N(r, j), N(r - 1, j - q), N(r - 1, j - 1 - q));
        emit(N(r - 1, j - q), N(r, j), N(r, j + 1));
      }
    }
  for (int r = nrows + 1; r <= 2 * nrows; ++r)
    for (int q = 0; q < 4; ++q) {
      const int span = 2 * nrows - r;
      const int j0 = q * span;
      emit(N(r, j0), N(r - 1, j0 + 1 + q), N(r - 1, j0 + q));
      for (int j = j0 + 1; j <= j0 + span; ++j) {
        emit(N(r, j), N(r - 1, j + q), N(r, j - 1));
        emit(N(r, j), N(r - 1, j + 1 + q), N(r - 1, j + q));
      }
    }
  return tri;
}

// Crout LU with implicit partial pivoting and back substitution (NR), on an
// (n+1)x(n+1) 1-based row-major matrix.
bool lu_solve(std::vector<double>& a, int n, std::vector<double>& b) {
  const int S = n + 1;
  std::vector<double> vv(S);
  std::vector<int> perm(S);
  for (int i = 1; i <= n; ++i) {
    double big = 0.0;
    for (int j = 1; j <= n; ++j) big = std::fabs(a[i * S + j]) > big ? std::fabs(a[i * S + j]) : big;
    if (big == 0.0) return false;
    vv[i] = 1.0 / big;
  }
  for (int j = 1; j <= n; ++j) {
    for (int i = 1; i < j; ++i) {
      double s = a[i * S + j];
      for (int k = 1; k < i; ++k) s -= a[i * S + k] * a[k * S + j];
      a[i * S + j] = s;
    }
    double big = 0.0;
    int imax = -1;
    for (int i = j; i <= n; ++i) {
      double s = a[i * S + j];
      for (int k = 1; k < j; ++k) s -= a[i * S + k] * a[k * S + j];
      a[i * S + j] = s;
      const double d = vv[i] * std::fabs(s);
      if (d >= big) {
        big = d;
        imax = i;
      }
    }
    if (j != imax) {
      for (int k = 1; k <= n; ++k) {
        const double tmp = a[imax * S + k];
        a[imax * S + k] = a[j * S + k];
        a[j * S + k] = tmp;
      }
      vv[imax] = vv[j];
    }
    perm[j] = imax;
    if (a[j * S + j] == 0.0) a[j * S + j] = 1.0e-20;
    if (j != n) {
      const double d = 1.0 / a[j * S + j];
      for (int i = j + 1; i <= n; ++i) a[i * S + j] *= d;
    }
  }
  int first = 0;
  for (int i = 1; i <= n; ++i) {
    const int ip = perm[i];
    double s = b[ip];
    b[ip] = b[i];
    if (first)
      for (int j = first; j <= i - 1; ++j) s -= a[i * S + j] * b[j];
    else if (s)
      first = i;
    b[i] = s;
  }
  for (int i = n; i >= 1; --i) {
    double s = b[i];
    for (int j = i + 1; j <= n; ++j) s -= a[i * S + j] * b[j];
    b[i] = s / a[i * S + i];
  }
  return true;
}

int build_geometry(InputFile& in, double a, double b, double c, std::string& err) {
  const int nrows = in.nrows, L = in.lmax, M = in.mmax;
  in.ncoef = 0;
  for (int m = 0; m <= M; ++m) in.ncoef += (L - m + 1) * (m ? 2 : 1);
  // vertex directions (period_search_BOINC.cpp:790-825)
  std::vector<double> th(1, 0.0), ph(1, 0.0);
  th.push_back(0.0);
  ph.push_back(0.0);
  const double dth = kPi / (2 * nrows);
  for (int i = 1; i <= nrows; ++i) {
    const double dph = kPi / (2 * i);
    for (int j = 0; j < 4 * i; ++j) {
      th.push_back(i * dth);
      ph.push_back(j * dph);
    }
  }
  for (int i = nrows - 1; i >= 1; --i) {
    const double dph = kPi / (2 * i);
    for (int j = 0; j < 4 * i; ++j) {
      th.push_back(kPi - i * dth);
      ph.push_back(j * dph);
    }
  }
  th.push_back(kPi);
  ph.push_back(0.0);
  const int nvert = (int)th.size() - 1;
  const int nf = 8 * nrows * nrows;
  in.nfac = nf;
  const std::vector<int> tri = triangulate(nrows);
  // normals and areas (areanorm.c)
  std::vector<double> vx(nvert + 1), vy(nvert + 1), vz(nvert + 1);
  for (int i = 1; i <= nvert; ++i) {
    const double st = std::sin(th[i]);
    vx[i] = st * std::cos(ph[i]);
    vy[i] = st * std::sin(ph[i]);
    vz[i] = std::cos(th[i]);
  }
  in.normal.assign(3 * nf, 0.0);
  in.d_area.assign(nf, 0.0);
  std::vector<double> nth(nf), nph(nf);
  for (int f = 1; f <= nf; ++f) {
    const int p0 = tri[3 * f], p1 = tri[3 * f + 1], p2 = tri[3 * f + 2];
    const double ax = vx[p1] - vx[p0], ay = vy[p1] - vy[p0], az = vz[p1] - vz[p0];
    const double bx = vx[p2] - vx[p0], by = vy[p2] - vy[p0], bz = vz[p2] - vz[p0];
    const double cx = ay * bz - by * az;
    const double cy = az * bx - bz * ax;
    const double cz = ax * by - bx * ay;
    const double len = std::sqrt(cx * cx + cy * cy + cz * cz);
    const double nx = cx / len, ny = cy / len, nz = cz / len;
    in.normal[0 * nf + f - 1] = nx;
    in.normal[1 * nf + f - 1] = ny;
    in.normal[2 * nf + f - 1] = nz;
    nth[f - 1] = std::acos(nz);
    nph[f - 1] = std::atan2(ny, nx);
    in.d_area[f - 1] = 0.5 * len;
  }
  // Laplace basis (sphfunc.c): aleg[n][l][m] coefficients of
  // cos^n(theta) sin^m(theta) in the unnormalised P_l^m
  const int S = L + 1;
  std::vector<double> aleg(S * S * S, 0.0);
  auto AL = [&](int n, int l, int m) -> double& { return aleg[(n * S + l) * S + m]; };
  AL(0, 0, 0) = 1;
  if (L >= 1) AL(1, 1, 0) = 1;
  for (int l = 1; l <= L; ++l) AL(0, l, l) = AL(0, l - 1, l - 1) * (2 * l - 1);
  for (int m = 0; m <= M; ++m)
    for (int l = m + 1; l <= L; ++l) {
      const bool even = ((l - m) % 2) == 0;
      if (even) AL(0, l, m) = -(l + m - 1) * AL(0, l - 2, m) / (1 * (l - m));
      if (l != 1)
        for (int n = even ? 2 : 1; n <= l - m; n += 2)
          AL(n, l, m) = ((2 * l - 1) * AL(n - 1, l - 1, m) - (l + m - 1) * AL(n, l - 2, m)) / (1 * (l - m));
    }
  in.fc.assign((M + 1) * nf, 0.0);
  in.fs.assign((M + 1) * nf, 0.0);
  in.pleg.assign((M + 1) * (L + 1) * nf, 0.0);
  in.d_sphere.assign(in.ncoef * nf, 0.0);
  std::vector<double> sp(S), cp(S);
  for (int f = 0; f < nf; ++f) {
    sp[0] = 1;
    cp[0] = 1;
    if (L >= 1) {
      sp[1] = std::sin(nth[f]);
      cp[1] = std::cos(nth[f]);
    }
    for (int j = 2; j <= L; ++j) {
      sp[j] = sp[1] * sp[j - 1];
      cp[j] = cp[1] * cp[j - 1];
    }
    in.fs[f] = 0;
    in.fc[f] = 1;
    for (int m = 1; m <= M; ++m) {
      in.fs[m * nf + f] = std::sin(m * nph[f]);
      in.fc[m * nf + f] = std::cos(m * nph[f]);
    }
    int k = 0;
    for (int m = 0; m <= M; ++m)
      for (int l = m; l <= L; ++l) {
        double p = 0;
        for (int n = ((l - m) % 2 == 0) ? 0 : 1; n <= l - m; n += 2) p = p + AL(n, l, m) * cp[n] * sp[m];
        in.pleg[(m * (L + 1) + l) * nf + f] = p;
        in.d_sphere[k++ * nf + f] = in.fc[m * nf + f] * p;
        if (m) in.d_sphere[k++ * nf + f] = in.fs[m * nf + f] * p;
      }
  }
  // least-squares log-curvature fit of the a,b,c ellipsoid (ellfit.c)
  const int nc = in.ncoef;
  std::vector<double> er(nf), fm(nf * nc);
  for (int f = 0; f < nf; ++f) {
    const double st = std::sin(nth[f]);
    const double q1 = a * st * std::cos(nph[f]);
    const double q2 = b * st * std::sin(nph[f]);
    const double q3 = c * std::cos(nth[f]);
    er[f] = 2 * (std::log(a * b * c) - std::log(q1 * q1 + q2 * q2 + q3 * q3));
    int n = 0;
    for (int m = 0; m <= M; ++m)
      for (int l = m; l <= L; ++l) {
        const double p = in.pleg[(m * (L + 1) + l) * nf + f];
        if (m) {
          fm[f * nc + n++] = p * std::cos(m * nph[f]);
          fm[f * nc + n++] = p * std::sin(m * nph[f]);
        } else {
          fm[f * nc + n++] = p;
        }
      }
  }
  std::vector<double> nm((nc + 1) * (nc + 1), 0.0), rhs(nc + 1, 0.0);
  for (int i = 0; i < nc; ++i) {
    for (int j = 0; j < nc; ++j) {
      double s = 0;
      for (int f = 0; f < nf; ++f) s = s + fm[f * nc + i] * fm[f * nc + j];
      nm[(i + 1) * (nc + 1) + j + 1] = s;
    }
    double s = 0;
    for (int f = 0; f < nf; ++f) s = s + fm[f * nc + i] * er[f];
    rhs[i + 1] = s;
  }
  if (!lu_solve(nm, nc, rhs)) {
    err = "singular ellipsoid fit";
    return PS_E_INPUT;
  }
  in.cg_first.assign(rhs.begin() + 1, rhs.end());
  return PS_OK;
}

}  // namespace

ps_problem InputFile::view() const {
  ps_problem p;
  std::memset(&p, 0, sizeof p);
  p.ndata = ndata;
  p.n_curves = (int)l_points.size();
  p.l_points = l_points.data();
  p.in_rel = in_rel.data();
  p.tim = tim.data();
  p.ee = ee.data();
  p.ee0 = ee0.data();
  p.brightness = brightness.data();
  p.sig = sig.data();
  p.sigr2 = sigr2.data();
  p.weight = weight.data();
  p.nfac = nfac;
  p.lmax = lmax;
  p.mmax = mmax;
  p.ncoef = ncoef;
  p.normal = normal.data();
  p.d_area = d_area.data();
  p.pleg = pleg.data();
  p.fc = fc.data();
  p.fs = fs.data();
  p.d_sphere = d_sphere.data();
  p.cg_first = cg_first.data();
  p.ia = ia.data();
  for (int i = 0; i < 3; ++i) p.par[i] = par[i];
  p.cl = cl;
  p.phi0 = phi0_deg * kDeg;
  p.alamda_start = alamda_start;
  p.alamda_incr = alamda_incr;
  p.alamda_incrr = alamda_incrr;
  p.stop_condition = stop_condition;
  p.n_iter_min = n_iter_min;
  p.f_start = f_start;
  p.f_end = f_end;
  p.f_step = f_step;
  p.escl = escl;
  // the ten initial poles (period_search_BOINC.cpp:229-238)
  static const double lam[PS_N_POLES] = {0, 90, 180, 270, 60, 180, 300, 60, 180, 300};
  static const double bet[PS_N_POLES] = {0, 0, 0, 0, 60, 60, 60, -60, -60, -60};
  for (int i = 0; i < PS_N_POLES; ++i) {
    p.lambda_pole[i] = lam[i];
    p.beta_pole[i] = bet[i];
  }
  return p;
}

int read_input(const char* path, InputFile& in, std::string& err) {
  Reader r{std::fopen(path, "r"), {0}};
  if (!r.f) {
    err = std::string("cannot open ") + path;
    return PS_E_INPUT;
  }
  FILE* f = r.f;
  bool ok = true;
  ok = ok && std::fscanf(f, "%lf %lf %lf %d", &in.p_start, &in.step_coef, &in.p_end, &in.ia_prd) == 4;
  r.skip_line();
  ok = ok && std::fscanf(f, "%lf", &in.jd00) == 1;
  r.skip_line();
  ok = ok && std::fscanf(f, "%lf", &in.phi0_deg) == 1;
  r.skip_line();
  ok = ok && std::fscanf(f, "%lf", &in.conw) == 1;
  r.skip_line();
  ok = ok && std::fscanf(f, "%d %d", &in.lmax, &in.mmax) == 2;
  r.skip_line();
  ok = ok && std::fscanf(f, "%d", &in.nrows) == 1;
  r.skip_line();
  for (int i = 0; i < 3; ++i) {
    ok = ok && std::fscanf(f, "%lf %d", &in.par[i], &in.ia_par[i]) == 2;
    r.skip_line();
  }
  ok = ok && std::fscanf(f, "%lf %d", &in.cl, &in.ia_cl) == 2;
  r.skip_line();
  ok = ok && std::fscanf(f, "%lf", &in.stop_condition) == 1;
  r.skip_line();
  ok = ok && std::fscanf(f, "%d", &in.n_iter_min) == 1;
  r.skip_line();
  ok = ok && std::fscanf(f, "%lf", &in.alamda_incr) == 1;
  r.skip_line();
  in.alamda_incrr = in.alamda_incr != 0.0 ? 1.0 / in.alamda_incr : DBL_MAX;
  ok = ok && std::fscanf(f, "%lf", &in.alamda_start) == 1;
  r.skip_line();
  int ncurves = 0;
  ok = ok && std::fscanf(f, "%d", &ncurves) == 1;
  if (!ok || ncurves <= 0) {
    std::fclose(f);
    err = "malformed header";
    return PS_E_INPUT;
  }
  if (in.lmax < 0 || in.mmax < 0 || in.mmax > in.lmax || in.lmax > 15 || in.nrows < 1) {
    std::fclose(f);
    err = "unsupported lmax/mmax/nrows";
    return PS_E_INPUT;
  }
  double al0 = kPi, al0_abs = kPi;
  int ial0 = -1, ial0_abs = -1;
  in.jd_min = 1e20;
  in.jd_max = -1e40;
  in.onlyrel = 1;
  std::vector<double> ex, ey, ez, sx, sy, sz;
  for (int c = 0; c < ncurves; ++c) {
    int np = 0, flag = 0;
    if (std::fscanf(f, "%d %d", &np, &flag) != 2 || np <= 0) {
      std::fclose(f);
      err = "malformed lightcurve header";
      return PS_E_INPUT;
    }
    r.skip_line();
    in.l_points.push_back(np);
    in.in_rel.push_back(1 - flag);
    if (flag) in.onlyrel = 0;
    double mean = 0;
    for (int j = 0; j < np; ++j) {
      double t, b, s[3], e[3];
      if (std::fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf", &t, &b, &s[0], &s[1], &s[2], &e[0], &e[1], &e[2]) != 8) {
        std::fclose(f);
        err = "malformed data line";
        return PS_E_INPUT;
      }
      ++in.ndata;
      in.tim.push_back(t);
      in.brightness.push_back(b);
      in.jd_min = t < in.jd_min ? t : in.jd_min;
      in.jd_max = t > in.jd_max ? t : in.jd_max;
      const double slen = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
      const double elen = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
      mean += b;
      ex.push_back(e[0] / elen);
      ey.push_back(e[1] / elen);
      ez.push_back(e[2] / elen);
      sx.push_back(s[0] / slen);
      sy.push_back(s[1] / slen);
      sz.push_back(s[2] / slen);
      if (j == 0) {
        const double al = std::acos((e[0] * s[0] + e[1] * s[1] + e[2] * s[2]) / (elen * slen));
        if (al < al0) {
          al0 = al;
          ial0 = in.ndata;
        }
        if (al < al0_abs && flag) {
          al0_abs = al;
          ial0_abs = in.ndata;
        }
      }
    }
    mean /= np;
    for (int j = 0; j < np; ++j) {
      in.sig.push_back(mean);
      in.sigr2.push_back(1.0 / (mean * mean));
    }
  }
  std::vector<double> wlc(ncurves + 1, -1.0);
  for (;;) {
    int lc;
    double w;
    if (std::fscanf(f, "%d", &lc) != 1 || std::fscanf(f, "%lf", &w) != 1) break;
    if (lc >= 1 && lc <= ncurves) wlc[lc] = w;
  }
  std::fclose(f);
  in.jd0 = in.jd00;
  if (in.jd0 <= 0) in.jd0 = (int)in.jd_min;
  for (double& t : in.tim) t = t - in.jd0;
  for (int c = 0; c < ncurves; ++c)
    for (int j = 0; j < in.l_points[c]; ++j) in.weight.push_back(wlc[c + 1] == -1 ? 1.0 : wlc[c + 1]);
  in.ee = ex;
  in.ee.insert(in.ee.end(), ey.begin(), ey.end());
  in.ee.insert(in.ee.end(), ez.begin(), ez.end());
  in.ee0 = sx;
  in.ee0.insert(in.ee0.end(), sy.begin(), sy.end());
  in.ee0.insert(in.ee0.end(), sz.begin(), sz.end());
  if (!in.onlyrel) {
    al0 = al0_abs;
    ial0 = ial0_abs;
  }
  if (ial0 < 1) {
    err = "no phase reference point";
    return PS_E_INPUT;
  }
  const double a0 = 1.05, b0 = 1.00, c0 = 0.95;
  const double rfit = std::sqrt(2 * in.sig[ial0 - 1] / (0.5 * kPi * (1 + std::cos(al0))));
  in.escl = rfit / std::sqrt((a0 * b0 + b0 * c0 + a0 * c0) / 3);
  if (!in.onlyrel) in.escl *= 0.8;
  int rc = build_geometry(in, a0 * in.escl, b0 * in.escl, c0 * in.escl, err);
  if (rc) return rc;
  // free/fixed mask (period_search_BOINC.cpp:843-857, start_CUDA.cu:850-853)
  const int C = in.ncoef;
  in.ma = C + 8;
  in.ia.assign(in.ma, 1);
  if (in.onlyrel) in.ia[0] = 0;
  in.ia[C + 2] = in.ia_prd;
  for (int i = 0; i < 3; ++i) in.ia[C + 3 + i] = in.ia_par[i];
  in.ia[C + 6] = in.ia_cl;
  in.ia[C + 7] = 0;
  in.f_start = 1 / in.p_start;
  in.f_end = 1 / in.p_end;
  in.f_step = 0.5 / (in.jd_max - in.jd_min) / 24 * in.step_coef;
  return PS_OK;
}

}  // namespace ps
