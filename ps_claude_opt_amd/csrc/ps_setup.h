// Host-side application setup of the period scan: period_search_in parsing
// and the shape-basis tables (the reference's main(), trifac.c, areanorm.c,
// sphfunc.c, ellfit.c, ludcmp.c, lubksb.c).  Compiled with g++ and
// -ffp-contract=off so the tables are bit-identical to the CPU oracle's.
#pragma once

#include <string>
#include <vector>

#include "../../include/ps_mi355x.h"

namespace ps {

struct InputFile {
  // header (period_search_BOINC.cpp:343-398)
  double p_start = 0, step_coef = 0, p_end = 0;
  int ia_prd = 1;
  double jd00 = 0, phi0_deg = 0, conw = 0;
  int lmax = 0, mmax = 0, nrows = 0;
  double par[3] = {0, 0, 0};
  int ia_par[3] = {0, 0, 0};
  double cl = 0;
  int ia_cl = 0;
  double stop_condition = 0;
  int n_iter_min = 0;
  double alamda_incr = 0, alamda_incrr = 0, alamda_start = 0;
  // observations
  std::vector<int> l_points, in_rel;
  std::vector<double> tim, brightness, sig, sigr2, weight, ee, ee0;  // ee/ee0: [3][N]
  int ndata = 0;
  int onlyrel = 1;
  double jd_min = 0, jd_max = 0, jd0 = 0, escl = 0;
  // geometry and fit controls
  int nfac = 0, ncoef = 0, ma = 0;
  std::vector<double> normal, d_area, pleg, fc, fs, d_sphere, cg_first;  // component-/row-major
  std::vector<int> ia;  // [ma], ia[i] == parameter i+1
  double f_start = 0, f_end = 0, f_step = 0;

  ps_problem view() const;  // pointers into this object
};

// Returns PS_OK or a PS_E_* code with a message in err.
int read_input(const char* path, InputFile& out, std::string& err);

}  // namespace ps
