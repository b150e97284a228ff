// Internal state of a ps_ctx (shared by the C ABI translation units).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ps_mi355x.h"
#include "ps_kernels.h"
#include "ps_setup.h"

struct ps_input {
  ps::InputFile f;
};

// Host-side facts of one prepared problem (with a copy of the observations
// the warm-start features are computed from).
struct HostProb {
  int ndata = 0, ncurves = 0, n_max = 0, n_iter_max = 0;
  double f_start = 0, f_end = 0, f_step = 0, escl = 0, conw = 0;
  bool have_file = false;
  std::vector<int> l_points;
  std::vector<double> tim, bright, ee, ee0;  // ee / ee0: [3][N]
  double beta_pole[PS_N_POLES], lambda_pole[PS_N_POLES];
  std::vector<double> cg_first;
  double par[3], lcl, phi0;
};

struct ps_ctx {
  int device = 0;
  int n_cu = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  // problems (one, or a batch sharing the shape basis)
  bool prepared = false;
  std::vector<HostProb> hp;
  std::vector<PsDevProblem> dps;  // host copies of the device descriptors
  PsDevProblem* d_probs = nullptr;
  int nfac = 0, ncoef = 0, ma = 0, mfit = 0;
  std::vector<int> ia;  // [ma] free-parameter mask shared by the problems
  std::vector<void*> dbufs;
  long long stride = 0;
  // launch configuration: wavefronts per workgroup (one pair per workgroup:
  // 1, 2, 4 or 8, compiled variants psw1 .. psw8), resident workgroups per
  // CU (0 = as many as fit) and 16-point facet-pass tiles a wave claims at
  // a time (1, 4, 8, 16: 16 to 256 points)
  int waves = 4;
  int wg_per_cu = 0;
  int claim_tiles = 1;
  size_t lds[4] = {0, 0, 0, 0};   // per variant (index: ps_variant_index)
  int per_cu[4] = {0, 0, 0, 0};  // resident workgroups per CU, 0 = unusable
  double* static_scratch = nullptr;  // this device's address of the file-scope scratch array
  // work-list buffers
  double* scratch = nullptr;
  size_t scratch_elems = 0;
  int* counter = nullptr;
  int* d_items_p = nullptr;
  int* d_items_n = nullptr;
  int* d_items_m = nullptr;
  double* d_items_cw = nullptr;
  double* d_out = nullptr;
  size_t items_cap = 0;
  double* d_a0 = nullptr;      // [items][ma] initial parameters (warm starts)
  double* d_afinal = nullptr;  // [items][ma] final parameters
  size_t a_cap = 0;
  // HIP graph of a scan step (work-counter reset + persistent fit launch),
  // replayed while the launch arguments stay the same
  bool use_graphs = false;
  hipGraphExec_t gexec = nullptr;
  PsFitArgs gargs{};
  int ggrid = 0;
  int gwaves = 0;
  int graph_captures = 0, graph_replays = 0;
  // statistics of the last fit call
  double last_ms = 0;
  int last_launches = 0;
  long long last_pairs = 0, last_vis = 0, last_iters = 0, last_evals = 0;
};
