// period_search: drop-in command line of the reference application
// (period_search_BOINC.cpp:171-907) on the MI355X kernels.
//
//   period_search [--device N] [--N k | --batch F] [--state PATH]
//                 [--max-batches K] [period_search_in [period_search_out]]
//
// Reads period_search_in, runs the convexity-weight search and the scan,
// writes one line per trial frequency:
//   "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n" = period[h], rms, chi^2,
//   dark facet % (first line: the final convexity weight), lambda, beta.
//
// Checkpoints (period_search_BOINC.cpp:102-118, start_CUDA.cu:1185-1191):
// after each conw round and each batch of frequencies the state file
// ("period_search_state" beside the output) is replaced by
// "nlines newConw conwR" (the reference's three fields: frequencies done,
// 1 once the convexity weight is final, the weight) and, fourth, the byte
// length of the output at that point.  A run that finds a state file
// resumes there (the reference writes the file but has its resume switched
// off, period_search_BOINC.cpp:293-312): the conw search continues from the
// saved weight unless it was final, the output is cut back to the saved
// length and the scan continues at frequency nlines + 1.  The lines of a
// resumed run are byte-identical to an uninterrupted one.
//
// --device N selects the GPU unless a BOINC client's init_data.xml names
// one (gpu_device_num, period_search_BOINC.cpp:598-614).  --N k sets the
// batch to 250 k frequencies, --batch F to F frequencies.  The default,
// k = 10, fits 2,500 frequencies (25,000 pair-fits) per launch: smaller
// launches end with a longer tail in which some workgroups have no pair
// left (about 4 % of the time at 500 frequencies).
// --max-batches K stops after K scan batches, leaving a resumable state
// (for time-sliced runs).
// Exit status: 0 done, 1 output not writable, 2 bad input, 3 device
// failure during the scan, 999 no usable device (as start_CUDA.cu:216;
// the parent process reads it modulo 256, as 231).
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ps_boinc.h"
#include "../../include/ps_mi355x.h"

namespace {

struct State {
  int nlines = 0, new_conw = 0;
  double conw_r = 0.0;
  long long out_bytes = -1;  // -1: not recorded
};

bool read_state(const std::string& path, State* s) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  const int n = std::fscanf(f, "%d %d %lf", &s->nlines, &s->new_conw, &s->conw_r);
  long long b = -1;
  if (n == 3 && std::fscanf(f, "%lld", &b) == 1) s->out_bytes = b;
  std::fclose(f);
  return n == 3;
}

// Written to a temporary file and renamed over the state (as
// period_search_BOINC.cpp:102-118 does), so a crash leaves the old or the
// new state, never a torn one.
bool write_state(const std::string& path, const State& s) {
  const std::string tmp = path + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "w");
  if (!f) return false;
  std::fprintf(f, "%d %d %.17g %lld\n", s.nlines, s.new_conw, s.conw_r, s.out_bytes);
  const bool ok = std::fflush(f) == 0 && fsync(fileno(f)) == 0;
  std::fclose(f);
  return ok && std::rename(tmp.c_str(), path.c_str()) == 0;
}

std::string dir_of(const std::string& p) {
  const size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

}  // namespace

int main(int argc, char** argv) {
  int device = -1, nfactor = 10, max_batches = 0, batch_freqs = 0;
  std::string in_name = "period_search_in", out_name = "period_search_out", state_path;
  int pos = 0;
  for (int i = 1; i < argc; ++i) {
    if (std::strcmp(argv[i], "--device") == 0 && i + 1 < argc) {
      device = std::atoi(argv[++i]);
    } else if (std::strcmp(argv[i], "--N") == 0 && i + 1 < argc) {
      nfactor = std::max(1, std::atoi(argv[++i]));
    } else if (std::strcmp(argv[i], "--state") == 0 && i + 1 < argc) {
      state_path = argv[++i];
    } else if (std::strcmp(argv[i], "--batch") == 0 && i + 1 < argc) {
      batch_freqs = std::max(1, std::atoi(argv[++i]));
    } else if (std::strcmp(argv[i], "--max-batches") == 0 && i + 1 < argc) {
      max_batches = std::max(0, std::atoi(argv[++i]));
    } else if (pos == 0) {
      in_name = argv[i];
      ++pos;
    } else {
      out_name = argv[i];
      ++pos;
    }
  }
  ps_boinc_init_data bd;
  ps_boinc_read_init_data(nullptr, &bd);
  if (bd.gpu_device_num >= 0) device = bd.gpu_device_num;
  if (device < 0) device = 0;
  char in_path[4096], out_path[4096];
  if (ps_boinc_resolve_filename(in_name.c_str(), in_path, sizeof in_path) ||
      ps_boinc_resolve_filename(out_name.c_str(), out_path, sizeof out_path)) {
    std::fprintf(stderr, "period_search: file name too long\n");
    return ps_boinc_finish(&bd, nullptr, 2);
  }
  if (state_path.empty()) state_path = dir_of(out_path) + "period_search_state";
  State st;
  const bool resume = read_state(state_path, &st);

  ps_ctx* ctx = nullptr;
  if (ps_open(device, &ctx) != PS_OK) {
    std::fprintf(stderr, "period_search: cannot open HIP device %d\n", device);
    return ps_boinc_finish(&bd, nullptr, 999);
  }
  auto t0 = std::chrono::steady_clock::now();
  if (ps_prepare_file(ctx, in_path) != PS_OK) {
    std::fprintf(stderr, "period_search: %s\n", ps_last_error(ctx));
    ps_close(ctx);
    return ps_boinc_finish(&bd, nullptr, 2);
  }
  int ints[8];
  double dbls[4];
  ps_info(ctx, ints, dbls);
  const int n_max = ints[6];
  const double escl = dbls[2], conw = dbls[3];
  auto fail_scan = [&]() {
    std::fprintf(stderr, "period_search: %s\n", ps_last_error(ctx));
    ps_close(ctx);
    return ps_boinc_finish(&bd, nullptr, 3);
  };

  // convexity-weight search (period_search_BOINC.cpp:647-771), one round
  // per ps_precalc, a checkpoint after each
  double conw_r = resume ? st.conw_r : conw / escl / escl;
  int new_conw = resume ? st.new_conw : 0;
  int nlines = resume && st.new_conw == 1 ? st.nlines : 0;
  ps_boinc_fraction_done(0.0001);
  while (new_conw != 1 && conw_r * escl * escl < 10.0) {
    if (ps_precalc(ctx, &conw_r, &new_conw) != PS_OK) return fail_scan();
    State s;
    s.nlines = 0;
    s.new_conw = new_conw;
    s.conw_r = conw_r;
    s.out_bytes = 0;
    if (ps_boinc_time_to_checkpoint(&bd)) {
      if (!write_state(state_path, s)) std::fprintf(stderr, "period_search: cannot write %s\n", state_path.c_str());
      ps_boinc_checkpoint_completed();
    }
  }

  // the output: cut back to the checkpointed length when resuming the scan
  FILE* out = nullptr;
  if (resume && st.new_conw == 1 && nlines > 0) {
    if (st.out_bytes >= 0 && truncate(out_path, (off_t)st.out_bytes) != 0) {
      std::fprintf(stderr, "period_search: cannot resume %s\n", out_path);
      ps_close(ctx);
      return ps_boinc_finish(&bd, nullptr, 1);
    }
    out = std::fopen(out_path, "a");
  } else {
    nlines = 0;
    out = std::fopen(out_path, "w");
  }
  if (!out) {
    std::fprintf(stderr, "period_search: cannot write %s\n", out_path);
    ps_close(ctx);
    return ps_boinc_finish(&bd, nullptr, 1);
  }

  // the scan in batches of frequencies (CUDAStart, start_CUDA.cu:792-1245)
  const int batch = batch_freqs > 0 ? batch_freqs : 250 * nfactor;
  std::vector<ps_line> lines(batch);
  int batches = 0;
  for (int n = nlines + 1; n <= n_max; n += batch) {
    const int cnt = std::min(batch, n_max - n + 1);
    if (ps_start(ctx, n, cnt, conw_r, lines.data()) != PS_OK) {
      std::fclose(out);
      return fail_scan();
    }
    for (int k = 0; k < cnt; ++k) {
      const ps_line& L = lines[k];
      if (std::isnan(L.period_h)) continue;  // no pole reported (start_CUDA.cu:1211)
      std::fprintf(out, "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n", L.period_h, L.rms, L.chi2, L.dark_or_conw,
                   std::round(L.lambda_deg), std::round(L.beta_deg));
    }
    std::fflush(out);
    nlines = n + cnt - 1;
    ps_boinc_fraction_done((double)nlines / n_max);
    if (ps_boinc_time_to_checkpoint(&bd)) {
      State s;
      s.nlines = nlines;
      s.new_conw = 1;
      s.conw_r = conw_r;
      s.out_bytes = (long long)std::ftell(out);
      if (fsync(fileno(out)) != 0 || !write_state(state_path, s))
        std::fprintf(stderr, "period_search: cannot write %s\n", state_path.c_str());
      ps_boinc_checkpoint_completed();
    }
    ++batches;
    std::fprintf(stderr, "period_search: frequencies 1..%d of %d done (%.2f%%)\n", nlines, n_max,
                 100.0 * ps_boinc_fraction_done_get());
    if (max_batches > 0 && batches >= max_batches && nlines < n_max) {
      // stop with a resumable state, whatever the checkpoint period
      State s;
      s.nlines = nlines;
      s.new_conw = 1;
      s.conw_r = conw_r;
      s.out_bytes = (long long)std::ftell(out);
      std::fclose(out);
      write_state(state_path, s);
      ps_close(ctx);
      std::fprintf(stderr, "period_search: stopped after %d batches at frequency %d\n", batches, nlines);
      return ps_boinc_finish(&bd, nullptr, 0);
    }
  }
  std::fclose(out);
  ps_boinc_fraction_done(1.0);
  auto t1 = std::chrono::steady_clock::now();
  std::fprintf(stderr, "period_search: %d frequencies x %d poles, conwR %.6g, %.3f s\n", n_max, PS_N_POLES, conw_r,
               std::chrono::duration<double>(t1 - t0).count());
  ps_close(ctx);
  return ps_boinc_finish(&bd, nullptr, 0);
}
