// period_search: drop-in command line of the reference application
// (period_search_BOINC.cpp:171-907) on the MI355X kernels.
//
//   period_search [--device N] [period_search_in [period_search_out]]
//
// Reads period_search_in, runs the convexity-weight search and the scan,
// writes one line per trial frequency:
//   "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n" = period[h], rms, chi^2,
//   dark facet % (first line: the final convexity weight), lambda, beta.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ps_mi355x.h"

int main(int argc, char** argv) {
  int device = 0;
  const char* in = "period_search_in";
  const char* outp = "period_search_out";
  int pos = 0;
  for (int i = 1; i < argc; ++i) {
    if (std::strcmp(argv[i], "--device") == 0 && i + 1 < argc) {
      device = std::atoi(argv[++i]);
    } else if (std::strcmp(argv[i], "--N") == 0 && i + 1 < argc) {
      ++i;  // accepted for compatibility, unused (as in the reference)
    } else if (pos == 0) {
      in = argv[i];
      ++pos;
    } else {
      outp = argv[i];
      ++pos;
    }
  }
  ps_ctx* ctx = nullptr;
  if (ps_open(device, &ctx) != PS_OK) {
    std::fprintf(stderr, "period_search: cannot open HIP device %d\n", device);
    return 999;
  }
  auto t0 = std::chrono::steady_clock::now();
  int rc = ps_prepare_file(ctx, in);
  if (rc != PS_OK) {
    std::fprintf(stderr, "period_search: %s\n", ps_last_error(ctx));
    ps_close(ctx);
    return 2;
  }
  int ints[8];
  double dbls[4];
  ps_info(ctx, ints, dbls);
  double conw_r = 0;
  rc = ps_conw_search(ctx, &conw_r);
  if (rc == PS_OK) {
    std::vector<ps_line> lines(ints[6]);
    rc = ps_start(ctx, 1, ints[6], conw_r, lines.data());
    if (rc == PS_OK) {
      FILE* f = std::fopen(outp, "w");
      if (!f) {
        std::fprintf(stderr, "period_search: cannot write %s\n", outp);
        ps_close(ctx);
        return 1;
      }
      for (const ps_line& L : lines) {
        if (std::isnan(L.period_h)) continue;
        std::fprintf(f, "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n", L.period_h, L.rms, L.chi2, L.dark_or_conw,
                     std::round(L.lambda_deg), std::round(L.beta_deg));
      }
      std::fclose(f);
    }
  }
  if (rc != PS_OK) {
    std::fprintf(stderr, "period_search: %s\n", ps_last_error(ctx));
    ps_close(ctx);
    return 3;
  }
  auto t1 = std::chrono::steady_clock::now();
  std::fprintf(stderr, "period_search: %d frequencies x %d poles, conwR %.6g, %.3f s\n", ints[6], PS_N_POLES,
               conw_r, std::chrono::duration<double>(t1 - t0).count());
  ps_close(ctx);
  return 0;
}
