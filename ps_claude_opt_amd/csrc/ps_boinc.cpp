// Host-only BOINC shim of the drop-in binary (include/ps_boinc.h): the
// slot-directory files a BOINC client provides, read without the client
// library.  Follows the calls of the reference's main()
// (period_search_BOINC.cpp:171-907).
#include "../../include/ps_boinc.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {

std::string slurp(const std::string& path, size_t cap) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return std::string();
  std::string s;
  char buf[4096];
  size_t n;
  while (s.size() < cap && (n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
  std::fclose(f);
  return s;
}

// The text between <tag> and </tag>, or "" if absent.
std::string tag_text(const std::string& xml, const char* tag) {
  const std::string open = std::string("<") + tag + ">", close = std::string("</") + tag + ">";
  const size_t a = xml.find(open);
  if (a == std::string::npos) return std::string();
  const size_t b = xml.find(close, a + open.size());
  if (b == std::string::npos) return std::string();
  return xml.substr(a + open.size(), b - a - open.size());
}

std::string in_dir(const char* dir, const char* name) {
  if (!dir || !*dir) return name;
  std::string d(dir);
  if (d.back() != '/') d += '/';
  return d + name;
}

using clock_t_ = std::chrono::steady_clock;
clock_t_::time_point g_last_checkpoint = clock_t_::now();
double g_fraction = 0.0;

}  // namespace

extern "C" {

int ps_boinc_read_init_data(const char* slot_dir, ps_boinc_init_data* out) {
  if (!out) return -1;
  out->under_client = 0;
  out->gpu_device_num = -1;
  out->checkpoint_period = 0.0;
  out->gpu_usage = 0.0;
  const std::string xml = slurp(in_dir(slot_dir, "init_data.xml"), 1 << 20);
  if (xml.empty()) return 0;
  out->under_client = 1;
  const std::string dev = tag_text(xml, "gpu_device_num");
  if (!dev.empty()) out->gpu_device_num = std::atoi(dev.c_str());
  const std::string cp = tag_text(xml, "checkpoint_period");
  if (!cp.empty()) out->checkpoint_period = std::atof(cp.c_str());
  const std::string gu = tag_text(xml, "gpu_usage");
  if (!gu.empty()) out->gpu_usage = std::atof(gu.c_str());
  return 0;
}

int ps_boinc_resolve_filename(const char* logical, char* physical, int cap) {
  if (!logical || !physical || cap <= 0) return -1;
  std::string target(logical);
  const std::string s = slurp(logical, 4096);
  if (s.compare(0, 11, "<soft_link>") == 0) {
    std::string t = tag_text(s, "soft_link");
    while (!t.empty() && (t.back() == '\n' || t.back() == '\r' || t.back() == ' ')) t.pop_back();
    if (!t.empty()) target = t;
  }
  if ((int)target.size() + 1 > cap) return -1;
  std::memcpy(physical, target.c_str(), target.size() + 1);
  return 0;
}

int ps_boinc_time_to_checkpoint(const ps_boinc_init_data* d) {
  if (!d || !d->under_client || d->checkpoint_period <= 0.0) return 1;
  const double dt = std::chrono::duration<double>(clock_t_::now() - g_last_checkpoint).count();
  return dt >= d->checkpoint_period ? 1 : 0;
}

void ps_boinc_checkpoint_completed(void) { g_last_checkpoint = clock_t_::now(); }

double ps_boinc_fraction_done(double f) {
  g_fraction = f < 0.0 ? 0.0 : f > 1.0 ? 1.0 : f;
  return g_fraction;
}

double ps_boinc_fraction_done_get(void) { return g_fraction; }

int ps_boinc_finish(const ps_boinc_init_data* d, const char* slot_dir, int status) {
  if (d && d->under_client) {
    const std::string path = in_dir(slot_dir, "boinc_finish_called");
    if (FILE* f = std::fopen(path.c_str(), "w")) {
      std::fprintf(f, "%d\n", status);
      std::fclose(f);
    }
  }
  return status;
}

}  // extern "C"
