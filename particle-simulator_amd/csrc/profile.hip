// Optional per-kernel timing: while enabled, the library records a HIP event
// pair on the caller's stream around every launch of its hot kernels, so a
// benchmark can read each kernel's own duration (not just whole-call wall
// time).  Process-wide; meant for single-threaded benchmark drivers.
#include <string.h>

#include <mutex>
#include <vector>

#include "dem_common.h"
#include "profile.h"

namespace {
struct Rec {
  const char* name;
  hipEvent_t a, b;
};
std::mutex g_mu;
std::vector<Rec> g_recs;
bool g_on = false;

void clear_locked() {
  for (auto& r : g_recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g_recs.clear();
}
}  // namespace

namespace dem {

int prof_begin(const char* name, hipStream_t st) {
  if (!g_on) return -1;
  std::lock_guard<std::mutex> lk(g_mu);
  Rec r{name, nullptr, nullptr};
  if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return -1;
  if (hipEventRecord(r.a, st) != hipSuccess) return -1;
  g_recs.push_back(r);
  return (int)g_recs.size() - 1;
}

void prof_end(int slot, hipStream_t st) {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (slot < (int)g_recs.size()) (void)hipEventRecord(g_recs[slot].b, st);
}

}  // namespace dem

extern "C" {

int dem_profile_start(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  clear_locked();
  g_on = true;
  return DEM_OK;
}

int dem_profile_stop(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_on = false;
  return DEM_OK;
}

int dem_profile_read(const char* kernel, double* total_ms, int* launches) {
  if (!kernel || !total_ms || !launches)
    return dem::set_error(DEM_ERR_INVALID, "dem_profile_read: null argument");
  std::lock_guard<std::mutex> lk(g_mu);
  double tot = 0.0;
  int cnt = 0;
  for (auto& r : g_recs) {
    if (strcmp(r.name, kernel) != 0) continue;
    DEM_CHECK_HIP(hipEventSynchronize(r.b));
    float ms = 0.f;
    DEM_CHECK_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    tot += ms;
    ++cnt;
  }
  *total_ms = tot;
  *launches = cnt;
  return DEM_OK;
}

}  // extern "C"
