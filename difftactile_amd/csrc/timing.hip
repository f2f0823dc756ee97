// Optional per-stage launch timing of the fused substep path: when enabled,
// every stage launch of dt_mpm_run_fwd / dt_mpm_run_adj / dt_mpm_advance is
// bracketed by a pair of HIP events on the launch stream, so a benchmark can
// read the average duration of each stage over the exact launches of its
// timed region.  Disabled (the default) it costs one branch per launch.
#include <vector>

#include "dt_common.h"
#include "dt_timing.h"

namespace dt {

namespace {
struct Rec {
  int stage;
  hipEvent_t e0, e1;
};
bool g_on = false;
std::vector<hipEvent_t> g_pool;   // events reused across enable() calls
size_t g_used = 0;
std::vector<Rec> g_recs;

hipEvent_t take_event() {
  if (g_used == g_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    g_pool.push_back(e);
  }
  return g_pool[g_used++];
}
}  // namespace

bool timing_enabled() { return g_on; }

StageTimer::StageTimer(int stage, hipStream_t st) : stage_(stage), st_(st), e1_(nullptr) {
  if (!g_on) return;
  hipEvent_t e0 = take_event(), e1 = take_event();
  if (!e0 || !e1) return;
  hipEventRecord(e0, st);
  e1_ = e1;
  g_recs.push_back({stage, e0, e1});
}

StageTimer::~StageTimer() {
  if (e1_) hipEventRecord(e1_, st_);
}

}  // namespace dt

using namespace dt;

extern "C" {

int dt_timing_enable(int32_t on) {
  g_on = on != 0;
  g_used = 0;
  g_recs.clear();
  return DT_OK;
}

int dt_timing_read(int32_t stage, double* total_ms, int64_t* launches) {
  DT_REQUIRE(total_ms && launches, "NULL output");
  double t = 0.0;
  int64_t k = 0;
  for (const Rec& r : g_recs) {
    if (r.stage != stage) continue;
    if (hipEventSynchronize(r.e1) != hipSuccess) {
      set_error("dt_timing_read: event synchronize failed");
      return DT_ELAUNCH;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.e0, r.e1) != hipSuccess) {
      set_error("dt_timing_read: elapsed time failed");
      return DT_ELAUNCH;
    }
    t += ms;
    ++k;
  }
  *total_ms = t;
  *launches = k;
  return DT_OK;
}

}  // extern "C"
