// HIP-graph cache for the trajectory loops.  A 50-substep segment issues
// ~900 small launches; replaying them as one graph removes the per-launch
// host and dispatch gaps.  Graphs are keyed by every argument of the loop, so
// a replay is exactly the launch sequence that was captured; the cache keeps
// the most recently used entries.  The captured sequences contain kernels
// only (no memset or memcpy nodes).
#include <atomic>
#include <list>
#include <mutex>
#include <utility>

#include "dt_graph.h"

namespace dt {

namespace {
struct Entry {
  std::vector<uint64_t> key;
  hipGraphExec_t exec;
};
// a 1000-substep checkpointed trajectory replays ~50 distinct segment graphs
constexpr size_t kMaxGraphs = 256;
std::mutex g_mu;
std::list<Entry> g_cache;   // front = most recently used
bool g_enabled = true;
long long g_hits = 0, g_misses = 0;
// launch shapes: atomics, read by launches that run inside run_graphed's lock
std::atomic<int32_t> g_brick_blocks{0}, g_node_blocks{0}, g_sort_blocks{0};
std::vector<std::pair<int, hipStream_t>> g_capture_streams;   // per device

hipStream_t capture_stream(int dev) {
  for (auto& d : g_capture_streams)
    if (d.first == dev) return d.second;
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
  g_capture_streams.push_back({dev, s});
  return s;
}
}  // namespace

int run_graphed(std::vector<uint64_t> key, hipStream_t st,
                const std::function<int(hipStream_t)>& body) {
  if (!g_enabled || timing_enabled()) return body(st);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return body(st);
  key.push_back((uint64_t)dev);
  std::lock_guard<std::mutex> lock(g_mu);
  // the launch shapes the body reads
  const dt_launch_config lc = launch_config();
  key.push_back(((uint64_t)(uint32_t)lc.brick_blocks << 32) | (uint32_t)lc.node_blocks);
  key.push_back((uint64_t)(uint32_t)lc.sort_blocks);
  for (auto it = g_cache.begin(); it != g_cache.end(); ++it) {
    if (it->key == key) {
      g_cache.splice(g_cache.begin(), g_cache, it);
      ++g_hits;
      if (hipGraphLaunch(g_cache.front().exec, st) != hipSuccess) {
        set_error("hipGraphLaunch failed");
        return DT_ELAUNCH;
      }
      return DT_OK;
    }
  }
  ++g_misses;
  hipStream_t cs = capture_stream(dev);
  if (!cs) return body(st);
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return body(st);
  }
  int rc = body(cs);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(cs, &g);
  if (rc != DT_OK || e != hipSuccess || g == nullptr) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    if (rc != DT_OK) return rc;
    set_error("graph capture failed");
    return DT_ELAUNCH;
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    set_error("hipGraphInstantiate failed");
    return DT_ELAUNCH;
  }
  g_cache.push_front({std::move(key), exec});
  while (g_cache.size() > kMaxGraphs) {
    (void)hipGraphExecDestroy(g_cache.back().exec);
    g_cache.pop_back();
  }
  if (hipGraphLaunch(exec, st) != hipSuccess) {
    set_error("hipGraphLaunch failed");
    return DT_ELAUNCH;
  }
  return DT_OK;
}

namespace {
// graphs captured around a caller's own launch sequence (dt_graph_begin /
// dt_graph_end): keys start with this tag, apart from run_graphed's
constexpr uint64_t kCallerTag = 0x43414c4c45524b59ull;
std::vector<uint64_t> caller_key(const uint64_t* key, int n, int dev) {
  std::vector<uint64_t> k;
  k.reserve(n + 4);
  k.push_back(kCallerTag);
  for (int i = 0; i < n; ++i) k.push_back(key[i]);
  k.push_back((uint64_t)dev);
  const dt_launch_config lc = launch_config();
  k.push_back(((uint64_t)(uint32_t)lc.brick_blocks << 32) | (uint32_t)lc.node_blocks);
  k.push_back((uint64_t)(uint32_t)lc.sort_blocks);
  return k;
}
}  // namespace

dt_launch_config launch_config() {
  dt_launch_config c;
  c.brick_blocks = g_brick_blocks.load();
  c.node_blocks = g_node_blocks.load();
  c.sort_blocks = g_sort_blocks.load();
  if (c.brick_blocks <= 0) c.brick_blocks = 2048;
  if (c.node_blocks <= 0) c.node_blocks = 1024;
  if (c.sort_blocks <= 0) c.sort_blocks = 2048;
  return c;
}

}  // namespace dt

using namespace dt;

extern "C" {

int dt_graphs_enable(int32_t on) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_enabled = on != 0;
  return DT_OK;
}

int dt_graph_replay(void* stream, const uint64_t* key, int32_t n_key, int32_t* hit) {
  DT_REQUIRE(hit && (key || n_key == 0) && n_key >= 0, "bad key");
  *hit = 0;
  if (!g_enabled || timing_enabled()) return DT_OK;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return DT_OK;
  const std::vector<uint64_t> k = caller_key(key, n_key, dev);
  std::lock_guard<std::mutex> lock(g_mu);
  for (auto it = g_cache.begin(); it != g_cache.end(); ++it) {
    if (it->key == k) {
      g_cache.splice(g_cache.begin(), g_cache, it);
      ++g_hits;
      if (hipGraphLaunch(g_cache.front().exec, (hipStream_t)stream) != hipSuccess) {
        set_error("hipGraphLaunch failed");
        return DT_ELAUNCH;
      }
      *hit = 1;
      return DT_OK;
    }
  }
  return DT_OK;
}

int dt_graph_begin(void* stream, int32_t* capturing) {
  DT_REQUIRE(capturing, "NULL output");
  *capturing = 0;
  if (!g_enabled || timing_enabled()) return DT_OK;
  if (hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return DT_OK;
  }
  *capturing = 1;
  return DT_OK;
}

int dt_graph_end(void* stream, void* launch_stream, const uint64_t* key, int32_t n_key) {
  DT_REQUIRE((key || n_key == 0) && n_key >= 0, "bad key");
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g);
  if (e != hipSuccess || g == nullptr) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    set_error("graph capture failed");
    return DT_ELAUNCH;
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    set_error("hipGraphInstantiate failed");
    return DT_ELAUNCH;
  }
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(g_mu);
  ++g_misses;
  g_cache.push_front({caller_key(key, n_key, dev), exec});
  while (g_cache.size() > kMaxGraphs) {
    (void)hipGraphExecDestroy(g_cache.back().exec);
    g_cache.pop_back();
  }
  // the captured work has not run yet
  if (hipGraphLaunch(exec, (hipStream_t)launch_stream) != hipSuccess) {
    set_error("hipGraphLaunch failed");
    return DT_ELAUNCH;
  }
  return DT_OK;
}

int dt_graph_abort(void* stream) {
  hipGraph_t g = nullptr;
  (void)hipStreamEndCapture((hipStream_t)stream, &g);
  if (g) (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  return DT_OK;
}

int dt_set_launch_config(const dt_launch_config* cfg) {
  DT_REQUIRE(cfg, "NULL config");
  DT_REQUIRE(cfg->brick_blocks >= 0 && cfg->brick_blocks <= 65536 && cfg->node_blocks >= 0 &&
                 cfg->node_blocks <= 65536 && cfg->sort_blocks >= 0 && cfg->sort_blocks <= 65536,
             "launch shapes must be in [0, 65536]");
  g_brick_blocks.store(cfg->brick_blocks);
  g_node_blocks.store(cfg->node_blocks);
  g_sort_blocks.store(cfg->sort_blocks);
  return DT_OK;
}

int dt_get_launch_config(dt_launch_config* cfg) {
  DT_REQUIRE(cfg, "NULL config");
  *cfg = launch_config();
  return DT_OK;
}

int dt_graph_stats(int64_t* hits, int64_t* misses) {
  DT_REQUIRE(hits && misses, "NULL output");
  std::lock_guard<std::mutex> lock(g_mu);
  *hits = g_hits;
  *misses = g_misses;
  return DT_OK;
}

}  // extern "C"
