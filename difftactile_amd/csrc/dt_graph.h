// HIP-graph replay of the trajectory loops (see graph.hip).
#pragma once
#include <functional>
#include <vector>

#include "dt_common.h"

namespace dt {

bool timing_enabled();

// Runs body(stream) on st: captured once into a HIP graph (on a private
// capture stream) and replayed with one hipGraphLaunch for every later call
// with the same key, or directly when graphs are disabled or timing is on.
// The key must contain every value the launches depend on (pointers, sizes,
// parameters); the graph replays them verbatim.
int run_graphed(std::vector<uint64_t> key, hipStream_t st,
                const std::function<int(hipStream_t)>& body);

// Appends the bytes of a trivially copyable value to a key.
template <class T>
void key_add(std::vector<uint64_t>& key, const T& v) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
  for (size_t i = 0; i < sizeof(T); i += 8) {
    uint64_t w = 0;
    for (size_t j = 0; j < 8 && i + j < sizeof(T); ++j) w |= (uint64_t)p[i + j] << (8 * j);
    key.push_back(w);
  }
}

}  // namespace dt
