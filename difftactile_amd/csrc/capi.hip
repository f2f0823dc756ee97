// C-ABI plumbing: thread-local error state, launch checks, version, and the
// stable radix sort of particle cell keys.
#include <hipcub/hipcub.hpp>
#include <string>

#include "dt_common.h"

namespace dt {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return DT_ELAUNCH;
  }
  return DT_OK;
}

__global__ void k_iota(int n, int32_t* __restrict__ a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

static size_t cub_bytes(int n, int bits) {
  size_t bytes = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const int32_t*)nullptr, (int32_t*)nullptr,
                                     (const int32_t*)nullptr, (int32_t*)nullptr, n, 0, bits);
  return bytes;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace dt

using namespace dt;

extern "C" {

const char* dt_last_error(void) { return g_last_error.c_str(); }

int dt_version(void) { return 1; }

int64_t dt_sort_tmp_bytes(int32_t n) {
  if (n <= 0) return 0;
  return (int64_t)(align256(sizeof(int32_t) * (size_t)n) + align256(cub_bytes(n, 32)));
}

// Stable LSD radix sort (rocPRIM via hipCUB) of (key, original index) pairs.
int dt_sort_by_key(int32_t n, int32_t key_bits, const int32_t* keys, int32_t* keys_out,
                   int32_t* perm, void* tmp, int64_t tmp_bytes, void* stream) {
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(key_bits > 0 && key_bits <= 31, "key_bits must be in [1, 31]");
  DT_REQUIRE(keys && keys_out && perm && tmp, "NULL pointer");
  DT_REQUIRE(tmp_bytes >= dt_sort_tmp_bytes(n), "tmp buffer smaller than dt_sort_tmp_bytes(n)");
  hipStream_t st = (hipStream_t)stream;
  int32_t* iota = (int32_t*)tmp;
  char* cub_tmp = (char*)tmp + align256(sizeof(int32_t) * (size_t)n);
  size_t cb = (size_t)tmp_bytes - align256(sizeof(int32_t) * (size_t)n);
  hipLaunchKernelGGL(k_iota, dim3(nblocks(n)), dim3(kBlock), 0, st, n, iota);
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(cub_tmp, cb, keys, keys_out, iota, perm, n, 0,
                                                    key_bits, st);
  if (e != hipSuccess) {
    set_error(std::string("dt_sort_by_key: ") + hipGetErrorString(e));
    return DT_ELAUNCH;
  }
  return check_launch("dt_sort_by_key");
}

}  // extern "C"
