// Per-substep binning of particles by the brick of their stencil base
// (counting sort over bricks).
//
// k_bin_count gives every particle its rank inside its brick with one global
// atomic per run of equal keys in a wavefront (cell-sorted particles make long
// runs), an exclusive scan over the brick counts gives the brick starts, and
// k_bin_fill writes the permutation.  k_brick_flags derives the active-brick
// flags (occupied bricks and their +1 neighbours, which a stencil can reach)
// and the occupied-brick list from the counts.  The order inside a brick
// follows the atomics; the brick passes sort each chunk by cell themselves.
#include <hipcub/hipcub.hpp>

#include "dt_bin.h"
#include "dt_device.h"

namespace dt {

// bricks of 4x4x4 nodes over the (n_env n_grid) x n_grid x n_grid grid
long long brick_total(const dt_mpm_params* prm) {
  const long long nb = (prm->n_grid + 3) >> 2;
  const long long nx = (long long)(prm->n_env > 1 ? prm->n_env : 1) * prm->n_grid;
  return ((nx + 3) >> 2) * nb * nb;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t bin_ws_layout(const dt_mpm_params* prm, int n, void* base, BinWs* ws) {
  const long long nbk = brick_total(prm);
  size_t scan_bytes = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const int32_t*)nullptr,
                                   (int32_t*)nullptr, (int)(nbk + 1));
  const size_t sz[11] = {sizeof(int32_t) * (size_t)nbk,
                        sizeof(int32_t) * (size_t)(nbk + 2),
                        sizeof(int32_t) * (size_t)(nbk + 1),
                        sizeof(int32_t) * (size_t)nbk,
                        sizeof(int32_t) * (size_t)n,
                        sizeof(int32_t) * (size_t)n,
                        sizeof(int32_t) * (size_t)n,
                        sizeof(float4) * (size_t)nbk * kTileNodes,
                        scan_bytes,
                        sizeof(int32_t) * (size_t)nbk,
                        sizeof(int32_t)};
  size_t off[11], total = 0;
  for (int i = 0; i < 11; ++i) {
    off[i] = total;
    total += align256(sz[i] ? sz[i] : 1);
  }
  if (base && ws) {
    char* b = static_cast<char*>(base);
    ws->flags = reinterpret_cast<int32_t*>(b + off[0]);
    ws->count = reinterpret_cast<int32_t*>(b + off[1]);
    ws->n_list = ws->count + nbk + 1;
    ws->start = reinterpret_cast<int32_t*>(b + off[2]);
    ws->list = reinterpret_cast<int32_t*>(b + off[3]);
    ws->key = reinterpret_cast<int32_t*>(b + off[4]);
    ws->rank = reinterpret_cast<int32_t*>(b + off[5]);
    ws->perm = reinterpret_cast<int32_t*>(b + off[6]);
    ws->tiles = reinterpret_cast<float4*>(b + off[7]);
    ws->scan_tmp = b + off[8];
    ws->alist = reinterpret_cast<int32_t*>(b + off[9]);
    ws->n_alist = reinterpret_cast<int32_t*>(b + off[10]);
    ws->scan_bytes = scan_bytes;
    ws->n_bricks = nbk;
  }
  return total;
}

__global__ __launch_bounds__(kBlock) void k_bin_zero(long long n, int32_t* __restrict__ a,
                                                     int32_t* __restrict__ one,
                                                     int32_t* __restrict__ two) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
  if (i == 0) {
    *one = 0;
    *two = 0;
  }
}

size_t bin_history_slot_bytes(const dt_mpm_params* prm, int n) {
  const long long nbk = brick_total(prm);
  return align256(sizeof(int32_t) * (size_t)n) + align256(sizeof(int32_t) * (size_t)(nbk + 1)) +
         3 * align256(sizeof(int32_t) * (size_t)nbk) + 512;
}

BinWs bin_slot(const BinWs& ws, int s, int n) {
  BinWs v = ws;
  char* b = ws.hist + (size_t)s * ws.hist_slot_bytes;
  const long long nbk = ws.n_bricks;
  v.perm = reinterpret_cast<int32_t*>(b);
  b += align256(sizeof(int32_t) * (size_t)n);
  v.start = reinterpret_cast<int32_t*>(b);
  b += align256(sizeof(int32_t) * (size_t)(nbk + 1));
  v.list = reinterpret_cast<int32_t*>(b);
  b += align256(sizeof(int32_t) * (size_t)nbk);
  v.flags = reinterpret_cast<int32_t*>(b);
  b += align256(sizeof(int32_t) * (size_t)nbk);
  v.alist = reinterpret_cast<int32_t*>(b);
  b += align256(sizeof(int32_t) * (size_t)nbk);
  v.n_list = reinterpret_cast<int32_t*>(b);
  v.n_alist = reinterpret_cast<int32_t*>(b + 256);
  return v;
}

// Active lanes form a prefix of every wavefront (p = block * 256 + thread).
__global__ __launch_bounds__(kBlock) void k_bin_count(MpmP P, int n, const float* __restrict__ x,
                                                      int32_t* __restrict__ key,
                                                      int32_t* __restrict__ rank,
                                                      int32_t* __restrict__ count) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = p < n;
  int k = -1;
  if (active) {
    float xp[3];
    ld_v3(x, p, pl_vec(P.soa, n), xp);
    k = bin_key(xp, P.inv_dx, P.n_grid, env_x0(P, p)) >> 6;
  }
  const int r = wave_run_rank(active, k, count);
  if (active) {
    key[p] = k;
    rank[p] = r;
  }
}

__global__ __launch_bounds__(kBlock) void k_bin_fill(int n, const int32_t* __restrict__ key,
                                                     const int32_t* __restrict__ rank,
                                                     const int32_t* __restrict__ start,
                                                     int32_t* __restrict__ perm) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) perm[start[key[p]] + rank[p]] = p;
}

// Appends the lanes with pred set to list (one atomic per wavefront).
__device__ __forceinline__ void wave_append(bool pred, int32_t value, int32_t* __restrict__ list,
                                            int32_t* __restrict__ n_list) {
  const unsigned long long m = __ballot(pred);
  if (m == 0ull) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(n_list, __popcll(m));
  base = __shfl(base, leader);
  if (pred) list[base + __popcll(m & ((1ull << lane) - 1ull))] = value;
}

// Per brick: flags = the brick or one of its -1 neighbours holds particles (a
// stencil reaches at most one brick past its base's); list += occupied
// bricks, alist += flagged bricks.
__global__ __launch_bounds__(kBlock) void k_brick_flags(long long n_bricks, int nb,
                                                        const int32_t* __restrict__ count,
                                                        int32_t* __restrict__ flags,
                                                        int32_t* __restrict__ list,
                                                        int32_t* __restrict__ n_list,
                                                        int32_t* __restrict__ alist,
                                                        int32_t* __restrict__ n_alist) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  bool occ = false, act = false;
  if (b < n_bricks) {
    const int bk = (int)(b % nb), bj = (int)((b / nb) % nb), bi = (int)(b / ((long long)nb * nb));
    int any = 0;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      const int i = bi - (d >> 2), j = bj - ((d >> 1) & 1), k = bk - (d & 1);
      if (i >= 0 && j >= 0 && k >= 0) any |= count[((long long)i * nb + j) * nb + k];
    }
    act = any != 0;
    flags[b] = act ? 1 : 0;
    occ = count[b] > 0;
  }
  wave_append(occ, (int32_t)b, list, n_list);
  wave_append(act, (int32_t)b, alist, n_alist);
}

int launch_bin(const MpmP& P, int n, const float* x, const BinWs& ws, hipStream_t st) {
  const long long nbk = ws.n_bricks;
  const int nb = (P.n_grid + 3) >> 2;
  hipLaunchKernelGGL(k_bin_zero, dim3(nblocks(nbk + 2)), dim3(kBlock), 0, st, nbk + 2, ws.count,
                     ws.n_list, ws.n_alist);
  if (n > 0)
    hipLaunchKernelGGL(k_bin_count, dim3(nblocks(n)), dim3(kBlock), 0, st, P, n, x, ws.key,
                       ws.rank, ws.count);
  size_t bytes = ws.scan_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(ws.scan_tmp, bytes, ws.count, ws.start, (int)(nbk + 1),
                                       st) != hipSuccess) {
    set_error("binning scan failed");
    return DT_ELAUNCH;
  }
  hipLaunchKernelGGL(k_brick_flags, dim3(nblocks(nbk)), dim3(kBlock), 0, st, nbk, nb,
                     (const int32_t*)ws.count, ws.flags, ws.list, ws.n_list, ws.alist,
                     ws.n_alist);
  if (n > 0)
    hipLaunchKernelGGL(k_bin_fill, dim3(nblocks(n)), dim3(kBlock), 0, st, n, ws.key, ws.rank,
                       ws.start, ws.perm);
  return check_launch("particle binning");
}

}  // namespace dt
