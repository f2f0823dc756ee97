// Per-substep binning of particles by the brick of their stencil base
// (counting sort over bricks).
//
// k_bin_count gives every particle its rank inside its brick with one global
// atomic per run of equal keys in a wavefront (cell-sorted particles make long
// runs), an exclusive scan over the brick counts gives the brick starts, and
// k_bin_fill writes the permutation.  The active-brick list (occupied bricks
// and their +1 neighbours, which a stencil can reach) and the occupied-brick
// list come from the counts.  The order inside a brick
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
  const size_t sz[10] = {sizeof(int32_t) * (size_t)(nbk + 2),
                        sizeof(int32_t) * (size_t)(nbk + 1),
                        sizeof(int32_t) * (size_t)nbk,
                        sizeof(int32_t) * (size_t)n,
                        sizeof(int32_t) * (size_t)n,
                        sizeof(int32_t) * (size_t)n,
                        sizeof(float4) * (size_t)nbk * kTileNodes,
                        scan_bytes,
                        sizeof(int32_t) * (size_t)nbk,
                        sizeof(int32_t)};
  size_t off[10], total = 0;
  for (int i = 0; i < 10; ++i) {
    off[i] = total;
    total += align256(sz[i] ? sz[i] : 1);
  }
  if (base && ws) {
    char* b = static_cast<char*>(base);
    ws->count = reinterpret_cast<int32_t*>(b + off[0]);
    ws->n_list = ws->count + nbk + 1;
    ws->start = reinterpret_cast<int32_t*>(b + off[1]);
    ws->list = reinterpret_cast<int32_t*>(b + off[2]);
    ws->key = reinterpret_cast<int32_t*>(b + off[3]);
    ws->rank = reinterpret_cast<int32_t*>(b + off[4]);
    ws->perm = reinterpret_cast<int32_t*>(b + off[5]);
    ws->tiles = reinterpret_cast<float4*>(b + off[6]);
    ws->scan_tmp = b + off[7];
    ws->alist = reinterpret_cast<int32_t*>(b + off[8]);
    ws->n_alist = reinterpret_cast<int32_t*>(b + off[9]);
    ws->scan_bytes = scan_bytes;
    ws->n_bricks = nbk;
    ws->nb = (prm->n_grid + 3) >> 2;
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
         2 * align256(sizeof(int32_t) * (size_t)nbk) + 512;
}

BinWs bin_slot(const BinWs& ws, int s, int n) {
  return bin_view(ws, ws.hist + (size_t)s * ws.hist_slot_bytes, n);
}

BinWs bin_view(const BinWs& ws, char* b, int n) {
  BinWs v = ws;
  const long long nbk = ws.n_bricks;
  v.perm = reinterpret_cast<int32_t*>(b);
  b += align256(sizeof(int32_t) * (size_t)n);
  v.start = reinterpret_cast<int32_t*>(b);
  b += align256(sizeof(int32_t) * (size_t)(nbk + 1));
  v.list = reinterpret_cast<int32_t*>(b);
  b += align256(sizeof(int32_t) * (size_t)nbk);
  v.alist = reinterpret_cast<int32_t*>(b);
  b += align256(sizeof(int32_t) * (size_t)nbk);
  v.n_list = reinterpret_cast<int32_t*>(b);
  v.n_alist = reinterpret_cast<int32_t*>(b + 256);
  return v;
}

// Copies one binning into another view: perm, brick starts, and the
// occupied- and active-brick lists up to their device-side counts.
__global__ __launch_bounds__(kBlock) void k_bin_copy(
    int n, long long nbk, const int32_t* __restrict__ perm, const int32_t* __restrict__ start,
    const int32_t* __restrict__ list, const int32_t* __restrict__ n_list,
    const int32_t* __restrict__ alist, const int32_t* __restrict__ n_alist,
    int32_t* __restrict__ perm2, int32_t* __restrict__ start2, int32_t* __restrict__ list2,
    int32_t* __restrict__ n_list2, int32_t* __restrict__ alist2, int32_t* __restrict__ n_alist2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int no = *n_list, na = *n_alist;
  if (i < n) perm2[i] = perm[i];
  if (i <= nbk) start2[i] = start[i];
  if (i < no) list2[i] = list[i];
  if (i < na) alist2[i] = alist[i];
  if (i == 0) {
    *n_list2 = no;
    *n_alist2 = na;
  }
}

void launch_bin_copy(const BinWs& from, const BinWs& to, int n, hipStream_t st) {
  const long long nbk = from.n_bricks;
  const long long m = (long long)n > nbk + 1 ? (long long)n : nbk + 1;
  hipLaunchKernelGGL(k_bin_copy, dim3(nblocks(m)), dim3(kBlock), 0, st, n, nbk,
                     (const int32_t*)from.perm, (const int32_t*)from.start,
                     (const int32_t*)from.list, (const int32_t*)from.n_list,
                     (const int32_t*)from.alist, (const int32_t*)from.n_alist, to.perm, to.start,
                     to.list, to.n_list, to.alist, to.n_alist);
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

// Per brick: active = the brick or one of its -1 neighbours holds particles (a
// stencil reaches at most one brick past its base's); list += occupied
// bricks, alist += active bricks.
__global__ __launch_bounds__(kBlock) void k_brick_flags(long long n_bricks, int nb,
                                                        const int32_t* __restrict__ count,
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
    occ = count[b] > 0;
  }
  wave_append(occ, (int32_t)b, list, n_list);
  wave_append(act, (int32_t)b, alist, n_alist);
}

// One workgroup for grids of at most 32 bricks per axis: the brick counts
// (plus the zero sentinel) go to LDS, each thread scans a run of 33 of them
// after a block scan of the run totals, and the occupied / active lists come
// from one 32-bit occupancy word per brick row (i, j) (bit k), so a brick's
// active test (it or a -1 neighbour occupied) is a shift and an OR of four
// words.  The counts are zeroed behind the read, so the next binning on this
// workspace needs no separate clear.
constexpr int kScanThreads = 1024;
constexpr int kScanRun = 33;
constexpr int kScanEdge = 32;   // bricks per axis

__global__ __launch_bounds__(kScanThreads) void k_bin_scan_lists(
    int nb, int32_t* __restrict__ count, int32_t* __restrict__ start, int32_t* __restrict__ list,
    int32_t* __restrict__ n_list, int32_t* __restrict__ alist, int32_t* __restrict__ n_alist) {
  // element e at e + e / 33: the 33 elements of a run are consecutive and
  // consecutive runs start two banks apart
  __shared__ int32_t c[kScanThreads * (kScanRun + 1)];
  __shared__ uint32_t word[kScanThreads];
  __shared__ int32_t wsum[kScanThreads / 64], wocc[kScanThreads / 64], wact[kScanThreads / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int nbk = nb * nb * nb, ne = nbk + 1;
  {
    int v[kScanRun];
#pragma unroll
    for (int r = 0; r < kScanRun; ++r) {
      const int e = t + r * kScanThreads;
      v[r] = e < ne ? count[e] : 0;
    }
#pragma unroll
    for (int r = 0; r < kScanRun; ++r) {
      const int e = t + r * kScanThreads;
      if (e < ne) count[e] = 0;
      c[e + e / kScanRun] = v[r];
    }
  }
  __syncthreads();
  // run totals, block exclusive scan over the 1024 runs
  int* run = c + t * (kScanRun + 1);
  int tot = 0;
#pragma unroll
  for (int r = 0; r < kScanRun; ++r) tot += run[r];
  int incl = tot;
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(incl, d);
    if (lane >= d) incl += u;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int base = incl - tot;
  for (int q = 0; q < w; ++q) base += wsum[q];
#pragma unroll
  for (int r = 0; r < kScanRun; ++r) {
    const int v = run[r];
    run[r] = base;
    base += v;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kScanRun; ++r) {
    const int e = t + r * kScanThreads;
    if (e < ne) start[e] = c[e + e / kScanRun];
  }
  // occupancy word of brick row t = (i, j): bit k set when brick (i, j, k)
  // holds particles
  const int nrow = nb * nb;
  uint32_t occ = 0u;
  if (t < nrow) {
    const int b0 = t * nb;
    int prev = c[b0 + b0 / kScanRun];
    for (int k = 0; k < nb; ++k) {
      const int e = b0 + k + 1;
      const int nxt = c[e + e / kScanRun];
      if (nxt > prev) occ |= 1u << k;
      prev = nxt;
    }
    word[t] = occ;
  }
  __syncthreads();
  uint32_t act = 0u;
  if (t < nrow) {
    const int i = t / nb, j = t - i * nb;
    uint32_t x = occ;
    if (j > 0) x |= word[t - 1];
    if (i > 0) x |= word[t - nb];
    if (i > 0 && j > 0) x |= word[t - nb - 1];
    act = x | (x << 1);
    if (nb < 32) act &= (1u << nb) - 1u;
  }
  // both lists ascending: per-row counts, a block scan, then each row writes
  const int no = __popc(occ), na = __popc(act);
  int io = no, ia = na;
  for (int d = 1; d < 64; d <<= 1) {
    const int uo = __shfl_up(io, d), ua = __shfl_up(ia, d);
    if (lane >= d) {
      io += uo;
      ia += ua;
    }
  }
  if (lane == 63) {
    wocc[w] = io;
    wact[w] = ia;
  }
  __syncthreads();
  int bo = io - no, ba = ia - na;
  for (int q = 0; q < w; ++q) {
    bo += wocc[q];
    ba += wact[q];
  }
  const int rb = t * nb;
  for (uint32_t m = occ; m; m &= m - 1u) list[bo++] = rb + __ffs((int)m) - 1;
  for (uint32_t m = act; m; m &= m - 1u) alist[ba++] = rb + __ffs((int)m) - 1;
  if (t == kScanThreads - 1) {
    *n_list = bo;
    *n_alist = ba;
  }
}

// one workgroup when the grid is a single environment of at most 32 bricks
// per axis; that scan leaves the counts zero
static bool one_block_scan(const BinWs& ws, int nb) {
  return nb <= kScanEdge && ws.n_bricks == (long long)nb * nb * nb;
}

bool bin_counts_fusable(const BinWs& ws) { return one_block_scan(ws, ws.nb); }

int launch_bin(const MpmP& P, int n, const float* x, const BinWs& ws, hipStream_t st,
               bool count_clear, bool counted) {
  const long long nbk = ws.n_bricks;
  const int nb = (P.n_grid + 3) >> 2;
  const bool one_block = one_block_scan(ws, nb);
  if (counted && !one_block) {
    set_error("binning: fused counts need the one-workgroup scan");
    return DT_EINVAL;
  }
  if (!counted) {
    if (!one_block || !count_clear)
      hipLaunchKernelGGL(k_bin_zero, dim3(nblocks(nbk + 2)), dim3(kBlock), 0, st, nbk + 2,
                         ws.count, ws.n_list, ws.n_alist);
    if (n > 0)
      hipLaunchKernelGGL(k_bin_count, dim3(nblocks(n)), dim3(kBlock), 0, st, P, n, x, ws.key,
                         ws.rank, ws.count);
  }
  if (one_block) {
    hipLaunchKernelGGL(k_bin_scan_lists, dim3(1), dim3(kScanThreads), 0, st, nb, ws.count,
                       ws.start, ws.list, ws.n_list, ws.alist, ws.n_alist);
  } else {
    size_t bytes = ws.scan_bytes;
    if (hipcub::DeviceScan::ExclusiveSum(ws.scan_tmp, bytes, ws.count, ws.start, (int)(nbk + 1),
                                         st) != hipSuccess) {
      set_error("binning scan failed");
      return DT_ELAUNCH;
    }
    hipLaunchKernelGGL(k_brick_flags, dim3(nblocks(nbk)), dim3(kBlock), 0, st, nbk, nb,
                       (const int32_t*)ws.count, ws.list, ws.n_list, ws.alist,
                       ws.n_alist);
  }
  if (n > 0)
    hipLaunchKernelGGL(k_bin_fill, dim3(nblocks(n)), dim3(kBlock), 0, st, n, ws.key, ws.rank,
                       ws.start, ws.perm);
  return check_launch("particle binning");
}

}  // namespace dt
