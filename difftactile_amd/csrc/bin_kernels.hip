// Per-substep binning of particles by the brick of their stencil base
// (counting sort over bricks).
//
// k_bin_count (or a G2P's fused count) counts every brick's particles with
// one global atomic per distinct brick of a wavefront and records that
// (wavefront, brick) run in the brick's run table; an exclusive scan over the
// brick counts gives the brick starts, and k_bin_fill writes the permutation.
// The active-brick list (occupied bricks and their +1 neighbours, which a
// stencil can reach) and the occupied-brick list come from the counts.  The
// fill orders a brick's particles by index from its run table; a brick whose
// runs overflow the table is filled in the atomics' order and k_bin_sort
// sorts it by index, so the permutation is the same on every run.
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
  const size_t sz[18] = {sizeof(int32_t) * (size_t)(nbk + 2),
                        sizeof(int32_t) * (size_t)(nbk + 1),
                        sizeof(int32_t) * (size_t)nbk,
                        sizeof(int32_t) * (size_t)n,
                        sizeof(int32_t) * (size_t)n,
                        sizeof(int32_t) * (size_t)n,
                        sizeof(float4) * (size_t)nbk * kTileNodes,
                        scan_bytes,
                        sizeof(int32_t) * (size_t)nbk,
                        sizeof(int32_t),
                        sizeof(int32_t) * (size_t)(kSlabMax + kSlabMax * kSlabMax),
                        sizeof(float) * 32 * (size_t)nblocks(n),
                        sizeof(int32_t) * (size_t)nbk,
                        sizeof(int32_t) * (size_t)nbk,
                        sizeof(int32_t) * (size_t)nbk * kRuns,
                        sizeof(int32_t) * (size_t)nbk,
                        sizeof(int32_t) * 2,
                        sizeof(int32_t) * (size_t)nbk};
  size_t off[18], total = 0;
  for (int i = 0; i < 18; ++i) {
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
    ws->slab_sum = reinterpret_cast<int32_t*>(b + off[10]);
    ws->row_occ = reinterpret_cast<uint32_t*>(ws->slab_sum + kSlabMax);
    ws->mat_part = reinterpret_cast<float*>(b + off[11]);
    ws->nruns = reinterpret_cast<int32_t*>(b + off[12]);
    ws->nsnap = reinterpret_cast<int32_t*>(b + off[13]);
    ws->runs = reinterpret_cast<int32_t*>(b + off[14]);
    ws->ovf = reinterpret_cast<int32_t*>(b + off[15]);
    ws->n_ovf = reinterpret_cast<int32_t*>(b + off[16]);
    ws->lstart = reinterpret_cast<int32_t*>(b + off[17]);
    ws->scan_bytes = scan_bytes;
    ws->n_bricks = nbk;
    ws->nb = (prm->n_grid + 3) >> 2;
  }
  return total;
}

// clears a[0, n), b[0, nb) and three counters
__global__ __launch_bounds__(kBlock) void k_bin_zero(long long n, int32_t* __restrict__ a,
                                                     long long nb, int32_t* __restrict__ b,
                                                     int32_t* __restrict__ one,
                                                     int32_t* __restrict__ two,
                                                     int32_t* __restrict__ three) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
  if (i < nb) b[i] = 0;
  if (i == 0) {
    *one = 0;
    *two = 0;
    *three = 0;
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

__global__ __launch_bounds__(kBlock) void k_bin_count(MpmP P, int n, const float* __restrict__ x,
                                                      int32_t* __restrict__ key,
                                                      int32_t* __restrict__ rank,
                                                      int32_t* __restrict__ count,
                                                      BinRuns R) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = p < n;
  int k = -1;
  if (active) {
    float xp[3];
    ld_v3(x, p, pl_vec(P.soa, n), xp);
    k = bin_key(xp, at_particle(P, p).inv_dx, P.n_grid, env_x0(P, p)) >> 6;
  }
  const int r = wave_key_rank_runs(active, k, count, R, p >> 6);
  if (active) {
    key[p] = k;
    rank[p] = r;
  }
}

// perm[start[key] + place] = p.  A brick of at most kRuns runs: place = the
// lanes of its runs from earlier wavefronts + the particle's rank among its
// wavefront's lanes of the brick (ascending index order).  More runs: place =
// the counting atomics' rank, and k_bin_sort orders the brick afterwards.
// nsnap: runs per brick of this binning.  Wavefront w covers indices
// [64 w, 64 w + 64), as in the counting pass.  add: added to start[key].
__device__ __forceinline__ void fill_particle(int p, int k, int lr, int add,
                                              const int32_t* __restrict__ rank,
                                              const int32_t* __restrict__ start,
                                              const int32_t* __restrict__ nsnap,
                                              const int32_t* __restrict__ runs,
                                              int32_t* __restrict__ perm) {
  // the row's first eight entries are loaded with the run count and the
  // brick start (entries past the count are left over from earlier
  // binnings and masked out)
  const int32_t* rt = runs + (long long)k * kRuns;
  const int nr = nsnap[k];
  const int s0 = start[k] + add;
  int e8[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) e8[q] = rt[q];
  const int w = p >> 6;
  int place;
  if (nr <= kRuns) {
    place = lr;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < nr && (e8[q] >> 6) < w) place += (e8[q] & 63) + 1;
    for (int q = 8; q < nr; ++q) {
      const int e = rt[q];
      if ((e >> 6) < w) place += (e & 63) + 1;
    }
  } else {
    place = rank[p];
  }
  perm[s0 + place] = p;
}

__global__ __launch_bounds__(kBlock) void k_bin_fill(int n, const int32_t* __restrict__ key,
                                                     const int32_t* __restrict__ rank,
                                                     const int32_t* __restrict__ start,
                                                     const int32_t* __restrict__ nsnap,
                                                     const int32_t* __restrict__ runs,
                                                     int32_t* __restrict__ perm) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = p < n;
  const int k = active ? key[p] : -1;
  const int lr = wave_lane_rank(active, k);
  if (active) fill_particle(p, k, lr, 0, rank, start, nsnap, runs, perm);
}

// Sorts the particles perm[start[b] .. start[b+1]) of each brick on the
// overflow list by index (ascending).  Blocks stride over the list.  A brick of at most
// kSortCap particles: every run of 64 is sorted in a wavefront's registers (a
// bitonic network over the lanes, no barriers), then each particle's place is
// its rank in its run plus, for every other run, the number of its entries
// below it (a binary search; indices are distinct).  A larger brick is sorted
// so in chunks of kSortCap, and merged by the same rule through tmp (n
// entries) and copied back.
constexpr int kSortCap = 2048;

// ascending sort of one value per lane over the 64 lanes
__device__ __forceinline__ int wave_sort64(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int o = __shfl_xor(v, j);
      const bool up = (lane & k) == 0, lower = (lane & j) == 0;
      v = (lower == up) ? min(v, o) : max(v, o);
    }
  }
  return v;
}

// entries of the ascending a[0, len) below v
__device__ __forceinline__ int count_below(const int32_t* a, int len, int v) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// sorts src[0, m) (m <= kSortCap) into dst[0, m); a, b: LDS of kSortCap
__device__ __forceinline__ void block_sort(const int32_t* src, int32_t* dst, int m, int32_t* a,
                                           int32_t* b) {
  // runs of 64 sorted in registers (padding INT_MAX sorts to the end)
  for (int i0 = 0; i0 < m; i0 += blockDim.x) {
    const int i = i0 + threadIdx.x;
    const int v = wave_sort64(i < m ? src[i] : 0x7fffffff);
    if (i < ((m + 63) & ~63)) a[i] = v;
  }
  __syncthreads();
  const int runs = (m + 63) >> 6;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const int v = a[i], r = i >> 6;
    int place = i & 63;
    for (int q = 0; q < runs; ++q)
      if (q != r) place += count_below(a + 64 * q, min(64, m - 64 * q), v);
    b[place] = v;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < m; i += blockDim.x) dst[i] = b[i];
  __syncthreads();
}

// The fill gives each (wavefront, brick) pair one run of consecutive slots
// holding that wavefront's particles of the brick in lane order (ascending
// indices; wavefront w covers indices [64 w, 64 w + 64)), with the runs in the
// order their atomics ran.  Ordering the runs by wavefront therefore sorts the
// brick: the runs (at most 64 here) are found by a block scan of the run
// heads, sorted by wavefront in one wavefront's registers, and moved to their
// places.  Returns false, doing nothing, when the brick has more than 64 runs.
__device__ __forceinline__ bool block_order_runs(int32_t* seg, int m, int32_t* a, int32_t* b,
                                                 int32_t* run_at, int32_t* run_to,
                                                 int32_t* wsum) {
  constexpr int kPer = kSortCap / kBlock;   // elements per thread
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int i = t; i < m; i += blockDim.x) a[i] = seg[i];
  __syncthreads();
  // heads of the thread's kPer consecutive elements
  const int i0 = t * kPer;
  int heads = 0;
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int i = i0 + u;
    if (i < m && (i == 0 || (a[i] >> 6) != (a[i - 1] >> 6))) ++heads;
  }
  int incl = heads;
  for (int d = 1; d < 64; d <<= 1) {
    const int v = __shfl_up(incl, d);
    if (lane >= d) incl += v;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int base = incl - heads;
  for (int q = 0; q < w; ++q) base += wsum[q];
  const int runs = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  if (runs > 64) {
    __syncthreads();
    return false;
  }
  // run r starts at run_at[r]
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int i = i0 + u;
    if (i < m && (i == 0 || (a[i] >> 6) != (a[i - 1] >> 6))) run_at[base++] = i;
  }
  if (t == 0) run_at[runs] = m;
  __syncthreads();
  if (w == 0) {
    // sort (wavefront, run) pairs, then each run's new start = the lengths of
    // the runs of earlier wavefronts
    const int key = lane < runs ? ((a[run_at[lane]] >> 6) << 6) | lane : 0x7fffffff;
    const int sk = wave_sort64(key);
    const int r = sk & 63;
    const int len = lane < runs ? run_at[r + 1] - run_at[r] : 0;
    int in2 = len;
    for (int d = 1; d < 64; d <<= 1) {
      const int v = __shfl_up(in2, d);
      if (lane >= d) in2 += v;
    }
    if (lane < runs) run_to[r] = in2 - len;
  }
  __syncthreads();
  // element i of run r moves to run_to[r] + (i - run_at[r])
  for (int r = w; r < runs; r += kBlock / 64) {
    const int s = run_at[r], e = run_at[r + 1], d = run_to[r] - s;
    for (int i = s + lane; i < e; i += 64) b[i + d] = a[i];
  }
  __syncthreads();
  for (int i = t; i < m; i += blockDim.x) seg[i] = b[i];
  __syncthreads();
  return true;
}

__global__ __launch_bounds__(kBlock) void k_bin_sort(const int32_t* __restrict__ start,
                                                     const int32_t* __restrict__ list,
                                                     const int32_t* __restrict__ n_list,
                                                     int32_t* __restrict__ perm,
                                                     int32_t* __restrict__ tmp) {
  __shared__ int32_t a[kSortCap], b[kSortCap];
  __shared__ int32_t run_at[65], run_to[64], wsum[kBlock / 64];
  const int n_occ = __builtin_amdgcn_readfirstlane(*n_list);
  for (int li = blockIdx.x; li < n_occ; li += gridDim.x) {
    const int bk = list[li];
    const int s0 = start[bk], m = start[bk + 1] - s0;
    if (m < 2) continue;
    if (m <= kSortCap) {
      if (!block_order_runs(perm + s0, m, a, b, run_at, run_to, wsum))
        block_sort(perm + s0, perm + s0, m, a, b);
      continue;
    }
    // chunks sorted in place, then merged by rank through tmp
    for (int c0 = 0; c0 < m; c0 += kSortCap)
      block_sort(perm + s0 + c0, perm + s0 + c0, min(kSortCap, m - c0), a, b);
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      const int v = perm[s0 + i];
      const int own = i / kSortCap;
      int place = i - own * kSortCap;
      for (int c0 = 0; c0 < m; c0 += kSortCap)
        if (c0 / kSortCap != own) place += count_below(perm + s0 + c0, min(kSortCap, m - c0), v);
      tmp[s0 + place] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += blockDim.x) perm[s0 + i] = tmp[s0 + i];
    __syncthreads();
  }
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

// Grids of at most 32 bricks per axis (one environment) scan in two
// launches of one workgroup per x-slab of nb x nb bricks (brick (i, j, k) =
// (i nb + j) nb + k, slab i, row (i, j)):
//   k_bin_slab_scan        slab-local exclusive scan of the counts into
//                          lstart, the slab's total, and one 32-bit occupancy
//                          word per row (bit k: brick (i, j, k) holds
//                          particles); the counts and run counts are zeroed
//                          behind the read, so the next binning on this
//                          workspace needs no separate clear;
//   k_bin_slab_lists_fill  its first nb workgroups add the earlier slabs'
//                          totals to lstart into start; every one of them
//                          scans the per-row counts of occupied and active
//                          bricks (active: it or a -1 neighbour occupied, a
//                          shift and an OR of four words) over all rows, and
//                          writes its own rows' entries of both lists, so
//                          both come out ascending.  Its other workgroups
//                          fill the permutation meanwhile.
constexpr int kSlabThreads = kSlabMax * kSlabMax;

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d);
    if (lane >= d) v += u;
  }
  return v;
}

__global__ __launch_bounds__(kSlabThreads) void k_bin_slab_scan(int nb, int32_t* __restrict__ count,
                                                                int32_t* __restrict__ lstart,
                                                                int32_t* __restrict__ slab_sum,
                                                                uint32_t* __restrict__ row_occ,
                                                                int32_t* __restrict__ nruns,
                                                                int32_t* __restrict__ nsnap,
                                                                int32_t* __restrict__ n_ovf) {
  __shared__ int32_t wsum[kSlabThreads / 64];
  __shared__ uint32_t word[kSlabMax];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int per = nb * nb;
  const long long e = (long long)blockIdx.x * per + t;
  if (t < kSlabMax) word[t] = 0u;
  int v = 0;
  if (t < per) {
    v = count[e];
    count[e] = 0;
    nsnap[e] = nruns[e];
    nruns[e] = 0;
  }
  if (blockIdx.x == 0 && t == 0) {
    n_ovf[1] = n_ovf[0];
    n_ovf[0] = 0;
  }
  const int incl = wave_incl_scan(v);
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  if (v > 0) atomicOr(&word[t / nb], 1u << (t % nb));
  int base = 0;
  for (int q = 0; q < w; ++q) base += wsum[q];
  if (t < per) lstart[e] = base + incl - v;
  if (t == kSlabThreads - 1) slab_sum[blockIdx.x] = base + incl;
  __syncthreads();
  if (t < nb) row_occ[blockIdx.x * nb + t] = word[t];
}

// Workgroups [0, nb) make the lists and the final starts (slab lists);
// the others fill the permutation from the slab-local starts plus their
// slab's offset (the exclusive prefix of the slab totals), so the fill does
// not wait for the lists.
__global__ __launch_bounds__(kSlabThreads) void k_bin_slab_lists_fill(
    int nb, const int32_t* __restrict__ slab_sum, const uint32_t* __restrict__ row_occ,
    const int32_t* __restrict__ lstart, int32_t* __restrict__ start,
    int32_t* __restrict__ list, int32_t* __restrict__ n_list, int32_t* __restrict__ alist,
    int32_t* __restrict__ n_alist, int n, const int32_t* __restrict__ key,
    const int32_t* __restrict__ rank, const int32_t* __restrict__ nsnap,
    const int32_t* __restrict__ runs, int32_t* __restrict__ perm) {
  __shared__ int32_t wocc[kSlabThreads / 64], wact[kSlabThreads / 64];
  __shared__ int32_t slab_off;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int per = nb * nb;
  if ((int)blockIdx.x >= nb) {
    const int p = ((int)blockIdx.x - nb) * kSlabThreads + t;
    const bool active = p < n;
    const int k = active ? key[p] : -1;
    const int sv = lane < nb ? slab_sum[lane] : 0;
    const int excl = wave_incl_scan(sv) - sv;
    const int add = __shfl(excl, active ? k / per : 0);
    const int lr = wave_lane_rank(active, k);
    if (active) fill_particle(p, k, lr, add, rank, lstart, nsnap, runs, perm);
    return;
  }
  const int slab = blockIdx.x;
  if (w == 0) {
    int v = lane < slab ? slab_sum[lane] : 0;
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    if (lane == 0) slab_off = v;
  }
  // row t = (i, j) over the whole grid
  uint32_t occ = 0u, act = 0u;
  if (t < per) {
    const int i = t / nb, j = t - i * nb;
    occ = row_occ[t];
    uint32_t x = occ;
    if (j > 0) x |= row_occ[t - 1];
    if (i > 0) x |= row_occ[t - nb];
    if (i > 0 && j > 0) x |= row_occ[t - nb - 1];
    act = x | (x << 1);
    if (nb < 32) act &= (1u << nb) - 1u;
  }
  const int no = __popc(occ), na = __popc(act);
  const int io = wave_incl_scan(no), ia = wave_incl_scan(na);
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
  if (slab == 0 && t == kSlabThreads - 1) {
    *n_list = bo + no;
    *n_alist = ba + na;
  }
  if (t < per) start[(long long)slab * per + t] = lstart[(long long)slab * per + t] + slab_off;
  if (slab == nb - 1 && t == 0) start[(long long)nb * per] = slab_off + slab_sum[slab];
  if (t < per && t / nb == slab) {
    const int rb = t * nb;
    for (uint32_t m = occ; m; m &= m - 1u) list[bo++] = rb + __ffs((int)m) - 1;
    for (uint32_t m = act; m; m &= m - 1u) alist[ba++] = rb + __ffs((int)m) - 1;
  }
}

// the slab scan when the grid is a single environment of at most 32 bricks
// per axis; that scan leaves the counts zero
static bool slab_scan_fits(const BinWs& ws, int nb) {
  return nb <= kSlabMax && ws.n_bricks == (long long)nb * nb * nb;
}

bool bin_counts_fusable(const BinWs& ws) { return slab_scan_fits(ws, ws.nb); }

int launch_bin(const MpmP& P, int n, const float* x, const BinWs& ws, hipStream_t st,
               bool count_clear, bool counted) {
  const long long nbk = ws.n_bricks;
  const int nb = (P.n_grid + 3) >> 2;
  const bool slab = slab_scan_fits(ws, nb);
  if (counted && !slab) {
    set_error("binning: fused counts need the slab scan");
    return DT_EINVAL;
  }
  if (!counted) {
    if (!slab || !count_clear)
      hipLaunchKernelGGL(k_bin_zero, dim3(nblocks(nbk + 2)), dim3(kBlock), 0, st, nbk + 2,
                         ws.count, nbk, ws.nruns, ws.n_list, ws.n_alist, ws.n_ovf);
    if (n > 0)
      hipLaunchKernelGGL(k_bin_count, dim3(nblocks(n)), dim3(kBlock), 0, st, P, n, x, ws.key,
                         ws.rank, ws.count, bin_runs(ws));
  }
  // the slab scan clears the run counts behind it (for a G2P's fused count)
  // and leaves this binning's in nsnap / n_ovf[1]
  const int32_t* nsnap = slab ? ws.nsnap : ws.nruns;
  const int32_t* n_ovf = slab ? ws.n_ovf + 1 : ws.n_ovf;
  if (slab) {
    hipLaunchKernelGGL(k_bin_slab_scan, dim3(nb), dim3(kSlabThreads), 0, st, nb, ws.count,
                       ws.lstart, ws.slab_sum, ws.row_occ, ws.nruns, ws.nsnap, ws.n_ovf);
    hipLaunchKernelGGL(k_bin_slab_lists_fill, dim3(nb + nblocks(n, kSlabThreads)),
                       dim3(kSlabThreads), 0, st, nb, (const int32_t*)ws.slab_sum,
                       (const uint32_t*)ws.row_occ, (const int32_t*)ws.lstart, ws.start, ws.list,
                       ws.n_list, ws.alist, ws.n_alist, n, (const int32_t*)ws.key,
                       (const int32_t*)ws.rank, (const int32_t*)ws.nsnap,
                       (const int32_t*)ws.runs, ws.perm);
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
  if (n > 0) {
    if (!slab)
      hipLaunchKernelGGL(k_bin_fill, dim3(nblocks(n)), dim3(kBlock), 0, st, n, ws.key, ws.rank,
                         ws.start, nsnap, (const int32_t*)ws.runs, ws.perm);
    // rank is dead after the fill: the large-brick path of the sort uses it
    const long long sort_cap = launch_config().sort_blocks;
    const long long nb_sort = ws.n_bricks < sort_cap ? ws.n_bricks : sort_cap;
    hipLaunchKernelGGL(k_bin_sort, dim3((unsigned)(nb_sort > 0 ? nb_sort : 1)), dim3(kBlock), 0,
                       st, (const int32_t*)ws.start, (const int32_t*)ws.ovf, n_ovf, ws.perm,
                       ws.rank);
  }
  return check_launch("particle binning");
}

}  // namespace dt
