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
  const size_t sz[11] = {sizeof(int32_t) * (size_t)(nbk + 2),
                        sizeof(int32_t) * (size_t)(nbk + 1),
                        sizeof(int32_t) * (size_t)nbk,
                        sizeof(int32_t) * (size_t)n,
                        sizeof(int32_t) * (size_t)n,
                        sizeof(int32_t) * (size_t)n,
                        sizeof(float4) * (size_t)nbk * kTileNodes,
                        scan_bytes,
                        sizeof(int32_t) * (size_t)nbk,
                        sizeof(int32_t),
                        sizeof(int32_t) * (size_t)(kSlabMax + kSlabMax * kSlabMax)};
  size_t off[11], total = 0;
  for (int i = 0; i < 11; ++i) {
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

// Grids of at most 32 bricks per axis (one environment) scan in two
// launches of one workgroup per x-slab of nb x nb bricks (brick (i, j, k) =
// (i nb + j) nb + k, slab i, row (i, j)):
//   k_bin_slab_scan   slab-local exclusive scan of the counts into start, the
//                     slab's total, and one 32-bit occupancy word per row
//                     (bit k: brick (i, j, k) holds particles); the counts are
//                     zeroed behind the read, so the next binning on this
//                     workspace needs no separate clear;
//   k_bin_slab_lists  adds the earlier slabs' totals to start; every
//                     workgroup scans the per-row counts of occupied and
//                     active bricks (active: it or a -1 neighbour occupied, a
//                     shift and an OR of four words) over all rows, and writes
//                     its own rows' entries of both lists, so both come out
//                     ascending.
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
                                                                int32_t* __restrict__ start,
                                                                int32_t* __restrict__ slab_sum,
                                                                uint32_t* __restrict__ row_occ) {
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
  }
  const int incl = wave_incl_scan(v);
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  if (v > 0) atomicOr(&word[t / nb], 1u << (t % nb));
  int base = 0;
  for (int q = 0; q < w; ++q) base += wsum[q];
  if (t < per) start[e] = base + incl - v;
  if (t == kSlabThreads - 1) slab_sum[blockIdx.x] = base + incl;
  __syncthreads();
  if (t < nb) row_occ[blockIdx.x * nb + t] = word[t];
}

__global__ __launch_bounds__(kSlabThreads) void k_bin_slab_lists(
    int nb, const int32_t* __restrict__ slab_sum, const uint32_t* __restrict__ row_occ,
    int32_t* __restrict__ start, int32_t* __restrict__ list, int32_t* __restrict__ n_list,
    int32_t* __restrict__ alist, int32_t* __restrict__ n_alist) {
  __shared__ int32_t wocc[kSlabThreads / 64], wact[kSlabThreads / 64];
  __shared__ int32_t slab_off;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int per = nb * nb, slab = blockIdx.x;
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
  if (t < per) start[(long long)slab * per + t] += slab_off;
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
                         ws.count, ws.n_list, ws.n_alist);
    if (n > 0)
      hipLaunchKernelGGL(k_bin_count, dim3(nblocks(n)), dim3(kBlock), 0, st, P, n, x, ws.key,
                         ws.rank, ws.count);
  }
  if (slab) {
    hipLaunchKernelGGL(k_bin_slab_scan, dim3(nb), dim3(kSlabThreads), 0, st, nb, ws.count,
                       ws.start, ws.slab_sum, ws.row_occ);
    hipLaunchKernelGGL(k_bin_slab_lists, dim3(nb), dim3(kSlabThreads), 0, st, nb,
                       (const int32_t*)ws.slab_sum, (const uint32_t*)ws.row_occ, ws.start,
                       ws.list, ws.n_list, ws.alist, ws.n_alist);
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
