// Per-substep particle binning for the brick-owned scatter passes.
//
// Particles are binned by the 4x4x4-cell brick of their stencil base.  After
// launch_bin, perm[start[b] .. start[b+1]) lists the particles of brick b in
// ascending index order; the brick passes order each chunk of a brick by cell
// in LDS (stably, so the order is the same on every run).
//
// The counting pass records every (wavefront, brick) run it counts in a run
// table of kRuns entries per brick.  The fill then places a particle at the
// lengths of its brick's runs of earlier wavefronts plus its rank among its
// wavefront's lanes of the brick: ascending index order with no sort.  A
// brick of more than kRuns runs (particles in no spatial order) is placed in
// the counting atomics' order and sorted afterwards.
#pragma once
#include "dt_common.h"

namespace dt {

constexpr int kTileEdge = 6;   // a brick's 4 cells plus the stencil's 2-node reach
constexpr int kTileNodes = kTileEdge * kTileEdge * kTileEdge;
// bricks per axis up to which a single-environment grid uses the slab scan
constexpr int kSlabMax = 32;
// run-table entries per brick
constexpr int kRuns = 32;

// Views into one caller-provided workspace buffer.
struct BinWs {
  int32_t* count;   // [n_bricks + 2] particles per brick ([n_bricks] stays 0);
                    //   [n_bricks + 1] counts the occupied bricks
  int32_t* n_list;  // = count + n_bricks + 1
  int32_t* start;   // [n_bricks + 1] exclusive prefix sum of count
  int32_t* list;    // [n_bricks]     occupied bricks (any order)
  int32_t* alist;   // [n_bricks]     active bricks: occupied bricks and their +1
                    //   neighbours, which a stencil can reach
  int32_t* n_alist; // count of alist
  int32_t* key;     // [n]            brick of each particle
  int32_t* rank;    // [n]            position of the particle inside its brick
  int32_t* perm;    // [n]            particles in brick order
  float4* tiles;    // [n_bricks * kTileNodes] per-brick partial node sums
  float* mat_part;  // [blocks of n particles][32] per-block material-adjoint partials
  void* scan_tmp;
  size_t scan_bytes;
  int32_t* slab_sum;  // [kSlabMax]            slab scan: particles per x-slab
  uint32_t* row_occ;  // [kSlabMax * kSlabMax] slab scan: occupancy word per brick row
  int32_t* nruns;   // [n_bricks]          runs counted per brick (zero between binnings)
  int32_t* nsnap;   // [n_bricks]          the slab scan's copy of nruns (it clears nruns)
  int32_t* runs;    // [n_bricks * kRuns]  (wavefront << 6) | (lanes - 1) of each run
  int32_t* ovf;     // [n_bricks]          bricks of more than kRuns runs
  int32_t* n_ovf;   // [2]                 count of ovf; [1] the slab scan's copy
  int32_t* lstart;  // [n_bricks]          slab scan: slab-local brick starts
  long long n_bricks;
  int nb;            // bricks along y and z
  // Optional binning history (one slot per substep of a trajectory): the
  // perm / start / list / n_list / alist / n_alist of substep s live in slot s so the
  // adjoint sweep can reuse the forward pass's binning.
  char* hist = nullptr;
  size_t hist_slot_bytes = 0;
  int hist_slots = 0;
};

// Bytes of one history slot for n particles.
size_t bin_history_slot_bytes(const dt_mpm_params* prm, int n);
// View of ws whose per-substep arrays live in history slot s.
BinWs bin_slot(const BinWs& ws, int s, int n);
// View of ws whose per-substep arrays (perm, start, list, alist and the two
// counts) live at base, laid out as one history slot.
BinWs bin_view(const BinWs& ws, char* base, int n);
// Copies the binning of view from into view to (lists up to their counts).
void launch_bin_copy(const BinWs& from, const BinWs& to, int n, hipStream_t st);

long long brick_total(const dt_mpm_params* prm);
// Total bytes of the workspace for n particles; fills *ws when base != NULL.
size_t bin_ws_layout(const dt_mpm_params* prm, int n, void* base, BinWs* ws);
// Rank of this lane among the active lanes of the wavefront that share its
// key (any int; the lanes of a key need not be adjacent), and one atomic per
// distinct key of the wavefront adding that key's lane count to counts[key]
// (LDS or global).  Returns the count before the wavefront's atomic plus the
// lane's rank among its key's lanes in lane order.  The loop runs once per
// distinct key, with a wave-uniform trip count.
__device__ __forceinline__ int wave_key_rank(bool active, int key, int* counts) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long todo = __ballot(active);
  int result = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int cur = __shfl(key, leader);
    const bool mine = active && key == cur;
    const unsigned long long peers = __ballot(mine);
    int old = 0;
    if (lane == leader) old = atomicAdd(&counts[cur], __popcll(peers));
    old = __shfl(old, leader);
    if (mine) result = old + __popcll(peers & below);
    todo &= ~peers;
  }
  return result;
}

// Mask of the active lanes whose 6-bit key equals this lane's (six ballots);
// meaningful for active lanes only.
__device__ __forceinline__ unsigned long long wave_match6(bool active, int key) {
  unsigned long long m = __ballot(active);
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    const bool bit = (key >> b) & 1;
    const unsigned long long bal = __ballot(active && bit);
    m &= bit ? bal : ~bal;
  }
  return m;
}

// The run-table side of a counting pass.
struct BinRuns {
  int32_t* nruns;
  int32_t* runs;
  int32_t* ovf;
  int32_t* n_ovf;
};
__host__ __device__ inline BinRuns bin_runs(const BinWs& ws) {
  return BinRuns{ws.nruns, ws.runs, ws.ovf, ws.n_ovf};
}

// wave_key_rank that also records each (wavefront, key) run in the run table
// (w: the wavefront's index, particle index >> 6)
__device__ __forceinline__ int wave_key_rank_runs(bool active, int key, int* counts,
                                                  const BinRuns& R, int w) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long todo = __ballot(active);
  int result = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int cur = __shfl(key, leader);
    const bool mine = active && key == cur;
    const unsigned long long peers = __ballot(mine);
    int old = 0;
    if (lane == leader) {
      const int m = __popcll(peers);
      old = atomicAdd(&counts[cur], m);
      const int slot = atomicAdd(&R.nruns[cur], 1);
      if (slot < kRuns) R.runs[(long long)cur * kRuns + slot] = (w << 6) | (m - 1);
      else if (slot == kRuns) R.ovf[atomicAdd(R.n_ovf, 1)] = cur;
    }
    old = __shfl(old, leader);
    if (mine) result = old + __popcll(peers & below);
    todo &= ~peers;
  }
  return result;
}

// This lane's rank among the active lanes of its wavefront sharing its key
// (the lane part of wave_key_rank, without atomics).
__device__ __forceinline__ int wave_lane_rank(bool active, int key) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long todo = __ballot(active);
  int result = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int cur = __shfl(key, leader);
    const bool mine = active && key == cur;
    const unsigned long long peers = __ballot(mine);
    if (mine) result = __popcll(peers & below);
    todo &= ~peers;
  }
  return result;
}

// Persistent grid of the brick-owned passes (blocks stride over the list).
constexpr int kBrickBlocks = 2048;
// count -> brick scan -> occupied / active brick lists -> fill -> sort of
// the bricks with more runs than the run table holds, on st.  The counts and
// the brick starts are integers and exact; perm[start[b] .. start[b+1]) is
// the brick's particles in ascending index order whatever order the counting
// atomics ran in, so the brick passes see the same particles in the same
// order on every run.
// count_clear: the workspace's brick counts are known to be zero (a previous
// binning on it used the slab scan, which clears them behind it).
// counted: key / rank / count already hold x's counts (made by the previous
// substep's G2P); needs bin_counts_fusable.
int launch_bin(const MpmP& P, int n, const float* x, const BinWs& ws, hipStream_t st,
               bool count_clear = false, bool counted = false);
// Whether this workspace's binning clears the counts behind its scan, so a
// G2P may make the next binning's counts.
bool bin_counts_fusable(const BinWs& ws);

}  // namespace dt
