// Host+device shared definitions of the HIP library: C-ABI structs, error
// state, launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/difftactile.h"

namespace dt {

void set_error(const std::string& msg);
int check_launch(const char* what);
// the process-wide launch shapes (dt_set_launch_config; zeros replaced by the defaults)
dt_launch_config launch_config();

#define DT_REQUIRE(cond, msg)                 \
  do {                                        \
    if (!(cond)) {                            \
      ::dt::set_error(msg);                   \
      return DT_EINVAL;                       \
    }                                         \
  } while (0)

constexpr int kBlock = 256;

inline int nblocks(long long n, int bs = kBlock) { return (int)((n + bs - 1) / bs); }

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Remap so that each XCD works on one contiguous range of logical blocks:
// neighbouring particles (after the cell sort) then share an L2.
__device__ __forceinline__ int xcd_block() {
  int b = blockIdx.x, nb = gridDim.x;
  int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

struct MpmP {
  int n_grid, bound, sticky_floor, clamp_grad, soa, svd_exact;
  int n_env, nx;   // environments stacked along x; nx = n_env n_grid x-nodes
  const int32_t* env_id;
  const int32_t* alist;     // node set of the grid-node passes (NULL: all nodes)
  const int32_t* n_alist;
  float dx, inv_dx, dt, p_vol, p_mass, eps;
  float g[3];
  float grad_clamp;
  // ragged environments: per-environment grid size and (dx, inv_dx, p_mass,
  // p_vol); n_wall is the grid size the walls of this view act at
  const int32_t* env_ngrid;
  const float4* env_scale;
  int n_wall;
  // always 0: the struct has no padding bytes, so its bytes can key the graph
  // cache (uninitialised padding would make equal parameters look different)
  int key_pad;
};
static_assert(sizeof(MpmP) == 120, "MpmP must have no padding bytes (graph keys hash them)");

inline MpmP to_dev(const dt_mpm_params& p) {
  MpmP d;
  d.n_grid = p.n_grid;
  d.bound = p.bound;
  d.sticky_floor = p.sticky_floor;
  d.clamp_grad = p.clamp_grad;
  d.soa = p.soa;
  d.svd_exact = p.svd_exact;
  d.n_env = p.n_env > 1 ? p.n_env : 1;
  d.nx = d.n_env * p.n_grid;
  d.env_id = d.n_env > 1 ? p.env_id : nullptr;
  // the node set is honoured only by the entry points documented to take it
  d.alist = nullptr;
  d.n_alist = nullptr;
  d.dx = p.dx;
  d.inv_dx = p.inv_dx;
  d.dt = p.dt;
  d.p_vol = p.p_vol;
  d.p_mass = p.p_mass;
  d.eps = p.eps;
  d.g[0] = p.gravity[0];
  d.g[1] = p.gravity[1];
  d.g[2] = p.gravity[2];
  d.grad_clamp = p.grad_clamp;
  d.env_ngrid = (d.n_env > 1 && p.env_scale) ? p.env_ngrid : nullptr;
  d.env_scale = (d.n_env > 1 && p.env_ngrid) ? reinterpret_cast<const float4*>(p.env_scale)
                                             : nullptr;
  if (!d.env_scale) d.env_ngrid = nullptr;
  d.n_wall = p.n_grid;
  d.key_pad = 0;
  return d;
}

// First x-node of particle p's environment (environments stacked along x).
__device__ __forceinline__ int env_x0(const MpmP& P, long long p) {
  return P.env_id ? P.env_id[p] * P.n_grid : 0;
}

// The view of P for environment e: its dx, inv_dx, p_mass, p_vol and wall
// grid size when the environments are ragged, P itself otherwise.
__device__ __forceinline__ MpmP at_env(const MpmP& P, int e) {
  MpmP Q = P;
  if (P.env_scale) {
    const float4 s = P.env_scale[e];
    Q.dx = s.x;
    Q.inv_dx = s.y;
    Q.p_mass = s.z;
    Q.p_vol = s.w;
    Q.n_wall = P.env_ngrid[e];
  }
  return Q;
}
__device__ __forceinline__ MpmP at_particle(const MpmP& P, long long p) {
  return P.env_scale ? at_env(P, P.env_id[p]) : P;
}
// environment of x-node i (environments stacked along x)
__device__ __forceinline__ MpmP at_xnode(const MpmP& P, int i) {
  return P.env_scale ? at_env(P, i / P.n_grid) : P;
}

// Node of lane threadIdx&63 in 4x4x4-node brick b (-1 outside the grid).
__device__ __forceinline__ long long active_node(const MpmP& P, int b) {
  const int n = P.n_grid, nb = (n + 3) >> 2;
  const int l = threadIdx.x & 63;
  const int bz = b % nb, by = (b / nb) % nb, bx = b / (nb * nb);
  const int i = bx * 4 + (l >> 4), j = by * 4 + ((l >> 2) & 3), k = bz * 4 + (l & 3);
  if (i >= P.nx || j >= n || k >= n) return -1;
  return ((long long)i * n + j) * n + k;
}

// Node loops of the grid-node passes: every node of the n-node grid (one per
// thread), or the nodes of the active bricks (P.alist: a persistent grid, four
// bricks per block and round).  f(nd) is called with nd = -1 for idle lanes
// (node_loop_all: every thread calls f the same number of times, so f may
// hold block barriers).
template <class F>
__device__ __forceinline__ void node_loop(const MpmP& P, long long nn, const F& f) {
  if (P.alist) {
    const int n_act = *P.n_alist;
    for (int ai = blockIdx.x * 4 + (threadIdx.x >> 6); ai < n_act; ai += gridDim.x * 4) {
      const long long nd = active_node(P, P.alist[ai]);
      if (nd >= 0) f(nd);
    }
  } else {
    const long long nd = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (nd < nn) f(nd);
  }
}
template <class F>
__device__ __forceinline__ void node_loop_all(const MpmP& P, long long nn, const F& f) {
  if (P.alist) {
    const int n_act = *P.n_alist;
    for (int b0 = blockIdx.x * 4; b0 < n_act; b0 += gridDim.x * 4) {
      const int ai = b0 + (threadIdx.x >> 6);
      f(ai < n_act ? active_node(P, P.alist[ai]) : -1ll);
    }
  } else {
    const long long nd = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    f(nd < nn ? nd : -1ll);
  }
}

// P with the node set of prm, and the launch grid of the node loops
inline MpmP to_dev_nodes(const dt_mpm_params& p) {
  MpmP d = to_dev(p);
  if (p.active_bricks && p.n_active_bricks) {
    d.alist = p.active_bricks;
    d.n_alist = p.n_active_bricks;
  }
  return d;
}
inline int node_blocks(const dt_mpm_params& p, long long nn) {
  if (p.active_bricks && p.n_active_bricks) {
    const long long nb = (p.n_grid + 3) >> 2;
    const long long nbx = ((long long)(p.n_env > 1 ? p.n_env : 1) * p.n_grid + 3) >> 2;
    const long long b = (nbx * nb * nb + 3) / 4;
    return (int)(b < 1024 ? b : 1024);
  }
  return nblocks(nn);
}

int validate_params(const dt_mpm_params* p);

}  // namespace dt
