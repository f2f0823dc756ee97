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
  float dx, inv_dx, dt, p_vol, p_mass, eps;
  float g[3];
  float grad_clamp;
};

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
  return d;
}

// First x-node of particle p's environment (environments stacked along x).
__device__ __forceinline__ int env_x0(const MpmP& P, long long p) {
  return P.env_id ? P.env_id[p] * P.n_grid : 0;
}

int validate_params(const dt_mpm_params* p);

}  // namespace dt
