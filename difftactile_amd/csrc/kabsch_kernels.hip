// Kabsch rotation angle of the masked ("rigid") particles (gfx950).
//
// Reference (difftactile/):
//   compute_COM        object_model/multi_obj.py:309-314  (divides by all N)
//   compute_H          object_model/multi_obj.py:316-328
//   compute_H_svd, R   object_model/multi_obj.py:330-340  (R = V U^T)
//   compute_angle      tasks/box_open.py:400-406          (atan2(R10, R11))
//
// Per substep: two fixed-order masked reductions over the particles (sums of
// x and y = x + dt v_next with the count, then the centred products H), one
// lane for the 3x3 SVD and the angle.  The adjoint differentiates R through
// the polar rotation of H exactly (the rotation part of an SVD adjoint,
// K_ij = (M_ij - M_ji) / (s_i + s_j)), then one pass over the particles.
//
// State (DT_KABSCH_STATE floats): [0:3] sum x, [3:6] sum y, [6] count, [8:17]
// H, [17:26] U, [26:29] s, [29:38] V, [38:47] R, [47] angle, [48:57] dL/dH,
// [57:60] dL/dc0, [60:63] dL/dc1.
#include "dt_common.h"
#include "dt_device.h"

namespace dt {

enum {
  kSx = 0, kSy = 3, kCnt = 6, kH = 8, kU = 17, kS = 26, kV = 29, kR = 38, kAng = 47, kGH = 48,
  kGc0 = 57, kGc1 = 60
};

__device__ __forceinline__ void rigid_y(const float* __restrict__ x, const float* __restrict__ v,
                                        float dt, long long p, float (&xp)[3], float (&yp)[3]) {
  for (int a = 0; a < 3; ++a) {
    xp[a] = x[3 * p + a];
    yp[a] = xp[a] + dt * v[3 * p + a];
  }
}

template <int kTerms>
__device__ __forceinline__ void block_sum_store(float (&v)[kTerms], float* out) {
  __shared__ float part[kBlock / 64][kTerms];
  for (int q = 0; q < kTerms; ++q)
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o);
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < kTerms; ++q) part[threadIdx.x >> 6][q] = v[q];
  __syncthreads();
  if (threadIdx.x < kTerms) {
    float s = part[0][threadIdx.x];
    for (int w = 1; w < kBlock / 64; ++w) s += part[w][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// per block: sum x, sum y, count over the masked particles
__global__ __launch_bounds__(kBlock) void k_rigid_sums(int n, const float* __restrict__ x,
                                                       const float* __restrict__ v, float dt,
                                                       const int32_t* __restrict__ mask,
                                                       float* __restrict__ partials) {
  float s[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    if (!mask[p]) continue;
    float xp[3], yp[3];
    rigid_y(x, v, dt, p, xp, yp);
    for (int a = 0; a < 3; ++a) s[a] += xp[a], s[3 + a] += yp[a];
    s[6] += 1.f;
  }
  block_sum_store<7>(s, partials + 7ll * blockIdx.x);
}

// per block: H = sum (x - c0)(y - c1)^T, c = sums / n (all particles)
__global__ __launch_bounds__(kBlock) void k_rigid_H(int n, const float* __restrict__ x,
                                                    const float* __restrict__ v, float dt,
                                                    const int32_t* __restrict__ mask,
                                                    const float* __restrict__ state,
                                                    float* __restrict__ partials) {
  const float inv_n = 1.f / (float)n;
  float c0[3], c1[3];
  for (int a = 0; a < 3; ++a) c0[a] = state[kSx + a] * inv_n, c1[a] = state[kSy + a] * inv_n;
  float h[9];
  for (int q = 0; q < 9; ++q) h[q] = 0.f;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    if (!mask[p]) continue;
    float xp[3], yp[3];
    rigid_y(x, v, dt, p, xp, yp);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) h[i * 3 + j] += (xp[i] - c0[i]) * (yp[j] - c1[j]);
  }
  block_sum_store<9>(h, partials + 9ll * blockIdx.x);
}

// state[off + q] = sum over blocks of partials[terms b + q]
__global__ __launch_bounds__(64) void k_sum_partials(int nb, int terms,
                                                     const float* __restrict__ partials,
                                                     float* __restrict__ state, int off) {
  if ((int)threadIdx.x >= terms) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += partials[(long long)terms * b + threadIdx.x];
  state[off + threadIdx.x] = s;
}

__global__ __launch_bounds__(64) void k_kabsch_finish(float* __restrict__ state) {
  if (threadIdx.x != 0) return;
  M3 H = ld_m3(state + kH), U, V;
  float s[3];
  svd3(H, U, s, V);
  const M3 R = m3_mul_bt(V, U);   // V U^T
  for (int q = 0; q < 9; ++q) {
    state[kU + q] = U.a[q];
    state[kV + q] = V.a[q];
    state[kR + q] = R.a[q];
  }
  for (int a = 0; a < 3; ++a) state[kS + a] = s[a];
  state[kAng] = atan2f(R.a[1 * 3 + 0], R.a[1 * 3 + 1]);
}

// dL/dH from dL/dangle; dL/dc0, dL/dc1 through the centring
__global__ __launch_bounds__(64) void k_kabsch_adj_small(int n, float* __restrict__ state,
                                                         const float* __restrict__ g_angle) {
  if (threadIdx.x != 0) return;
  const float a = state[kR + 3], b = state[kR + 4];   // R10, R11
  const float r2 = a * a + b * b;
  const float g = *g_angle;
  // R = V U^T = Q^T with Q = U V^T the polar rotation of H: dL/dQ = (dL/dR)^T
  M3 gQ = m3_zero();
  gQ.a[0 * 3 + 1] = g * b / r2;    // (dL/dR)_10
  gQ.a[1 * 3 + 1] = -g * a / r2;   // (dL/dR)_11
  const M3 U = ld_m3(state + kU), V = ld_m3(state + kV);
  const M3 M = m3_mul(m3_mul_at(U, gQ), V);
  M3 K = m3_zero();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (i == j) continue;
      float den = state[kS + i] + state[kS + j];
      if (fabsf(den) < 1e-12f) den = den < 0.f ? -1e-12f : 1e-12f;
      K.a[i * 3 + j] = (M.a[i * 3 + j] - M.a[j * 3 + i]) / den;
    }
  const M3 gH = m3_mul_bt(m3_mul(U, K), V);
  const float inv_n = 1.f / (float)n, cnt = state[kCnt];
  float dx[3], dy[3];   // sum over the masked particles of (x - c0), (y - c1)
  for (int k = 0; k < 3; ++k) {
    dx[k] = state[kSx + k] - cnt * state[kSx + k] * inv_n;
    dy[k] = state[kSy + k] - cnt * state[kSy + k] * inv_n;
  }
  for (int i = 0; i < 3; ++i) {
    float g0 = 0.f, g1 = 0.f;
    for (int j = 0; j < 3; ++j) {
      g0 -= gH.a[i * 3 + j] * dy[j];
      g1 -= gH.a[j * 3 + i] * dx[j];
    }
    state[kGc0 + i] = g0;
    state[kGc1 + i] = g1;
  }
  for (int q = 0; q < 9; ++q) state[kGH + q] = gH.a[q];
}

// gx += dL/dx + dL/dy, gv += dt dL/dy for the masked particles
__global__ __launch_bounds__(kBlock) void k_kabsch_adj(int n, const float* __restrict__ x,
                                                       const float* __restrict__ v, float dt,
                                                       const int32_t* __restrict__ mask,
                                                       const float* __restrict__ state,
                                                       float* __restrict__ gx,
                                                       float* __restrict__ gv) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || !mask[p]) return;
  const float inv_n = 1.f / (float)n;
  float xp[3], yp[3];
  rigid_y(x, v, dt, p, xp, yp);
  float c0[3], c1[3];
  for (int a = 0; a < 3; ++a) c0[a] = state[kSx + a] * inv_n, c1[a] = state[kSy + a] * inv_n;
  for (int i = 0; i < 3; ++i) {
    float g_x = state[kGc0 + i] * inv_n, g_y = state[kGc1 + i] * inv_n;
    for (int j = 0; j < 3; ++j) {
      g_x += state[kGH + i * 3 + j] * (yp[j] - c1[j]);
      g_y += state[kGH + j * 3 + i] * (xp[j] - c0[j]);
    }
    gx[3ll * p + i] += g_x + g_y;
    gv[3ll * p + i] += dt * g_y;
  }
}

static unsigned kabsch_blocks(int n) {
  const int b = (n + kBlock - 1) / kBlock;
  return (unsigned)(b < DT_REDUCE_BLOCKS ? (b < 1 ? 1 : b) : DT_REDUCE_BLOCKS);
}

}  // namespace dt

using namespace dt;

extern "C" {

int dt_kabsch_angle(int32_t n, const float* x, const float* v_next, float dt, const int32_t* mask,
                    float* scratch, float* state, void* stream) {
  DT_REQUIRE(n > 0, "n must be > 0");
  DT_REQUIRE(x && v_next && mask && scratch && state, "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = kabsch_blocks(n);
  hipLaunchKernelGGL(k_rigid_sums, dim3(nb), dim3(kBlock), 0, st, (int)n, x, v_next, dt, mask,
                     scratch);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, (int)nb, 7,
                     (const float*)scratch, state, (int)kSx);
  hipLaunchKernelGGL(k_rigid_H, dim3(nb), dim3(kBlock), 0, st, (int)n, x, v_next, dt, mask,
                     (const float*)state, scratch);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, (int)nb, 9,
                     (const float*)scratch, state, (int)kH);
  hipLaunchKernelGGL(k_kabsch_finish, dim3(1), dim3(64), 0, st, state);
  return check_launch("dt_kabsch_angle");
}

int dt_kabsch_angle_adj(int32_t n, const float* x, const float* v_next, float dt,
                        const int32_t* mask, float* state, const float* g_angle, float* gx,
                        float* gv_next, void* stream) {
  DT_REQUIRE(n > 0, "n must be > 0");
  DT_REQUIRE(x && v_next && mask && state && g_angle && gx && gv_next, "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_kabsch_adj_small, dim3(1), dim3(64), 0, st, (int)n, state, g_angle);
  hipLaunchKernelGGL(k_kabsch_adj, dim3(nblocks(n)), dim3(kBlock), 0, st, (int)n, x, v_next, dt,
                     mask, (const float*)state, gx, gv_next);
  return check_launch("dt_kabsch_angle_adj");
}

}  // extern "C"
