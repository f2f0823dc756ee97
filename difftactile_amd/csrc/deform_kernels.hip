// Grid-mass deformation losses (gfx950).
//
// Reference (difftactile/):
//   compute_grid_m_kernel   object_model/mpm_plastic.py:168-178  (mass-only p2g)
//   update_target_sdf       tasks/deform_loss.py:61-81           (nearest-point SDF)
//   compute_density_loss    tasks/deform_loss.py:99-102
//   compute_sdf_loss        tasks/deform_loss.py:104-107
//   iou                     tasks/deform_loss.py:139-153
//
// The mass-only scatter is per-particle global atomics (a loss evaluation,
// not a substep stage); its adjoint gathers.  The SDF propagation is one
// Jacobi sweep per launch over every node with the reference's neighbour
// order, ping-ponging two buffers where the reference copies.  Sums are
// fixed-order block reductions into a caller scratch, then one block, so the
// losses are reproducible bit for bit.
#include "dt_common.h"
#include "dt_device.h"

namespace dt {

constexpr float kSdfInf = 1000.0f;   // deform_loss.py:27

__global__ __launch_bounds__(kBlock) void k_grid_mass(MpmP P, int n, const float* __restrict__ x,
                                                      float* __restrict__ gm) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float xp[3] = {x[3ll * p], x[3ll * p + 1], x[3ll * p + 2]};
  Stencil s;
  make_stencil(xp, P.inv_dx, P.n_grid, s);
  const int nn = P.n_grid;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        const float w = s.w[i][0] * s.w[j][1] * s.w[k][2];
        const long long nd =
            ((long long)(s.base[0] + i) * nn + (s.base[1] + j)) * nn + (s.base[2] + k);
        atomic_add_f(gm + nd, w * P.p_mass);
      }
}

// gx += p_mass sum_nodes g_m(node) d w(node) / dx  (the base is held fixed)
__global__ __launch_bounds__(kBlock) void k_grid_mass_adj(MpmP P, int n,
                                                          const float* __restrict__ x,
                                                          const float* __restrict__ g_gm,
                                                          float* __restrict__ gx) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float xp[3] = {x[3ll * p], x[3ll * p + 1], x[3ll * p + 2]};
  Stencil s;
  make_stencil(xp, P.inv_dx, P.n_grid, s);
  const int nn = P.n_grid;
  float g[3] = {0.f, 0.f, 0.f};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        const long long nd =
            ((long long)(s.base[0] + i) * nn + (s.base[1] + j)) * nn + (s.base[2] + k);
        const float gm = g_gm[nd] * P.p_mass;
        g[0] += gm * s.dw[i][0] * s.w[j][1] * s.w[k][2];
        g[1] += gm * s.w[i][0] * s.dw[j][1] * s.w[k][2];
        g[2] += gm * s.w[i][0] * s.w[j][1] * s.dw[k][2];
      }
  for (int a = 0; a < 3; ++a) gx[3ll * p + a] += g[a] * P.inv_dx;
}

// One update_target_sdf sweep: node I takes distance 0 where the target
// density exceeds 1e-4, else the nearest point of the first neighbour (offsets
// -3..2 per axis in x-y-z lexicographic order, the node itself excluded) whose
// nearest point is closest, by sqrt(|d|^2 + 1e-8) with d = I dx - point and a
// strict <; a node with no neighbour holding a point keeps distance 1000 and
// its previous point.
__global__ __launch_bounds__(kBlock) void k_sdf_sweep(int n, float dx,
                                                      const float* __restrict__ density,
                                                      const float* __restrict__ sdf_in,
                                                      const float* __restrict__ np_in,
                                                      float* __restrict__ sdf_out,
                                                      float* __restrict__ np_out) {
#pragma clang fp contract(off)
  const long long nd = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nn = (long long)n * n * n;
  if (nd >= nn) return;
  const int i = (int)(nd / ((long long)n * n)), j = (int)((nd / n) % n), k = (int)(nd % n);
  const float gp[3] = {(float)i * dx, (float)j * dx, (float)k * dx};
  float best = kSdfInf;
  float bp[3] = {np_in[3 * nd], np_in[3 * nd + 1], np_in[3 * nd + 2]};
  if (density[nd] > 1e-4f) {
    best = 0.f;
    bp[0] = gp[0], bp[1] = gp[1], bp[2] = gp[2];
  } else {
    for (int ox = -3; ox < 3; ++ox)
      for (int oy = -3; oy < 3; ++oy)
        for (int oz = -3; oz < 3; ++oz) {
          const int vi = i + ox, vj = j + oy, vk = k + oz;
          if (vi < 0 || vj < 0 || vk < 0 || vi >= n || vj >= n || vk >= n) continue;
          if (ox == 0 && oy == 0 && oz == 0) continue;
          const long long v = ((long long)vi * n + vj) * n + vk;
          if (!(sdf_in[v] < kSdfInf)) continue;
          const float q[3] = {np_in[3 * v], np_in[3 * v + 1], np_in[3 * v + 2]};
          const float d0 = gp[0] - q[0], d1 = gp[1] - q[1], d2 = gp[2] - q[2];
          const float dist = sqrtf(((d0 * d0 + d1 * d1) + d2 * d2) + 1e-8f);
          if (dist < best) {
            best = dist;
            bp[0] = q[0], bp[1] = q[1], bp[2] = q[2];
          }
        }
  }
  sdf_out[nd] = best;
  np_out[3 * nd] = bp[0];
  np_out[3 * nd + 1] = bp[1];
  np_out[3 * nd + 2] = bp[2];
}

__global__ __launch_bounds__(kBlock) void k_fill(long long n, float* __restrict__ a, float v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// fixed-order block reduction of kTerms per-thread partial values
template <int kTerms, bool kMax0 = false>
__device__ __forceinline__ void block_reduce_store(float (&v)[kTerms], float* out) {
  __shared__ float part[kBlock / 64][kTerms];
  for (int q = 0; q < kTerms; ++q)
    for (int o = 32; o > 0; o >>= 1) {
      const float u = __shfl_xor(v[q], o);
      v[q] = (kMax0 && q < 2) ? fmaxf(v[q], u) : v[q] + u;
    }
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < kTerms; ++q) part[threadIdx.x >> 6][q] = v[q];
  __syncthreads();
  if (threadIdx.x < kTerms) {
    float s = part[0][threadIdx.x];
    for (int w = 1; w < kBlock / 64; ++w)
      s = (kMax0 && threadIdx.x < 2) ? fmaxf(s, part[w][threadIdx.x]) : s + part[w][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// per block: (sum |m - t|, sum sdf m)
__global__ __launch_bounds__(kBlock) void k_deform_partials(long long nn,
                                                            const float* __restrict__ m,
                                                            const float* __restrict__ t,
                                                            const float* __restrict__ sdf,
                                                            float* __restrict__ partials) {
  float v[2] = {0.f, 0.f};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
       i += (long long)gridDim.x * blockDim.x) {
    v[0] += fabsf(m[i] - t[i]);
    v[1] += sdf[i] * m[i];
  }
  block_reduce_store<2>(v, partials + 2ll * blockIdx.x);
}

// per block: (max m, max t, sum m t, sum m, sum t)
__global__ __launch_bounds__(kBlock) void k_iou_partials(long long nn, const float* __restrict__ m,
                                                         const float* __restrict__ t,
                                                         float* __restrict__ partials) {
  float v[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
       i += (long long)gridDim.x * blockDim.x) {
    v[0] = fmaxf(v[0], m[i]);
    v[1] = fmaxf(v[1], t[i]);
    v[2] += m[i] * t[i];
    v[3] += m[i];
    v[4] += t[i];
  }
  block_reduce_store<5, true>(v, partials + 5ll * blockIdx.x);
}

// out[q] = reduction over the blocks' partials (term q; max for the first
// n_max terms)
__global__ __launch_bounds__(kBlock) void k_reduce_partials(int nb, int terms, int n_max,
                                                            const float* __restrict__ partials,
                                                            float* __restrict__ out) {
  if (threadIdx.x >= terms) return;
  const int q = threadIdx.x;
  float s = partials[q];
  for (int b = 1; b < nb; ++b)
    s = q < n_max ? fmaxf(s, partials[(long long)terms * b + q]) : s + partials[(long long)terms * b + q];
  out[q] = s;
}

// g_m += g_loss (w_density sign(m - t) + w_sdf sdf); sign(0) = 0
__global__ __launch_bounds__(kBlock) void k_deform_loss_adj(long long nn,
                                                            const float* __restrict__ m,
                                                            const float* __restrict__ t,
                                                            const float* __restrict__ sdf,
                                                            float wd, float ws,
                                                            const float* __restrict__ g_loss,
                                                            float* __restrict__ gm) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const float d = m[i] - t[i];
  const float sg = d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f);
  gm[i] += *g_loss * (wd * sg + ws * sdf[i]);
}

static unsigned reduce_blocks(long long nn) {
  const long long b = (nn + kBlock - 1) / kBlock;
  return (unsigned)(b < DT_REDUCE_BLOCKS ? (b < 1 ? 1 : b) : DT_REDUCE_BLOCKS);
}

}  // namespace dt

using namespace dt;

extern "C" {

int dt_mpm_grid_mass(const dt_mpm_params* prm, int32_t n, const float* x, float* grid_m,
                     void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(prm->n_env <= 1, "one environment only");
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && grid_m, "NULL pointer");
  hipLaunchKernelGGL(k_grid_mass, dim3(nblocks(n)), dim3(kBlock), 0, (hipStream_t)stream,
                     to_dev(*prm), (int)n, x, grid_m);
  return check_launch("dt_mpm_grid_mass");
}

int dt_mpm_grid_mass_adj(const dt_mpm_params* prm, int32_t n, const float* x,
                         const float* g_grid_m, float* gx, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(prm->n_env <= 1, "one environment only");
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && g_grid_m && gx, "NULL pointer");
  hipLaunchKernelGGL(k_grid_mass_adj, dim3(nblocks(n)), dim3(kBlock), 0, (hipStream_t)stream,
                     to_dev(*prm), (int)n, x, g_grid_m, gx);
  return check_launch("dt_mpm_grid_mass_adj");
}

int dt_deform_target_sdf(int32_t n_grid, float dx, const float* density, int32_t n_sweeps,
                         float* sdf, float* nearest, float* sdf_tmp, float* nearest_tmp,
                         void* stream) {
  DT_REQUIRE(n_grid > 0 && n_sweeps >= 0, "bad size");
  DT_REQUIRE(density && sdf && nearest && sdf_tmp && nearest_tmp, "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  const long long nn = (long long)n_grid * n_grid * n_grid;
  // the copies start at "no point" (update_target, deform_loss.py:83-86)
  hipLaunchKernelGGL(k_fill, dim3(nblocks(nn)), dim3(kBlock), 0, st, nn, sdf, kSdfInf);
  hipLaunchKernelGGL(k_fill, dim3(nblocks(3 * nn)), dim3(kBlock), 0, st, 3 * nn, nearest, 0.f);
  float *a = sdf, *an = nearest, *b = sdf_tmp, *bn = nearest_tmp;
  for (int it = 0; it < n_sweeps; ++it) {
    hipLaunchKernelGGL(k_sdf_sweep, dim3(nblocks(nn)), dim3(kBlock), 0, st, (int)n_grid, dx,
                       density, (const float*)a, (const float*)an, b, bn);
    float* t = a;
    a = b, b = t;
    t = an;
    an = bn, bn = t;
  }
  if (a != sdf) {   // odd sweep count: the result is in the scratch pair
    hipMemcpyAsync(sdf, a, sizeof(float) * nn, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(nearest, an, sizeof(float) * 3 * nn, hipMemcpyDeviceToDevice, st);
  }
  return check_launch("dt_deform_target_sdf");
}

int dt_deform_loss(int64_t n_nodes, const float* grid_m, const float* target_density,
                   const float* target_sdf, float* scratch, float* out2, void* stream) {
  DT_REQUIRE(n_nodes > 0, "n_nodes must be > 0");
  DT_REQUIRE(grid_m && target_density && target_sdf && scratch && out2, "NULL pointer");
  const unsigned nb = reduce_blocks(n_nodes);
  hipLaunchKernelGGL(k_deform_partials, dim3(nb), dim3(kBlock), 0, (hipStream_t)stream,
                     (long long)n_nodes, grid_m, target_density, target_sdf, scratch);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(64), 0, (hipStream_t)stream, (int)nb, 2, 0,
                     (const float*)scratch, out2);
  return check_launch("dt_deform_loss");
}

int dt_deform_loss_adj(int64_t n_nodes, const float* grid_m, const float* target_density,
                       const float* target_sdf, float w_density, float w_sdf, const float* g_loss,
                       float* g_grid_m, void* stream) {
  DT_REQUIRE(n_nodes > 0, "n_nodes must be > 0");
  DT_REQUIRE(grid_m && target_density && target_sdf && g_loss && g_grid_m, "NULL pointer");
  hipLaunchKernelGGL(k_deform_loss_adj, dim3(nblocks(n_nodes)), dim3(kBlock), 0,
                     (hipStream_t)stream, (long long)n_nodes, grid_m, target_density, target_sdf,
                     w_density, w_sdf, g_loss, g_grid_m);
  return check_launch("dt_deform_loss_adj");
}

int dt_grid_iou_terms(int64_t n_nodes, const float* grid_m, const float* target_density,
                      float* scratch, float* out5, void* stream) {
  DT_REQUIRE(n_nodes > 0, "n_nodes must be > 0");
  DT_REQUIRE(grid_m && target_density && scratch && out5, "NULL pointer");
  const unsigned nb = reduce_blocks(n_nodes);
  hipLaunchKernelGGL(k_iou_partials, dim3(nb), dim3(kBlock), 0, (hipStream_t)stream,
                     (long long)n_nodes, grid_m, target_density, scratch);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(64), 0, (hipStream_t)stream, (int)nb, 5, 2,
                     (const float*)scratch, out5);
  return check_launch("dt_grid_iou_terms");
}

}  // extern "C"
