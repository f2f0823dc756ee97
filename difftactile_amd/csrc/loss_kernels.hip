// Losses on the measured path (gfx950): centre of mass, tactile markers and
// the pads' contact force.
//
// Reference (difftactile/):
//   compute_obj_pos        tasks/grasp_elastic.py:240-243 (mean particle position)
//   extract_markers        sensor_model/fem_sensor.py:161-187
//   project_3d_2d          sensor_model/fisheye_model.py:13-30 (equidistant fisheye)
//   compute_marker_motion  tasks/grasp_plastic.py:240-246 (mean marker offset)
//   compute_marker_loss    tasks/grasp_plastic.py:254-256
//   compute_contact_force  tasks/grasp_plastic.py:218-230 (1/6 per triangle vertex)
//   compute_force_loss     tasks/grasp_plastic.py:232-238
//
// Every reduction is float32 in a fixed order (per-block partials, then one
// block over the partials), so a loss and its gradient repeat bit for bit;
// the adjoints gather (the marker interpolation through a surface-point CSR)
// or add equal values, never order-dependent float atomics.
#include "dt_common.h"
#include "dt_device.h"

namespace dt {

// ---------------------------------------------------------------------------
// fixed-order sums
// ---------------------------------------------------------------------------
// Sum of v over the block (a wave butterfly, then the block's waves in
// order); thread q < kTerms writes term q, times scale, to out[q].
template <int kTerms>
__device__ __forceinline__ void loss_block_sum(float (&v)[kTerms], float* out, float scale) {
  __shared__ float part[kBlock / 64][kTerms];
  for (int q = 0; q < kTerms; ++q)
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o);
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < kTerms; ++q) part[threadIdx.x >> 6][q] = v[q];
  __syncthreads();
  if ((int)threadIdx.x < kTerms) {
    float t = part[0][threadIdx.x];
    for (int w = 1; w < kBlock / 64; ++w) t += part[w][threadIdx.x];
    out[threadIdx.x] = t * scale;
  }
}

// blocks of a reduction over n items (a function of n only)
static inline unsigned reduce_grid(long long n) {
  const long long b = (n + kBlock - 1) / kBlock;
  return (unsigned)(b < DT_REDUCE_BLOCKS ? (b < 1 ? 1 : b) : DT_REDUCE_BLOCKS);
}

// out[q] = scale * sum over nb blocks of partials[kTerms b + q] (one block)
template <int kTerms>
__global__ __launch_bounds__(kBlock) void k_sum_block_partials(int nb,
                                                               const float* __restrict__ partials,
                                                               float scale, float* __restrict__ out) {
  float v[kTerms];
  for (int q = 0; q < kTerms; ++q)
    v[q] = (int)threadIdx.x < nb ? partials[(long long)kTerms * threadIdx.x + q] : 0.f;
  loss_block_sum<kTerms>(v, out, scale);
}

// ---------------------------------------------------------------------------
// centre of mass
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_com_partials(int n, int soa, const float* __restrict__ x,
                                                         float* __restrict__ partials) {
  PL L = pl_vec(soa, n);
  float s[3] = {0.f, 0.f, 0.f};
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x)
    for (int c = 0; c < 3; ++c) s[c] += x[p * L.sp + c * L.sc];
  loss_block_sum<3>(s, partials + 3ll * blockIdx.x, 1.f);
}

// gx[p] += g / n, g: device float[3]
__global__ __launch_bounds__(kBlock) void k_com_adj(int n, int soa, const float* __restrict__ g,
                                                    float* __restrict__ gx) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  PL L = pl_vec(soa, n);
  float inv = 1.f / (float)n;
  gx[p * L.sp] += g[0] * inv;
  gx[p * L.sp + L.sc] += g[1] * inv;
  gx[p * L.sp + 2 * L.sc] += g[2] * inv;
}

// ---------------------------------------------------------------------------
// contact force of a pad and the force loss
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_contact_force_partials(int nt,
                                                                   const int32_t* __restrict__ tris,
                                                                   const float* __restrict__ ext,
                                                                   float* __restrict__ partials) {
  float s[3] = {0.f, 0.f, 0.f};
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nt;
       t += (long long)gridDim.x * blockDim.x)
    for (int k = 0; k < 3; ++k) {
      const long long v = tris[3 * t + k];
      for (int c = 0; c < 3; ++c) s[c] += ext[3 * v + c];
    }
  loss_block_sum<3>(s, partials + 3ll * blockIdx.x, 1.f);
}

// g_ext[v] += g / 6 for each of the pad's triangles holding v: all of a
// vertex's addends are equal, so the sum does not depend on their order
__global__ __launch_bounds__(kBlock) void k_contact_force_adj(int nt,
                                                              const int32_t* __restrict__ tris,
                                                              const float* __restrict__ g,
                                                              float* __restrict__ g_ext) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const float e[3] = {g[0] * (1.f / 6.f), g[1] * (1.f / 6.f), g[2] * (1.f / 6.f)};
  for (int k = 0; k < 3; ++k) {
    const long long v = tris[3 * t + k];
    for (int c = 0; c < 3; ++c) atomic_add_f(g_ext + 3 * v + c, e[c]);
  }
}

constexpr int kMaxForcePads = 8;
struct ForceLossArgs {
  float inv_rot[kMaxForcePads][9];
  float mu, beta;
  int n;
};

// p_k = inv_rot_k cf_k; loss = beta sum_k |p_k.x| / |mu p_k.y| when every
// |p_k.x| > 1e-3, else 0.  out[0] = loss, out[1 + 3k + c] = d loss / d cf_k[c].
__global__ __launch_bounds__(64) void k_force_loss(ForceLossArgs A, const float* __restrict__ cf,
                                                   float* __restrict__ out) {
  if (threadIdx.x != 0) return;
  float p[kMaxForcePads][3];
  bool on = A.n > 0;
  for (int k = 0; k < A.n; ++k) {
    for (int r = 0; r < 3; ++r)
      p[k][r] = A.inv_rot[k][3 * r] * cf[3 * k] + A.inv_rot[k][3 * r + 1] * cf[3 * k + 1] +
                A.inv_rot[k][3 * r + 2] * cf[3 * k + 2];
    on = on && fabsf(p[k][0]) > 1e-3f;
  }
  float loss = 0.f;
  for (int k = 0; k < A.n; ++k) {
    float gp[3] = {0.f, 0.f, 0.f};
    if (on) {
      const float den = fabsf(A.mu * p[k][1]);
      loss += A.beta * (fabsf(p[k][0]) / den);
      // d/dx |x| / |mu y| = sign(x) / |mu y|; d/dy = -|x| sign(y) / (|mu| y^2)
      const float sx = p[k][0] > 0.f ? 1.f : (p[k][0] < 0.f ? -1.f : 0.f);
      const float sy = p[k][1] > 0.f ? 1.f : (p[k][1] < 0.f ? -1.f : 0.f);
      gp[0] = A.beta * sx / den;
      gp[1] = -A.beta * fabsf(p[k][0]) * sy / (fabsf(A.mu) * p[k][1] * p[k][1]);
    }
    // d loss / d cf = inv_rot^T d loss / d p
    for (int c = 0; c < 3; ++c)
      out[1 + 3 * k + c] = A.inv_rot[k][c] * gp[0] + A.inv_rot[k][3 + c] * gp[1] +
                           A.inv_rot[k][6 + c] * gp[2];
  }
  out[0] = loss;
}

// ---------------------------------------------------------------------------
// markers
// ---------------------------------------------------------------------------
struct Fisheye {
  float f, m, cx, cy, z_offset;
};

struct Proj {
  float a[3], an, c, cc, theta, om, r, co, so;
  bool clamped;
};

// project_3d_2d: a[2] += z_offset; theta = acos(clamp(a2/|a|_eps));
// omega = atan2(a1, a0 + 1e-8) + pi; r = m f theta; p = r (cos, sin)(omega) + (cx, cy)
__device__ __forceinline__ void project_fwd(const Fisheye& K, const float (&cam)[3], Proj& q,
                                            float (&p)[2]) {
  q.a[0] = cam[0];
  q.a[1] = cam[1];
  q.a[2] = cam[2] + K.z_offset;
  q.an = sqrtf(q.a[0] * q.a[0] + q.a[1] * q.a[1] + q.a[2] * q.a[2] + 1e-10f);
  q.c = q.a[2] / q.an;
  q.cc = fmaxf(-1.f, fminf(1.f, q.c));
  q.clamped = q.cc != q.c;
  q.theta = acosf(q.cc);
  q.om = atan2f(q.a[1], q.a[0] + 1e-8f) + 3.14159265358979323846f;
  q.r = K.m * K.f * q.theta;
  q.co = cosf(q.om);
  q.so = sinf(q.om);
  p[0] = q.r * q.co + K.cx;
  p[1] = q.r * q.so + K.cy;
}

__device__ __forceinline__ void project_adj(const Fisheye& K, const Proj& q, const float (&gp)[2],
                                            float (&gcam)[3]) {
  float gr = gp[0] * q.co + gp[1] * q.so;
  float gom = -gp[0] * q.r * q.so + gp[1] * q.r * q.co;
  float gth = K.m * K.f * gr;
  float gc = 0.f;
  if (!q.clamped) {
    float s = sqrtf(fmaxf(1.f - q.cc * q.cc, 1e-12f));
    gc = -gth / s;
  }
  // c = a2 / an
  float inv = 1.f / q.an, inv3 = inv * inv * inv;
  float ga[3];
  for (int k = 0; k < 3; ++k) ga[k] = -gc * q.a[2] * q.a[k] * inv3;
  ga[2] += gc * inv;
  // omega = atan2(a1, a0 + 1e-8)
  float x0 = q.a[0] + 1e-8f;
  float d = x0 * x0 + q.a[1] * q.a[1];
  if (d > 0.f) {
    ga[0] += -gom * q.a[1] / d;
    ga[1] += gom * x0 / d;
  }
  for (int k = 0; k < 3; ++k) gcam[k] = ga[k];
}

// camera-frame point of a world position: inv_T [p, 1], then (x, z, y)
__device__ __forceinline__ void to_cam(const float* __restrict__ T, const float (&p)[3],
                                       float (&cam)[3]) {
  float q[3];
  for (int r = 0; r < 3; ++r) q[r] = T[r * 4 + 0] * p[0] + T[r * 4 + 1] * p[1] + T[r * 4 + 2] * p[2] + T[r * 4 + 3];
  cam[0] = q[0];
  cam[1] = q[2];
  cam[2] = q[1];
}

struct MarkerArgs {
  float T[16];   // inverse pad transform (row-major 4x4)
  Fisheye K;
};

__global__ __launch_bounds__(kBlock) void k_surface_project(MarkerArgs A, int ns,
                                                            const int32_t* __restrict__ sid,
                                                            const float* __restrict__ pos,
                                                            const float* __restrict__ vpos,
                                                            float* __restrict__ loc,
                                                            float* __restrict__ vloc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  int v = sid[i];
  float p[3] = {pos[3ll * v], pos[3ll * v + 1], pos[3ll * v + 2]};
  float cam[3], out[2];
  Proj q;
  to_cam(A.T, p, cam);
  project_fwd(A.K, cam, q, out);
  loc[2 * i] = out[0];
  loc[2 * i + 1] = out[1];
  float vp[3] = {vpos[3ll * v], vpos[3ll * v + 1], vpos[3ll * v + 2]};
  to_cam(A.T, vp, cam);
  project_fwd(A.K, cam, q, out);
  vloc[2 * i] = out[0];
  vloc[2 * i + 1] = out[1];
}

__global__ __launch_bounds__(kBlock) void k_marker_interp(int nm, int kc,
                                                          const int32_t* __restrict__ idx,
                                                          const float* __restrict__ w,
                                                          const float* __restrict__ loc,
                                                          const float* __restrict__ vloc,
                                                          float* __restrict__ pred,
                                                          float* __restrict__ virt) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  float a0 = 0.f, a1 = 0.f, b0 = 0.f, b1 = 0.f;
  for (int j = 0; j < kc; ++j) {
    int s = idx[m * kc + j];
    float ww = w[m * kc + j];
    a0 += ww * loc[2 * s];
    a1 += ww * loc[2 * s + 1];
    b0 += ww * vloc[2 * s];
    b1 += ww * vloc[2 * s + 1];
  }
  pred[2 * m] = a0;
  pred[2 * m + 1] = a1;
  virt[2 * m] = b0;
  virt[2 * m + 1] = b1;
}

// adjoint of predict_markers w.r.t. the surface projections, gathered per
// surface point over its (marker, neighbour) entries in entry order:
// entries s_start[s] .. s_start[s+1] of s_entry hold m k_closest + j for every
// interp_idx[m][j] == s
__global__ __launch_bounds__(kBlock) void k_marker_interp_adj(int ns, int kc,
                                                              const int32_t* __restrict__ s_start,
                                                              const int32_t* __restrict__ s_entry,
                                                              const float* __restrict__ w,
                                                              const float* __restrict__ gpred,
                                                              float* __restrict__ gloc) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  float g0 = 0.f, g1 = 0.f;
  for (int e = s_start[s]; e < s_start[s + 1]; ++e) {
    const int mj = s_entry[e];
    const int m = mj / kc;
    g0 += w[mj] * gpred[2 * m];
    g1 += w[mj] * gpred[2 * m + 1];
  }
  gloc[2 * s] = g0;
  gloc[2 * s + 1] = g1;
}

// mean marker offset: out2 = scale * sum_m (pred[m] - virt[m]) (one block partial per block)
__global__ __launch_bounds__(kBlock) void k_marker_offset_partials(int nm,
                                                                   const float* __restrict__ pred,
                                                                   const float* __restrict__ virt,
                                                                   float* __restrict__ partials) {
  float s[2] = {0.f, 0.f};
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < nm;
       m += (long long)gridDim.x * blockDim.x) {
    s[0] += pred[2 * m] - virt[2 * m];
    s[1] += pred[2 * m + 1] - virt[2 * m + 1];
  }
  loss_block_sum<2>(s, partials + 2ll * blockIdx.x, 1.f);
}

// g_pred[m] += g2 / nm, g_virt[m] -= g2 / nm (g2: device float[2])
__global__ __launch_bounds__(kBlock) void k_marker_offset_adj(int nm, const float* __restrict__ g2,
                                                              float* __restrict__ gpred,
                                                              float* __restrict__ gvirt) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const float inv = 1.f / (float)nm;
  const float a = g2[0] * inv, b = g2[1] * inv;
  gpred[2 * m] += a;
  gpred[2 * m + 1] += b;
  if (gvirt) {
    gvirt[2 * m] -= a;
    gvirt[2 * m + 1] -= b;
  }
}

__global__ __launch_bounds__(kBlock) void k_surface_project_adj(MarkerArgs A, int ns,
                                                                const int32_t* __restrict__ sid,
                                                                const float* __restrict__ pos,
                                                                const float* __restrict__ gloc,
                                                                float* __restrict__ gpos) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  int v = sid[i];
  float p[3] = {pos[3ll * v], pos[3ll * v + 1], pos[3ll * v + 2]};
  float cam[3], out[2];
  Proj q;
  to_cam(A.T, p, cam);
  project_fwd(A.K, cam, q, out);
  float gp[2] = {gloc[2 * i], gloc[2 * i + 1]};
  float gcam[3];
  project_adj(A.K, q, gp, gcam);
  // cam = (q0, q2, q1), q = T[:3,:3] p + t
  float gq[3] = {gcam[0], gcam[2], gcam[1]};
  // surface ids are distinct: one thread per vertex
  for (int c = 0; c < 3; ++c) {
    float g = A.T[0 * 4 + c] * gq[0] + A.T[1 * 4 + c] * gq[1] + A.T[2 * 4 + c] * gq[2];
    gpos[3ll * v + c] += g;
  }
}

}  // namespace dt

using namespace dt;

extern "C" {

int dt_loss_com(int32_t n, int32_t soa, const float* x, float* scratch, float* out3,
                void* stream) {
  DT_REQUIRE(n >= 1, "n must be >= 1");
  DT_REQUIRE(x && scratch && out3, "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = reduce_grid(n);
  hipLaunchKernelGGL(k_com_partials, dim3(nb), dim3(kBlock), 0, st, n, soa, x, scratch);
  hipLaunchKernelGGL(k_sum_block_partials<3>, dim3(1), dim3(kBlock), 0, st, (int)nb, scratch,
                     1.f / (float)n, out3);
  return check_launch("dt_loss_com");
}

int dt_loss_com_adj(int32_t n, int32_t soa, const float* g_com3, float* gx, void* stream) {
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(g_com3 && gx, "NULL pointer");
  hipLaunchKernelGGL(k_com_adj, dim3(nblocks(n)), dim3(kBlock), 0, (hipStream_t)stream, n, soa,
                     g_com3, gx);
  return check_launch("dt_loss_com_adj");
}

int dt_contact_force_sum(int32_t n_tris, const int32_t* tris, const float* ext, float* scratch,
                         float* out3, void* stream) {
  DT_REQUIRE(n_tris >= 0, "n_tris must be >= 0");
  DT_REQUIRE(scratch && out3 && (n_tris == 0 || (tris && ext)), "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = reduce_grid(n_tris);
  hipLaunchKernelGGL(k_contact_force_partials, dim3(nb), dim3(kBlock), 0, st, n_tris, tris, ext,
                     scratch);
  hipLaunchKernelGGL(k_sum_block_partials<3>, dim3(1), dim3(kBlock), 0, st, (int)nb, scratch,
                     1.f / 6.f, out3);
  return check_launch("dt_contact_force_sum");
}

int dt_contact_force_sum_adj(int32_t n_tris, const int32_t* tris, const float* g3, float* g_ext,
                             void* stream) {
  DT_REQUIRE(n_tris >= 0, "n_tris must be >= 0");
  if (n_tris == 0) return DT_OK;
  DT_REQUIRE(tris && g3 && g_ext, "NULL pointer");
  hipLaunchKernelGGL(k_contact_force_adj, dim3(nblocks(n_tris)), dim3(kBlock), 0,
                     (hipStream_t)stream, n_tris, tris, g3, g_ext);
  return check_launch("dt_contact_force_sum_adj");
}

int dt_force_loss(int32_t n_pads, const float* cf, const float* inv_rot9, float mu, float beta,
                  float* out, void* stream) {
  DT_REQUIRE(n_pads >= 1 && n_pads <= kMaxForcePads, "n_pads must be in [1, 8]");
  DT_REQUIRE(cf && inv_rot9 && out, "NULL pointer");
  ForceLossArgs A;
  for (int k = 0; k < n_pads; ++k)
    for (int i = 0; i < 9; ++i) A.inv_rot[k][i] = inv_rot9[9 * k + i];
  A.mu = mu;
  A.beta = beta;
  A.n = n_pads;
  hipLaunchKernelGGL(k_force_loss, dim3(1), dim3(64), 0, (hipStream_t)stream, A, cf, out);
  return check_launch("dt_force_loss");
}

int dt_marker_offset(int32_t n_markers, const float* pred, const float* virt, float* scratch,
                     float* out2, void* stream) {
  DT_REQUIRE(n_markers >= 1, "n_markers must be >= 1");
  DT_REQUIRE(pred && virt && scratch && out2, "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = reduce_grid(n_markers);
  hipLaunchKernelGGL(k_marker_offset_partials, dim3(nb), dim3(kBlock), 0, st, n_markers, pred,
                     virt, scratch);
  hipLaunchKernelGGL(k_sum_block_partials<2>, dim3(1), dim3(kBlock), 0, st, (int)nb, scratch,
                     1.f / (float)n_markers, out2);
  return check_launch("dt_marker_offset");
}

int dt_marker_offset_adj(int32_t n_markers, const float* g2, float* g_pred, float* g_virt,
                         void* stream) {
  DT_REQUIRE(n_markers >= 0, "n_markers must be >= 0");
  if (n_markers == 0) return DT_OK;
  DT_REQUIRE(g2 && g_pred, "NULL pointer");
  hipLaunchKernelGGL(k_marker_offset_adj, dim3(nblocks(n_markers)), dim3(kBlock), 0,
                     (hipStream_t)stream, n_markers, g2, g_pred, g_virt);
  return check_launch("dt_marker_offset_adj");
}

static MarkerArgs marker_args(const float* inv_trans16, const float* fisheye5) {
  MarkerArgs A;
  for (int i = 0; i < 16; ++i) A.T[i] = inv_trans16[i];
  A.K.f = fisheye5[0];
  A.K.m = fisheye5[1];
  A.K.cx = fisheye5[2];
  A.K.cy = fisheye5[3];
  A.K.z_offset = fisheye5[4];
  return A;
}

int dt_markers_fwd(int32_t n_surface, const int32_t* surface_id, const float* pos,
                   const float* vpos, const float* inv_trans16, const float* fisheye5,
                   int32_t n_markers, int32_t k_closest, const int32_t* interp_idx,
                   const float* interp_w, float* surf_loc, float* surf_vloc, float* pred,
                   float* virt, void* stream) {
  DT_REQUIRE(n_surface >= 0 && n_markers >= 0 && k_closest > 0, "bad size");
  DT_REQUIRE(inv_trans16 && fisheye5, "NULL host parameter");
  DT_REQUIRE(surface_id && pos && vpos && interp_idx && interp_w && surf_loc && surf_vloc &&
                 pred && virt,
             "NULL pointer");
  MarkerArgs A = marker_args(inv_trans16, fisheye5);
  hipStream_t st = (hipStream_t)stream;
  if (n_surface > 0)
    hipLaunchKernelGGL(k_surface_project, dim3(nblocks(n_surface)), dim3(kBlock), 0, st, A,
                       n_surface, surface_id, pos, vpos, surf_loc, surf_vloc);
  if (n_markers > 0)
    hipLaunchKernelGGL(k_marker_interp, dim3(nblocks(n_markers)), dim3(kBlock), 0, st, n_markers,
                       k_closest, interp_idx, interp_w, surf_loc, surf_vloc, pred, virt);
  return check_launch("dt_markers_fwd");
}

int dt_markers_adj(int32_t n_surface, const int32_t* surface_id, const float* pos,
                   const float* inv_trans16, const float* fisheye5, int32_t n_markers,
                   int32_t k_closest, const int32_t* surf_start, const int32_t* surf_entry,
                   const float* interp_w, const float* g_pred, float* g_surf_loc, float* g_pos,
                   void* stream) {
  DT_REQUIRE(n_surface >= 0 && n_markers >= 0 && k_closest > 0, "bad size");
  DT_REQUIRE(inv_trans16 && fisheye5, "NULL host parameter");
  DT_REQUIRE(surface_id && pos && surf_start && surf_entry && interp_w && g_pred && g_surf_loc &&
                 g_pos,
             "NULL pointer");
  MarkerArgs A = marker_args(inv_trans16, fisheye5);
  hipStream_t st = (hipStream_t)stream;
  if (n_surface > 0)
    hipLaunchKernelGGL(k_marker_interp_adj, dim3(nblocks(n_surface)), dim3(kBlock), 0, st,
                       n_surface, k_closest, surf_start, surf_entry, interp_w, g_pred,
                       g_surf_loc);
  if (n_surface > 0)
    hipLaunchKernelGGL(k_surface_project_adj, dim3(nblocks(n_surface)), dim3(kBlock), 0, st, A,
                       n_surface, surface_id, pos, g_surf_loc, g_pos);
  return check_launch("dt_markers_adj");
}

}  // extern "C"
