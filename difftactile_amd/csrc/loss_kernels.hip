// Losses on the measured path (gfx950): centre of mass and tactile markers.
//
// Reference (difftactile/):
//   compute_obj_pos        tasks/grasp_elastic.py:240-243 (mean particle position)
//   extract_markers        sensor_model/fem_sensor.py:161-187
//   project_3d_2d          sensor_model/fisheye_model.py:13-30 (equidistant fisheye)
//   compute_marker_motion  tasks/grasp_plastic.py:240-246 (mean marker offset)
#include "dt_common.h"
#include "dt_device.h"

namespace dt {

// ---------------------------------------------------------------------------
// centre of mass: out[c] += sum_p x[p][c] (float64 accumulation)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_sum_positions(int n, int soa,
                                                          const float* __restrict__ x,
                                                          double* __restrict__ out) {
  __shared__ double red[3][kBlock];
  PL L = pl_vec(soa, n);
  double s[3] = {0.0, 0.0, 0.0};
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (long long)gridDim.x * blockDim.x)
    for (int c = 0; c < 3; ++c) s[c] += (double)x[p * L.sp + c * L.sc];
  for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) atomicAdd(&out[c], red[c][0]);
}

__global__ __launch_bounds__(kBlock) void k_com_adj(int n, int soa, float g0, float g1, float g2,
                                                    float* __restrict__ gx) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  PL L = pl_vec(soa, n);
  float inv = 1.f / (float)n;
  gx[p * L.sp] += g0 * inv;
  gx[p * L.sp + L.sc] += g1 * inv;
  gx[p * L.sp + 2 * L.sc] += g2 * inv;
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

// adjoint of predict_markers w.r.t. the surface projections (scatter)
__global__ __launch_bounds__(kBlock) void k_marker_interp_adj(int nm, int kc,
                                                              const int32_t* __restrict__ idx,
                                                              const float* __restrict__ w,
                                                              const float* __restrict__ gpred,
                                                              float* __restrict__ gloc) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  for (int j = 0; j < kc; ++j) {
    int s = idx[m * kc + j];
    float ww = w[m * kc + j];
    atomic_add_f(gloc + 2 * s, ww * gpred[2 * m]);
    atomic_add_f(gloc + 2 * s + 1, ww * gpred[2 * m + 1]);
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
  for (int c = 0; c < 3; ++c) {
    float g = A.T[0 * 4 + c] * gq[0] + A.T[1 * 4 + c] * gq[1] + A.T[2 * 4 + c] * gq[2];
    atomic_add_f(gpos + 3ll * v + c, g);
  }
}

}  // namespace dt

using namespace dt;

extern "C" {

int dt_loss_sum_positions(int32_t n, int32_t soa, const float* x, double* out3, void* stream) {
  DT_REQUIRE(n >= 0, "n must be >= 0");
  DT_REQUIRE(out3, "NULL pointer");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x, "NULL pointer");
  int blocks = nblocks(n);
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_sum_positions, dim3(blocks), dim3(kBlock), 0, (hipStream_t)stream, n, soa,
                     x, out3);
  return check_launch("dt_loss_sum_positions");
}

int dt_loss_com_adj(int32_t n, int32_t soa, const float* g_com3, float* gx, void* stream) {
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(g_com3 && gx, "NULL pointer");
  hipLaunchKernelGGL(k_com_adj, dim3(nblocks(n)), dim3(kBlock), 0, (hipStream_t)stream, n, soa,
                     g_com3[0], g_com3[1], g_com3[2], gx);
  return check_launch("dt_loss_com_adj");
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
                   int32_t k_closest, const int32_t* interp_idx, const float* interp_w,
                   const float* g_pred, float* g_surf_loc, float* g_pos, void* stream) {
  DT_REQUIRE(n_surface >= 0 && n_markers >= 0 && k_closest > 0, "bad size");
  DT_REQUIRE(inv_trans16 && fisheye5, "NULL host parameter");
  DT_REQUIRE(surface_id && pos && interp_idx && interp_w && g_pred && g_surf_loc && g_pos,
             "NULL pointer");
  MarkerArgs A = marker_args(inv_trans16, fisheye5);
  hipStream_t st = (hipStream_t)stream;
  if (n_markers > 0)
    hipLaunchKernelGGL(k_marker_interp_adj, dim3(nblocks(n_markers)), dim3(kBlock), 0, st,
                       n_markers, k_closest, interp_idx, interp_w, g_pred, g_surf_loc);
  if (n_surface > 0)
    hipLaunchKernelGGL(k_surface_project_adj, dim3(nblocks(n_surface)), dim3(kBlock), 0, st, A,
                       n_surface, surface_id, pos, g_surf_loc, g_pos);
  return check_launch("dt_markers_adj");
}

}  // extern "C"
