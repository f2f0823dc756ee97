// FEM tactile-pad kernels for gfx950 (forward + adjoint).
//
// Reference kernels (difftactile/sensor_model/):
//   init_pos B-matrix  fem_sensor.py:334-338 (gripper_fem.py:299-303)
//   update             fem_sensor.py:408-440 (gripper_fem.py:374-401)
//   update2            fem_sensor.py:443-457 (dome) / gripper_fem.py:403-418 (gripper)
// Per-tet forces are scattered into the vertex velocities with global float
// atomics, as the reference's `vel[f, v] += ...` does.
#include "dt_common.h"
#include "dt_device.h"

namespace dt {

struct FemP {
  float mu, lam, dt, mass;
};

__device__ __forceinline__ void ld3(const float* p, float (&v)[3]) {
  v[0] = p[0];
  v[1] = p[1];
  v[2] = p[2];
}

__device__ __forceinline__ void cross3(const float (&a)[3], const float (&b)[3], float (&c)[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

__global__ __launch_bounds__(kBlock) void k_fem_init_B(int nt, const int32_t* __restrict__ f2v,
                                                       const float* __restrict__ pos,
                                                       float* __restrict__ B) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int iv[4] = {f2v[4 * t], f2v[4 * t + 1], f2v[4 * t + 2], f2v[4 * t + 3]};
  float p[4][3];
  for (int k = 0; k < 4; ++k) ld3(pos + 3ll * iv[k], p[k]);
  M3 D;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) D.a[r * 3 + c] = p[c][r] - p[3][r];
  st_m3(B + 9ll * t, m3_inv(D));
}

// Forward per-tet quantities.
struct TetFwd {
  M3 D, F, dJdF, P;
  float det_D, vol, J, IC;
};

__device__ __forceinline__ void tet_fwd(const FemP& P, const float (&p)[4][3], const M3& B,
                                        TetFwd& t) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) t.D.a[r * 3 + c] = p[c][r] - p[3][r];
  t.det_D = m3_det(t.D);
  t.vol = fabsf(t.det_D) / 6.f;
  t.F = m3_mul(t.D, B);
  t.J = m3_det(t.F);
  t.IC = 0.f;
  for (int i = 0; i < 9; ++i) t.IC += t.F.a[i] * t.F.a[i];
  float f0[3] = {t.F.a[0], t.F.a[3], t.F.a[6]};
  float f1[3] = {t.F.a[1], t.F.a[4], t.F.a[7]};
  float f2[3] = {t.F.a[2], t.F.a[5], t.F.a[8]};
  float c0[3], c1[3], c2[3];
  cross3(f1, f2, c0);
  cross3(f2, f0, c1);
  cross3(f0, f1, c2);
  for (int r = 0; r < 3; ++r) {
    t.dJdF.a[r * 3 + 0] = c0[r];
    t.dJdF.a[r * 3 + 1] = c1[r];
    t.dJdF.a[r * 3 + 2] = c2[r];
  }
  float alpha = 1.f + 0.75f * P.mu / P.lam;
  float a = P.mu * (1.f - 1.f / (t.IC + 1.f));
  float b = P.lam * (t.J - alpha);
  for (int i = 0; i < 9; ++i) t.P.a[i] = a * t.F.a[i] + b * t.dJdF.a[i];
}

// kOut == 0: scatter into vel with float atomics (the reference's vel[v] +=);
// kOut == 1: write the tet's four vertex contributions to tbuf[4 t + k] for the
// per-vertex gather k_fem_vertex_sum (no atomics, fixed summation order).
template <int kOut>
__global__ __launch_bounds__(kBlock) void k_fem_update(FemP P, int nt,
                                                       const int32_t* __restrict__ f2v,
                                                       const float* __restrict__ B,
                                                       const float* __restrict__ pos,
                                                       float* __restrict__ vel,
                                                       float4* __restrict__ tbuf) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int iv[4] = {f2v[4 * t], f2v[4 * t + 1], f2v[4 * t + 2], f2v[4 * t + 3]};
  float p[4][3];
  for (int k = 0; k < 4; ++k) ld3(pos + 3ll * iv[k], p[k]);
  M3 Bm = ld_m3(B + 9ll * t);
  TetFwd tf;
  tet_fwd(P, p, Bm, tf);
  // H = -V P B^T; column k is the force on vertex k
  M3 H = m3_mul_bt(tf.P, Bm);
  float s = -tf.vol * P.dt / P.mass;
  float f3x = 0.f, f3y = 0.f, f3z = 0.f;
  for (int k = 0; k < 3; ++k) {
    float fx = s * H.a[0 * 3 + k], fy = s * H.a[1 * 3 + k], fz = s * H.a[2 * 3 + k];
    if (kOut == 0) {
      atomic_add_f(vel + 3ll * iv[k] + 0, fx);
      atomic_add_f(vel + 3ll * iv[k] + 1, fy);
      atomic_add_f(vel + 3ll * iv[k] + 2, fz);
      atomic_add_f(vel + 3ll * iv[3] + 0, -fx);
      atomic_add_f(vel + 3ll * iv[3] + 1, -fy);
      atomic_add_f(vel + 3ll * iv[3] + 2, -fz);
    } else {
      tbuf[4ll * t + k] = make_float4(fx, fy, fz, 0.f);
      f3x -= fx;
      f3y -= fy;
      f3z -= fz;
    }
  }
  if (kOut == 1) tbuf[4ll * t + 3] = make_float4(f3x, f3y, f3z, 0.f);
}

// out[v] += sum of tbuf[e] over the vertex's incident (tet, corner) entries e
// (CSR by vertex).  kVertexLanes lanes share a vertex: lane r sums entries
// r, r + kVertexLanes, ... in order, then a fixed butterfly adds the lanes'
// sums, so the summation order is fixed and the short per-lane chains of
// dependent loads overlap.
constexpr int kVertexLanes = 8;
__global__ __launch_bounds__(kBlock) void k_fem_vertex_sum(int nv,
                                                           const int32_t* __restrict__ vstart,
                                                           const int32_t* __restrict__ vtets,
                                                           const float4* __restrict__ tbuf,
                                                           float* __restrict__ out) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = (int)(g / kVertexLanes), r = (int)(g % kVertexLanes);
  float sx = 0.f, sy = 0.f, sz = 0.f;
  if (v < nv) {
    for (int e = vstart[v] + r; e < vstart[v + 1]; e += kVertexLanes) {
      const float4 c = tbuf[vtets[e]];
      sx += c.x;
      sy += c.y;
      sz += c.z;
    }
  }
  for (int o = kVertexLanes / 2; o > 0; o >>= 1) {
    sx += __shfl_xor(sx, o);
    sy += __shfl_xor(sy, o);
    sz += __shfl_xor(sz, o);
  }
  if (v < nv && r == 0) {
    out[3ll * v + 0] += sx;
    out[3ll * v + 1] += sy;
    out[3ll * v + 2] += sz;
  }
}

// Adjoint of update: given g_vel (adjoint of vel[f] after the kernel), add
// the adjoint of pos[f].  vel[f] itself passes its adjoint through unchanged.
template <int kOut>
__global__ __launch_bounds__(kBlock) void k_fem_update_adj(FemP P, int nt,
                                                           const int32_t* __restrict__ f2v,
                                                           const float* __restrict__ B,
                                                           const float* __restrict__ pos,
                                                           const float* __restrict__ gvel,
                                                           float* __restrict__ gpos,
                                                           float4* __restrict__ tbuf) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int iv[4] = {f2v[4 * t], f2v[4 * t + 1], f2v[4 * t + 2], f2v[4 * t + 3]};
  float p[4][3], g[4][3];
  for (int k = 0; k < 4; ++k) {
    ld3(pos + 3ll * iv[k], p[k]);
    ld3(gvel + 3ll * iv[k], g[k]);
  }
  M3 Bm = ld_m3(B + 9ll * t);
  TetFwd tf;
  tet_fwd(P, p, Bm, tf);
  // L = (dt/mass) sum_k<3 (g_k - g_3) . H[:,k],  H = -V P B^T
  float c = P.dt / P.mass;
  M3 gH;
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) gH.a[r * 3 + k] = c * (g[k][r] - g[3][r]);
  M3 PBt = m3_mul_bt(tf.P, Bm);
  float gvol = 0.f;
  for (int i = 0; i < 9; ++i) gvol -= gH.a[i] * PBt.a[i];
  M3 gP = m3_mul(gH, Bm);
  for (int i = 0; i < 9; ++i) gP.a[i] *= -tf.vol;
  // P = a F + b dJdF, a = mu (1 - 1/(IC+1)), b = lam (J - alpha)
  float alpha = 1.f + 0.75f * P.mu / P.lam;
  float a = P.mu * (1.f - 1.f / (tf.IC + 1.f));
  float b = P.lam * (tf.J - alpha);
  float gPF = 0.f, gPdJ = 0.f;
  for (int i = 0; i < 9; ++i) {
    gPF += gP.a[i] * tf.F.a[i];
    gPdJ += gP.a[i] * tf.dJdF.a[i];
  }
  float ga = gPF;                   // d/da
  float gb = gPdJ;                  // d/db
  float gIC = ga * P.mu / ((tf.IC + 1.f) * (tf.IC + 1.f));
  float gJ = gb * P.lam;
  M3 gF;
  for (int i = 0; i < 9; ++i) gF.a[i] = a * gP.a[i] + 2.f * gIC * tf.F.a[i] + gJ * tf.dJdF.a[i];
  // b * d(dJdF): columns c0 = f1 x f2, c1 = f2 x f0, c2 = f0 x f1
  float f0[3] = {tf.F.a[0], tf.F.a[3], tf.F.a[6]};
  float f1[3] = {tf.F.a[1], tf.F.a[4], tf.F.a[7]};
  float f2[3] = {tf.F.a[2], tf.F.a[5], tf.F.a[8]};
  float g0[3] = {b * gP.a[0], b * gP.a[3], b * gP.a[6]};
  float g1[3] = {b * gP.a[1], b * gP.a[4], b * gP.a[7]};
  float g2[3] = {b * gP.a[2], b * gP.a[5], b * gP.a[8]};
  float t0[3], t1[3];
  // d(a x b).g = da.(b x g) + db.(g x a)
  float gf0[3] = {0.f, 0.f, 0.f}, gf1[3] = {0.f, 0.f, 0.f}, gf2[3] = {0.f, 0.f, 0.f};
  cross3(f2, g0, t0); cross3(g0, f1, t1);             // c0 = f1 x f2
  for (int r = 0; r < 3; ++r) { gf1[r] += t0[r]; gf2[r] += t1[r]; }
  cross3(f0, g1, t0); cross3(g1, f2, t1);             // c1 = f2 x f0
  for (int r = 0; r < 3; ++r) { gf2[r] += t0[r]; gf0[r] += t1[r]; }
  cross3(f1, g2, t0); cross3(g2, f0, t1);             // c2 = f0 x f1
  for (int r = 0; r < 3; ++r) { gf0[r] += t0[r]; gf1[r] += t1[r]; }
  for (int r = 0; r < 3; ++r) {
    gF.a[r * 3 + 0] += gf0[r];
    gF.a[r * 3 + 1] += gf1[r];
    gF.a[r * 3 + 2] += gf2[r];
  }
  // F = D B;  vol = |det D| / 6
  M3 gD = m3_mul_bt(gF, Bm);
  M3 cof = m3_cof(tf.D);
  float sg = tf.det_D > 0.f ? 1.f : (tf.det_D < 0.f ? -1.f : 0.f);
  for (int i = 0; i < 9; ++i) gD.a[i] += gvol * sg / 6.f * cof.a[i];
  // D = [a-d, b-d, c-d]
  if (kOut == 0) {
    for (int r = 0; r < 3; ++r) {
      float s3 = 0.f;
      for (int k = 0; k < 3; ++k) {
        float v = gD.a[r * 3 + k];
        atomic_add_f(gpos + 3ll * iv[k] + r, v);
        s3 += v;
      }
      atomic_add_f(gpos + 3ll * iv[3] + r, -s3);
    }
  } else {
    float s3[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < 3; ++k) {
      tbuf[4ll * t + k] = make_float4(gD.a[0 * 3 + k], gD.a[1 * 3 + k], gD.a[2 * 3 + k], 0.f);
      for (int r = 0; r < 3; ++r) s3[r] += gD.a[r * 3 + k];
    }
    tbuf[4ll * t + 3] = make_float4(-s3[0], -s3[1], -s3[2], 0.f);
  }
}

// update2: v = vel + dt ext / mass; base layer takes the control velocity;
// dome: vel' = v; gripper (cv_next != NULL): vel' = v + cv_next - cv.
__global__ __launch_bounds__(kBlock) void k_fem_update2(FemP P, int nv,
                                                        const int32_t* __restrict__ layer,
                                                        const float* __restrict__ pos,
                                                        const float* __restrict__ vel,
                                                        const float* __restrict__ vpos,
                                                        const float* __restrict__ ext,
                                                        const float* __restrict__ cv,
                                                        const float* __restrict__ cv1,
                                                        float* __restrict__ pos1,
                                                        float* __restrict__ vel1,
                                                        float* __restrict__ vpos1) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  bool base = layer[i] == 0;
  for (int a = 0; a < 3; ++a) {
    long long o = 3ll * i + a;
    float v = vel[o] + P.dt * ext[o] / P.mass;
    if (base) v = cv[o];
    vel1[o] = cv1 ? v + (cv1[o] - cv[o]) : v;
    pos1[o] = pos[o] + P.dt * v;
    vpos1[o] = vpos[o] + P.dt * cv[o];
  }
}

__global__ __launch_bounds__(kBlock) void k_fem_update2_adj(
    FemP P, int nv, const int32_t* __restrict__ layer, const float* __restrict__ gpos1,
    const float* __restrict__ gvel1, const float* __restrict__ gvpos1, float* __restrict__ gpos,
    float* __restrict__ gvel, float* __restrict__ gvpos, float* __restrict__ gext,
    float* __restrict__ gcv, float* __restrict__ gcv1) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  bool base = layer[i] == 0;
  for (int a = 0; a < 3; ++a) {
    long long o = 3ll * i + a;
    float gv = gvel1[o] + P.dt * gpos1[o];
    gpos[o] += gpos1[o];
    gvpos[o] += gvpos1[o];
    float gc = P.dt * gvpos1[o];
    if (gcv1) {
      gcv1[o] += gvel1[o];
      gc -= gvel1[o];
    }
    if (base) {
      gc += gv;
    } else {
      gvel[o] += gv;
      gext[o] += P.dt / P.mass * gv;
    }
    if (gcv) gcv[o] += gc;
  }
}

__global__ __launch_bounds__(kBlock) void k_clamp(long long n, float* __restrict__ a, float lim) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = fminf(fmaxf(a[i], -lim), lim);
}

static FemP fem_dev(const dt_fem_params* p) {
  FemP d;
  d.mu = p->mu;
  d.lam = p->lam;
  d.dt = p->dt;
  d.mass = p->mass;
  return d;
}

static int check_fem(const dt_fem_params* p) {
  DT_REQUIRE(p != nullptr, "dt_fem_params is NULL");
  DT_REQUIRE(p->mass > 0.f && p->lam != 0.f, "fem params: mass must be > 0 and lam != 0");
  return DT_OK;
}

}  // namespace dt

using namespace dt;

extern "C" {

int dt_fem_init_B(int32_t n_tets, const int32_t* f2v, const float* pos0, float* B, void* stream) {
  DT_REQUIRE(n_tets >= 0, "n_tets must be >= 0");
  if (n_tets == 0) return DT_OK;
  DT_REQUIRE(f2v && pos0 && B, "NULL pointer");
  hipLaunchKernelGGL(k_fem_init_B, dim3(nblocks(n_tets)), dim3(kBlock), 0, (hipStream_t)stream,
                     n_tets, f2v, pos0, B);
  return check_launch("dt_fem_init_B");
}

int dt_fem_update(const dt_fem_params* prm, int32_t n_tets, const int32_t* f2v, const float* B,
                  const float* pos, float* vel, void* stream) {
  int rc = check_fem(prm);
  if (rc) return rc;
  DT_REQUIRE(n_tets >= 0, "n_tets must be >= 0");
  if (n_tets == 0) return DT_OK;
  DT_REQUIRE(f2v && B && pos && vel, "NULL pointer");
  hipLaunchKernelGGL(k_fem_update<0>, dim3(nblocks(n_tets)), dim3(kBlock), 0, (hipStream_t)stream,
                     fem_dev(prm), n_tets, f2v, B, pos, vel, (float4*)nullptr);
  return check_launch("dt_fem_update");
}

int dt_fem_update_csr(const dt_fem_params* prm, int32_t n_tets, int32_t n_verts,
                      const int32_t* f2v, const float* B, const float* pos, float* vel,
                      const int32_t* vert_start, const int32_t* vert_tets, float* tet_buf,
                      void* stream) {
  int rc = check_fem(prm);
  if (rc) return rc;
  DT_REQUIRE(n_tets >= 0 && n_verts >= 0, "n_tets and n_verts must be >= 0");
  if (n_tets == 0) return DT_OK;
  DT_REQUIRE(f2v && B && pos && vel && vert_start && vert_tets && tet_buf, "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fem_update<1>, dim3(nblocks(n_tets)), dim3(kBlock), 0, st, fem_dev(prm),
                     n_tets, f2v, B, pos, (float*)nullptr, (float4*)tet_buf);
  hipLaunchKernelGGL(k_fem_vertex_sum, dim3(nblocks((long long)n_verts * kVertexLanes)),
                     dim3(kBlock), 0, st, n_verts, vert_start, vert_tets, (const float4*)tet_buf,
                     vel);
  return check_launch("dt_fem_update_csr");
}

int dt_fem_update_adj(const dt_fem_params* prm, int32_t n_tets, const int32_t* f2v,
                      const float* B, const float* pos, const float* g_vel, float* g_pos,
                      void* stream) {
  int rc = check_fem(prm);
  if (rc) return rc;
  DT_REQUIRE(n_tets >= 0, "n_tets must be >= 0");
  if (n_tets == 0) return DT_OK;
  DT_REQUIRE(f2v && B && pos && g_vel && g_pos, "NULL pointer");
  hipLaunchKernelGGL(k_fem_update_adj<0>, dim3(nblocks(n_tets)), dim3(kBlock), 0,
                     (hipStream_t)stream, fem_dev(prm), n_tets, f2v, B, pos, g_vel, g_pos,
                     (float4*)nullptr);
  return check_launch("dt_fem_update_adj");
}

int dt_fem_update_adj_csr(const dt_fem_params* prm, int32_t n_tets, int32_t n_verts,
                          const int32_t* f2v, const float* B, const float* pos,
                          const float* g_vel, float* g_pos, const int32_t* vert_start,
                          const int32_t* vert_tets, float* tet_buf, void* stream) {
  int rc = check_fem(prm);
  if (rc) return rc;
  DT_REQUIRE(n_tets >= 0 && n_verts >= 0, "n_tets and n_verts must be >= 0");
  if (n_tets == 0) return DT_OK;
  DT_REQUIRE(f2v && B && pos && g_vel && g_pos && vert_start && vert_tets && tet_buf,
             "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fem_update_adj<1>, dim3(nblocks(n_tets)), dim3(kBlock), 0, st,
                     fem_dev(prm), n_tets, f2v, B, pos, g_vel, (float*)nullptr,
                     (float4*)tet_buf);
  hipLaunchKernelGGL(k_fem_vertex_sum, dim3(nblocks((long long)n_verts * kVertexLanes)),
                     dim3(kBlock), 0, st, n_verts, vert_start, vert_tets, (const float4*)tet_buf,
                     g_pos);
  return check_launch("dt_fem_update_adj_csr");
}

int dt_fem_update2(const dt_fem_params* prm, int32_t n_verts, const int32_t* layer_id,
                   const float* pos, const float* vel, const float* vpos, const float* ext_f,
                   const float* cv, const float* cv_next, float* pos_next, float* vel_next,
                   float* vpos_next, void* stream) {
  int rc = check_fem(prm);
  if (rc) return rc;
  DT_REQUIRE(n_verts >= 0, "n_verts must be >= 0");
  if (n_verts == 0) return DT_OK;
  DT_REQUIRE(layer_id && pos && vel && vpos && ext_f && cv && pos_next && vel_next && vpos_next,
             "NULL pointer");
  hipLaunchKernelGGL(k_fem_update2, dim3(nblocks(n_verts)), dim3(kBlock), 0, (hipStream_t)stream,
                     fem_dev(prm), n_verts, layer_id, pos, vel, vpos, ext_f, cv, cv_next, pos_next,
                     vel_next, vpos_next);
  return check_launch("dt_fem_update2");
}

int dt_fem_update2_adj(const dt_fem_params* prm, int32_t n_verts, const int32_t* layer_id,
                       const float* g_pos_next, const float* g_vel_next, const float* g_vpos_next,
                       float* g_pos, float* g_vel, float* g_vpos, float* g_ext_f, float* g_cv,
                       float* g_cv_next, void* stream) {
  int rc = check_fem(prm);
  if (rc) return rc;
  DT_REQUIRE(n_verts >= 0, "n_verts must be >= 0");
  if (n_verts == 0) return DT_OK;
  DT_REQUIRE(layer_id && g_pos_next && g_vel_next && g_vpos_next && g_pos && g_vel && g_vpos &&
                 g_ext_f,
             "NULL pointer");
  hipLaunchKernelGGL(k_fem_update2_adj, dim3(nblocks(n_verts)), dim3(kBlock), 0,
                     (hipStream_t)stream, fem_dev(prm), n_verts, layer_id, g_pos_next, g_vel_next,
                     g_vpos_next, g_pos, g_vel, g_vpos, g_ext_f, g_cv, g_cv_next);
  return check_launch("dt_fem_update2_adj");
}

int dt_clamp(int64_t n, float* a, float limit, void* stream) {
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(a, "NULL pointer");
  hipLaunchKernelGGL(k_clamp, dim3(nblocks(n)), dim3(kBlock), 0, (hipStream_t)stream,
                     (long long)n, a, limit);
  return check_launch("dt_clamp");
}

}  // extern "C"
