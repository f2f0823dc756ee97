// Parallel-jaw kinematic chain and the pads' control velocities (gfx950).
//
// Reference (difftactile/sensor_model/gripper_kinematics.py):
//   eul2mat                     289-303  Rx Ry Rz of degrees + translation
//   pos_transformation          186-192  end-effector step
//   grip_kinematic              169-183  finger opening step
//   fem_sensor{1,2}_set_control_vel 194-216  per-vertex control velocity
//   kinematic / kinematic_grad  219-229
//
// A trajectory's frames come from one sequential chain (one lane: a 4x4 step
// per substep, microseconds for a whole rollout), the control velocities of
// every (substep, vertex) pair from one launch, and the adjoints mirror both:
// per-frame pose gradients are block reductions over the vertices (fixed
// order, no atomics), then one reverse sweep of the chain.
//
// Frames: six row-major 4x4 matrices per frame (DT_JAW_FRAME_FLOATS floats):
// ee_local, ee, finger1_local, finger2_local, finger1_ee, finger2_ee.
// Controls: seven floats per step, (d_pos xyz, d_ori xyz in degrees per unit
// time, d_gripper).
#include "dt_common.h"

namespace dt {

enum { kEeLocal = 0, kEe = 1, kL1 = 2, kL2 = 3, kF1 = 4, kF2 = 5 };

struct M4 {
  float a[16];
};

__device__ __forceinline__ M4 m4_ld(const float* p) {
  M4 m;
  for (int i = 0; i < 16; ++i) m.a[i] = p[i];
  return m;
}
__device__ __forceinline__ void m4_st(float* p, const M4& m) {
  for (int i = 0; i < 16; ++i) p[i] = m.a[i];
}
__device__ __forceinline__ void m4_add_to(float* p, const M4& m) {
  for (int i = 0; i < 16; ++i) p[i] += m.a[i];
}
__device__ __forceinline__ M4 m4_mul(const M4& A, const M4& B) {
  M4 C;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      float s = 0.f;
      for (int k = 0; k < 4; ++k) s += A.a[r * 4 + k] * B.a[k * 4 + c];
      C.a[r * 4 + c] = s;
    }
  return C;
}
// A^T B
__device__ __forceinline__ M4 m4_mul_at(const M4& A, const M4& B) {
  M4 C;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      float s = 0.f;
      for (int k = 0; k < 4; ++k) s += A.a[k * 4 + r] * B.a[k * 4 + c];
      C.a[r * 4 + c] = s;
    }
  return C;
}
// A B^T
__device__ __forceinline__ M4 m4_mul_bt(const M4& A, const M4& B) {
  M4 C;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      float s = 0.f;
      for (int k = 0; k < 4; ++k) s += A.a[r * 4 + k] * B.a[c * 4 + k];
      C.a[r * 4 + c] = s;
    }
  return C;
}

// the three axis rotations of x-y-z angles in radians, row-major 3x3, and
// their derivatives
__device__ __forceinline__ void axis_rots(const float (&r)[3], float (&m)[3][9],
                                          float (&dm)[3][9]) {
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 9; ++i) m[a][i] = dm[a][i] = 0.f;
  const float cx = cosf(r[0]), sx = sinf(r[0]), cy = cosf(r[1]), sy = sinf(r[1]);
  const float cz = cosf(r[2]), sz = sinf(r[2]);
  // Rx = [1 0 0; 0 c -s; 0 s c]
  m[0][0] = 1.f, m[0][4] = cx, m[0][5] = -sx, m[0][7] = sx, m[0][8] = cx;
  dm[0][4] = -sx, dm[0][5] = -cx, dm[0][7] = cx, dm[0][8] = -sx;
  // Ry = [c 0 s; 0 1 0; -s 0 c]
  m[1][0] = cy, m[1][2] = sy, m[1][4] = 1.f, m[1][6] = -sy, m[1][8] = cy;
  dm[1][0] = -sy, dm[1][2] = cy, dm[1][6] = -cy, dm[1][8] = -sy;
  // Rz = [c -s 0; s c 0; 0 0 1]
  m[2][0] = cz, m[2][1] = -sz, m[2][3] = sz, m[2][4] = cz, m[2][8] = 1.f;
  dm[2][0] = -sz, dm[2][1] = -cz, dm[2][3] = cz, dm[2][4] = -sz;
}

__device__ __forceinline__ void m3_mul(const float* A, const float* B, float* C) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      float s = 0.f;
      for (int k = 0; k < 3; ++k) s += A[r * 3 + k] * B[k * 3 + c];
      C[r * 3 + c] = s;
    }
}

// eul2mat(angles in degrees, translation): [(Rx Ry) Rz, t; 0 1]
__device__ __forceinline__ M4 eul2mat(const float (&deg)[3], const float (&t)[3]) {
  const float k = 3.14159265358979323846f / 180.f;
  const float r[3] = {deg[0] * k, deg[1] * k, deg[2] * k};
  float m[3][9], dm[3][9], xy[9], R[9];
  axis_rots(r, m, dm);
  m3_mul(m[0], m[1], xy);
  m3_mul(xy, m[2], R);
  M4 M;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) M.a[i * 4 + j] = R[i * 3 + j];
    M.a[i * 4 + 3] = t[i];
  }
  M.a[12] = M.a[13] = M.a[14] = 0.f;
  M.a[15] = 1.f;
  return M;
}

// d/d(angles in degrees) of <gR, R(angles)>
__device__ __forceinline__ void eul2mat_adj(const float (&deg)[3], const float (&gR)[9],
                                            float (&g_deg)[3]) {
  const float k = 3.14159265358979323846f / 180.f;
  const float r[3] = {deg[0] * k, deg[1] * k, deg[2] * k};
  float m[3][9], dm[3][9];
  axis_rots(r, m, dm);
  for (int a = 0; a < 3; ++a) {
    const float* A = a == 0 ? dm[0] : m[0];
    const float* B = a == 1 ? dm[1] : m[1];
    const float* C = a == 2 ? dm[2] : m[2];
    float ab[9], dR[9];
    m3_mul(A, B, ab);
    m3_mul(ab, C, dR);
    float s = 0.f;
    for (int i = 0; i < 9; ++i) s += gR[i] * dR[i];
    g_deg[a] = s * k;
  }
}

struct JawP {
  float W[16], T1[16], T2[16], dt;
};

__device__ __forceinline__ M4 ld_const(const float (&p)[16]) {
  M4 m;
  for (int i = 0; i < 16; ++i) m.a[i] = p[i];
  return m;
}

// one step of the chain: frame f+1 from frame f and controls c
__device__ __forceinline__ void jaw_step(const JawP& P, const float* c, const float* fr0,
                                         float* fr1) {
  const float ang[3] = {c[3] * P.dt, c[4] * P.dt, c[5] * P.dt};
  const float tr[3] = {c[0] * P.dt, c[1] * P.dt, c[2] * P.dt};
  const M4 D = eul2mat(ang, tr);                       // pos_transformation
  const M4 el1 = m4_mul(D, m4_ld(fr0 + 16 * kEeLocal));
  const M4 ee1 = m4_mul(ld_const(P.W), el1);
  const float gd = 2.f * c[6] * P.dt;                  // grip_delta
  const float z3[3] = {0.f, 0.f, 0.f}, gt[3] = {0.f, gd / 2.f, 0.f};
  const M4 G = eul2mat(z3, gt);                        // grip_kinematic
  const M4 l1 = m4_mul(G, m4_ld(fr0 + 16 * kL1));
  const M4 l2 = m4_mul(G, m4_ld(fr0 + 16 * kL2));
  m4_st(fr1 + 16 * kEeLocal, el1);
  m4_st(fr1 + 16 * kEe, ee1);
  m4_st(fr1 + 16 * kL1, l1);
  m4_st(fr1 + 16 * kL2, l2);
  m4_st(fr1 + 16 * kF1, m4_mul(m4_mul(ee1, ld_const(P.T1)), l1));
  m4_st(fr1 + 16 * kF2, m4_mul(m4_mul(ee1, ld_const(P.T2)), l2));
}

__global__ __launch_bounds__(64) void k_jaw_chain(JawP P, int n_steps,
                                                  const float* __restrict__ controls,
                                                  float* __restrict__ frames) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int f = 0; f < n_steps; ++f)
    jaw_step(P, controls + 7ll * f, frames + (long long)DT_JAW_FRAME_FLOATS * f,
             frames + (long long)DT_JAW_FRAME_FLOATS * (f + 1));
}

// reverse sweep: g_frames[f+1] -> g_frames[f] and g_controls[f], f = n-1..0
__global__ __launch_bounds__(64) void k_jaw_chain_adj(JawP P, int n_steps,
                                                      const float* __restrict__ controls,
                                                      const float* __restrict__ frames,
                                                      float* __restrict__ g_frames,
                                                      float* __restrict__ g_controls) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const M4 W = ld_const(P.W), T1 = ld_const(P.T1), T2 = ld_const(P.T2);
  for (int f = n_steps - 1; f >= 0; --f) {
    const float* c = controls + 7ll * f;
    const float* fr0 = frames + (long long)DT_JAW_FRAME_FLOATS * f;
    const float* fr1 = frames + (long long)DT_JAW_FRAME_FLOATS * (f + 1);
    float* g0 = g_frames + (long long)DT_JAW_FRAME_FLOATS * f;
    float* g1 = g_frames + (long long)DT_JAW_FRAME_FLOATS * (f + 1);
    float* gc = g_controls + 7ll * f;
    const M4 ee1 = m4_ld(fr1 + 16 * kEe), l1 = m4_ld(fr1 + 16 * kL1), l2 = m4_ld(fr1 + 16 * kL2);
    // finger_k_ee = (ee T_k) L_k
    M4 gee = m4_ld(g1 + 16 * kEe);
    M4 gl1 = m4_ld(g1 + 16 * kL1), gl2 = m4_ld(g1 + 16 * kL2);
    {
      const M4 gF = m4_ld(g1 + 16 * kF1);
      const M4 P1 = m4_mul(ee1, T1);
      const M4 gP = m4_mul_bt(gF, l1);
      const M4 a = m4_mul_at(P1, gF), b = m4_mul_bt(gP, T1);
      for (int i = 0; i < 16; ++i) gl1.a[i] += a.a[i], gee.a[i] += b.a[i];
    }
    {
      const M4 gF = m4_ld(g1 + 16 * kF2);
      const M4 P2 = m4_mul(ee1, T2);
      const M4 gP = m4_mul_bt(gF, l2);
      const M4 a = m4_mul_at(P2, gF), b = m4_mul_bt(gP, T2);
      for (int i = 0; i < 16; ++i) gl2.a[i] += a.a[i], gee.a[i] += b.a[i];
    }
    // L_k(f+1) = G L_k(f), G = translation (0, gd/2, 0), gd = 2 d_gripper dt
    const float gd = 2.f * c[6] * P.dt;
    const float z3[3] = {0.f, 0.f, 0.f}, gt[3] = {0.f, gd / 2.f, 0.f};
    const M4 G = eul2mat(z3, gt);
    const M4 gG1 = m4_mul_bt(gl1, m4_ld(fr0 + 16 * kL1));
    const M4 gG2 = m4_mul_bt(gl2, m4_ld(fr0 + 16 * kL2));
    gc[6] += (gG1.a[1 * 4 + 3] + gG2.a[1 * 4 + 3]) * P.dt;
    m4_add_to(g0 + 16 * kL1, m4_mul_at(G, gl1));
    m4_add_to(g0 + 16 * kL2, m4_mul_at(G, gl2));
    // ee(f+1) = W ee_local(f+1);  ee_local(f+1) = D ee_local(f)
    M4 gel = m4_ld(g1 + 16 * kEeLocal);
    {
      const M4 a = m4_mul_at(W, gee);
      for (int i = 0; i < 16; ++i) gel.a[i] += a.a[i];
    }
    const float ang[3] = {c[3] * P.dt, c[4] * P.dt, c[5] * P.dt};
    const float tr[3] = {c[0] * P.dt, c[1] * P.dt, c[2] * P.dt};
    const M4 D = eul2mat(ang, tr);
    const M4 gD = m4_mul_bt(gel, m4_ld(fr0 + 16 * kEeLocal));
    m4_add_to(g0 + 16 * kEeLocal, m4_mul_at(D, gel));
    for (int i = 0; i < 3; ++i) gc[i] += gD.a[i * 4 + 3] * P.dt;
    float gR[9], g_deg[3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) gR[i * 3 + j] = gD.a[i * 4 + j];
    eul2mat_adj(ang, gR, g_deg);
    for (int i = 0; i < 3; ++i) gc[3 + i] += g_deg[i] * P.dt;
  }
}

// control_vel[f][i] = ((E[f+1] x) - (E[f] x))_{0:3} / dt, x = (init_x_i, 1),
// E = the finger pose, f < n_steps
__global__ __launch_bounds__(kBlock) void k_jaw_control_vel(int n_steps, int nv, int pose,
                                                            const float* __restrict__ frames,
                                                            const float* __restrict__ x0,
                                                            float dt, float* __restrict__ cv) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)n_steps * nv) return;
  const int f = (int)(g / nv), i = (int)(g % nv);
  const float* E0 = frames + (long long)DT_JAW_FRAME_FLOATS * f + 16 * pose;
  const float* E1 = E0 + DT_JAW_FRAME_FLOATS;
  const float x[4] = {x0[3ll * i], x0[3ll * i + 1], x0[3ll * i + 2], 1.f};
  for (int r = 0; r < 3; ++r) {
    float a = 0.f, b = 0.f;
    for (int k = 0; k < 4; ++k) {
      a += E1[r * 4 + k] * x[k];
      b += E0[r * 4 + k] * x[k];
    }
    cv[3 * g + r] = (a - b) / dt;
  }
}

// g_frames[f].pose rows 0..2 += sum_i (g_cv[f-1][i] - g_cv[f][i]) x_i^T / dt,
// one block per frame f = 0..n_steps (a term outside 0..n_steps-1 is zero)
__global__ __launch_bounds__(kBlock) void k_jaw_control_vel_adj(int n_steps, int nv, int pose,
                                                                const float* __restrict__ x0,
                                                                float dt,
                                                                const float* __restrict__ gcv,
                                                                float* __restrict__ g_frames) {
  __shared__ float part[kBlock / 64][12];
  const int f = blockIdx.x;
  float acc[12];
  for (int q = 0; q < 12; ++q) acc[q] = 0.f;
  for (int i = threadIdx.x; i < nv; i += blockDim.x) {
    float gv[3] = {0.f, 0.f, 0.f};
    if (f >= 1)
      for (int r = 0; r < 3; ++r) gv[r] += gcv[(3ll * (f - 1)) * nv + 3ll * i + r];
    if (f < n_steps)
      for (int r = 0; r < 3; ++r) gv[r] -= gcv[(3ll * f) * nv + 3ll * i + r];
    const float x[4] = {x0[3ll * i], x0[3ll * i + 1], x0[3ll * i + 2], 1.f};
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 4; ++k) acc[r * 4 + k] += gv[r] * x[k];
  }
  for (int q = 0; q < 12; ++q)
    for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 12; ++q) part[w][q] = acc[q];
  __syncthreads();
  if (threadIdx.x < 12) {
    float s = 0.f;
    for (int q = 0; q < kBlock / 64; ++q) s += part[q][threadIdx.x];
    g_frames[(long long)DT_JAW_FRAME_FLOATS * f + 16 * pose + threadIdx.x] += s / dt;
  }
}

static JawP jaw_dev(const dt_jaw_params* p) {
  JawP d;
  for (int i = 0; i < 16; ++i) {
    d.W[i] = p->ee_world[i];
    d.T1[i] = p->finger_t[0][i];
    d.T2[i] = p->finger_t[1][i];
  }
  d.dt = p->dt;
  return d;
}

}  // namespace dt

using namespace dt;

extern "C" {

int dt_jaw_chain(const dt_jaw_params* prm, int32_t n_steps, const float* controls, float* frames,
                 void* stream) {
  DT_REQUIRE(prm, "NULL params");
  DT_REQUIRE(n_steps >= 0, "n_steps must be >= 0");
  if (n_steps == 0) return DT_OK;
  DT_REQUIRE(controls && frames, "NULL pointer");
  hipLaunchKernelGGL(k_jaw_chain, dim3(1), dim3(64), 0, (hipStream_t)stream, jaw_dev(prm),
                     (int)n_steps, controls, frames);
  return check_launch("dt_jaw_chain");
}

int dt_jaw_chain_adj(const dt_jaw_params* prm, int32_t n_steps, const float* controls,
                     const float* frames, float* g_frames, float* g_controls, void* stream) {
  DT_REQUIRE(prm, "NULL params");
  DT_REQUIRE(n_steps >= 0, "n_steps must be >= 0");
  if (n_steps == 0) return DT_OK;
  DT_REQUIRE(controls && frames && g_frames && g_controls, "NULL pointer");
  hipLaunchKernelGGL(k_jaw_chain_adj, dim3(1), dim3(64), 0, (hipStream_t)stream, jaw_dev(prm),
                     (int)n_steps, controls, frames, g_frames, g_controls);
  return check_launch("dt_jaw_chain_adj");
}

int dt_jaw_control_vel(int32_t n_steps, int32_t n_verts, int32_t finger, const float* frames,
                       const float* init_x, float dt, float* control_vel, void* stream) {
  DT_REQUIRE(n_steps >= 0 && n_verts >= 0, "sizes must be >= 0");
  DT_REQUIRE(finger == 1 || finger == 2, "finger must be 1 or 2");
  DT_REQUIRE(dt != 0.f, "dt must be non-zero");
  if (n_steps == 0 || n_verts == 0) return DT_OK;
  DT_REQUIRE(frames && init_x && control_vel, "NULL pointer");
  const long long n = (long long)n_steps * n_verts;
  hipLaunchKernelGGL(k_jaw_control_vel, dim3(nblocks(n)), dim3(kBlock), 0, (hipStream_t)stream,
                     (int)n_steps, (int)n_verts, finger == 1 ? (int)kF1 : (int)kF2, frames,
                     init_x, dt, control_vel);
  return check_launch("dt_jaw_control_vel");
}

int dt_jaw_control_vel_adj(int32_t n_steps, int32_t n_verts, int32_t finger, const float* init_x,
                           float dt, const float* g_control_vel, float* g_frames, void* stream) {
  DT_REQUIRE(n_steps >= 0 && n_verts >= 0, "sizes must be >= 0");
  DT_REQUIRE(finger == 1 || finger == 2, "finger must be 1 or 2");
  DT_REQUIRE(dt != 0.f, "dt must be non-zero");
  if (n_steps == 0 || n_verts == 0) return DT_OK;
  DT_REQUIRE(init_x && g_control_vel && g_frames, "NULL pointer");
  hipLaunchKernelGGL(k_jaw_control_vel_adj, dim3(n_steps + 1), dim3(kBlock), 0,
                     (hipStream_t)stream, (int)n_steps, (int)n_verts,
                     finger == 1 ? (int)kF1 : (int)kF2, init_x, dt, g_control_vel, g_frames);
  return check_launch("dt_jaw_control_vel_adj");
}

}  // extern "C"
