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
//
// Also the dome pad's pose control (sensor_model/fem_sensor.py):
//   eul2mat                     196-210  Rz Ry Rx of degrees + translation
//   set_trans_h                 217-227  world-frame pose step
//   set_pose_control            229-245  local-frame pose step, dtrans_h
//   set_pose_control_bp         248-256  dtrans_h and the inverses only
//   set_control_vel             258-266  per-vertex control velocity
// One lane per pose step; the control velocity one lane per vertex, its
// adjoint a fixed-order block reduction.
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

// ---------------------------------------------------------------------------
// dome pad pose control (sensor_model/fem_sensor.py:196-266)
// ---------------------------------------------------------------------------
// Pose state: DT_DOME_POSE_FLOATS floats (offsets below), row-major matrices.
enum {
  kPW = DT_DOME_TRANS_WORLD, kPL = DT_DOME_TRANS_LOCAL, kPH = DT_DOME_TRANS_H,
  kPHi = DT_DOME_INV_TRANS_H, kPD = DT_DOME_DTRANS_H, kPRW = DT_DOME_ROT_WORLD,
  kPRL = DT_DOME_ROT_LOCAL, kPRH = DT_DOME_ROT_H, kPRi = DT_DOME_INV_ROT
};

struct M3 {
  float a[9];
};
__device__ __forceinline__ M3 m3_ld(const float* p) {
  M3 m;
  for (int i = 0; i < 9; ++i) m.a[i] = p[i];
  return m;
}
__device__ __forceinline__ void m3_st(float* p, const M3& m) {
  for (int i = 0; i < 9; ++i) p[i] = m.a[i];
}
__device__ __forceinline__ M3 m3_prod(const M3& A, const M3& B) {
  M3 C;
  m3_mul(A.a, B.a, C.a);
  return C;
}
// A^T B, A B^T
__device__ __forceinline__ M3 m3_prod_at(const M3& A, const M3& B) {
  M3 C;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      float s = 0.f;
      for (int k = 0; k < 3; ++k) s += A.a[k * 3 + r] * B.a[k * 3 + c];
      C.a[r * 3 + c] = s;
    }
  return C;
}
__device__ __forceinline__ M3 m3_prod_bt(const M3& A, const M3& B) {
  M3 C;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      float s = 0.f;
      for (int k = 0; k < 3; ++k) s += A.a[r * 3 + k] * B.a[c * 3 + k];
      C.a[r * 3 + c] = s;
    }
  return C;
}

// general inverses through the adjugate (the reference's .inverse() is general)
__device__ __forceinline__ M3 m3_inv(const M3& m) {
  const float* a = m.a;
  M3 c;
  c.a[0] = a[4] * a[8] - a[5] * a[7];
  c.a[1] = a[2] * a[7] - a[1] * a[8];
  c.a[2] = a[1] * a[5] - a[2] * a[4];
  c.a[3] = a[5] * a[6] - a[3] * a[8];
  c.a[4] = a[0] * a[8] - a[2] * a[6];
  c.a[5] = a[2] * a[3] - a[0] * a[5];
  c.a[6] = a[3] * a[7] - a[4] * a[6];
  c.a[7] = a[1] * a[6] - a[0] * a[7];
  c.a[8] = a[0] * a[4] - a[1] * a[3];
  const float det = a[0] * c.a[0] + a[1] * c.a[3] + a[2] * c.a[6];
  const float inv = 1.f / det;
  for (int i = 0; i < 9; ++i) c.a[i] *= inv;
  return c;
}

__device__ __forceinline__ M4 m4_inv(const M4& M) {
  const float* m = M.a;
  float inv[16];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] +
           m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] -
           m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] +
           m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] -
            m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] -
           m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] +
           m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] -
           m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] +
            m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] +
           m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] -
           m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] +
            m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] -
            m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] -
           m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] +
           m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] -
            m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] +
            m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  const float det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  const float r = 1.f / det;
  M4 o;
  for (int i = 0; i < 16; ++i) o.a[i] = inv[i] * r;
  return o;
}

// Y = X^-1: dL/dX = -Y^T dL/dY Y^T
__device__ __forceinline__ M4 m4_inv_adj(const M4& Y, const M4& gY) {
  M4 t = m4_mul_at(Y, gY), o = m4_mul_bt(t, Y);
  for (int i = 0; i < 16; ++i) o.a[i] = -o.a[i];
  return o;
}
__device__ __forceinline__ M3 m3_inv_adj(const M3& Y, const M3& gY) {
  M3 t = m3_prod_at(Y, gY), o = m3_prod_bt(t, Y);
  for (int i = 0; i < 9; ++i) o.a[i] = -o.a[i];
  return o;
}

// the dome's eul2mat (fem_sensor.py:196-210): R = Rz Ry Rx of degrees; the
// transform [R, t; 0 1]
__device__ __forceinline__ void eul2mat_zyx(const float (&deg)[3], const float (&t)[3], M4& T,
                                            M3& R) {
  const float k = 3.14159265358979323846f / 180.f;
  const float r[3] = {deg[0] * k, deg[1] * k, deg[2] * k};
  float m[3][9], dm[3][9], zy[9];
  axis_rots(r, m, dm);
  m3_mul(m[2], m[1], zy);
  m3_mul(zy, m[0], R.a);
  for (int i = 0; i < 16; ++i) T.a[i] = 0.f;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) T.a[i * 4 + j] = R.a[i * 3 + j];
    T.a[i * 4 + 3] = t[i];
  }
  T.a[15] = 1.f;
}

__device__ __forceinline__ void eul2mat_zyx_adj(const float (&deg)[3], const M3& gR,
                                                float (&g_deg)[3]) {
  const float k = 3.14159265358979323846f / 180.f;
  const float r[3] = {deg[0] * k, deg[1] * k, deg[2] * k};
  float m[3][9], dm[3][9];
  axis_rots(r, m, dm);
  for (int a = 0; a < 3; ++a) {
    const float* Z = a == 2 ? dm[2] : m[2];
    const float* Y = a == 1 ? dm[1] : m[1];
    const float* X = a == 0 ? dm[0] : m[0];
    float zy[9], dR[9];
    m3_mul(Z, Y, zy);
    m3_mul(zy, X, dR);
    float s = 0.f;
    for (int i = 0; i < 9; ++i) s += gR.a[i] * dR[i];
    g_deg[a] = s * k;
  }
}

__device__ __forceinline__ void dome_step_mats(const float* d6, float st, M4& T, M3& R,
                                               float (&ang)[3]) {
  const float tr[3] = {d6[0] * st, d6[1] * st, d6[2] * st};
  ang[0] = d6[3] * st, ang[1] = d6[4] * st, ang[2] = d6[5] * st;
  eul2mat_zyx(ang, tr, T, R);
}

// mode 0 set_pose_control, 1 set_trans_h, 2 set_pose_control_bp; in place is allowed
__global__ __launch_bounds__(64) void k_dome_pose(int mode, const float* __restrict__ d6, float st,
                                                  const float* pin, float* pout) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float P[DT_DOME_POSE_FLOATS];
  for (int i = 0; i < DT_DOME_POSE_FLOATS; ++i) P[i] = pin[i];
  M4 T;
  M3 R;
  float ang[3];
  dome_step_mats(d6, st, T, R, ang);
  const M4 W = m4_ld(P + kPW);
  if (mode != 1) {
    const M4 D = m4_mul(m4_mul(W, T), m4_inv(W));
    m4_st(P + kPD, D);
  }
  if (mode == 0) {
    const M4 L = m4_mul(T, m4_ld(P + kPL));
    m4_st(P + kPL, L);
    m4_st(P + kPH, m4_mul(W, L));
    const M3 RL = m3_prod(R, m3_ld(P + kPRL));
    m3_st(P + kPRL, RL);
    m3_st(P + kPRH, m3_prod(m3_ld(P + kPRW), RL));
  } else if (mode == 1) {
    m4_st(P + kPH, m4_mul(T, m4_ld(P + kPH)));
    m3_st(P + kPRH, m3_prod(R, m3_ld(P + kPRH)));
  }
  m4_st(P + kPHi, m4_inv(m4_ld(P + kPH)));
  m3_st(P + kPRi, m3_inv(m3_ld(P + kPRH)));
  for (int i = 0; i < DT_DOME_POSE_FLOATS; ++i) pout[i] = P[i];
}

// reverse of k_dome_pose: g_out holds d/d(every output field); adds d/d(the
// input fields read) into g_in and d/d(d_pos, d_ori) into g_d6.  g_in may be
// g_out (the fields an update overwrote are cleared there first).
__global__ __launch_bounds__(64) void k_dome_pose_adj(int mode, const float* __restrict__ d6,
                                                      float st, const float* __restrict__ pin,
                                                      const float* g_out, float* g_in,
                                                      float* __restrict__ g_d6) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float G[DT_DOME_POSE_FLOATS], P[DT_DOME_POSE_FLOATS];
  for (int i = 0; i < DT_DOME_POSE_FLOATS; ++i) G[i] = g_out[i], P[i] = pin[i];
  M4 T;
  M3 R;
  float ang[3];
  dome_step_mats(d6, st, T, R, ang);
  const M4 W = m4_ld(P + kPW), Wi = m4_inv(W);
  M4 gT;
  M3 gR;
  for (int i = 0; i < 16; ++i) gT.a[i] = 0.f;
  for (int i = 0; i < 9; ++i) gR.a[i] = 0.f;
  // output pose (recomputed) for the inverses
  M4 H;
  M3 RH;
  if (mode == 0) {
    H = m4_mul(W, m4_mul(T, m4_ld(P + kPL)));
    RH = m3_prod(m3_ld(P + kPRW), m3_prod(R, m3_ld(P + kPRL)));
  } else if (mode == 1) {
    H = m4_mul(T, m4_ld(P + kPH));
    RH = m3_prod(R, m3_ld(P + kPRH));
  } else {
    H = m4_ld(P + kPH);
    RH = m3_ld(P + kPRH);
  }
  // inverses
  M4 gH = m4_ld(G + kPH);
  {
    const M4 a = m4_inv_adj(m4_inv(H), m4_ld(G + kPHi));
    for (int i = 0; i < 16; ++i) gH.a[i] += a.a[i];
  }
  M3 gRH = m3_ld(G + kPRH);
  {
    const M3 a = m3_inv_adj(m3_inv(RH), m3_ld(G + kPRi));
    for (int i = 0; i < 9; ++i) gRH.a[i] += a.a[i];
  }
  // the fields this mode reads, before their gradients are added
  M4 gL_in, gH_in;
  M3 gRL_in, gRH_in;
  for (int i = 0; i < 16; ++i) gL_in.a[i] = gH_in.a[i] = 0.f;
  for (int i = 0; i < 9; ++i) gRL_in.a[i] = gRH_in.a[i] = 0.f;
  if (mode == 0) {
    // trans_h = W L', L' = T L;  rot_h = RW RL', RL' = R RL
    M4 gL1 = m4_ld(G + kPL);
    {
      const M4 a = m4_mul_at(W, gH);
      for (int i = 0; i < 16; ++i) gL1.a[i] += a.a[i];
    }
    const M4 L = m4_ld(P + kPL);
    gT = m4_mul_bt(gL1, L);
    gL_in = m4_mul_at(T, gL1);
    M3 gRL1 = m3_ld(G + kPRL);
    {
      const M3 a = m3_prod_at(m3_ld(P + kPRW), gRH);
      for (int i = 0; i < 9; ++i) gRL1.a[i] += a.a[i];
    }
    const M3 RL = m3_ld(P + kPRL);
    gR = m3_prod_bt(gRL1, RL);
    gRL_in = m3_prod_at(R, gRL1);
  } else if (mode == 1) {
    // trans_h' = T trans_h; rot_h' = R rot_h
    gT = m4_mul_bt(gH, m4_ld(P + kPH));
    gH_in = m4_mul_at(T, gH);
    gR = m3_prod_bt(gRH, m3_ld(P + kPRH));
    gRH_in = m3_prod_at(R, gRH);
  } else {
    gH_in = gH;
    gRH_in = gRH;
  }
  if (mode != 1) {
    // dtrans_h = W T W^-1
    const M4 a = m4_mul_bt(m4_mul_at(W, m4_ld(G + kPD)), Wi);
    for (int i = 0; i < 16; ++i) gT.a[i] += a.a[i];
  }
  // T = [R, t; 0 1] with R the rotation of the same step
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) gR.a[i * 3 + j] += gT.a[i * 4 + j];
  float g_deg[3];
  eul2mat_zyx_adj(ang, gR, g_deg);
  for (int i = 0; i < 3; ++i) {
    g_d6[i] += gT.a[i * 4 + 3] * st;
    g_d6[3 + i] += g_deg[i] * st;
  }
  // write: the overwritten fields' gradients are consumed, the read ones added
  float O[DT_DOME_POSE_FLOATS];
  for (int i = 0; i < DT_DOME_POSE_FLOATS; ++i) O[i] = G[i];
  for (int i = 0; i < 16; ++i) O[kPHi + i] = 0.f;
  for (int i = 0; i < 9; ++i) O[kPRi + i] = 0.f;
  if (mode != 1)
    for (int i = 0; i < 16; ++i) O[kPD + i] = 0.f;
  if (mode == 0) {
    for (int i = 0; i < 16; ++i) O[kPL + i] = gL_in.a[i], O[kPH + i] = 0.f;
    for (int i = 0; i < 9; ++i) O[kPRL + i] = gRL_in.a[i], O[kPRH + i] = 0.f;
  } else {
    for (int i = 0; i < 16; ++i) O[kPH + i] = gH_in.a[i];
    for (int i = 0; i < 9; ++i) O[kPRH + i] = gRH_in.a[i];
  }
  for (int i = 0; i < DT_DOME_POSE_FLOATS; ++i) g_in[i] = O[i];
}

// control_vel[i] = ((dtrans_h [v_i; 1]) - v_i)_{0:3} / st (set_control_vel, 259-266)
__global__ __launch_bounds__(kBlock) void k_dome_control_vel(int nv, const float* __restrict__ vpos,
                                                             const float* __restrict__ pose,
                                                             float st, float* __restrict__ cv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const float* D = pose + kPD;
  const float x[4] = {vpos[3ll * i], vpos[3ll * i + 1], vpos[3ll * i + 2], 1.f};
  for (int r = 0; r < 3; ++r) {
    float a = 0.f;
    for (int k = 0; k < 4; ++k) a += D[r * 4 + k] * x[k];
    cv[3ll * i + r] = (a - x[r]) / st;
  }
}

// per block: sum_i g_cv_i [v_i; 1]^T / st (rows 0..2 of d/d dtrans_h), and
// g_vpos_i += (D_{0:3,0:3}^T g_cv_i - g_cv_i) / st
__global__ __launch_bounds__(kBlock) void k_dome_control_vel_adj(
    int nv, const float* __restrict__ vpos, const float* __restrict__ pose,
    const float* __restrict__ gcv, float st, float* __restrict__ partials,
    float* __restrict__ g_vpos) {
  __shared__ float part[kBlock / 64][12];
  const float* D = pose + kPD;
  float acc[12];
  for (int q = 0; q < 12; ++q) acc[q] = 0.f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
    const float g[3] = {gcv[3ll * i] / st, gcv[3ll * i + 1] / st, gcv[3ll * i + 2] / st};
    const float x[4] = {vpos[3ll * i], vpos[3ll * i + 1], vpos[3ll * i + 2], 1.f};
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 4; ++k) acc[r * 4 + k] += g[r] * x[k];
    if (g_vpos)
      for (int k = 0; k < 3; ++k)
        g_vpos[3ll * i + k] += D[k] * g[0] + D[4 + k] * g[1] + D[8 + k] * g[2] - g[k];
  }
  for (int q = 0; q < 12; ++q)
    for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o);
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 12; ++q) part[threadIdx.x >> 6][q] = acc[q];
  __syncthreads();
  if (threadIdx.x < 12) {
    float t = part[0][threadIdx.x];
    for (int w = 1; w < kBlock / 64; ++w) t += part[w][threadIdx.x];
    partials[12ll * blockIdx.x + threadIdx.x] = t;
  }
}

// g_pose[dtrans_h rows 0..2] += sum of the nb block partials (one block, fixed order)
__global__ __launch_bounds__(64) void k_dome_partials_add(int nb, const float* __restrict__ partials,
                                                          float* __restrict__ g_pose) {
  if (threadIdx.x >= 12) return;
  float t = 0.f;
  for (int b = 0; b < nb; ++b) t += partials[12ll * b + threadIdx.x];
  g_pose[kPD + threadIdx.x] += t;
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

int dt_dome_pose_control(int32_t mode, const float* d6, float step_time, const float* pose_in,
                         float* pose_out, void* stream) {
  DT_REQUIRE(mode >= 0 && mode <= 2, "mode must be 0, 1 or 2");
  DT_REQUIRE(d6 && pose_in && pose_out, "NULL pointer");
  hipLaunchKernelGGL(k_dome_pose, dim3(1), dim3(64), 0, (hipStream_t)stream, (int)mode, d6,
                     step_time, pose_in, pose_out);
  return check_launch("dt_dome_pose_control");
}

int dt_dome_pose_control_adj(int32_t mode, const float* d6, float step_time, const float* pose_in,
                             const float* g_pose_out, float* g_pose_in, float* g_d6,
                             void* stream) {
  DT_REQUIRE(mode >= 0 && mode <= 2, "mode must be 0, 1 or 2");
  DT_REQUIRE(d6 && pose_in && g_pose_out && g_pose_in && g_d6, "NULL pointer");
  hipLaunchKernelGGL(k_dome_pose_adj, dim3(1), dim3(64), 0, (hipStream_t)stream, (int)mode, d6,
                     step_time, pose_in, g_pose_out, g_pose_in, g_d6);
  return check_launch("dt_dome_pose_control_adj");
}

int dt_dome_control_vel(int32_t n_verts, const float* vpos, const float* pose, float step_time,
                        float* control_vel, void* stream) {
  DT_REQUIRE(n_verts >= 0, "n_verts must be >= 0");
  DT_REQUIRE(step_time != 0.f, "step_time must be non-zero");
  if (n_verts == 0) return DT_OK;
  DT_REQUIRE(vpos && pose && control_vel, "NULL pointer");
  hipLaunchKernelGGL(k_dome_control_vel, dim3(nblocks(n_verts)), dim3(kBlock), 0,
                     (hipStream_t)stream, (int)n_verts, vpos, pose, step_time, control_vel);
  return check_launch("dt_dome_control_vel");
}

int dt_dome_control_vel_adj(int32_t n_verts, const float* vpos, const float* pose,
                            const float* g_control_vel, float step_time, float* scratch,
                            float* g_pose, float* g_vpos, void* stream) {
  DT_REQUIRE(n_verts >= 0, "n_verts must be >= 0");
  DT_REQUIRE(step_time != 0.f, "step_time must be non-zero");
  if (n_verts == 0) return DT_OK;
  DT_REQUIRE(vpos && pose && g_control_vel && scratch && g_pose, "NULL pointer");
  long long b = ((long long)n_verts + kBlock - 1) / kBlock;
  const int nb = (int)(b < DT_REDUCE_BLOCKS ? b : DT_REDUCE_BLOCKS);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_dome_control_vel_adj, dim3(nb), dim3(kBlock), 0, st, (int)n_verts, vpos,
                     pose, g_control_vel, step_time, scratch, g_vpos);
  hipLaunchKernelGGL(k_dome_partials_add, dim3(1), dim3(64), 0, st, nb, scratch, g_pose);
  return check_launch("dt_dome_control_vel_adj");
}

}  // extern "C"
