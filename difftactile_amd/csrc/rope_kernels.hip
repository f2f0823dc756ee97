// Position-based rope: one substep and its adjoint (gfx950).
//
// Reference: difftactile/object_model/PBD_rope.py, update (111-180):
//   v' = v + ext_f / m + g dt, the table's floor friction, p* = p + v' dt;
//   one Jacobi pass of stretch constraints (i, i+1), one of bending
//   constraints (i, i+2); particle 0 held at its initial position;
//   v_next = (p_next - p) / dt.
//
// A rope is a few thousand particles, so one workgroup runs the whole
// substep: phases separated by barriers, each particle's constraint
// corrections gathered from its (at most two) constraints rather than
// scattered with atomics, so the result is reproducible bit for bit (a sum of
// two terms is the same in either order).  The adjoint recomputes the
// phases into a caller scratch and runs them backwards the same way.
#include "dt_common.h"

namespace dt {

constexpr int kRopeThreads = 1024;

struct RopeP {
  float dt, inv_m, p_mass, table_height, friction, eps;
  float g[3];
  float rest, bend;             // stretch and bending lengths
  float k_stretch, k_bend;      // relaxation / (2 inv_m + compliance / dt^2)
  float fixed[3];
};

__device__ __forceinline__ void ld3r(const float* a, int i, float (&v)[3]) {
  v[0] = a[3 * i], v[1] = a[3 * i + 1], v[2] = a[3 * i + 2];
}
__device__ __forceinline__ void st3r(float* a, int i, const float (&v)[3]) {
  a[3 * i] = v[0], a[3 * i + 1] = v[1], a[3 * i + 2] = v[2];
}

// correction dp of the constraint between particles a and b (applied +w to a,
// -w to b): dp = -k (|n| - L) n / |n|, n = x_a - x_b, |n| = sqrt(n.n + eps)
__device__ __forceinline__ void constraint_dp(const float* x, int a, int b, float L, float k,
                                              float eps, float (&dp)[3]) {
  float xa[3], xb[3], n[3];
  ld3r(x, a, xa);
  ld3r(x, b, xb);
  for (int c = 0; c < 3; ++c) n[c] = xa[c] - xb[c];
  const float nn = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] + eps);
  const float s = -(nn - L) * k / nn;
  for (int c = 0; c < 3; ++c) dp[c] = s * n[c];
}

// J^T g of that correction (J = d dp / d n is symmetric):
// J = -k [u u^T + (|n| - L) / |n| (I - u u^T)], u = n / |n|
__device__ __forceinline__ void constraint_dp_adj(const float* x, int a, int b, float L, float k,
                                                  float eps, const float (&g)[3],
                                                  float (&gn)[3]) {
  float xa[3], xb[3], n[3];
  ld3r(x, a, xa);
  ld3r(x, b, xb);
  for (int c = 0; c < 3; ++c) n[c] = xa[c] - xb[c];
  const float nn = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] + eps);
  const float u[3] = {n[0] / nn, n[1] / nn, n[2] / nn};
  const float ug = u[0] * g[0] + u[1] * g[1] + u[2] * g[2];
  const float r = (nn - L) / nn;
  for (int c = 0; c < 3; ++c) gn[c] = -k * (u[c] * ug + r * (g[c] - u[c] * ug));
}

// particle i's correction from the constraints (i, i+d) and (i-d, i)
__device__ __forceinline__ void gather_dp(const float* x, int n, int i, int d, float L, float k,
                                          float w, float eps, float (&out)[3]) {
  out[0] = out[1] = out[2] = 0.f;
  float dp[3];
  if (i + d < n) {
    constraint_dp(x, i, i + d, L, k, eps, dp);
    for (int c = 0; c < 3; ++c) out[c] += dp[c] * w;
  }
  if (i - d >= 0) {
    constraint_dp(x, i - d, i, L, k, eps, dp);
    for (int c = 0; c < 3; ++c) out[c] -= dp[c] * w;
  }
}

// predicted velocity before / after the floor and whether the floor acted
__device__ __forceinline__ bool rope_predict(const RopeP& P, const float (&p)[3],
                                             const float (&v)[3], const float (&e)[3],
                                             float (&vn)[3]) {
  for (int c = 0; c < 3; ++c) vn[c] = v[c] + e[c] / P.p_mass + P.g[c] * P.dt;
  const bool floor = (p[1] + vn[1] * P.dt) < P.table_height && vn[1] < 0.f;
  if (floor) {
    // v_n = (N . v) N with N = (0, 1, 0); v = (v - v_n) friction
    vn[0] = vn[0] * P.friction;
    vn[1] = (vn[1] - vn[1]) * P.friction;
    vn[2] = vn[2] * P.friction;
  }
  return floor;
}

// forward phases into t0 (predicted), t1 (after stretch), t2 (after bending)
__device__ void rope_phases(const RopeP& P, int n, const float* pos, const float* vel,
                            const float* ext, float* t0, float* t1, float* t2) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    float p[3], v[3], e[3], vn[3];
    ld3r(pos, i, p);
    ld3r(vel, i, v);
    ld3r(ext, i, e);
    rope_predict(P, p, v, e, vn);
    const float t[3] = {p[0] + vn[0] * P.dt, p[1] + vn[1] * P.dt, p[2] + vn[2] * P.dt};
    st3r(t0, i, t);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    float d[3], t[3];
    gather_dp(t0, n, i, 1, P.rest, P.k_stretch, P.inv_m, P.eps, d);
    ld3r(t0, i, t);
    for (int c = 0; c < 3; ++c) t[c] += d[c];
    st3r(t1, i, t);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    float d[3], t[3];
    gather_dp(t1, n, i, 2, P.bend, P.k_bend, P.inv_m, P.eps, d);
    ld3r(t1, i, t);
    for (int c = 0; c < 3; ++c) t[c] += d[c];
    st3r(t2, i, t);
  }
  __syncthreads();
}

__global__ __launch_bounds__(kRopeThreads) void k_rope_step(RopeP P, int n,
                                                            const float* __restrict__ pos,
                                                            const float* __restrict__ vel,
                                                            const float* __restrict__ ext,
                                                            float* __restrict__ work,
                                                            float* __restrict__ pos1,
                                                            float* __restrict__ vel1) {
  float* t0 = work;
  float* t1 = work + 3ll * n;
  float* t2 = work + 6ll * n;
  rope_phases(P, n, pos, vel, ext, t0, t1, t2);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    float p1[3], p[3];
    if (i == 0) {
      p1[0] = P.fixed[0], p1[1] = P.fixed[1], p1[2] = P.fixed[2];
    } else {
      ld3r(t2, i, p1);
    }
    ld3r(pos, i, p);
    st3r(pos1, i, p1);
    const float v1[3] = {(p1[0] - p[0]) / P.dt, (p1[1] - p[1]) / P.dt, (p1[2] - p[2]) / P.dt};
    st3r(vel1, i, v1);
  }
}

// adjoint: g_pos, g_vel, g_ext accumulate; work holds 12 n floats
__global__ __launch_bounds__(kRopeThreads) void k_rope_step_adj(
    RopeP P, int n, const float* __restrict__ pos, const float* __restrict__ vel,
    const float* __restrict__ ext, const float* __restrict__ g_pos1,
    const float* __restrict__ g_vel1, float* __restrict__ work, float* __restrict__ g_pos,
    float* __restrict__ g_vel, float* __restrict__ g_ext) {
  float* t0 = work;
  float* t1 = work + 3ll * n;
  float* t2 = work + 6ll * n;
  float* gt = work + 9ll * n;     // adjoint of the current phase's input
  rope_phases(P, n, pos, vel, ext, t0, t1, t2);
  // pos1 = t2 (i > 0), vel1 = (pos1 - pos) / dt
  float gcur[3];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    float gp1[3], gv1[3];
    ld3r(g_pos1, i, gp1);
    ld3r(g_vel1, i, gv1);
    for (int c = 0; c < 3; ++c) {
      const float gpn = gp1[c] + gv1[c] / P.dt;
      g_pos[3 * i + c] -= gv1[c] / P.dt;
      gcur[c] = i == 0 ? 0.f : gpn;   // particle 0 is held: no dependence on t2
    }
    st3r(gt, i, gcur);
  }
  __syncthreads();
  // phase adjoints: x_out = x_in + gather_dp(x_in); g_in = g_out + J^T g_out,
  // with the constraint (a, b) contributing w J (g_a - g_b) to a and minus it to b
  for (int phase = 0; phase < 2; ++phase) {
    const float* x = phase == 0 ? t1 : t0;
    const int d = phase == 0 ? 2 : 1;
    const float L = phase == 0 ? P.bend : P.rest;
    const float k = phase == 0 ? P.k_bend : P.k_stretch;
    float acc[3];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      float g_i[3];
      ld3r(gt, i, g_i);
      for (int c = 0; c < 3; ++c) acc[c] = g_i[c];
      float gd[3], gn[3], g_o[3];
      if (i + d < n) {
        ld3r(gt, i + d, g_o);
        for (int c = 0; c < 3; ++c) gd[c] = P.inv_m * (g_i[c] - g_o[c]);
        constraint_dp_adj(x, i, i + d, L, k, P.eps, gd, gn);
        for (int c = 0; c < 3; ++c) acc[c] += gn[c];
      }
      if (i - d >= 0) {
        ld3r(gt, i - d, g_o);
        for (int c = 0; c < 3; ++c) gd[c] = P.inv_m * (g_o[c] - g_i[c]);
        constraint_dp_adj(x, i - d, i, L, k, P.eps, gd, gn);
        for (int c = 0; c < 3; ++c) acc[c] -= gn[c];
      }
      // staged in t2 (no longer needed): gt is still being read by others
      st3r(t2, i, acc);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      float a[3];
      ld3r(t2, i, a);
      st3r(gt, i, a);
    }
    __syncthreads();
  }
  // predict: t0 = p + v' dt, v' = F(v + e / m + g dt)
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    float p[3], v[3], e[3], vn[3], g0[3];
    ld3r(pos, i, p);
    ld3r(vel, i, v);
    ld3r(ext, i, e);
    const bool floor = rope_predict(P, p, v, e, vn);
    ld3r(gt, i, g0);
    float gvp[3] = {g0[0] * P.dt, g0[1] * P.dt, g0[2] * P.dt};
    if (floor) {
      gvp[0] *= P.friction;
      gvp[1] = 0.f;
      gvp[2] *= P.friction;
    }
    for (int c = 0; c < 3; ++c) {
      g_pos[3 * i + c] += g0[c];
      g_vel[3 * i + c] += gvp[c];
      g_ext[3 * i + c] += gvp[c] / P.p_mass;
    }
  }
}

}  // namespace dt

using namespace dt;

static RopeP rope_dev(const dt_rope_params* p, const float* fixed) {
  RopeP d;
  d.dt = p->dt;
  d.p_mass = p->p_mass;
  d.inv_m = 1.f / p->p_mass;
  d.table_height = p->table_height;
  d.friction = p->floor_friction;
  d.eps = 1e-10f;
  for (int c = 0; c < 3; ++c) d.g[c] = p->gravity[c], d.fixed[c] = fixed[c];
  d.rest = p->rest_length;
  d.bend = p->bending_length;
  const float dt2 = p->dt * p->dt;
  d.k_stretch = p->stretch_relaxation / (2.f * d.inv_m + p->stretch_compliance / dt2);
  d.k_bend = p->bending_relaxation / (2.f * d.inv_m + p->bending_compliance / dt2);
  return d;
}

extern "C" {

int dt_rope_step(const dt_rope_params* prm, int32_t n, const float* fixed_pos, const float* pos,
                 const float* vel, const float* ext_impulse, float* work, float* pos_next,
                 float* vel_next, void* stream) {
  DT_REQUIRE(prm && fixed_pos, "NULL params");
  DT_REQUIRE(n >= 1, "n must be >= 1");
  DT_REQUIRE(prm->dt > 0.f && prm->p_mass > 0.f, "dt and p_mass must be > 0");
  DT_REQUIRE(pos && vel && ext_impulse && work && pos_next && vel_next, "NULL pointer");
  hipLaunchKernelGGL(k_rope_step, dim3(1), dim3(kRopeThreads), 0, (hipStream_t)stream,
                     rope_dev(prm, fixed_pos), (int)n, pos, vel, ext_impulse, work, pos_next,
                     vel_next);
  return check_launch("dt_rope_step");
}

int dt_rope_step_adj(const dt_rope_params* prm, int32_t n, const float* fixed_pos,
                     const float* pos, const float* vel, const float* ext_impulse,
                     const float* g_pos_next, const float* g_vel_next, float* work, float* g_pos,
                     float* g_vel, float* g_ext_impulse, void* stream) {
  DT_REQUIRE(prm && fixed_pos, "NULL params");
  DT_REQUIRE(n >= 1, "n must be >= 1");
  DT_REQUIRE(prm->dt > 0.f && prm->p_mass > 0.f, "dt and p_mass must be > 0");
  DT_REQUIRE(pos && vel && ext_impulse && g_pos_next && g_vel_next && work && g_pos && g_vel &&
                 g_ext_impulse,
             "NULL pointer");
  hipLaunchKernelGGL(k_rope_step_adj, dim3(1), dim3(kRopeThreads), 0, (hipStream_t)stream,
                     rope_dev(prm, fixed_pos), (int)n, pos, vel, ext_impulse, g_pos_next,
                     g_vel_next, work, g_pos, g_vel, g_ext_impulse);
  return check_launch("dt_rope_step_adj");
}

}  // extern "C"
