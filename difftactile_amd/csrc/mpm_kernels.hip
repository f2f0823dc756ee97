// MLS-MPM object kernels for gfx950: per-reference-kernel entry points and the
// fused substep path.
//
// Reference kernels (difftactile/object_model/):
//   compute_new_F mpm_elastic.py:134-137   svd        mpm_elastic.py:139-142
//   svd_grad      mpm_elastic.py:144-173   p2g        mpm_elastic.py:175-208,
//                                                     mpm_plastic.py:221-256,
//                                                     multi_obj.py:227-252
//   check_grid_occupy mpm_elastic.py:210-214  grid_op mpm_elastic.py:216-240
//   g2p           mpm_elastic.py:243-260
//
// Scatter strategy (p2g and g2p adjoint).  Every particle's contribution to
// stencil node base + o (o in {0,1,2}^3) has the form
//     w_o(fx) * (b + M o),     w_o = wx[o0] wy[o1] wz[o2],
// with b, M per particle (p2g: b = p_mass v - A fx dx, M = A dx; g2p adjoint:
// b = gvt - 4 inv_dx gC fx, M = 4 inv_dx gC).  A block stages (w, b, M) of its
// particles in LDS and splits them into runs of consecutive particles that
// share a stencil base (cell).  One work item per (run, o1, o2) sums the run's
// contributions to the three nodes o0 = 0..2 and adds each sum to the grid
// with one global float atomic per component.  With cell-sorted particles a
// run holds all of a cell's particles in the block (about 8 at two particles
// per cell edge), so a node receives one atomic per neighbouring cell instead
// of one per particle, and no LDS float atomics are used (those execute per
// lane on gfx950 and dominated the per-particle scatter).  Unsorted input
// gives runs of one particle and stays correct.
#include <climits>
#include <mutex>
#include <string>
#include <vector>

#include "dt_bin.h"
#include "dt_common.h"
#include "dt_device.h"
#include "dt_graph.h"
#include "dt_timing.h"

namespace dt {

struct Grid {
  float* mv;   // momentum (or any 3-vector)
  float* m;    // mass (may be null for 3-component scatters)
  int smv, sm; // strides in floats
};

constexpr int kStage = kBlock + 1;   // padded row stride of the LDS staging rows

// LDS staging of one block's scatter: row r of w/b/M holds component r of every
// particle (w rows: wx0..2, wy0..2, wz0..2; M rows: row-major 3x3).
struct RunStage {
  float w[9 * kStage];
  float b[3 * kStage];
  float M[9 * kStage];
  int key[kBlock];           // linear node index of the stencil base, -1 when inactive
  int start[kBlock + 1];     // first particle of each run, start[R] = active count
  int wtot[kBlock / 64];
};

__device__ __forceinline__ int node_index(int i, int j, int k, int n) { return (i * n + j) * n + k; }

// weight / weight derivative of stencil offset o (runtime) along axis 0,
// recomputed from the scalar fx with the same expressions as make_stencil
// (selecting from the weight array by a runtime index would demote the stencil
// to scratch memory)
__device__ __forceinline__ float w_axis0(const Stencil& s, int o) {
  float f = s.fx[0];
  float d0 = 1.5f - f, d1 = f - 1.f, d2 = f - 0.5f;
  float a = 0.5f * (d0 * d0), b = 0.75f - d1 * d1, c = 0.5f * (d2 * d2);
  return o == 0 ? a : (o == 1 ? b : c);
}
__device__ __forceinline__ float dw_axis0(const Stencil& s, int o) {
  float f = s.fx[0];
  float a = -(1.5f - f), b = -2.f * (f - 1.f), c = f - 0.5f;
  return o == 0 ? a : (o == 1 ? b : c);
}

// Stencil node sources.  at(base) binds a particle's stencil base; the bound
// accessor returns node base + (i, j, k) as (x, y, z, w).
// kBatch: global-memory sources, whose nine loads of a slice are issued
// together; LDS sources are read where used.
struct NodeRef4 {              // float4 nodes, row-major n^3
  static constexpr bool kBatch = true;
  const float4* p;
  int n;
  __device__ __forceinline__ float4 operator()(int i, int j, int k) const {
    return p[(i * n + j) * n + k];
  }
};
struct Nodes4 {
  const float4* g;
  int n;
  __device__ __forceinline__ NodeRef4 at(const int (&b)[3]) const {
    return NodeRef4{g + node_index(b[0], b[1], b[2], n), n};
  }
};
template <bool kW>
struct NodeRefStrided {        // 3 floats at stride s (+ a 4th array at stride sq)
  static constexpr bool kBatch = true;
  const float* p;
  const float* q;
  int s, sq, n;
  __device__ __forceinline__ float4 operator()(int i, int j, int k) const {
    const int o = (i * n + j) * n + k;
    return make_float4(p[(long long)o * s], p[(long long)o * s + 1], p[(long long)o * s + 2],
                       kW ? q[(long long)o * sq] : 0.f);
  }
};
template <bool kW>
struct NodesStrided {
  const float* g;
  const float* gw;
  int s, sw, n;
  __device__ __forceinline__ NodeRefStrided<kW> at(const int (&b)[3]) const {
    const long long o = node_index(b[0], b[1], b[2], n);
    return NodeRefStrided<kW>{g + o * s, kW ? gw + o * sw : nullptr, s, sw, n};
  }
};

// Sum-factorised gather over the 27 stencil nodes o = (i, j, k) of grid values
// g_o with weights w_o = wx_i wy_j wz_k:
//   G      = sum_o w_o g_o.xyz
//   GO[r]  = sum_o w_o g_o.xyz o_r                           (kMoments)
//   gfx[a] = sum_o (d w_o / d fx_a) (g_o.xyz . (b + M o) + g_o.w m3)   (kGw)
// The k sums are formed per (i, j) row and the j sums per i, so a node costs
// about fifteen FMAs instead of the forty-odd of the expanded per-node form.
template <bool kGw, bool kMoments, class Ref>
__device__ __forceinline__ void stencil_gather(const Stencil& s, const Ref& g, const float (&b)[3],
                                               const M3& M, float m3, float (&G)[3],
                                               float (&GO)[3][3], float (&gfx)[3]) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    G[q] = 0.f;
    gfx[q] = 0.f;
#pragma unroll
    for (int r = 0; r < 3; ++r) GO[r][q] = 0.f;
  }
#pragma unroll 1
  for (int i = 0; i < 3; ++i) {
    const float wi = w_axis0(s, i), dwi = dw_axis0(s, i), fi = (float)i;
    float ci[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) ci[q] = b[q] + fi * M.a[3 * q];
    float Pi = 0.f, Qi = 0.f, Ri = 0.f;
    float Gi[3] = {0.f, 0.f, 0.f}, GJ[3] = {0.f, 0.f, 0.f}, GK[3] = {0.f, 0.f, 0.f};
    // global sources: the slice's nine nodes are loaded before any is used,
    // so the loads are in flight together
    float4 vs[3][3];
    if constexpr (Ref::kBatch) {
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) vs[j][k] = g(i, j, k);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float wj = s.w[j][1], dwj = s.dw[j][1];
      float cij[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) cij[q] = j == 0 ? ci[q] : ci[q] + (float)j * M.a[3 * q + 1];
      float Pk = 0.f, Rk = 0.f, H[3] = {0.f, 0.f, 0.f}, HK[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float wk = s.w[k][2], dwk = s.dw[k][2];
        float4 v;
        if constexpr (Ref::kBatch) v = vs[j][k];
        else v = g(i, j, k);
        if (kGw) {
          float c[3];
#pragma unroll
          for (int q = 0; q < 3; ++q) c[q] = k == 0 ? cij[q] : cij[q] + (float)k * M.a[3 * q + 2];
          const float gw = v.x * c[0] + v.y * c[1] + v.z * c[2] + v.w * m3;
          Pk += wk * gw;
          Rk += dwk * gw;
        }
        H[0] += wk * v.x;
        H[1] += wk * v.y;
        H[2] += wk * v.z;
        if (kMoments && k > 0) {
          const float kw = (float)k * wk;
          HK[0] += kw * v.x;
          HK[1] += kw * v.y;
          HK[2] += kw * v.z;
        }
      }
      if (kGw) {
        Pi += wj * Pk;
        Qi += dwj * Pk;
        Ri += wj * Rk;
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        Gi[q] += wj * H[q];
        if (kMoments) {
          if (j > 0) GJ[q] += ((float)j * wj) * H[q];
          GK[q] += wj * HK[q];
        }
      }
    }
    if (kGw) {
      gfx[0] += dwi * Pi;
      gfx[1] += wi * Qi;
      gfx[2] += wi * Ri;
    }
    const float iw = fi * wi;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      G[q] += wi * Gi[q];
      if (kMoments) {
        GO[0][q] += iw * Gi[q];
        GO[1][q] += wi * GJ[q];
        GO[2][q] += wi * GK[q];
      }
    }
  }
}

// ----------------------------------------------------------------------------
// per-particle helpers
// ----------------------------------------------------------------------------
__device__ __forceinline__ Mat mat_row(const dt_material* mats, int id) {
  const dt_material& m = mats[id];
  Mat r;
  r.mu = m.mu;
  r.lam = m.lam;
  r.yield_stress = m.yield_stress;
  r.dp_alpha = m.dp_alpha;
  r.kind = m.kind;
  return r;
}
__device__ __forceinline__ Mat load_mat(const dt_material* mats, const int32_t* mat_id, int p) {
  return mat_row(mats, mat_id ? mat_id[p] : 0);
}

// affine = stress + p_mass C with stress = (-dt p_vol 4 inv_dx^2) tau
__device__ __forceinline__ M3 affine_of(const MpmP& P, const M3& tau, const M3& C) {
  float k = -P.dt * P.p_vol * 4.f * P.inv_dx * P.inv_dx;
  M3 A;
  for (int i = 0; i < 9; ++i) A.a[i] = k * tau.a[i] + P.p_mass * C.a[i];
  return A;
}

// Stage one particle (thread) and split the block into runs; returns the run
// count.  Active threads must be a prefix of the block (p < n).
__device__ __forceinline__ int stage_runs(RunStage& sh, bool active, int n_act, const Stencil& s,
                                          int n_grid, const float (&b)[3], const M3& M) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int key = active ? (s.base[0] * n_grid + s.base[1]) * n_grid + s.base[2] : -1;
  sh.key[t] = key;
  if (active) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int o = 0; o < 3; ++o) sh.w[(3 * a + o) * kStage + t] = s.w[o][a];
#pragma unroll
    for (int c = 0; c < 3; ++c) sh.b[c * kStage + t] = b[c];
#pragma unroll
    for (int r = 0; r < 9; ++r) sh.M[r * kStage + t] = M.a[r];
  }
  __syncthreads();
  const bool head = active && (t == 0 || sh.key[t - 1] != key);
  const unsigned long long hm = __ballot(head);
  const int below = __popcll(hm & ((1ull << lane) - 1ull));
  if (lane == 0) sh.wtot[wv] = __popcll(hm);
  __syncthreads();
  int off = 0, R = 0;
#pragma unroll
  for (int i = 0; i < kBlock / 64; ++i) {
    if (i < wv) off += sh.wtot[i];
    R += sh.wtot[i];
  }
  if (head) sh.start[off + below] = t;
  if (t == 0) sh.start[R] = n_act;
  __syncthreads();
  return R;
}

// Sum each run's contributions per node and add them to the grid (see the
// file header).  kMass adds p_mass * sum(w) to the mass channel.
template <bool kMass>
__device__ __forceinline__ void run_scatter(const RunStage& sh, int R, int n_grid, Grid g,
                                            float p_mass) {
  const long long n2 = (long long)n_grid * n_grid;
  for (int it = threadIdx.x; it < R * 9; it += blockDim.x) {
    const int run = it / 9, jk = it - run * 9;
    const int j = jk / 3, k = jk - 3 * j;
    const float fj = (float)j, fk = (float)k;
    const int q0 = sh.start[run], q1 = sh.start[run + 1];
    float acc[3][3], accm[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      accm[i] = 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[i][c] = 0.f;
    }
    for (int q = q0; q < q1; ++q) {
      const float wyz = sh.w[(3 + j) * kStage + q] * sh.w[(6 + k) * kStage + q];
      float u[3], m0[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        m0[c] = sh.M[(3 * c) * kStage + q];
        u[c] = sh.b[c * kStage + q] + sh.M[(3 * c + 1) * kStage + q] * fj +
               sh.M[(3 * c + 2) * kStage + q] * fk;
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const float wi = sh.w[i * kStage + q] * wyz;
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[i][c] += wi * (u[c] + m0[c] * (float)i);
        accm[i] += wi;
      }
    }
    const long long nd0 = (long long)sh.key[q0] + (long long)j * n_grid + k;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const long long nd = nd0 + i * n2;
      atomic_add_f(g.mv + nd * g.smv + 0, acc[i][0]);
      atomic_add_f(g.mv + nd * g.smv + 1, acc[i][1]);
      atomic_add_f(g.mv + nd * g.smv + 2, acc[i][2]);
      if (kMass) atomic_add_f(g.m + nd * g.sm, p_mass * accm[i]);
    }
  }
}

__device__ __forceinline__ int block_active(int n) {
  const int first = xcd_block() * blockDim.x;
  return n - first >= (int)blockDim.x ? (int)blockDim.x : (n > first ? n - first : 0);
}

// Shared body of the p2g kernels once the per-particle stress is known:
// contribution w_o (p_mass v + A (o - fx) dx), mass w_o p_mass.
__device__ __forceinline__ void p2g_block(const MpmP& P, bool active, int n, const Stencil& s,
                                          const float (&pv)[3], const M3& A, Grid g,
                                          RunStage& sh) {
  float b[3] = {0.f, 0.f, 0.f};
  M3 M = m3_zero();
  if (active) {
    for (int c = 0; c < 3; ++c) {
      b[c] = pv[c] - (A.a[3 * c] * s.fx[0] + A.a[3 * c + 1] * s.fx[1] + A.a[3 * c + 2] * s.fx[2]) * P.dx;
      for (int a = 0; a < 3; ++a) M.a[3 * c + a] = A.a[3 * c + a] * P.dx;
    }
  }
  const int R = stage_runs(sh, active, block_active(n), s, P.n_grid, b, M);
  run_scatter<true>(sh, R, P.n_grid, g, P.p_mass);
}

// ----------------------------------------------------------------------------
// kernels: compute_new_F, svd, svd_grad
// ----------------------------------------------------------------------------
// init_object: x = R p + t, v = ball_vel, F = I (one thread per particle)
struct InitObj {
  float T[12], v[3];
};
__global__ __launch_bounds__(kBlock) void k_init_object(int n, InitObj A,
                                                        const float* __restrict__ pts,
                                                        float* __restrict__ x,
                                                        float* __restrict__ v,
                                                        float* __restrict__ F) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  for (int r = 0; r < 3; ++r) {
    x[3 * i + r] = A.T[4 * r] * p[0] + A.T[4 * r + 1] * p[1] + A.T[4 * r + 2] * p[2] + A.T[4 * r + 3];
    v[3 * i + r] = A.v[r];
  }
  for (int k = 0; k < 9; ++k) F[9 * i + k] = (k % 4 == 0) ? 1.f : 0.f;
}

__global__ __launch_bounds__(kBlock) void k_compute_new_F(int n, const float* __restrict__ C,
                                                          const float* __restrict__ F,
                                                          float* __restrict__ Fn, float dt) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  M3 c = ld_m3(C + 9ll * p), f = ld_m3(F + 9ll * p);
  M3 a = m3_eye();
  for (int i = 0; i < 9; ++i) a.a[i] += dt * c.a[i];
  st_m3(Fn + 9ll * p, m3_mul(a, f));
}

__global__ __launch_bounds__(kBlock) void k_compute_new_F_adj(int n, const float* __restrict__ C,
                                                              const float* __restrict__ F,
                                                              const float* __restrict__ gFn,
                                                              float* __restrict__ gC,
                                                              float* __restrict__ gF, float dt) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  M3 c = ld_m3(C + 9ll * p), f = ld_m3(F + 9ll * p), g = ld_m3(gFn + 9ll * p);
  M3 a = m3_eye();
  for (int i = 0; i < 9; ++i) a.a[i] += dt * c.a[i];
  M3 gc = m3_mul_bt(g, f);
  for (int i = 0; i < 9; ++i) gc.a[i] *= dt;
  add_m3(gC + 9ll * p, gc);
  add_m3(gF + 9ll * p, m3_mul_at(a, g));
}

__global__ __launch_bounds__(kBlock) void k_svd(int n, const float* __restrict__ Fn,
                                                float* __restrict__ U, float* __restrict__ S,
                                                float* __restrict__ V) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  M3 u, v;
  float s[3];
  svd3(ld_m3(Fn + 9ll * p), u, s, v);
  st_m3(U + 9ll * p, u);
  st_m3(V + 9ll * p, v);
  M3 sm = m3_zero();
  sm.a[0] = s[0];
  sm.a[4] = s[1];
  sm.a[8] = s[2];
  st_m3(S + 9ll * p, sm);
}

__device__ __forceinline__ void svd_adj_core(const M3& U, const float (&sig)[3], const M3& V,
                                             const M3d& gU, const M3d& gV, const double (&gsig)[3],
                                             M3& out) {
  M3d Ud = m3d_from(U), Vd = m3d_from(V);
  double s2[3] = {(double)sig[0] * sig[0], (double)sig[1] * sig[1], (double)sig[2] * sig[2]};
  M3d Xu = m3d_mul_at(Ud, gU);
  M3d Xv = m3d_mul_at(Vd, gV);
  M3d Mi;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double v = 0.0;
      if (i != j) {
        double ff = 1.0 / clampd(s2[j] - s2[i]);
        v = ff * (Xu.a[i * 3 + j] - Xu.a[j * 3 + i]) * (double)sig[j] +
            (double)sig[i] * ff * (Xv.a[i * 3 + j] - Xv.a[j * 3 + i]);
      } else {
        v = gsig[i];
      }
      Mi.a[i * 3 + j] = v;
    }
  M3d G = m3d_mul_bt(m3d_mul(Ud, Mi), Vd);
  for (int i = 0; i < 9; ++i) out.a[i] = (float)G.a[i];
}

__global__ __launch_bounds__(kBlock) void k_svd_adj(int n, const float* __restrict__ U,
                                                    const float* __restrict__ S,
                                                    const float* __restrict__ V,
                                                    const float* __restrict__ gU,
                                                    const float* __restrict__ gS,
                                                    const float* __restrict__ gV,
                                                    float* __restrict__ gFn) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  M3 u = ld_m3(U + 9ll * p), v = ld_m3(V + 9ll * p), s = ld_m3(S + 9ll * p);
  float sig[3] = {s.a[0], s.a[4], s.a[8]};
  M3 gs = ld_m3(gS + 9ll * p);
  double gsig[3] = {gs.a[0], gs.a[4], gs.a[8]};
  M3 out;
  svd_adj_core(u, sig, v, m3d_from(ld_m3(gU + 9ll * p)), m3d_from(ld_m3(gV + 9ll * p)), gsig, out);
  add_m3(gFn + 9ll * p, out);
}

// ----------------------------------------------------------------------------
// p2g (reference kernel form: reads F_new, U, S, V)
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_p2g(MpmP P, const dt_material* __restrict__ mats,
                                                int n, const int32_t* __restrict__ mat_id,
                                                const float* __restrict__ x,
                                                const float* __restrict__ v,
                                                const float* __restrict__ C,
                                                const float* __restrict__ Fn,
                                                const float* __restrict__ U,
                                                const float* __restrict__ S,
                                                const float* __restrict__ V, Grid g,
                                                float* __restrict__ F_next) {
  __shared__ RunStage sh;
  int p = xcd_block() * blockDim.x + threadIdx.x;
  bool active = p < n;
  Stencil s;
  M3 A;
  float pv[3] = {0.f, 0.f, 0.f};
  s.base[0] = s.base[1] = s.base[2] = 0;
  if (active) {
    M3 fn = ld_m3(Fn + 9ll * p), u = ld_m3(U + 9ll * p), vv = ld_m3(V + 9ll * p);
    M3 sm = ld_m3(S + 9ll * p);
    float sig[3] = {sm.a[0], sm.a[4], sm.a[8]};
    Mat m = load_mat(mats, mat_id, p);
    StressFwd st;
    M3 tau;
    stress_fwd(fn, u, sig, vv, m, st, tau);
    A = affine_of(P, tau, ld_m3(C + 9ll * p));
    st_m3(F_next + 9ll * p, st.Fp);
    float xp[3] = {x[3ll * p], x[3ll * p + 1], x[3ll * p + 2]};
    make_stencil(xp, P.inv_dx, P.n_grid, s, env_x0(P, p));
    for (int a = 0; a < 3; ++a) pv[a] = P.p_mass * v[3ll * p + a];
  }
  p2g_block(P, active, n, s, pv, A, g, sh);
}

// ----------------------------------------------------------------------------
// fused compute_new_F + svd + return map + p2g: reads (x, v, C, F)
// ----------------------------------------------------------------------------
// inputs of the fused p2g of one particle
struct P2GIn {
  float x[3], v[3];
  M3 c, f;
  int mid, env;   // material row; environment (0 unless stacked)
};
__device__ __forceinline__ void p2g_load(const MpmP& P, int n, const int32_t* mat_id,
                                         const float* x, const float* v, const float* C,
                                         const float* F, int p, P2GIn& in) {
  PL Lv = pl_vec(P.soa, n), Lm = pl_mat(P.soa, n);
  in.c = ld_m3L(C, p, Lm);
  in.f = ld_m3L(F, p, Lm);
  ld_v3(x, p, Lv, in.x);
  ld_v3(v, p, Lv, in.v);
  in.mid = mat_id ? mat_id[p] : 0;
  in.env = P.env_id ? P.env_id[p] : 0;
}

// compute_new_F + svd + return map of particle p; writes F_next and returns
// the stencil and the scatter coefficients (b, M) of the p2g contribution
// w_o (p_mass v + A (o - fx) dx) = w_o (b + M o)
// P: the view of the parameters for the particle's environment, x0 the
// environment's first x-node
__device__ __forceinline__ void p2g_compute_at(const MpmP& P, int x0, const dt_material* mats,
                                               int n, const P2GIn& in, float* F_next, int p,
                                               Stencil& s, float (&b)[3], M3& M) {
  PL Lm = pl_mat(P.soa, n);
  M3 a = m3_eye();
  for (int i = 0; i < 9; ++i) a.a[i] += P.dt * in.c.a[i];
  M3 fn = m3_mul(a, in.f);
  M3 u, vv;
  float sig[3];
  svd3(fn, u, sig, vv);
  Mat m = mat_row(mats, in.mid);
  StressFwd st;
  M3 tau;
  stress_fwd(fn, u, sig, vv, m, st, tau);
  const M3 A = affine_of(P, tau, in.c);
  st_m3L(F_next, p, Lm, st.Fp);
  make_stencil(in.x, P.inv_dx, P.n_grid, s, x0);
  for (int q = 0; q < 3; ++q) {
    b[q] = P.p_mass * in.v[q] -
           (A.a[3 * q] * s.fx[0] + A.a[3 * q + 1] * s.fx[1] + A.a[3 * q + 2] * s.fx[2]) * P.dx;
    for (int r = 0; r < 3; ++r) M.a[3 * q + r] = A.a[3 * q + r] * P.dx;
  }
}

__device__ __forceinline__ void p2g_fused_compute(const MpmP& P, const dt_material* mats, int n,
                                                  const P2GIn& in, float* F_next, int p,
                                                  Stencil& s, float (&b)[3], M3& M) {
  p2g_compute_at(at_particle(P, p), in.env * P.n_grid, mats, n, in, F_next, p, s, b, M);
}

__device__ __forceinline__ void p2g_fused_particle(const MpmP& P, const dt_material* mats, int n,
                                                   const int32_t* mat_id, const float* x,
                                                   const float* v, const float* C, const float* F,
                                                   float* F_next, int p, Stencil& s,
                                                   float (&b)[3], M3& M) {
  P2GIn in;
  p2g_load(P, n, mat_id, x, v, C, F, p, in);
  p2g_fused_compute(P, mats, n, in, F_next, p, s, b, M);
}

__global__ __launch_bounds__(kBlock) void k_p2g_fused(MpmP P, const dt_material* __restrict__ mats,
                                                      int n, const int32_t* __restrict__ mat_id,
                                                      const float* __restrict__ x,
                                                      const float* __restrict__ v,
                                                      const float* __restrict__ C,
                                                      const float* __restrict__ F, Grid g,
                                                      float* __restrict__ F_next) {
  __shared__ RunStage sh;
  int p = xcd_block() * blockDim.x + threadIdx.x;
  bool active = p < n;
  Stencil s;
  s.base[0] = s.base[1] = s.base[2] = 0;
  float b[3] = {0.f, 0.f, 0.f};
  M3 M = m3_zero();
  if (active) p2g_fused_particle(P, mats, n, mat_id, x, v, C, F, F_next, p, s, b, M);
  const int R = stage_runs(sh, active, block_active(n), s, P.n_grid, b, M);
  run_scatter<true>(sh, R, P.n_grid, g, P.p_mass);
}

// ----------------------------------------------------------------------------
// p2g adjoint core: gathers the grid adjoints (d/d mv in xyz, d/d m in w) over
// the 27 nodes.  The forward contribution to node o is w_o (pv + A d_o, m) with
// d_o = (o - fx) dx = w_o (b + M o, m), b = pv - A fx dx, M = A dx.
//   outputs gx (+= through weights and dpos), gv, gA (adjoint of affine)
// ----------------------------------------------------------------------------
template <class Ref>
__device__ __forceinline__ void p2g_gather_adj(const MpmP& P, const Stencil& s,
                                               const float (&pv)[3], const M3& A, const Ref& g,
                                               float (&gx)[3], float (&gv)[3], M3& gA) {
  float b[3];
  M3 M;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    b[q] = pv[q] - (A.a[3 * q] * s.fx[0] + A.a[3 * q + 1] * s.fx[1] + A.a[3 * q + 2] * s.fx[2]) *
                       P.dx;
#pragma unroll
    for (int r = 0; r < 3; ++r) M.a[3 * q + r] = A.a[3 * q + r] * P.dx;
  }
  float G[3], GO[3][3], gfx[3];
  stencil_gather<true, true>(s, g, b, M, P.p_mass, G, GO, gfx);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    gv[q] = P.p_mass * G[q];
#pragma unroll
    for (int r = 0; r < 3; ++r) gA.a[3 * q + r] = P.dx * (GO[r][q] - G[q] * s.fx[r]);
  }
  // fx = x inv_dx - base;  dpos = (o - fx) dx:  d/dx through dpos is -A^T G
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float gd = A.a[a] * G[0] + A.a[3 + a] * G[1] + A.a[6 + a] * G[2];
    gx[a] += P.inv_dx * gfx[a] - P.inv_dx * P.dx * gd;
  }
}

// affine adjoint -> tau adjoint and gC
__device__ __forceinline__ void affine_adj(const MpmP& P, const M3& gA, M3& gtau, M3& gC) {
  float k = -P.dt * P.p_vol * 4.f * P.inv_dx * P.inv_dx;
  for (int i = 0; i < 9; ++i) {
    gtau.a[i] = k * gA.a[i];
    gC.a[i] = P.p_mass * gA.a[i];
  }
}

// Adds the block's per-particle material adjoints (gmu, glam) of material
// ids below kMatSlots: per wavefront one shuffle sum per material present,
// stored in the wavefront's own LDS slots; then the sums over the block's
// wavefronts, in wavefront order, are added to part[blockIdx.x][...] (the
// block's own slots, accumulated over the launches of a trajectory: no
// atomics, the same sums on every run) or, with part NULL, to g_mats by one
// global atomic per block and material.  Ids at or beyond kMatSlots add per
// wavefront to g_mats (atomics, order not fixed).  Every thread of the block
// calls it.
constexpr int kMatSlots = 16;
__device__ __forceinline__ void block_mat_grad(bool active, int id, float gmu, float glam,
                                               int n_mats, float* __restrict__ g_mats,
                                               float* __restrict__ part) {
  __shared__ float red[kBlock / 64][2 * kMatSlots];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  // each wavefront clears its own row (its later writes follow in program order)
  if (lane < 2 * kMatSlots) red[w][lane] = 0.f;
  unsigned long long todo = __ballot(active);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int cur = __shfl(id, leader);
    const bool mine = active && id == cur;
    float a = mine ? gmu : 0.f, b = mine ? glam : 0.f;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off);
      b += __shfl_xor(b, off);
    }
    if (lane == leader) {
      if (cur < kMatSlots) {
        red[w][2 * cur] = a;
        red[w][2 * cur + 1] = b;
      } else {
        atomicAdd(&g_mats[2 * cur], a);
        atomicAdd(&g_mats[2 * cur + 1], b);
      }
    }
    todo &= ~__ballot(mine);
  }
  __syncthreads();
  if (t < 2 * min(n_mats, kMatSlots)) {
    float sum = 0.f;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) sum += red[q][t];
    if (part)
      part[(long long)blockIdx.x * (2 * kMatSlots) + t] += sum;
    else if (sum != 0.f)
      atomicAdd(&g_mats[t], sum);
  }
}

// g_mats[v] += the sum over blocks b of part[b][v] (v < n_vals), in a fixed
// order: thread t sums blocks t, t + 256, ..., then a pairwise tree over the
// threads.  One block.
__global__ __launch_bounds__(kBlock) void k_mat_part_reduce(int n_blocks, int n_vals,
                                                            const float* __restrict__ part,
                                                            float* __restrict__ g_mats) {
  __shared__ float red[kBlock][2 * kMatSlots + 1];
  const int t = threadIdx.x;
  float acc[2 * kMatSlots];
#pragma unroll
  for (int v = 0; v < 2 * kMatSlots; ++v) acc[v] = 0.f;
  for (int b = t; b < n_blocks; b += kBlock) {
    const float4* q = reinterpret_cast<const float4*>(part + (long long)b * (2 * kMatSlots));
#pragma unroll
    for (int v = 0; v < 2 * kMatSlots / 4; ++v) {
      const float4 f = q[v];
      acc[4 * v] += f.x;
      acc[4 * v + 1] += f.y;
      acc[4 * v + 2] += f.z;
      acc[4 * v + 3] += f.w;
    }
  }
#pragma unroll
  for (int v = 0; v < 2 * kMatSlots; ++v) red[t][v] = acc[v];
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int v = 0; v < 2 * kMatSlots; ++v) red[t][v] += red[t + h][v];
    __syncthreads();
  }
  if (t < n_vals) g_mats[t] += red[0][t];
}

__global__ __launch_bounds__(kBlock) void k_zero_f(long long n, float* __restrict__ a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.f;
}

// p2g.grad (reference kernel form): emits gU, gS, gV for svd_grad, the direct
// F_new adjoint and, when g_mats is non-null, the material adjoints.
__global__ __launch_bounds__(kBlock) void k_p2g_adj(
    MpmP P, const dt_material* __restrict__ mats, int n, const int32_t* __restrict__ mat_id,
    const float* __restrict__ x, const float* __restrict__ v, const float* __restrict__ C,
    const float* __restrict__ Fn, const float* __restrict__ U, const float* __restrict__ S,
    const float* __restrict__ V, const float* __restrict__ gmv, const float* __restrict__ gm,
    const float* __restrict__ gFnext, float* __restrict__ gx, float* __restrict__ gv,
    float* __restrict__ gC, float* __restrict__ gFn, float* __restrict__ gU,
    float* __restrict__ gS, float* __restrict__ gV, int n_mats, float* __restrict__ g_mats) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = p < n;
  float gmu = 0.f, glam = 0.f;
  if (active) {
  M3 fn = ld_m3(Fn + 9ll * p), u = ld_m3(U + 9ll * p), vv = ld_m3(V + 9ll * p);
  M3 sm = ld_m3(S + 9ll * p);
  float sig[3] = {sm.a[0], sm.a[4], sm.a[8]};
  Mat m = load_mat(mats, mat_id, p);
  StressFwd st;
  M3 tau;
  stress_fwd(fn, u, sig, vv, m, st, tau);
  M3 c = ld_m3(C + 9ll * p);
  M3 A = affine_of(P, tau, c);
  Stencil s;
  float xp[3] = {x[3ll * p], x[3ll * p + 1], x[3ll * p + 2]};
  make_stencil(xp, P.inv_dx, P.n_grid, s, env_x0(P, p));
  float pv[3];
  for (int a = 0; a < 3; ++a) pv[a] = P.p_mass * v[3ll * p + a];
  float lgx[3] = {0.f, 0.f, 0.f}, lgv[3] = {0.f, 0.f, 0.f};
  M3 gA;
  p2g_gather_adj(P, s, pv, A, NodesStrided<true>{gmv, gm, 3, 1, P.n_grid}.at(s.base), lgx, lgv,
                 gA);
  M3 gtau, lgC;
  affine_adj(P, gA, gtau, lgC);
  for (int a = 0; a < 3; ++a) {
    gx[3ll * p + a] += lgx[a];
    gv[3ll * p + a] += lgv[a];
  }
  add_m3(gC + 9ll * p, lgC);

  // stress adjoint, split as the reference splits it: direct F_new part here,
  // U/S/V parts stored for svd_grad.
  M3 gFnext_p = ld_m3(gFnext + 9ll * p);
  M3 D;
  for (int i = 0; i < 9; ++i) D.a[i] = st.Fp.a[i] - st.R.a[i];
  M3 gFp = m3_mul(gtau, st.Fp);
  M3 t2 = m3_mul_at(gtau, D);
  for (int i = 0; i < 9; ++i) gFp.a[i] = 2.f * m.mu * (gFp.a[i] + t2.a[i]) + gFnext_p.a[i];
  M3 gR = m3_mul(gtau, st.Fp);
  for (int i = 0; i < 9; ++i) gR.a[i] *= -2.f * m.mu;
  float trg = gtau.a[0] + gtau.a[4] + gtau.a[8];
  float gJ = m.lam * (2.f * st.J - 1.f) * trg;
  // d tau / d mu = 2 (Fp - R) Fp^T, d tau / d lam = J (J - 1) I
  {
    const M3 gtF = m3_mul(gtau, st.Fp);
    for (int i = 0; i < 9; ++i) gmu += 2.f * gtF.a[i] * D.a[i];
    glam = trg * st.J * (st.J - 1.f);
  }
  if (m.kind != KIND_ELASTIC && st.mode == 3) {
    // cone tip: tau = 0 whatever F (F_p = R, J = 1); F_next = R's adjoint
    // reaches U and V through the rotation alone
    gR = gFnext_p;
    gFp = m3_zero();
    gJ = 0.f;
  }
  M3d gUd = m3d_mul(m3d_from(gR), m3d_from(vv));
  M3d gVd = m3d_mul_at(m3d_from(gR), m3d_from(u));
  double gsig[3] = {0.0, 0.0, 0.0};
  M3 gFdirect = m3_zero();
  if (m.kind == KIND_ELASTIC) {
    gsig[0] = (double)gJ * sig[1] * sig[2];
    gsig[1] = (double)gJ * sig[0] * sig[2];
    gsig[2] = (double)gJ * sig[0] * sig[1];
    gFdirect = gFp;
  } else {
    M3 cof = m3_cof(st.Fp);
    for (int i = 0; i < 9; ++i) gFp.a[i] += gJ * cof.a[i];
    if (st.mode == 0) {
      gFdirect = gFp;
    } else {
      // reuse stress_adj's return-map branch via a zero-gtau call would
      // recompute everything; the plastic branch is restated inline.
      double d[3] = {exp((double)st.epsn[0]), exp((double)st.epsn[1]), exp((double)st.epsn[2])};
      M3d Ud = m3d_from(u), Vd = m3d_from(vv), gFpd = m3d_from(gFp);
      M3d UtgV = m3d_mul(m3d_mul_at(Ud, gFpd), Vd);
      M3d gFpV = m3d_mul(gFpd, Vd);
      M3d gFptU = m3d_mul_at(gFpd, Ud);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          gUd.a[i * 3 + j] += gFpV.a[i * 3 + j] * d[j];
          gVd.a[i * 3 + j] += gFptU.a[i * 3 + j] * d[j];
        }
      if (st.mode != 3) {
        double ge[3];
        for (int i = 0; i < 3; ++i) ge[i] = UtgV.a[i * 4] * d[i];
        double eh[3] = {st.eh[0], st.eh[1], st.eh[2]};
        double nn = st.ehn;
        double gdot = ge[0] * eh[0] + ge[1] * eh[1] + ge[2] * eh[2];
        double geps[3], geh[3], gtr = 0.0;
        if (st.mode == 1) {
          double cc = (double)m.yield_stress / (2.0 * m.mu);
          for (int i = 0; i < 3; ++i) {
            geps[i] = ge[i];
            geh[i] = (cc / nn - 1.0) * ge[i] - cc * gdot / (nn * nn * nn) * eh[i];
          }
          // threshold cc = tau_y / (2 mu): d eps' / d cc = eh / n
          gmu += (float)(-gdot / nn * cc / m.mu);
        } else {
          double r = (3.0 * m.lam + 2.0 * m.mu) / (2.0 * m.mu) * m.dp_alpha;
          double tr = st.trv;
          for (int i = 0; i < 3; ++i) {
            geps[i] = ge[i];
            geh[i] = -ge[i] - r * tr * (ge[i] / nn - gdot * eh[i] / (nn * nn * nn));
          }
          gtr = -r * gdot / nn;
          // cone slope r = (3 lam + 2 mu) / (2 mu) alpha: d eps' / d r = -tr eh / n
          const double gr = -tr * gdot / nn, h = 3.0 * m.dp_alpha / (2.0 * m.mu);
          glam += (float)(gr * h);
          gmu += (float)(-gr * h * m.lam / m.mu);
        }
        double mean = (geh[0] + geh[1] + geh[2]) / 3.0;
        for (int i = 0; i < 3; ++i) {
          geps[i] += geh[i] - mean + gtr;
          if (sig[i] > 0.05f) gsig[i] += geps[i] / (double)st.s[i];
        }
      }
    }
  }
  add_m3(gFn + 9ll * p, gFdirect);
  M3 gu, gvv, gs = m3_zero();
  for (int i = 0; i < 9; ++i) {
    gu.a[i] = (float)gUd.a[i];
    gvv.a[i] = (float)gVd.a[i];
  }
  gs.a[0] = (float)gsig[0];
  gs.a[4] = (float)gsig[1];
  gs.a[8] = (float)gsig[2];
  add_m3(gU + 9ll * p, gu);
  add_m3(gV + 9ll * p, gvv);
  add_m3(gS + 9ll * p, gs);
  }
  if (g_mats)
    block_mat_grad(active, active ? (mat_id ? mat_id[p] : 0) : 0, gmu, glam, n_mats, g_mats,
                   nullptr);
}

// fused adjoint of (compute_new_F + svd + return map + p2g): writes the
// complete adjoint of (x, v, C, F); gx_io holds the g2p-adjoint partial on
// entry.
// Opaque copy: the compiler may not reuse values computed from the original
// (forces the SVD below to be recomputed instead of kept in registers).
__device__ __forceinline__ void opaque(M3& m) {
#pragma unroll
  for (int i = 0; i < 9; ++i) asm volatile("" : "+v"(m.a[i]));
}

// Two passes over the particle keep the register footprint small: the first
// recomputes the SVD and the stress to gather the affine adjoint from the grid
// and parks U, sigma, V in LDS (one column per thread); the second reloads them
// and recomputes only the return map for the stress adjoint.
template <bool kExact>
__device__ __forceinline__ void p2g_fused_adj_particle(
    const MpmP& P0, const dt_material* __restrict__ mats, int n, const int32_t* __restrict__ mat_id,
    const float* __restrict__ x, const float* __restrict__ v, const float* __restrict__ C,
    const float* __restrict__ F, const float* __restrict__ ggrid4,
    const float* __restrict__ gFnext, float* __restrict__ gx_io, float* __restrict__ gv_out,
    float* __restrict__ gC_out, float* __restrict__ gF_out, int p, float (*park)[kBlock],
    float& gmu, float& glam) {
  const MpmP P = at_particle(P0, p);
  const int t = threadIdx.x;
  PL Lv = pl_vec(P.soa, n), Lm = pl_mat(P.soa, n);
  const Mat m = load_mat(mats, mat_id, p);
  M3 gtau, lgC;
  {
    // pass 1: affine A, gather d/d(mv, m) over the stencil
    const M3 c = ld_m3L(C, p, Lm), f = ld_m3L(F, p, Lm);
    M3 a = m3_eye();
    for (int i = 0; i < 9; ++i) a.a[i] += P.dt * c.a[i];
    const M3 fn = m3_mul(a, f);
    M3 u, vv;
    float sig[3];
    svd3(fn, u, sig, vv);
    for (int r = 0; r < 9; ++r) {
      park[r][t] = u.a[r];
      park[12 + r][t] = vv.a[r];
    }
    for (int r = 0; r < 3; ++r) park[9 + r][t] = sig[r];
    StressFwd st;
    M3 tau;
    stress_fwd(fn, u, sig, vv, m, st, tau);
    const M3 A = affine_of(P, tau, c);
    Stencil s;
    float xp[3], vp[3];
    ld_v3(x, p, Lv, xp);
    ld_v3(v, p, Lv, vp);
    make_stencil(xp, P.inv_dx, P.n_grid, s, env_x0(P, p));
    float pv[3];
    for (int q = 0; q < 3; ++q) pv[q] = P.p_mass * vp[q];
    float lgx[3];
    ld_v3(gx_io, p, Lv, lgx);
    float lgv[3] = {0.f, 0.f, 0.f};
    M3 gA;
    p2g_gather_adj(P, s, pv, A, Nodes4{(const float4*)ggrid4, P.n_grid}.at(s.base), lgx, lgv,
                   gA);
    st_v3(gx_io, p, Lv, lgx);
    st_v3(gv_out, p, Lv, lgv);
    affine_adj(P, gA, gtau, lgC);
  }
  // pass 2: return map again from the parked SVD, stress adjoint,
  // F_new = (I + dt C) F
  M3 c = ld_m3L(C, p, Lm), f = ld_m3L(F, p, Lm);
  opaque(c);
  opaque(f);
  M3 a = m3_eye();
  for (int i = 0; i < 9; ++i) a.a[i] += P.dt * c.a[i];
  const M3 fn = m3_mul(a, f);
  M3 u, vv;
  float sig[3];
  for (int r = 0; r < 9; ++r) {
    u.a[r] = park[r][t];
    vv.a[r] = park[12 + r][t];
  }
  for (int r = 0; r < 3; ++r) sig[r] = park[9 + r][t];
  StressFwd st;
  M3 tau;
  stress_fwd(fn, u, sig, vv, m, st, tau);
  M3 gFn;
  stress_adj<kExact>(fn, u, sig, vv, m, st, gtau, ld_m3L(gFnext, p, Lm), gFn, gmu, glam);
  M3 gc2 = m3_mul_bt(gFn, f);
  for (int i = 0; i < 9; ++i) lgC.a[i] += P.dt * gc2.a[i];
  M3 gf = m3_mul_at(a, gFn);
  st_m3L(gC_out, p, Lm, lgC);
  st_m3L(gF_out, p, Lm, gf);
}

// kMatGrad: also adds the material adjoints (mu, lam per material) into
// g_mats[n_mats][2]
template <bool kExact, bool kMatGrad>
__global__ __launch_bounds__(kBlock, 3) void k_p2g_fused_adj(
    MpmP P, const dt_material* __restrict__ mats, int n, const int32_t* __restrict__ mat_id,
    const float* __restrict__ x, const float* __restrict__ v, const float* __restrict__ C,
    const float* __restrict__ F, const float* __restrict__ ggrid4,
    const float* __restrict__ gFnext, float* __restrict__ gx_io, float* __restrict__ gv_out,
    float* __restrict__ gC_out, float* __restrict__ gF_out, int n_mats,
    float* __restrict__ g_mats, float* __restrict__ mat_part) {
  __shared__ float park[21][kBlock];   // U (9), sigma (3), V (9) of each thread's particle
  const int p = xcd_block() * blockDim.x + threadIdx.x;
  float gmu = 0.f, glam = 0.f;
  if (!kMatGrad) {
    if (p >= n) return;
    p2g_fused_adj_particle<kExact>(P, mats, n, mat_id, x, v, C, F, ggrid4, gFnext, gx_io, gv_out,
                                   gC_out, gF_out, p, park, gmu, glam);
    return;
  }
  const bool active = p < n;
  if (active)
    p2g_fused_adj_particle<kExact>(P, mats, n, mat_id, x, v, C, F, ggrid4, gFnext, gx_io, gv_out,
                                   gC_out, gF_out, p, park, gmu, glam);
  block_mat_grad(active, active ? (mat_id ? mat_id[p] : 0) : 0, gmu, glam, n_mats, g_mats,
                 mat_part);
}

// ----------------------------------------------------------------------------
// grid kernels
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_grid_occupy(MpmP P, long long nn,
                                                        const float* __restrict__ m,
                                                        int32_t* __restrict__ occ) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  if (m[i] > P.eps) occ[i] = 1;
}

// grid_op on one node; returns whether the node is occupied.  mask[a] = 1 when
// component a was zeroed by a wall condition.
__device__ __forceinline__ void grid_node_fwd(const MpmP& P, long long nd, float mv0, float mv1,
                                              float mv2, float m, float f0, float f1, float f2,
                                              float (&vo)[3], bool (&mask)[3]) {
  // one rounding sequence wherever this is inlined (dense, node-set and tile
  // loops): no contraction, so hoisting dt*g out of a loop cannot change it
#pragma clang fp contract(off)
  const int n = P.n_grid;
  int i = (int)(nd / ((long long)n * n)), j = (int)((nd / n) % n), k = (int)(nd % n);
  float inv_m = 1.f / (m + P.eps);
  vo[0] = inv_m * mv0 + inv_m * f0 + P.dt * P.g[0];
  vo[1] = inv_m * mv1 + inv_m * f1 + P.dt * P.g[1];
  vo[2] = inv_m * mv2 + inv_m * f2 + P.dt * P.g[2];
  int idx[3] = {i % n, j, k};   // x index inside the node's environment
  // walls of the node's environment (its own grid size when ragged)
  const int nw = P.env_scale ? P.env_ngrid[i / n] : n;
  for (int a = 0; a < 3; ++a) {
    mask[a] = (idx[a] < P.bound && vo[a] < 0.f) || (idx[a] > nw - P.bound && vo[a] > 0.f);
    if (P.sticky_floor && j < P.bound) mask[a] = true;
    if (mask[a]) vo[a] = 0.f;
  }
}

// reference-layout grid_op (separate mv [n^3,3], m [n^3], f [n^3,3], occupy)
__global__ __launch_bounds__(kBlock) void k_grid_op(MpmP P, long long nn,
                                                    const float* __restrict__ mv,
                                                    const float* __restrict__ m,
                                                    const float* __restrict__ gf,
                                                    const int32_t* __restrict__ occ,
                                                    float* __restrict__ vout) {
  long long nd = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nn) return;
  float mm = m[nd];
  bool o = occ ? (occ[nd] == 1) : (mm > P.eps);
  if (!o) return;
  float f0 = 0.f, f1 = 0.f, f2 = 0.f;
  if (gf) { f0 = gf[3 * nd]; f1 = gf[3 * nd + 1]; f2 = gf[3 * nd + 2]; }
  float vo[3];
  bool mask[3];
  grid_node_fwd(P, nd, mv[3 * nd], mv[3 * nd + 1], mv[3 * nd + 2], mm, f0, f1, f2, vo, mask);
  vout[3 * nd] = vo[0];
  vout[3 * nd + 1] = vo[1];
  vout[3 * nd + 2] = vo[2];
}

__global__ __launch_bounds__(kBlock) void k_grid_op_adj(MpmP P, long long nn,
                                                        const float* __restrict__ mv,
                                                        const float* __restrict__ m,
                                                        const float* __restrict__ gf,
                                                        const int32_t* __restrict__ occ,
                                                        const float* __restrict__ gvout,
                                                        float* __restrict__ gmv,
                                                        float* __restrict__ gm,
                                                        float* __restrict__ ggf) {
  long long nd = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nn) return;
  float mm = m[nd];
  bool o = occ ? (occ[nd] == 1) : (mm > P.eps);
  float gmm = 0.f;
  if (o) {
    float f0 = 0.f, f1 = 0.f, f2 = 0.f;
    if (gf) { f0 = gf[3 * nd]; f1 = gf[3 * nd + 1]; f2 = gf[3 * nd + 2]; }
    float m0 = mv[3 * nd], m1 = mv[3 * nd + 1], m2 = mv[3 * nd + 2];
    float vo[3];
    bool mask[3];
    grid_node_fwd(P, nd, m0, m1, m2, mm, f0, f1, f2, vo, mask);
    float g[3];
    for (int a = 0; a < 3; ++a) g[a] = mask[a] ? 0.f : gvout[3 * nd + a];
    float inv_m = 1.f / (mm + P.eps);
    gmv[3 * nd] += inv_m * g[0];
    gmv[3 * nd + 1] += inv_m * g[1];
    gmv[3 * nd + 2] += inv_m * g[2];
    if (ggf) {
      ggf[3 * nd] += inv_m * g[0];
      ggf[3 * nd + 1] += inv_m * g[1];
      ggf[3 * nd + 2] += inv_m * g[2];
    }
    gmm = -inv_m * inv_m * (g[0] * (m0 + f0) + g[1] * (m1 + f1) + g[2] * (m2 + f2));
  }
  float r = gm[nd] + gmm;
  if (P.clamp_grad) r = fminf(fmaxf(r, -P.grad_clamp), P.grad_clamp);
  gm[nd] = r;
}

// packed grid_op: grid4 (mv, m) (+ optional contact impulse gf [n^3,3]) ->
// gv4 (v, 0); every node written (0 if empty)
__global__ __launch_bounds__(kBlock) void k_grid_op4(MpmP P, long long nn,
                                                     const float4* __restrict__ g4,
                                                     const float* __restrict__ gf,
                                                     float4* __restrict__ v4) {
  node_loop(P, nn, [&](long long nd) {
    float4 q = g4[nd];
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (q.w > P.eps) {
      float f0 = 0.f, f1 = 0.f, f2 = 0.f;
      if (gf) { f0 = gf[3 * nd]; f1 = gf[3 * nd + 1]; f2 = gf[3 * nd + 2]; }
      float vo[3];
      bool mask[3];
      grid_node_fwd(P, nd, q.x, q.y, q.z, q.w, f0, f1, f2, vo, mask);
      out = make_float4(vo[0], vo[1], vo[2], 0.f);
    }
    v4[nd] = out;
  });
}

// packed grid_op adjoint: gv4 -> ggrid4 (overwritten); optional contact
// impulse gf and its adjoint ggf ([n^3,3], overwritten)
__global__ __launch_bounds__(kBlock) void k_grid_op4_adj(MpmP P, long long nn,
                                                         const float4* __restrict__ g4,
                                                         const float* __restrict__ gf,
                                                         const float4* __restrict__ gv4,
                                                         float4* __restrict__ gg4,
                                                         float* __restrict__ ggf) {
  node_loop(P, nn, [&](long long nd) {
#pragma clang fp contract(off)
    float4 q = g4[nd];
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 outf = make_float4(0.f, 0.f, 0.f, 0.f);
    if (q.w > P.eps) {
      float f0 = 0.f, f1 = 0.f, f2 = 0.f;
      if (gf) { f0 = gf[3 * nd]; f1 = gf[3 * nd + 1]; f2 = gf[3 * nd + 2]; }
      float vo[3];
      bool mask[3];
      grid_node_fwd(P, nd, q.x, q.y, q.z, q.w, f0, f1, f2, vo, mask);
      float4 gi = gv4[nd];
      float g0 = mask[0] ? 0.f : gi.x, g1 = mask[1] ? 0.f : gi.y, g2 = mask[2] ? 0.f : gi.z;
      float inv_m = 1.f / (q.w + P.eps);
      float gmm = -inv_m * inv_m * (g0 * (q.x + f0) + g1 * (q.y + f1) + g2 * (q.z + f2));
      if (P.clamp_grad) gmm = fminf(fmaxf(gmm, -P.grad_clamp), P.grad_clamp);
      out = make_float4(inv_m * g0, inv_m * g1, inv_m * g2, gmm);
      outf = make_float4(inv_m * g0, inv_m * g1, inv_m * g2, 0.f);
    }
    gg4[nd] = out;
    if (ggf) {
      ggf[3 * nd] = outf.x;
      ggf[3 * nd + 1] = outf.y;
      ggf[3 * nd + 2] = outf.z;
    }
  });
}

// zero width floats per node of the node set
__global__ __launch_bounds__(kBlock) void k_zero_nodes(MpmP P, long long nn, float* __restrict__ a,
                                                       int width) {
  node_loop(P, nn, [&](long long nd) {
    for (int c = 0; c < width; ++c) a[nd * width + c] = 0.f;
  });
}

// Up to four node arrays moved together by k_brick_nodes_copy
struct NodeArrays {
  const uint32_t* src[4];
  uint32_t* dst[4];
  int width[4];
  int n;
};
constexpr int kMaxNodeArrays = 4;

// Nodes of the active bricks of P (64 per brick, in list order) between dense
// node arrays and compact ones, width[a] 32-bit words per node of array a:
// kPack copies dense (src) -> compact (dst), else compact (src) -> dense
// (dst).  Only the first cap bricks are copied; a pack of more sets *overflow
// (lane 0 of block 0, a vector store).  Lanes of nodes outside the grid copy
// nothing.
template <bool kPack>
__global__ __launch_bounds__(kBlock) void k_brick_nodes_copy(MpmP P, NodeArrays A, int cap,
                                                             int32_t* __restrict__ overflow) {
  const int n_act = *P.n_alist;
  if (kPack && overflow && blockIdx.x == 0 && threadIdx.x == 0 && n_act > cap) *overflow = 1;
  const int n = n_act < cap ? n_act : cap;
  const int l = threadIdx.x & 63;
  for (int ai = blockIdx.x * 4 + (threadIdx.x >> 6); ai < n; ai += gridDim.x * 4) {
    const long long nd = active_node(P, P.alist[ai]);
    if (nd < 0) continue;
    const long long cn = (long long)ai * 64 + l;
    for (int a = 0; a < A.n; ++a) {
      const int w = A.width[a];
      const uint32_t* __restrict__ src = A.src[a];
      uint32_t* __restrict__ dst = A.dst[a];
      for (int c = 0; c < w; ++c) {
        if (kPack) dst[cn * w + c] = src[nd * w + c];
        else dst[nd * w + c] = src[cn * w + c];
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_zero4(long long nn, float4* __restrict__ a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nn) a[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ----------------------------------------------------------------------------
// active 4x4x4-node bricks: the grid passes of the fused path visit only the
// bricks some particle stencil touches (flags[brick] != 0)
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_mark_bricks(MpmP P, int n, const float* __restrict__ x,
                                                        int32_t* __restrict__ flags) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float xp[3];
  ld_v3(x, p, pl_vec(P.soa, n), xp);
  Stencil s;
  make_stencil(xp, P.inv_dx, P.n_grid, s, env_x0(P, p));
  const int nb = (P.n_grid + 3) >> 2;
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = s.base[a] >> 2;
    hi[a] = (s.base[a] + 2) >> 2;
  }
  for (int i = lo[0]; i <= hi[0]; ++i)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int k = lo[2]; k <= hi[2]; ++k) flags[(i * nb + j) * nb + k] = 1;
}

// node of thread t in the brick-pass launch (4 bricks per 256-thread block);
// returns -1 when the brick is inactive or the node lies outside the grid
__device__ __forceinline__ long long brick_node(const MpmP& P, const int32_t* __restrict__ flags) {
  const int n = P.n_grid, nb = (n + 3) >> 2, nbx = (P.nx + 3) >> 2;
  long long b = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= (long long)nbx * nb * nb || flags[b] == 0) return -1;
  int l = threadIdx.x & 63;
  int bz = (int)(b % nb), by = (int)((b / nb) % nb), bx = (int)(b / ((long long)nb * nb));
  int i = bx * 4 + (l >> 4), j = by * 4 + ((l >> 2) & 3), k = bz * 4 + (l & 3);
  if (i >= P.nx || j >= n || k >= n) return -1;
  return ((long long)i * n + j) * n + k;
}

__global__ __launch_bounds__(kBlock) void k_zero4_bricks(MpmP P, const int32_t* __restrict__ flags,
                                                         float4* __restrict__ a) {
  long long nd = brick_node(P, flags);
  if (nd >= 0) a[nd] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(kBlock) void k_grid_op4_bricks(MpmP P,
                                                            const int32_t* __restrict__ flags,
                                                            const float4* __restrict__ g4,
                                                            const float* __restrict__ gf,
                                                            float4* __restrict__ v4) {
  long long nd = brick_node(P, flags);
  if (nd < 0) return;
  float4 q = g4[nd];
  float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
  if (q.w > P.eps) {
    float f0 = 0.f, f1 = 0.f, f2 = 0.f;
    if (gf) { f0 = gf[3 * nd]; f1 = gf[3 * nd + 1]; f2 = gf[3 * nd + 2]; }
    float vo[3];
    bool mask[3];
    grid_node_fwd(P, nd, q.x, q.y, q.z, q.w, f0, f1, f2, vo, mask);
    out = make_float4(vo[0], vo[1], vo[2], 0.f);
  }
  v4[nd] = out;
}

__global__ __launch_bounds__(kBlock) void k_grid_op4_adj_bricks(MpmP P,
                                                                const int32_t* __restrict__ flags,
                                                                const float4* __restrict__ g4,
                                                                const float4* __restrict__ gv4,
                                                                float4* __restrict__ gg4) {
  long long nd = brick_node(P, flags);
  if (nd < 0) return;
  float4 q = g4[nd];
  float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
  if (q.w > P.eps) {
    float vo[3];
    bool mask[3];
    grid_node_fwd(P, nd, q.x, q.y, q.z, q.w, 0.f, 0.f, 0.f, vo, mask);
    float4 gi = gv4[nd];
    float g0 = mask[0] ? 0.f : gi.x, g1 = mask[1] ? 0.f : gi.y, g2 = mask[2] ? 0.f : gi.z;
    float inv_m = 1.f / (q.w + P.eps);
    float gmm = -inv_m * inv_m * (g0 * q.x + g1 * q.y + g2 * q.z);
    if (P.clamp_grad) gmm = fminf(fmaxf(gmm, -P.grad_clamp), P.grad_clamp);
    out = make_float4(inv_m * g0, inv_m * g1, inv_m * g2, gmm);
  }
  gg4[nd] = out;
}

// Node value from the brick tiles: node (brick bk, local l) receives the tile
// of brick bk - d for every d in {0,1}^3 with l + 4 d < 6 whose brick holds
// particles, summed in the order of d.  bk is one brick per wavefront: the
// occupancy of the up to eight bricks whose tiles reach into it is read once
// per wavefront (scalar loads), and a lane's tile loads are issued together.
__device__ __forceinline__ float4 tile_sum(const MpmP& P, int bk,
                                           const float4* __restrict__ tiles,
                                           const int32_t* __restrict__ start) {
  const int nb = (P.n_grid + 3) >> 2;
  const int bz = bk % nb, by = (bk / nb) % nb, bx = bk / (nb * nb);
  const int l = threadIdx.x & 63;
  const int l0 = l >> 4, l1 = (l >> 2) & 3, l2 = l & 3;
  float4 q[8];
  bool use[8];
#pragma unroll
  for (int d = 0; d < 8; ++d) {
    const int d0 = d >> 2, d1 = (d >> 1) & 1, d2 = d & 1;
    const int B0 = bx - d0, B1 = by - d1, B2 = bz - d2;
    long long B = 0;
    bool occ = B0 >= 0 && B1 >= 0 && B2 >= 0;
    if (occ) {
      B = ((long long)B0 * nb + B1) * nb + B2;
      occ = start[B] != start[B + 1];
    }
    use[d] = occ && (d0 == 0 || l0 < 2) && (d1 == 0 || l1 < 2) && (d2 == 0 || l2 < 2);
    q[d] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (use[d])
      q[d] = tiles[B * kTileNodes + ((l0 + 4 * d0) * kTileEdge + (l1 + 4 * d1)) * kTileEdge +
                   (l2 + 4 * d2)];
  }
  float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int d = 0; d < 8; ++d)
    if (use[d]) {
      r.x += q[d].x;
      r.y += q[d].y;
      r.z += q[d].z;
      r.w += q[d].w;
    }
  return r;
}

// Grid passes over the active-brick list: a persistent grid, 4 bricks (one per
// wavefront) per block and iteration; bk is the wavefront's brick.
#define DT_ACTIVE_NODES(nd)                                                          \
  const int n_act = *n_alist;                                                        \
  for (int ai = blockIdx.x * 4 + (threadIdx.x >> 6); ai < n_act; ai += gridDim.x * 4) \
    for (int bk = __builtin_amdgcn_readfirstlane(alist[ai]); bk >= 0; bk = -1)       \
      for (long long nd = active_node(P, bk); nd >= 0; nd = -1)

// node values of the active bricks summed from the tiles (no grid_op)
__global__ __launch_bounds__(kBlock) void k_tiles_sum(MpmP P, const int32_t* __restrict__ alist,
                                                      const int32_t* __restrict__ n_alist,
                                                      const float4* __restrict__ tiles,
                                                      const int32_t* __restrict__ start,
                                                      float4* __restrict__ out4) {
  DT_ACTIVE_NODES(nd) { out4[nd] = tile_sum(P, bk, tiles, start); }
}

// the same, also clearing width floats per node of zbuf on the same nodes
__global__ __launch_bounds__(kBlock) void k_tiles_sum_zero(MpmP P, const int32_t* __restrict__ alist,
                                                           const int32_t* __restrict__ n_alist,
                                                           const float4* __restrict__ tiles,
                                                           const int32_t* __restrict__ start,
                                                           float4* __restrict__ out4,
                                                           float* __restrict__ zbuf, int width) {
  DT_ACTIVE_NODES(nd) {
    out4[nd] = tile_sum(P, bk, tiles, start);
    for (int c = 0; c < width; ++c) zbuf[nd * width + c] = 0.f;
  }
}

// grid_op over the active bricks with (mv, m) summed from the p2g tiles;
// writes the summed grid4 and grid_v4
__global__ __launch_bounds__(kBlock) void k_grid_op4_tiles(MpmP P, const int32_t* __restrict__ alist,
                                                           const int32_t* __restrict__ n_alist,
                                                           const float4* __restrict__ tiles,
                                                           const int32_t* __restrict__ start,
                                                           float4* __restrict__ g4,
                                                           float4* __restrict__ v4) {
  DT_ACTIVE_NODES(nd) {
    const float4 q = tile_sum(P, bk, tiles, start);
    g4[nd] = q;
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (q.w > P.eps) {
      float vo[3];
      bool mask[3];
      grid_node_fwd(P, nd, q.x, q.y, q.z, q.w, 0.f, 0.f, 0.f, vo, mask);
      out = make_float4(vo[0], vo[1], vo[2], 0.f);
    }
    v4[nd] = out;
  }
}

// grid_op adjoint over the active bricks with d/d grid_v_out summed from the
// g2p-adjoint tiles; writes the summed gv4 and gg4
__global__ __launch_bounds__(kBlock) void k_grid_op4_adj_tiles(MpmP P,
                                                               const int32_t* __restrict__ alist,
                                                               const int32_t* __restrict__ n_alist,
                                                               const float4* __restrict__ g4,
                                                               const float4* __restrict__ tiles,
                                                               const int32_t* __restrict__ start,
                                                               float4* __restrict__ gv4,
                                                               float4* __restrict__ gg4) {
  DT_ACTIVE_NODES(nd) {
#pragma clang fp contract(off)
    const float4 gi = tile_sum(P, bk, tiles, start);
    gv4[nd] = gi;
    const float4 q = g4[nd];
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (q.w > P.eps) {
      float vo[3];
      bool mask[3];
      grid_node_fwd(P, nd, q.x, q.y, q.z, q.w, 0.f, 0.f, 0.f, vo, mask);
      float g0 = mask[0] ? 0.f : gi.x, g1 = mask[1] ? 0.f : gi.y, g2 = mask[2] ? 0.f : gi.z;
      float inv_m = 1.f / (q.w + P.eps);
      float gmm = -inv_m * inv_m * (g0 * q.x + g1 * q.y + g2 * q.z);
      if (P.clamp_grad) gmm = fminf(fmaxf(gmm, -P.grad_clamp), P.grad_clamp);
      out = make_float4(inv_m * g0, inv_m * g1, inv_m * g2, gmm);
    }
    gg4[nd] = out;
  }
}

// The same with the contact force grid gf (as dt_mpm_grid_op4_adj): writes
// gg4 and ggf (d/d gf), and the summed gv4 when it is non-NULL.
__global__ __launch_bounds__(kBlock) void k_grid_op4_adj_tiles_f(
    MpmP P, const int32_t* __restrict__ alist, const int32_t* __restrict__ n_alist,
    const float4* __restrict__ g4, const float* __restrict__ gf, const float4* __restrict__ tiles,
    const int32_t* __restrict__ start, float4* __restrict__ gv4, float4* __restrict__ gg4,
    float* __restrict__ ggf) {
  DT_ACTIVE_NODES(nd) {
#pragma clang fp contract(off)
    const float4 gi = tile_sum(P, bk, tiles, start);
    if (gv4) gv4[nd] = gi;
    const float4 q = g4[nd];
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    float of0 = 0.f, of1 = 0.f, of2 = 0.f;
    if (q.w > P.eps) {
      const float f0 = gf[3 * nd], f1 = gf[3 * nd + 1], f2 = gf[3 * nd + 2];
      float vo[3];
      bool mask[3];
      grid_node_fwd(P, nd, q.x, q.y, q.z, q.w, f0, f1, f2, vo, mask);
      float g0 = mask[0] ? 0.f : gi.x, g1 = mask[1] ? 0.f : gi.y, g2 = mask[2] ? 0.f : gi.z;
      float inv_m = 1.f / (q.w + P.eps);
      float gmm = -inv_m * inv_m * (g0 * (q.x + f0) + g1 * (q.y + f1) + g2 * (q.z + f2));
      if (P.clamp_grad) gmm = fminf(fmaxf(gmm, -P.grad_clamp), P.grad_clamp);
      out = make_float4(inv_m * g0, inv_m * g1, inv_m * g2, gmm);
      of0 = inv_m * g0;
      of1 = inv_m * g1;
      of2 = inv_m * g2;
    }
    gg4[nd] = out;
    ggf[3 * nd] = of0;
    ggf[3 * nd + 1] = of1;
    ggf[3 * nd + 2] = of2;
  }
}

#undef DT_ACTIVE_NODES

__global__ __launch_bounds__(kBlock) void k_zero_i32(long long n, int32_t* __restrict__ a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// ----------------------------------------------------------------------------
// g2p and its adjoint
// ----------------------------------------------------------------------------
// new v = sum_o w_o gv_o, new C = 4 inv_dx sum_o w_o gv_o (o - fx)^T
template <class Ref>
__device__ __forceinline__ void g2p_core(const MpmP& P, const Stencil& s, const Ref& g,
                                         float (&nv)[3], M3& nC) {
  const float zb[3] = {0.f, 0.f, 0.f};
  float GO[3][3], gfx[3];
  stencil_gather<false, true>(s, g, zb, m3_zero(), 0.f, nv, GO, gfx);
  const float k4 = 4.f * P.inv_dx;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) nC.a[r * 3 + q] = k4 * (GO[q][r] - nv[r] * s.fx[q]);
}

// kCount: also makes the brick counts of the next binning from x1 (the key,
// rank and count arrays k_bin_count would write)
template <bool kCount>
__global__ __launch_bounds__(kBlock) void k_g2p(MpmP P, int n, const float* __restrict__ x,
                                                const float* __restrict__ gv, int sgv,
                                                float* __restrict__ x1, float* __restrict__ v1,
                                                float* __restrict__ C1, int32_t* __restrict__ bkey,
                                                int32_t* __restrict__ brank,
                                                int32_t* __restrict__ bcount, BinRuns R) {
  const int p = xcd_block() * blockDim.x + threadIdx.x;
  const bool active = p < n;
  if (!kCount && !active) return;
  int k = -1;
  if (active) {
    const MpmP Q = at_particle(P, p);
    PL Lv = pl_vec(P.soa, n), Lm = pl_mat(P.soa, n);
    Stencil s;
    float xp[3];
    ld_v3(x, p, Lv, xp);
    make_stencil(xp, Q.inv_dx, P.n_grid, s, env_x0(P, p));
    float nv[3];
    M3 nC;
    if (sgv == 4)
      g2p_core(Q, s, Nodes4{(const float4*)gv, P.n_grid}.at(s.base), nv, nC);
    else
      g2p_core(Q, s, NodesStrided<false>{gv, nullptr, sgv, 0, P.n_grid}.at(s.base), nv, nC);
    float nx[3];
    for (int a = 0; a < 3; ++a) nx[a] = xp[a] + P.dt * nv[a];
    st_v3(v1, p, Lv, nv);
    st_v3(x1, p, Lv, nx);
    st_m3L(C1, p, Lm, nC);
    if (kCount) k = bin_key(nx, Q.inv_dx, P.n_grid, env_x0(P, p)) >> 6;
  }
  if (kCount) {
    const int r = wave_key_rank_runs(active, k, bcount, R, p >> 6);
    if (active) {
      bkey[p] = k;
      brank[p] = r;
    }
  }
}

// g2p adjoint of particle p: gathers the grid-velocity adjoint into gx
// (kGxAssign 1: gx =, 0: gx +=) and returns the stencil and the scatter
// coefficients of d/d grid_v_out[node] += w_o (gvt + 4 inv_dx gC (o - fx))
// = w_o (b + M o)
// Grid-velocity sources of the g2p adjoint: the global grid (Nodes4 or
// NodesStrided), or the LDS copy of a brick's 6x6x6-node tile.
struct NodeRefTile {
  static constexpr bool kBatch = false;
  const float4* p;
  __device__ __forceinline__ float4 operator()(int i, int j, int k) const {
    return p[(i * kTileEdge + j) * kTileEdge + k];
  }
};
struct GvTile {
  const float4* tile;
  int o0, o1, o2;
  __device__ __forceinline__ NodeRefTile at(const int (&b)[3]) const {
    return NodeRefTile{tile + ((b[0] - o0) * kTileEdge + (b[1] - o1)) * kTileEdge + (b[2] - o2)};
  }
};

// inputs of the g2p adjoint of one particle
struct G2PAdjIn {
  float x[3], g1[3], v1[3];
  M3 gc;
  int env;   // environment (0 unless stacked)
};
__device__ __forceinline__ void g2p_adj_load(const MpmP& P, int n, const float* x, const float* gx1,
                                             const float* gv1, const float* gC1, int p,
                                             G2PAdjIn& in) {
  PL Lv = pl_vec(P.soa, n), Lm = pl_mat(P.soa, n);
  ld_v3(x, p, Lv, in.x);
  ld_v3(gx1, p, Lv, in.g1);
  ld_v3(gv1, p, Lv, in.v1);
  in.gc = ld_m3L(gC1, p, Lm);
  in.env = P.env_id ? P.env_id[p] : 0;
}

// P: the view of the parameters for the particle's environment, x0 the
// environment's first x-node
template <int kGxAssign, class Gv>
__device__ __forceinline__ void g2p_adj_compute_at(const MpmP& P, int x0, int n,
                                                   const G2PAdjIn& in, const Gv& gvsrc, float* gx,
                                                   int p, Stencil& s, float (&b)[3], M3& M) {
  PL Lv = pl_vec(P.soa, n);
  float gvt[3], lgx[3];
  make_stencil(in.x, P.inv_dx, P.n_grid, s, x0);
  for (int a = 0; a < 3; ++a) {
    gvt[a] = in.v1[a] + P.dt * in.g1[a];
    lgx[a] = in.g1[a];
  }
  const M3& gc = in.gc;
  const float k4 = 4.f * P.inv_dx;
  for (int c = 0; c < 3; ++c) {
    b[c] = gvt[c] - k4 * (gc.a[3 * c] * s.fx[0] + gc.a[3 * c + 1] * s.fx[1] +
                          gc.a[3 * c + 2] * s.fx[2]);
    for (int a = 0; a < 3; ++a) M.a[3 * c + a] = k4 * gc.a[3 * c + a];
  }
  // d/d w_o of w_o gv_o . (gvt + 4 inv_dx gC (o - fx)) = gv_o . (b + M o)
  float G[3], GO[3][3], gfx[3];
  stencil_gather<true, false>(s, gvsrc.at(s.base), b, M, 0.f, G, GO, gfx);
  // fx = x inv_dx - base; d = o - fx: d/dx through d is -M^T G
  for (int a = 0; a < 3; ++a) {
    const float gd = M.a[a] * G[0] + M.a[3 + a] * G[1] + M.a[6 + a] * G[2];
    lgx[a] += P.inv_dx * gfx[a] - P.inv_dx * gd;
  }
  for (int a = 0; a < 3; ++a) {
    long long o = p * Lv.sp + a * Lv.sc;
    if (kGxAssign) gx[o] = lgx[a];
    else gx[o] += lgx[a];
  }
}

template <int kGxAssign, class Gv>
__device__ __forceinline__ void g2p_adj_compute(const MpmP& P, int n, const G2PAdjIn& in,
                                                const Gv& gvsrc, float* gx, int p, Stencil& s,
                                                float (&b)[3], M3& M) {
  g2p_adj_compute_at<kGxAssign>(at_particle(P, p), in.env * P.n_grid, n, in, gvsrc, gx, p, s, b,
                                M);
}

template <int kGxAssign, class Gv>
__device__ __forceinline__ void g2p_adj_particle(const MpmP& P, int n, const float* x,
                                                 const Gv& gvsrc, const float* gx1,
                                                 const float* gv1, const float* gC1, float* gx,
                                                 int p, Stencil& s, float (&b)[3], M3& M) {
  G2PAdjIn in;
  g2p_adj_load(P, n, x, gx1, gv1, gC1, p, in);
  g2p_adj_compute<kGxAssign>(P, n, in, gvsrc, gx, p, s, b, M);
}

// g2p adjoint: gather gx, scatter to the grid-velocity adjoint by runs
template <int kGxAssign>
__global__ __launch_bounds__(kBlock) void k_g2p_adj(MpmP P, int n, const float* __restrict__ x,
                                                    const float* __restrict__ gv, int sgv,
                                                    const float* __restrict__ gx1,
                                                    const float* __restrict__ gv1,
                                                    const float* __restrict__ gC1,
                                                    float* __restrict__ gx,
                                                    float* __restrict__ ggv, int sggv) {
  __shared__ RunStage sh;
  int p = xcd_block() * blockDim.x + threadIdx.x;
  bool active = p < n;
  Stencil s;
  s.base[0] = s.base[1] = s.base[2] = 0;
  float b[3] = {0.f, 0.f, 0.f};
  M3 M = m3_zero();
  if (active) {
    if (sgv == 4)
      g2p_adj_particle<kGxAssign>(P, n, x, Nodes4{(const float4*)gv, P.n_grid}, gx1, gv1, gC1, gx,
                                  p, s, b, M);
    else
      g2p_adj_particle<kGxAssign>(P, n, x, NodesStrided<false>{gv, nullptr, sgv, 0, P.n_grid}, gx1,
                                  gv1, gC1, gx, p, s, b, M);
  }
  const int R = stage_runs(sh, active, block_active(n), s, P.n_grid, b, M);
  Grid g{ggv, nullptr, sggv, 0};
  run_scatter<false>(sh, R, P.n_grid, g, 0.f);
}

// ----------------------------------------------------------------------------
// brick-owned scatter over binned particles: block b processes the particles
// of brick b in chunks of 256 and owns the 6x6x6-node tile of partial sums its
// stencils reach.
//   per chunk  1. thread t computes particle t; a stable LDS counting sort by
//                 cell (64 bins, per-wavefront counts) gives its staging slot,
//                 where it stages (w, b, M);
//              2. work item (cy, cz, j, k) owns the x-row of six tile nodes
//                 (x, cy + j, cz + k): it sums, over the particles of the four
//                 cells (cx, cy, cz), cx = 0..3, the contributions
//                 w_o (b + M o) to the nodes x = cx + i, i = 0..2, and adds
//                 the row to its own LDS slots (the nine lanes of one (cy, cz)
//                 read the same staged particle: LDS broadcast);
//   per brick  3. tile node thread t < 216 adds the at most nine row slots
//                 that can name it.
// The row slots (864 float4) keep the block within 40 KB of LDS, so four
// blocks share a CU.
// No atomics anywhere; the tiles are summed per node by the grid passes below.
// ----------------------------------------------------------------------------
constexpr int kRowItems = 16 * 9;                  // (cy, cz, j, k)
constexpr int kSlots = kRowItems * kTileEdge;      // (cy, cz, j, k, x)
struct BrickStage {
  // particle q at row r = stage_row(q): pd[r+0] = (wx0, wx1, wx2, b0),
  // pd[r+1] = (wy0.., b1), pd[r+2] = (wz0.., b2), pd[r+3..5] = M row-major
  // (3 floats of padding).
  // Before the staging of a chunk the same LDS holds the brick's 6x6x6-node
  // grid-velocity tile read by the g2p adjoint's gather (gtile).
  union {
    float4 pd[6 * kBlock + kBlock / 8];
    float4 gtile[kTileNodes];
  };
  float4 slot[kSlots + 1];   // (cy, cz, j, k, x): (sum w u0, w u1, w u2, w); last stays 0
  uint16_t cs[66];           // cell starts of the chunk (staging slots)
  uint16_t hist[kBlock / 64][64]; // per wavefront: particles per cell of the chunk,
                                  //   then its first staging slot in each cell
};

// First float4 of staged particle q.  Six float4 per particle put particles
// q and q + 8 on the same LDS banks; the lanes of one wavefront in
// brick_items read the i-th particle of several cells at once, and cells
// hold ~8 particles, so one float4 of skew per 8 particles spreads those
// reads over the banks.
__device__ __forceinline__ int stage_row(int q) { return 6 * q + (q >> 3); }

__device__ __forceinline__ int local_cell(const Stencil& s) {
  return ((s.base[0] & 3) * 4 + (s.base[1] & 3)) * 4 + (s.base[2] & 3);
}

// wavefront 0: cell starts (exclusive scan of the per-cell totals), each
// wavefront's first slot in each cell (its particles follow those of the
// wavefronts before it: the staging order is the chunk order, cell by cell,
// so it is the same on every run)
__device__ __forceinline__ void chunk_scan(BrickStage& sh) {
  constexpr int kW = kBlock / 64;
  const int l = threadIdx.x;   // 0..63, one full wavefront
  int hw[kW], h = 0;
#pragma unroll
  for (int w = 0; w < kW; ++w) {
    hw[w] = sh.hist[w][l];
    h += hw[w];
  }
  int incl = h;
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(incl, d);
    if (l >= d) incl += u;
  }
  int at = incl - h;
  sh.cs[l] = (uint16_t)at;
  if (l == 63) sh.cs[64] = (uint16_t)incl;
#pragma unroll
  for (int w = 0; w < kW; ++w) {
    sh.hist[w][l] = (uint16_t)at;
    at += hw[w];
  }
}

__device__ __forceinline__ void stage_particle(BrickStage& sh, int t, const Stencil& s,
                                               const float (&b)[3], const M3& M) {
  float4* q = sh.pd + stage_row(t);
  q[0] = make_float4(s.w[0][0], s.w[1][0], s.w[2][0], b[0]);
  q[1] = make_float4(s.w[0][1], s.w[1][1], s.w[2][1], b[1]);
  q[2] = make_float4(s.w[0][2], s.w[1][2], s.w[2][2], b[2]);
  q[3] = make_float4(M.a[0], M.a[1], M.a[2], M.a[3]);
  q[4] = make_float4(M.a[4], M.a[5], M.a[6], M.a[7]);
  q[5] = make_float4(M.a[8], 0.f, 0.f, 0.f);
}

// step 2: x-row items (cy, cz, j, k) over the chunk's particles; thread t < 144
// takes item t.  kMass: also sum the weights (the mass channel); the g2p
// adjoint has none.
template <bool kMass>
__device__ __forceinline__ void brick_items(BrickStage& sh) {
  const int it = threadIdx.x;
  if (it >= kRowItems) return;
  const int yz = it / 9, jk = it - 9 * yz;
  const int cy = yz >> 2, cz = yz & 3;
  const int j = jk / 3, k = jk - 3 * j;
  const float fj = (float)j, fk = (float)k;
  const float* pf = reinterpret_cast<const float*>(sh.pd);
  float4 acc[kTileEdge];
#pragma unroll
  for (int x = 0; x < kTileEdge; ++x) acc[x] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int cx = 0; cx < 4; ++cx) {
    const int cell = (cx * 4 + cy) * 4 + cz;
    const int q0 = sh.cs[cell], q1 = sh.cs[cell + 1];
    for (int q = q0; q < q1; ++q) {
      const float4* pq = sh.pd + stage_row(q);
      const float* fq = pf + 4 * stage_row(q);
      const float4 X = pq[0], M0 = pq[3], M1 = pq[4];
      const float wy = fq[4 + j], b1 = fq[7], wz = fq[8 + k], b2 = fq[11], m22 = fq[20];
      const float wyz = wy * wz;
      // u = b + M (0, j, k); node (cx + i, j, k) adds w_i (u + M (i, 0, 0))
      const float u0 = X.w + M0.y * fj + M0.z * fk;
      const float u1 = b1 + M1.x * fj + M1.y * fk;
      const float u2 = b2 + M1.w * fj + m22 * fk;
      const float m00 = M0.x, m10 = M0.w, m20 = M1.z;
      const float w0 = X.x * wyz, w1 = X.y * wyz, w2 = X.z * wyz;
      float4& a0 = acc[cx];
      float4& a1 = acc[cx + 1];
      float4& a2 = acc[cx + 2];
      a0.x += w0 * u0;
      a0.y += w0 * u1;
      a0.z += w0 * u2;
      a1.x += w1 * (u0 + m00);
      a1.y += w1 * (u1 + m10);
      a1.z += w1 * (u2 + m20);
      a2.x += w2 * (u0 + 2.f * m00);
      a2.y += w2 * (u1 + 2.f * m10);
      a2.z += w2 * (u2 + 2.f * m20);
      if (kMass) {
        a0.w += w0;
        a1.w += w1;
        a2.w += w2;
      }
    }
  }
  float4* sl = sh.slot + it * kTileEdge;
#pragma unroll
  for (int x = 0; x < kTileEdge; ++x) {
    const float4 o = sl[x];
    sl[x] = make_float4(o.x + acc[x].x, o.y + acc[x].y, o.z + acc[x].z, o.w + acc[x].w);
  }
}

// step 3: tile node thread t adds the row slots of the items that reach it.
// The thread index goes through an empty asm so its nine slot indices are
// computed here each time, not hoisted out of the brick loop and held (or
// spilled) across the whole kernel.
__device__ __forceinline__ float4 brick_gather(const BrickStage& sh) {
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  if (t >= kTileNodes) return acc;
  const int a0 = t / (kTileEdge * kTileEdge), a1 = (t / kTileEdge) % kTileEdge,
            a2 = t % kTileEdge;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int cy = a1 - j, cz = a2 - k;
      const bool in = (unsigned)cy < 4u && (unsigned)cz < 4u;
      const int idx = in ? ((cy * 4 + cz) * 9 + j * 3 + k) * kTileEdge + a0 : kSlots;
      const float4 v = sh.slot[idx];
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
  return acc;
}

__device__ __forceinline__ float uniform_f(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

// the view of P for brick b's environment: every particle of a brick lives in
// one environment (the slots of a ragged batch are multiples of 4 nodes).
// The environment is uniform over the block, so its fields are kept in scalar
// registers.
__device__ __forceinline__ MpmP at_brick(const MpmP& P, long long b) {
  if (!P.env_scale) return P;
  const int nb = (P.n_grid + 3) >> 2;
  const int e = (int)(b / ((long long)nb * nb)) * 4 / P.n_grid;
  MpmP Q = at_env(P, __builtin_amdgcn_readfirstlane(e));
  Q.dx = uniform_f(Q.dx);
  Q.inv_dx = uniform_f(Q.inv_dx);
  Q.p_mass = uniform_f(Q.p_mass);
  Q.p_vol = uniform_f(Q.p_vol);
  Q.n_wall = __builtin_amdgcn_readfirstlane(Q.n_wall);
  return Q;
}

// Blocks stride over the occupied-brick list (a persistent grid: unoccupied
// bricks cost nothing), chunk by chunk, software-pipelined: while chunk k is
// sorted and scattered, the inputs of chunk k + 1 (the next of the brick, or
// the first of the block's next brick) are in flight, and perm of chunk k + 1
// is read at the start of chunk k so those loads need not wait for it.
//   load(p, in)                 issues the loads of particle p's inputs
//   compute(in, p, Q, b, s, bb, M)  particle p of brick b (Q: the parameters
//                               of its environment, read once per brick):
//                               stencil and (b, M)
//   mass_of(Q)                  the particle mass of a brick's tile
//   fetch(b)                    (kTile) node t of brick b's 6x6x6 grid tile,
//                               stored to sh.gtile at every chunk start of b
//                               (the staging of a chunk overwrites it)
template <bool kTile, class In, class MassOf, class Load, class Compute, class Fetch>
__device__ __forceinline__ void brick_loop(const MpmP& P, BrickStage& sh,
                                           const int32_t* __restrict__ perm,
                                           const int32_t* __restrict__ start,
                                           const int32_t* __restrict__ list,
                                           const int32_t* __restrict__ n_list,
                                           float4* __restrict__ tiles, const MassOf& mass_of,
                                           const Load& load, const Compute& compute,
                                           const Fetch& fetch) {
  const int t = threadIdx.x;
  const int n_occ = *n_list;
  int li = blockIdx.x;
  if (li >= n_occ) return;
  long long b = list[li];
  int s0 = start[b], s1 = start[b + 1];
  In in;
  int p = 0;
  float4 tq = make_float4(0.f, 0.f, 0.f, 0.f);
  if (kTile) tq = fetch(b);
  if (s0 + t < min(s0 + (int)kBlock, s1)) {
    p = perm[s0 + t];
    load(p, in);
  }
  for (;;) {
    const int li2 = li + gridDim.x;
    long long b2 = -1;
    int s02 = 0, s12 = 0;
    if (li2 < n_occ) {
      b2 = list[li2];
      s02 = start[b2];
      s12 = start[b2 + 1];
    }
    const MpmP Q = at_brick(P, b);
    for (int l = t; l < kSlots + 1; l += kBlock) sh.slot[l] = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 tq2 = tq;
    for (int c0 = s0; c0 < s1; c0 += kBlock) {
      const int c1 = min(c0 + (int)kBlock, s1);
      const int i = c0 + t;
      // the chunk after this one
      const bool last = c1 >= s1;
      const int nc0 = last ? s02 : c1;
      const int nc1 = last ? min(s02 + (int)kBlock, s12) : min(c1 + (int)kBlock, s1);
      const bool has_next = nc0 + t < nc1;
      int pn = 0;
      if (has_next) pn = perm[nc0 + t];
      if (kTile) {
        if (t < kTileNodes) sh.gtile[t] = tq;
        if (last && b2 >= 0) tq2 = fetch(b2);
      }
      sh.hist[t >> 6][t & 63] = 0;
      __syncthreads();
      Stencil s;
      float bb[3];
      M3 M;
      int lc = 0;
      if (i < c1) {
        compute(in, p, Q, b, s, bb, M);
        lc = local_cell(s);
      }
      // rank among this wavefront's particles of the same cell; the first of
      // them writes the wavefront's count of that cell (no atomics)
      const unsigned long long peers = wave_match6(i < c1, lc);
      const int wv = t >> 6;
      int r = 0;
      if (i < c1) {
        r = __popcll(peers & ((1ull << (t & 63)) - 1ull));
        if (r == 0) sh.hist[wv][lc] = (uint16_t)__popcll(peers);
      }
      if (has_next) {
        p = pn;
        load(p, in);
      }
      __syncthreads();
      if (t < 64) chunk_scan(sh);
      __syncthreads();
      if (i < c1) stage_particle(sh, sh.hist[wv][lc] + r, s, bb, M);
      __syncthreads();
      brick_items<!kTile>(sh);
      __syncthreads();
    }
    float4 acc = brick_gather(sh);
    if (t < kTileNodes) {
      acc.w *= mass_of(Q);
      tiles[b * kTileNodes + t] = acc;
    }
    __syncthreads();
    if (b2 < 0) break;
    li = li2;
    b = b2;
    s0 = s02;
    s1 = s12;
    tq = tq2;
  }
}

__global__ __launch_bounds__(kBlock, 4) void k_p2g_binned(MpmP P, const dt_material* __restrict__ mats,
                                                       int n, const int32_t* __restrict__ mat_id,
                                                       const float* __restrict__ x,
                                                       const float* __restrict__ v,
                                                       const float* __restrict__ C,
                                                       const float* __restrict__ F,
                                                       float* __restrict__ F_next,
                                                       const int32_t* __restrict__ perm,
                                                       const int32_t* __restrict__ start,
                                                       const int32_t* __restrict__ list,
                                                       const int32_t* __restrict__ n_list,
                                                       float4* __restrict__ tiles) {
  __shared__ BrickStage sh;
  brick_loop<false, P2GIn>(
      P, sh, perm, start, list, n_list, tiles, [](const MpmP& Q) { return Q.p_mass; },
      [&](int p, P2GIn& in) { p2g_load(P, n, mat_id, x, v, C, F, p, in); },
      [&](const P2GIn& in, int p, const MpmP& Q, long long, Stencil& s, float(&bb)[3], M3& M) {
        p2g_compute_at(Q, in.env * P.n_grid, mats, n, in, F_next, p, s, bb, M);
      },
      [](long long) { return make_float4(0.f, 0.f, 0.f, 0.f); });
}

// node t of the 6x6x6-node tile of brick b (zero outside the grid and for
// t >= 216)
__device__ __forceinline__ float4 fetch_gv_tile(const MpmP& P, long long b, const float* gv,
                                                int sgv) {
  const int n = P.n_grid, nb = (n + 3) >> 2;
  const int bz = (int)(b % nb), by = (int)((b / nb) % nb), bx = (int)(b / ((long long)nb * nb));
  const int t = threadIdx.x;
  float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
  if (t < kTileNodes) {
    const int i = bx * 4 + t / (kTileEdge * kTileEdge), j = by * 4 + (t / kTileEdge) % kTileEdge,
              k = bz * 4 + t % kTileEdge;
    if (i < P.nx && j < n && k < n) {
      const long long nd = node_index(i, j, k, n);
      q = make_float4(gv[nd * sgv], gv[nd * sgv + 1], gv[nd * sgv + 2], 0.f);
    }
  }
  return q;
}

// node t of brick b's 6x6x6-node tile of grid velocities made from the kept
// (mv, m) grid g4 by the node update of the forward pass (no contact force):
// the same function of the same inputs as k_grid_op4_tiles, so the same bits
__device__ __forceinline__ float4 fetch_gv_tile_op(const MpmP& P, long long b,
                                                   const float4* __restrict__ g4) {
  const int n = P.n_grid, nb = (n + 3) >> 2;
  const int bz = (int)(b % nb), by = (int)((b / nb) % nb), bx = (int)(b / ((long long)nb * nb));
  const int t = threadIdx.x;
  float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
  if (t < kTileNodes) {
    const int i = bx * 4 + t / (kTileEdge * kTileEdge), j = by * 4 + (t / kTileEdge) % kTileEdge,
              k = bz * 4 + t % kTileEdge;
    if (i < P.nx && j < n && k < n) {
      const long long nd = node_index(i, j, k, n);
      const float4 g = g4[nd];
      if (g.w > P.eps) {
        float vo[3];
        bool mask[3];
        grid_node_fwd(P, nd, g.x, g.y, g.z, g.w, 0.f, 0.f, 0.f, vo, mask);
        q = make_float4(vo[0], vo[1], vo[2], 0.f);
      }
    }
  }
  return q;
}

// kFromMass: gv is the kept (mv, m) grid (float4 per node) and each brick's
// velocity tile is made from it; else gv is the velocity grid (stride sgv)
template <bool kFromMass>
__global__ __launch_bounds__(kBlock, 4) void k_g2p_adj_binned(MpmP P, int n,
                                                           const float* __restrict__ x,
                                                           const float* __restrict__ gv, int sgv,
                                                           const float* __restrict__ gx1,
                                                           const float* __restrict__ gv1,
                                                           const float* __restrict__ gC1,
                                                           float* __restrict__ gx,
                                                           const int32_t* __restrict__ perm,
                                                           const int32_t* __restrict__ start,
                                                           const int32_t* __restrict__ list,
                                                           const int32_t* __restrict__ n_list,
                                                           float4* __restrict__ tiles) {
  // the brick's 6x6x6 grid-velocity tile sits in LDS (aliasing the staging
  // area): the stencils of its particles read only these nodes
  __shared__ BrickStage sh;
  const int nb = (P.n_grid + 3) >> 2;
  brick_loop<true, G2PAdjIn>(
      P, sh, perm, start, list, n_list, tiles, [](const MpmP&) { return 0.f; },
      [&](int p, G2PAdjIn& in) { g2p_adj_load(P, n, x, gx1, gv1, gC1, p, in); },
      [&](const G2PAdjIn& in, int p, const MpmP& Q, long long b, Stencil& s, float(&bb)[3],
          M3& M) {
        const GvTile src{sh.gtile, (int)(b / ((long long)nb * nb)) * 4, (int)((b / nb) % nb) * 4,
                         (int)(b % nb) * 4};
        g2p_adj_compute_at<1>(Q, in.env * P.n_grid, n, in, src, gx, p, s, bb, M);
      },
      [&](long long b) {
        return kFromMass ? fetch_gv_tile_op(P, b, reinterpret_cast<const float4*>(gv))
                         : fetch_gv_tile(P, b, gv, sgv);
      });
}

// ----------------------------------------------------------------------------
// ordering helpers
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_cell_keys(MpmP P, int n, const float* __restrict__ x,
                                                      int32_t* __restrict__ keys) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float xp[3];
  ld_v3(x, p, pl_vec(P.soa, n), xp);
  keys[p] = cell_key(xp, at_particle(P, p).inv_dx, P.n_grid, env_x0(P, p));
}

__global__ __launch_bounds__(kBlock) void k_gather_rows(long long total, int width,
                                                        const int32_t* __restrict__ perm,
                                                        const float* __restrict__ src,
                                                        float* __restrict__ dst) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  long long i = e / width;
  int c = (int)(e % width);
  dst[e] = src[(long long)perm[i] * width + c];
}

// [n][w] (rows picked by perm, or identity when perm is NULL) -> [w][n]
__global__ __launch_bounds__(kBlock) void k_gather_soa(long long n, int width,
                                                       const int32_t* __restrict__ perm,
                                                       const float* __restrict__ src,
                                                       float* __restrict__ dst) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * width) return;
  int c = (int)(e / n);
  long long i = e % n;
  long long r = perm ? (long long)perm[i] : i;
  dst[e] = src[r * width + c];
}

// [w][n] -> [n][w] rows placed at perm[i] (or i when perm is NULL)
__global__ __launch_bounds__(kBlock) void k_scatter_soa(long long n, int width,
                                                        const int32_t* __restrict__ perm,
                                                        const float* __restrict__ src,
                                                        float* __restrict__ dst) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * width) return;
  int c = (int)(e / n);
  long long i = e % n;
  long long r = perm ? (long long)perm[i] : i;
  dst[r * width + c] = src[e];
}

// [w][n] -> [w][n] reordered: gather dst[c][i] = src[c][perm[i]], or scatter
// dst[c][perm[i]] = src[c][i]
template <bool kScatter>
__global__ __launch_bounds__(kBlock) void k_permute_soa(long long n, int width,
                                                        const int32_t* __restrict__ perm,
                                                        const float* __restrict__ src,
                                                        float* __restrict__ dst) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * width) return;
  const long long c = e / n, i = e % n;
  const long long r = c * n + perm[i];
  if (kScatter)
    dst[r] = src[e];
  else
    dst[e] = src[r];
}

__global__ __launch_bounds__(kBlock) void k_scatter_rows(long long total, int width,
                                                         const int32_t* __restrict__ perm,
                                                         const float* __restrict__ src,
                                                         float* __restrict__ dst) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  long long i = e / width;
  int c = (int)(e % width);
  dst[(long long)perm[i] * width + c] = src[e];
}

// ----------------------------------------------------------------------------
// host-side launchers (used by the C-ABI layer)
// ----------------------------------------------------------------------------
static long long grid_nodes(const dt_mpm_params* prm) {
  const long long e = prm->n_env > 1 ? prm->n_env : 1;
  return e * prm->n_grid * prm->n_grid * prm->n_grid;
}

int validate_params(const dt_mpm_params* p) {
  DT_REQUIRE(p != nullptr, "dt_mpm_params is NULL");
  DT_REQUIRE(p->n_grid >= 3 && p->n_grid <= 1024, "n_grid must be in [3, 1024]");
  DT_REQUIRE(p->n_env >= 0, "n_env must be >= 0");
  DT_REQUIRE((long long)(p->n_env > 1 ? p->n_env : 1) * p->n_grid * p->n_grid * p->n_grid <
                 (1ll << 31),
             "n_env * n_grid^3 must stay below 2^31 nodes");
  DT_REQUIRE(p->dx > 0.f && p->inv_dx > 0.f, "dx and inv_dx must be positive");
  DT_REQUIRE(p->n_env <= 1 || !p->env_scale || !p->env_ngrid || p->n_grid % 4 == 0,
             "ragged environments need a slot grid size that is a multiple of 4");
  return DT_OK;
}

static int check_mats(const dt_material* mats, int n_mats) {
  DT_REQUIRE(mats != nullptr && n_mats > 0, "material table is empty");
  return DT_OK;
}

#define DT_STREAM(s) ((hipStream_t)(s))

static long long n_bricks(const dt_mpm_params* prm) { return brick_total(prm); }

// Marks the bricks touched by the particles at x (flags cleared first).
static void launch_mark(const MpmP& P, const dt_mpm_params* prm, int n, const float* x,
                        int32_t* flags, hipStream_t st) {
  long long nbk = n_bricks(prm);
  hipLaunchKernelGGL(k_zero_i32, dim3(nblocks(nbk)), dim3(kBlock), 0, st, nbk, flags);
  if (n > 0)
    hipLaunchKernelGGL(k_mark_bricks, dim3(nblocks(n)), dim3(kBlock), 0, st, P, n, x, flags);
}

static void launch_p2g_fused_adj(const MpmP& P, hipStream_t st, const dt_material* mats,
                                 int n_mats, int n, const int32_t* mat_id, const float* x,
                                 const float* v, const float* C, const float* F,
                                 const float* ggrid4, const float* gF1, float* gx, float* gv,
                                 float* gC, float* gF, float* g_mats,
                                 float* mat_part = nullptr) {
#define DT_P2G_ADJ(E, G)                                                                       \
  hipLaunchKernelGGL((k_p2g_fused_adj<E, G>), dim3(nblocks(n)), dim3(kBlock), 0, st, P, mats, n, \
                     mat_id, x, v, C, F, ggrid4, gF1, gx, gv, gC, gF, n_mats, g_mats, mat_part)
  if (P.svd_exact) {
    if (g_mats) DT_P2G_ADJ(true, true);
    else DT_P2G_ADJ(true, false);
  } else {
    if (g_mats) DT_P2G_ADJ(false, true);
    else DT_P2G_ADJ(false, false);
  }
#undef DT_P2G_ADJ
}

static unsigned active_blocks(const BinWs& ws) {
  const long long b = (ws.n_bricks + 3) / 4;
  const long long cap = launch_config().node_blocks;
  return (unsigned)(b < cap ? b : cap);
}

static unsigned brick_blocks(const BinWs& ws) {
  const long long cap = launch_config().brick_blocks;
  return (unsigned)(ws.n_bricks < cap ? ws.n_bricks : cap);
}

// bin the particles of x (also sets the active-brick flags)
// (count_clear: an earlier binning in the same launch sequence left the
// workspace's brick counts zero; counted: the previous substep's G2P already
// made the counts of x)
static int launch_mark_bin(const MpmP& P, const dt_mpm_params*, int n, const float* x,
                           const BinWs& ws, hipStream_t st, bool count_clear = false,
                           bool counted = false) {
  return launch_bin(P, n, x, ws, st, count_clear, counted);
}

// One forward substep.  ws == NULL: dense grid passes and the run scatter;
// otherwise active-brick passes and the brick-owned binned scatter.
int launch_substep_fwd(const dt_mpm_params* prm, const dt_material* mats, int n,
                       const int32_t* mat_id, const float* x, const float* v, const float* C,
                       const float* F, float* grid4, float* grid_v4, float* x1, float* v1,
                       float* C1, float* F1, hipStream_t st, const BinWs* ws = nullptr,
                       bool count_clear = false, bool counted = false,
                       bool count_next = false) {
  MpmP P = to_dev(*prm);
  long long nn = grid_nodes(prm);
  if (ws) {
    {
      StageTimer tm(STAGE_BIN, st);
      int rc = launch_mark_bin(P, prm, n, x, *ws, st, count_clear, counted);
      if (rc) return rc;
    }
    {
      StageTimer tm(STAGE_P2G, st);
      hipLaunchKernelGGL(k_p2g_binned, dim3(brick_blocks(*ws)), dim3(kBlock), 0, st, P, mats, n,
                       mat_id, x, v, C, F, F1, (const int32_t*)ws->perm,
                       (const int32_t*)ws->start, (const int32_t*)ws->list,
                       (const int32_t*)ws->n_list, ws->tiles);
    }
    StageTimer tm(STAGE_GRID_OP, st);
    hipLaunchKernelGGL(k_grid_op4_tiles, dim3(active_blocks(*ws)), dim3(kBlock), 0, st, P,
                       (const int32_t*)ws->alist, (const int32_t*)ws->n_alist,
                       (const float4*)ws->tiles, (const int32_t*)ws->start, (float4*)grid4,
                       (float4*)grid_v4);
  } else {
    DT_REQUIRE(!P.env_scale, "ragged environments need the binned path (a workspace)");
    hipLaunchKernelGGL(k_zero4, dim3(nblocks(nn)), dim3(kBlock), 0, st, nn, (float4*)grid4);
    Grid g{grid4, grid4 + 3, 4, 4};
    if (n > 0)
      hipLaunchKernelGGL(k_p2g_fused, dim3(nblocks(n)), dim3(kBlock), 0, st, P, mats, n, mat_id,
                         x, v, C, F, g, F1);
    hipLaunchKernelGGL(k_grid_op4, dim3(nblocks(nn)), dim3(kBlock), 0, st, P, nn,
                       (const float4*)grid4, (const float*)nullptr, (float4*)grid_v4);
  }
  if (n > 0) {
    StageTimer tm(STAGE_G2P, st);
    if (count_next)
      hipLaunchKernelGGL(k_g2p<true>, dim3(nblocks(n)), dim3(kBlock), 0, st, P, n, x,
                         (const float*)grid_v4, 4, x1, v1, C1, ws->key, ws->rank, ws->count,
                         bin_runs(*ws));
    else
      hipLaunchKernelGGL(k_g2p<false>, dim3(nblocks(n)), dim3(kBlock), 0, st, P, n, x,
                         (const float*)grid_v4, 4, x1, v1, C1, (int32_t*)nullptr,
                         (int32_t*)nullptr, (int32_t*)nullptr, BinRuns{});
  }
  return check_launch("dt_mpm_substep_fwd");
}

int launch_substep_adj(const dt_mpm_params* prm, const dt_material* mats, int n_mats, int n,
                       const int32_t* mat_id, const float* x, const float* v, const float* C,
                       const float* F, const float* grid4, const float* grid_v4,
                       const float* gx1, const float* gv1, const float* gC1, const float* gF1,
                       float* ggrid4, float* ggv4, float* gx, float* gv, float* gC, float* gF,
                       float* g_mats, hipStream_t st, const BinWs* ws = nullptr,
                       bool binned = false, bool count_clear = false,
                       float* mat_part = nullptr) {
  MpmP P = to_dev(*prm);
  long long nn = grid_nodes(prm);
  if (ws) {
    if (!binned) {
      StageTimer tm(STAGE_BIN, st);
      int rc = launch_mark_bin(P, prm, n, x, *ws, st, count_clear);
      if (rc) return rc;
    }
    {
      StageTimer tm(STAGE_G2P_ADJ, st);
      // grid_v4 NULL: the velocities come from grid4 (no contact force)
      if (grid_v4)
        hipLaunchKernelGGL(k_g2p_adj_binned<false>, dim3(brick_blocks(*ws)), dim3(kBlock), 0, st,
                           P, n, x, grid_v4, 4, gx1, gv1, gC1, gx, (const int32_t*)ws->perm,
                           (const int32_t*)ws->start, (const int32_t*)ws->list,
                           (const int32_t*)ws->n_list, ws->tiles);
      else
        hipLaunchKernelGGL(k_g2p_adj_binned<true>, dim3(brick_blocks(*ws)), dim3(kBlock), 0, st,
                           P, n, x, grid4, 4, gx1, gv1, gC1, gx, (const int32_t*)ws->perm,
                           (const int32_t*)ws->start, (const int32_t*)ws->list,
                           (const int32_t*)ws->n_list, ws->tiles);
    }
    StageTimer tm(STAGE_GRID_OP_ADJ, st);
    hipLaunchKernelGGL(k_grid_op4_adj_tiles, dim3(active_blocks(*ws)), dim3(kBlock), 0, st, P,
                       (const int32_t*)ws->alist, (const int32_t*)ws->n_alist,
                       (const float4*)grid4, (const float4*)ws->tiles,
                       (const int32_t*)ws->start, (float4*)ggv4, (float4*)ggrid4);
  } else {
    DT_REQUIRE(!P.env_scale, "ragged environments need the binned path (a workspace)");
    hipLaunchKernelGGL(k_zero4, dim3(nblocks(nn)), dim3(kBlock), 0, st, nn, (float4*)ggv4);
    if (n > 0)
      hipLaunchKernelGGL(k_g2p_adj<1>, dim3(nblocks(n)), dim3(kBlock), 0, st, P, n, x, grid_v4, 4,
                         gx1, gv1, gC1, gx, ggv4, 4);
    hipLaunchKernelGGL(k_grid_op4_adj, dim3(nblocks(nn)), dim3(kBlock), 0, st, P, nn,
                       (const float4*)grid4, (const float*)nullptr, (const float4*)ggv4,
                       (float4*)ggrid4, (float*)nullptr);
  }
  if (n > 0) {
    StageTimer tm(STAGE_P2G_ADJ, st);
    launch_p2g_fused_adj(P, st, mats, n_mats, n, mat_id, x, v, C, F, ggrid4, gF1, gx, gv, gC, gF,
                         g_mats, mat_part);
  }
  return check_launch("dt_mpm_substep_adj");
}

// Workspace views for n particles; error when the buffer is missing or short.
// With history_substeps > 0 and room for that many history slots after the
// base layout, the views also carry the binning history.
static int workspace_views(const dt_mpm_params* prm, int n, void* workspace, size_t bytes,
                           BinWs& ws, int history_substeps = 0) {
  const size_t need = bin_ws_layout(prm, n, nullptr, nullptr);
  DT_REQUIRE(workspace != nullptr, "NULL workspace");
  if (bytes < need) {
    set_error("workspace too small: " + std::to_string(bytes) + " bytes given, " +
              std::to_string(need) + " needed (dt_mpm_workspace_bytes)");
    return DT_EINVAL;
  }
  DT_REQUIRE(((uintptr_t)workspace & 255) == 0, "workspace must be 256-byte aligned");
  bin_ws_layout(prm, n, workspace, &ws);
  const size_t slot = bin_history_slot_bytes(prm, n);
  if (history_substeps > 0 && bytes >= need + slot * (size_t)history_substeps) {
    ws.hist = static_cast<char*>(workspace) + need;
    ws.hist_slot_bytes = slot;
    ws.hist_slots = history_substeps;
  }
  return DT_OK;
}

// Which trajectory the binning history of a workspace describes: set by a
// history-keeping dt_mpm_run_fwd, checked by dt_mpm_run_adj.
namespace {
struct HistRecord {
  const void* ws;
  const float* x;
  int n, n_substeps;
};
std::vector<HistRecord> g_hist;
std::mutex g_hist_mu;

void hist_set(const void* ws, const float* x, int n, int k) {
  std::lock_guard<std::mutex> lock(g_hist_mu);
  for (auto& r : g_hist)
    if (r.ws == ws) {
      r = {ws, x, n, k};
      return;
    }
  g_hist.push_back({ws, x, n, k});
}

bool hist_matches(const void* ws, const float* x, int n, int k) {
  std::lock_guard<std::mutex> lock(g_hist_mu);
  for (const auto& r : g_hist)
    if (r.ws == ws) return r.x == x && r.n == n && r.n_substeps == k;
  return false;
}

void hist_clear(const void* ws) {
  std::lock_guard<std::mutex> lock(g_hist_mu);
  for (auto& r : g_hist)
    if (r.ws == ws) r.x = nullptr;
}
}  // namespace

}  // namespace dt

// ============================================================================
// Static scratch of dt_mpm_advance_static (callers that cannot allocate)
// ============================================================================
// The ping-pong state, the two single-substep grids and the binning workspace
// of dt_mpm_advance_static live in this array.  Every GPU the library runs on
// holds its own copy (kStaticScratchBytes, zero-filled at load) until the
// process ends.  Host code never passes the array's name to a kernel: its
// device address is looked up once per device (hipGetSymbolAddress, with that
// device current) and that address is used instead.
namespace dt_static {
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
constexpr int kMaxDevices = 64;
void* g_scratch_addr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of the current device's copy of scratch.
int scratch_address(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    dt::set_error("dt_mpm_advance_static: no current HIP device (or device id >= 64)");
    return DT_ELAUNCH;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      dt::set_error(std::string("dt_mpm_advance_static: hipGetSymbolAddress failed: ") +
                    hipGetErrorString(e));
      return DT_ELAUNCH;
    }
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return DT_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets of (x2, v2, C2, F2, grid4, grid_v4, workspace) in the scratch
// and the total for n particles on prm's grid.
size_t scratch_layout(const dt_mpm_params* prm, int n, size_t (&off)[7]) {
  const long long nn = (long long)(prm->n_env > 1 ? prm->n_env : 1) * prm->n_grid *
                       prm->n_grid * prm->n_grid;
  const size_t sz[7] = {sizeof(float) * 3 * (size_t)n, sizeof(float) * 3 * (size_t)n,
                        sizeof(float) * 9 * (size_t)n, sizeof(float) * 9 * (size_t)n,
                        sizeof(float) * 4 * (size_t)nn, sizeof(float) * 4 * (size_t)nn,
                        dt::bin_ws_layout(prm, n, nullptr, nullptr)};
  size_t total = 0;
  for (int i = 0; i < 7; ++i) {
    off[i] = total;
    total += align256(sz[i] ? sz[i] : 1);
  }
  return total;
}
}  // namespace dt_static

// ============================================================================
// C-ABI: MPM entry points
// ============================================================================
using namespace dt;

extern "C" {

int dt_mpm_init_object(int32_t n, const float* particles, const float* trans16,
                       const float* ball_vel3, float* x, float* v, float* F, void* stream) {
  DT_REQUIRE(n >= 0, "n must be >= 0");
  DT_REQUIRE(trans16 && ball_vel3, "NULL host parameter");
  if (n == 0) return DT_OK;
  DT_REQUIRE(particles && x && v && F, "NULL pointer");
  InitObj A;
  for (int i = 0; i < 12; ++i) A.T[i] = trans16[i];
  for (int i = 0; i < 3; ++i) A.v[i] = ball_vel3[i];
  hipLaunchKernelGGL(k_init_object, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream), n, A,
                     particles, x, v, F);
  return check_launch("dt_mpm_init_object");
}

int dt_mpm_compute_new_F(int32_t n, const float* C, const float* F, float* F_new, float dt,
                         void* stream) {
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(C && F && F_new, "NULL pointer");
  hipLaunchKernelGGL(k_compute_new_F, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream), n, C,
                     F, F_new, dt);
  return check_launch("dt_mpm_compute_new_F");
}

int dt_mpm_compute_new_F_adj(int32_t n, const float* C, const float* F, const float* gF_new,
                             float* gC, float* gF, float dt, void* stream) {
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(C && F && gF_new && gC && gF, "NULL pointer");
  hipLaunchKernelGGL(k_compute_new_F_adj, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream), n,
                     C, F, gF_new, gC, gF, dt);
  return check_launch("dt_mpm_compute_new_F_adj");
}

int dt_mpm_svd(int32_t n, const float* F_new, float* U, float* S, float* V, void* stream) {
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(F_new && U && S && V, "NULL pointer");
  hipLaunchKernelGGL(k_svd, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream), n, F_new, U, S,
                     V);
  return check_launch("dt_mpm_svd");
}

int dt_mpm_svd_adj(int32_t n, const float* U, const float* S, const float* V, const float* gU,
                   const float* gS, const float* gV, float* gF_new, void* stream) {
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(U && S && V && gU && gS && gV && gF_new, "NULL pointer");
  hipLaunchKernelGGL(k_svd_adj, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream), n, U, S, V,
                     gU, gS, gV, gF_new);
  return check_launch("dt_mpm_svd_adj");
}

int dt_mpm_p2g(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
               const int32_t* mat_id, const float* x, const float* v, const float* C,
               const float* F_new, const float* U, const float* S, const float* V, float* grid_mv,
               float* grid_m, float* F_next, void* stream) {
  DT_REQUIRE(!prm || !prm->env_scale, "ragged environments need the binned path");
  int rc = validate_params(prm);
  if (rc) return rc;
  if ((rc = check_mats(mats, n_mats))) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(prm->soa == 0, "dt_mpm_p2g reads the reference [n][k] layout (soa must be 0)");
  DT_REQUIRE(x && v && C && F_new && U && S && V && grid_mv && grid_m && F_next, "NULL pointer");
  Grid g{grid_mv, grid_m, 3, 1};
  hipLaunchKernelGGL(k_p2g, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream), to_dev(*prm),
                     mats, n, mat_id, x, v, C, F_new, U, S, V, g, F_next);
  return check_launch("dt_mpm_p2g");
}

int dt_mpm_p2g_adj(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
                   const int32_t* mat_id, const float* x, const float* v, const float* C,
                   const float* F_new, const float* U, const float* S, const float* V,
                   const float* g_grid_mv, const float* g_grid_m, const float* gF_next, float* gx,
                   float* gv, float* gC, float* gF_new, float* gU, float* gS, float* gV,
                   float* g_mats, void* stream) {
  DT_REQUIRE(!prm || !prm->env_scale, "ragged environments need the binned path");
  int rc = validate_params(prm);
  if (rc) return rc;
  if ((rc = check_mats(mats, n_mats))) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(prm->soa == 0, "dt_mpm_p2g_adj reads the reference [n][k] layout (soa must be 0)");
  DT_REQUIRE(x && v && C && F_new && U && S && V && g_grid_mv && g_grid_m && gF_next && gx && gv &&
                 gC && gF_new && gU && gS && gV,
             "NULL pointer");
  hipLaunchKernelGGL(k_p2g_adj, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), mats, n, mat_id, x, v, C, F_new, U, S, V, g_grid_mv, g_grid_m,
                     gF_next, gx, gv, gC, gF_new, gU, gS, gV, n_mats, g_mats);
  return check_launch("dt_mpm_p2g_adj");
}

int dt_mpm_grid_occupy(const dt_mpm_params* prm, const float* grid_m, int32_t* occupy,
                       void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(grid_m && occupy, "NULL pointer");
  long long nn = grid_nodes(prm);
  hipLaunchKernelGGL(k_grid_occupy, dim3(nblocks(nn)), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), nn, grid_m, occupy);
  return check_launch("dt_mpm_grid_occupy");
}

int dt_mpm_grid_op(const dt_mpm_params* prm, const float* grid_mv, const float* grid_m,
                   const float* grid_f, const int32_t* occupy, float* grid_v_out, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(grid_mv && grid_m && grid_v_out, "NULL pointer");
  long long nn = grid_nodes(prm);
  hipLaunchKernelGGL(k_grid_op, dim3(nblocks(nn)), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), nn, grid_mv, grid_m, grid_f, occupy, grid_v_out);
  return check_launch("dt_mpm_grid_op");
}

int dt_mpm_grid_op_adj(const dt_mpm_params* prm, const float* grid_mv, const float* grid_m,
                       const float* grid_f, const int32_t* occupy, const float* g_grid_v_out,
                       float* g_grid_mv, float* g_grid_m, float* g_grid_f, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(grid_mv && grid_m && g_grid_v_out && g_grid_mv && g_grid_m, "NULL pointer");
  long long nn = grid_nodes(prm);
  hipLaunchKernelGGL(k_grid_op_adj, dim3(nblocks(nn)), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), nn, grid_mv, grid_m, grid_f, occupy, g_grid_v_out, g_grid_mv,
                     g_grid_m, g_grid_f);
  return check_launch("dt_mpm_grid_op_adj");
}

int dt_mpm_g2p(const dt_mpm_params* prm, int32_t n, const float* x, const float* grid_v_out,
               float* x_next, float* v_next, float* C_next, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && grid_v_out && x_next && v_next && C_next, "NULL pointer");
  hipLaunchKernelGGL(k_g2p<false>, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), n, x, grid_v_out, 3, x_next, v_next, C_next, (int32_t*)nullptr,
                     (int32_t*)nullptr, (int32_t*)nullptr, BinRuns{});
  return check_launch("dt_mpm_g2p");
}

int dt_mpm_g2p_adj(const dt_mpm_params* prm, int32_t n, const float* x, const float* grid_v_out,
                   const float* gx_next, const float* gv_next, const float* gC_next, float* gx,
                   float* g_grid_v_out, void* stream) {
  DT_REQUIRE(!prm || !prm->env_scale, "ragged environments need the binned path");
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && grid_v_out && gx_next && gv_next && gC_next && gx && g_grid_v_out,
             "NULL pointer");
  hipLaunchKernelGGL(k_g2p_adj<0>, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), n, x, grid_v_out, 3, gx_next, gv_next, gC_next, gx,
                     g_grid_v_out, 3);
  return check_launch("dt_mpm_g2p_adj");
}

int dt_mpm_substep_fwd(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats,
                       int32_t n, const int32_t* mat_id, const float* x, const float* v,
                       const float* C, const float* F, float* grid4, float* grid_v4,
                       float* x_next, float* v_next, float* C_next, float* F_next, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  if ((rc = check_mats(mats, n_mats))) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  DT_REQUIRE(grid4 && grid_v4, "NULL grid pointer");
  DT_REQUIRE(n == 0 || (x && v && C && F && x_next && v_next && C_next && F_next), "NULL pointer");
  return launch_substep_fwd(prm, mats, n, mat_id, x, v, C, F, grid4, grid_v4, x_next, v_next,
                            C_next, F_next, DT_STREAM(stream));
}

int dt_mpm_substep_adj(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats,
                       int32_t n, const int32_t* mat_id, const float* x, const float* v,
                       const float* C, const float* F, const float* grid4, const float* grid_v4,
                       const float* gx_next, const float* gv_next, const float* gC_next,
                       const float* gF_next, float* g_grid4, float* g_grid_v4, float* gx,
                       float* gv, float* gC, float* gF, float* g_mats, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  if ((rc = check_mats(mats, n_mats))) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  DT_REQUIRE(grid4 && grid_v4 && g_grid4 && g_grid_v4, "NULL grid pointer");
  DT_REQUIRE(n == 0 || (x && v && C && F && gx_next && gv_next && gC_next && gF_next && gx && gv &&
                        gC && gF),
             "NULL pointer");
  return launch_substep_adj(prm, mats, n_mats, n, mat_id, x, v, C, F, grid4, grid_v4, gx_next,
                            gv_next, gC_next, gF_next, g_grid4, g_grid_v4, gx, gv, gC, gF, g_mats,
                            DT_STREAM(stream));
}

int dt_mpm_run_fwd(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
                   const int32_t* mat_id, float* x, float* v, float* C, float* F,
                   int32_t n_substeps, float* grids4, float* grids_v4, float* scratch4,
                   float* scratch_v4, void* workspace, size_t workspace_bytes, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  if ((rc = check_mats(mats, n_mats))) return rc;
  DT_REQUIRE(n >= 0 && n_substeps >= 0, "n and n_substeps must be >= 0");
  // kept grids: grids4 and grids_v4, or on the sparse path grids4 alone (the
  // adjoint remakes the velocities from it) with scratch_v4 as the
  // substep's velocity grid
  const bool keep = grids4 != nullptr && (grids_v4 != nullptr || workspace != nullptr);
  const bool keep_v = keep && grids_v4 != nullptr;
  DT_REQUIRE(keep || (scratch4 && scratch_v4), "need kept grids or scratch grids");
  DT_REQUIRE(keep_v || scratch_v4, "need kept velocity grids or a scratch velocity grid");
  DT_REQUIRE(grids4 || !grids_v4, "grids_v4 without grids4");
  BinWs ws{};
  if (workspace &&
      (rc = workspace_views(prm, n, workspace, workspace_bytes, ws, keep ? n_substeps : 0)))
    return rc;
  const bool hist = workspace && keep && ws.hist_slots >= n_substeps && n_substeps > 0;
  // G2P makes the next substep's brick counts when the binning clears them
  const bool fuse = workspace && bin_counts_fusable(ws);
  long long nn = grid_nodes(prm);
  auto body = [&](hipStream_t st) -> int {
    for (int s = 0; s < n_substeps; ++s) {
      long long o = (long long)s * n, o1 = (long long)(s + 1) * n;
      float* g4 = keep ? grids4 + (long long)s * nn * 4 : scratch4;
      float* gv4 = keep_v ? grids_v4 + (long long)s * nn * 4 : scratch_v4;
      BinWs slot = hist ? bin_slot(ws, s, n) : ws;
      const BinWs* wp = workspace ? &slot : nullptr;
      int r = launch_substep_fwd(prm, mats, n, mat_id, x + 3 * o, v + 3 * o, C + 9 * o, F + 9 * o,
                                 g4, gv4, x + 3 * o1, v + 3 * o1, C + 9 * o1, F + 9 * o1, st, wp,
                                 s > 0, fuse && s > 0, fuse && s + 1 < n_substeps);
      if (r) return r;
    }
    return DT_OK;
  };
  std::vector<uint64_t> key{1};
  key_add(key, to_dev(*prm));   // no padding bytes
  for (const void* q : {(const void*)mats, (const void*)mat_id, (const void*)x, (const void*)v,
                        (const void*)C, (const void*)F, (const void*)grids4,
                        (const void*)grids_v4, (const void*)scratch4, (const void*)scratch_v4,
                        (const void*)workspace})
    key_add(key, q);
  key_add(key, (long long)n_mats << 32 | (uint32_t)n);
  key_add(key, (long long)n_substeps);
  key_add(key, workspace_bytes);
  rc = run_graphed(std::move(key), DT_STREAM(stream), body);
  if (workspace) {
    if (hist && rc == DT_OK) hist_set(workspace, x, n, n_substeps);
    else hist_clear(workspace);
  }
  return rc;
}

int dt_mpm_run_adj(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
                   const int32_t* mat_id, const float* x, const float* v, const float* C,
                   const float* F, int32_t n_substeps, const float* grids4, const float* grids_v4,
                   float* g_grid4, float* g_grid_v4, float* gx, float* gv, float* gC, float* gF,
                   float* gx2, float* gv2, float* gC2, float* gF2, void* workspace,
                   size_t workspace_bytes, float* g_mats, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  if ((rc = check_mats(mats, n_mats))) return rc;
  DT_REQUIRE(n >= 0 && n_substeps >= 0, "n and n_substeps must be >= 0");
  BinWs ws{};
  if (workspace && (rc = workspace_views(prm, n, workspace, workspace_bytes, ws, n_substeps)))
    return rc;
  // reuse the forward pass's binning when this workspace holds it for these frames
  const bool hist = workspace && ws.hist_slots >= n_substeps && n_substeps > 0 &&
                    hist_matches(workspace, x, n, n_substeps);
  DT_REQUIRE(grids4 && g_grid4 && g_grid_v4, "NULL grid pointer");
  DT_REQUIRE(grids_v4 || workspace, "grids_v4 NULL needs the sparse path (a workspace)");
  long long nn = grid_nodes(prm);
  hipStream_t st = DT_STREAM(stream);
  float *a[4] = {gx, gv, gC, gF}, *b[4] = {gx2, gv2, gC2, gF2};
  // material adjoints with a workspace: per-block partials accumulated over
  // the sweep, reduced in a fixed order at its end (the same sums every run)
  const bool mat_det = workspace && g_mats && n > 0 && n_substeps > 0;
  const long long part_n = (long long)nblocks(n) * 2 * kMatSlots;
  auto body = [&](hipStream_t cs) -> int {
    float *p[4] = {gx, gv, gC, gF}, *q[4] = {gx2, gv2, gC2, gF2};
    if (mat_det)
      hipLaunchKernelGGL(k_zero_f, dim3(nblocks(part_n)), dim3(kBlock), 0, cs, part_n,
                         ws.mat_part);
    for (int s = n_substeps - 1; s >= 0; --s) {
      long long o = (long long)s * n;
      BinWs slot = hist ? bin_slot(ws, s, n) : ws;
      const BinWs* wp = workspace ? &slot : nullptr;
      int r = launch_substep_adj(prm, mats, n_mats, n, mat_id, x + 3 * o, v + 3 * o, C + 9 * o,
                                 F + 9 * o, grids4 + (long long)s * nn * 4,
                                 grids_v4 ? grids_v4 + (long long)s * nn * 4 : nullptr, p[0], p[1],
                                 p[2], p[3], g_grid4,
                                 g_grid_v4, q[0], q[1], q[2], q[3], g_mats, cs, wp, hist,
                                 s != n_substeps - 1, mat_det ? ws.mat_part : nullptr);
      if (r) return r;
      for (int k = 0; k < 4; ++k) { float* t = p[k]; p[k] = q[k]; q[k] = t; }
    }
    if (mat_det)
      hipLaunchKernelGGL(k_mat_part_reduce, dim3(1), dim3(kBlock), 0, cs, nblocks(n),
                         2 * (n_mats < kMatSlots ? n_mats : kMatSlots), (const float*)ws.mat_part,
                         g_mats);
    return DT_OK;
  };
  std::vector<uint64_t> key{hist ? 4u : 2u};
  key_add(key, to_dev(*prm));   // no padding bytes
  for (const void* ptr : {(const void*)mats, (const void*)mat_id, (const void*)x, (const void*)v,
                          (const void*)C, (const void*)F, (const void*)grids4,
                          (const void*)grids_v4, (const void*)g_grid4, (const void*)g_grid_v4,
                          (const void*)gx, (const void*)gv, (const void*)gC, (const void*)gF,
                          (const void*)gx2, (const void*)gv2, (const void*)gC2, (const void*)gF2,
                          (const void*)workspace, (const void*)g_mats})
    key_add(key, ptr);
  key_add(key, (long long)n_mats << 32 | (uint32_t)n);
  key_add(key, (long long)n_substeps);
  key_add(key, workspace_bytes);
  if ((rc = run_graphed(std::move(key), st, body))) return rc;
  // the result sits in gx.. after an even number of substeps, else in gx2..
  if (n_substeps % 2 == 1)
    for (int k = 0; k < 4; ++k) { float* t = a[k]; a[k] = b[k]; b[k] = t; }
  // result sits in a[]; move it into the caller's seed buffers when needed
  if (a[0] != gx) {
    hipMemcpyAsync(gx, a[0], sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(gv, a[1], sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(gC, a[2], sizeof(float) * 9 * (size_t)n, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(gF, a[3], sizeof(float) * 9 * (size_t)n, hipMemcpyDeviceToDevice, st);
  }
  return check_launch("dt_mpm_run_adj");
}

int dt_mpm_advance(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
                   const int32_t* mat_id, float* x, float* v, float* C, float* F, float* x2,
                   float* v2, float* C2, float* F2, int32_t n_substeps, float* scratch4,
                   float* scratch_v4, void* workspace, size_t workspace_bytes, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  if ((rc = check_mats(mats, n_mats))) return rc;
  DT_REQUIRE(n >= 0 && n_substeps >= 0, "n and n_substeps must be >= 0");
  DT_REQUIRE(scratch4 && scratch_v4, "NULL grid scratch");
  BinWs ws{};
  if (workspace && (rc = workspace_views(prm, n, workspace, workspace_bytes, ws))) return rc;
  const BinWs* wp = workspace ? &ws : nullptr;
  const bool fuse = workspace && bin_counts_fusable(ws);
  DT_REQUIRE(n == 0 || (x && v && C && F && x2 && v2 && C2 && F2), "NULL pointer");
  hipStream_t st = DT_STREAM(stream);
  float *a[4] = {x, v, C, F}, *b[4] = {x2, v2, C2, F2};
  auto body = [&](hipStream_t cs) -> int {
    float *p[4] = {x, v, C, F}, *q[4] = {x2, v2, C2, F2};
    for (int s = 0; s < n_substeps; ++s) {
      int r = launch_substep_fwd(prm, mats, n, mat_id, p[0], p[1], p[2], p[3], scratch4,
                                 scratch_v4, q[0], q[1], q[2], q[3], cs, wp, s > 0,
                                 fuse && s > 0, fuse && s + 1 < n_substeps);
      if (r) return r;
      for (int k = 0; k < 4; ++k) { float* t = p[k]; p[k] = q[k]; q[k] = t; }
    }
    return DT_OK;
  };
  std::vector<uint64_t> key{3};
  key_add(key, to_dev(*prm));   // no padding bytes
  for (const void* ptr : {(const void*)mats, (const void*)mat_id, (const void*)x, (const void*)v,
                          (const void*)C, (const void*)F, (const void*)x2, (const void*)v2,
                          (const void*)C2, (const void*)F2, (const void*)scratch4,
                          (const void*)scratch_v4, (const void*)workspace})
    key_add(key, ptr);
  key_add(key, (long long)n_mats << 32 | (uint32_t)n);
  key_add(key, (long long)n_substeps);
  key_add(key, workspace_bytes);
  if ((rc = run_graphed(std::move(key), st, body))) return rc;
  if (n_substeps % 2 == 1)
    for (int k = 0; k < 4; ++k) { float* t = a[k]; a[k] = b[k]; b[k] = t; }
  if (a[0] != x && n > 0) {
    hipMemcpyAsync(x, a[0], sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(v, a[1], sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(C, a[2], sizeof(float) * 9 * (size_t)n, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(F, a[3], sizeof(float) * 9 * (size_t)n, hipMemcpyDeviceToDevice, st);
  }
  return check_launch("dt_mpm_advance");
}

int dt_mpm_advance_static(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats,
                          int32_t n, const int32_t* mat_id, float* x, float* v, float* C,
                          float* F, int32_t n_substeps, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  if ((rc = check_mats(mats, n_mats))) return rc;
  DT_REQUIRE(n >= 0 && n_substeps >= 0, "n and n_substeps must be >= 0");
  DT_REQUIRE(n == 0 || (x && v && C && F), "NULL pointer");
  size_t off[7];
  const size_t need = dt_static::scratch_layout(prm, n, off);
  if (need > dt_static::kStaticScratchBytes) {
    set_error("dt_mpm_advance_static: " + std::to_string(n) + " particles on this grid need " +
              std::to_string(need) + " bytes of scratch; the static scratch holds " +
              std::to_string(dt_static::kStaticScratchBytes) + " (use dt_mpm_advance)");
    return DT_EINVAL;
  }
  void* base = nullptr;
  if ((rc = dt_static::scratch_address(&base))) return rc;
  char* b = static_cast<char*>(base);
  float* x2 = reinterpret_cast<float*>(b + off[0]);
  float* v2 = reinterpret_cast<float*>(b + off[1]);
  float* C2 = reinterpret_cast<float*>(b + off[2]);
  float* F2 = reinterpret_cast<float*>(b + off[3]);
  float* g4 = reinterpret_cast<float*>(b + off[4]);
  float* gv4 = reinterpret_cast<float*>(b + off[5]);
  BinWs ws{};
  bin_ws_layout(prm, n, b + off[6], &ws);
  // the first binning clears the brick counts it accumulates into (k_bin_zero);
  // nothing else in the scratch is read before this call writes it
  const bool fuse = bin_counts_fusable(ws);
  hipStream_t st = DT_STREAM(stream);
  float *p[4] = {x, v, C, F}, *q[4] = {x2, v2, C2, F2};
  for (int s = 0; s < n_substeps; ++s) {
    int r = launch_substep_fwd(prm, mats, n, mat_id, p[0], p[1], p[2], p[3], g4, gv4, q[0],
                               q[1], q[2], q[3], st, &ws, s > 0, fuse && s > 0,
                               fuse && s + 1 < n_substeps);
    if (r) return r;
    for (int k = 0; k < 4; ++k) { float* t = p[k]; p[k] = q[k]; q[k] = t; }
  }
  if (p[0] != x && n > 0) {
    hipMemcpyAsync(x, p[0], sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(v, p[1], sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(C, p[2], sizeof(float) * 9 * (size_t)n, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(F, p[3], sizeof(float) * 9 * (size_t)n, hipMemcpyDeviceToDevice, st);
  }
  return check_launch("dt_mpm_advance_static");
}


// ---- fused-path stages --------------------------------------------------
int dt_mpm_zero4(int64_t n_nodes, float* buf4, void* stream) {
  DT_REQUIRE(n_nodes >= 0, "n_nodes must be >= 0");
  if (n_nodes == 0) return DT_OK;
  DT_REQUIRE(buf4, "NULL pointer");
  hipLaunchKernelGGL(k_zero4, dim3(nblocks(n_nodes)), dim3(kBlock), 0, DT_STREAM(stream),
                     (long long)n_nodes, (float4*)buf4);
  return check_launch("dt_mpm_zero4");
}

int dt_mpm_p2g_fused(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
                     const int32_t* mat_id, const float* x, const float* v, const float* C,
                     const float* F, float* grid4, float* F_next, void* stream) {
  DT_REQUIRE(!prm || !prm->env_scale, "ragged environments need the binned path");
  int rc = validate_params(prm);
  if (rc) return rc;
  if ((rc = check_mats(mats, n_mats))) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && v && C && F && grid4 && F_next, "NULL pointer");
  Grid g{grid4, grid4 + 3, 4, 4};
  hipLaunchKernelGGL(k_p2g_fused, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), mats, n, mat_id, x, v, C, F, g, F_next);
  return check_launch("dt_mpm_p2g_fused");
}

int dt_mpm_grid_op4(const dt_mpm_params* prm, const float* grid4, const float* grid_f,
                    float* grid_v4, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(grid4 && grid_v4, "NULL pointer");
  long long nn = grid_nodes(prm);
  hipLaunchKernelGGL(k_grid_op4, dim3(node_blocks(*prm, nn)), dim3(kBlock), 0,
                     DT_STREAM(stream), to_dev_nodes(*prm), nn, (const float4*)grid4, grid_f,
                     (float4*)grid_v4);
  return check_launch("dt_mpm_grid_op4");
}

int dt_mpm_g2p4(const dt_mpm_params* prm, int32_t n, const float* x, const float* grid_v4,
                float* x_next, float* v_next, float* C_next, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && grid_v4 && x_next && v_next && C_next, "NULL pointer");
  hipLaunchKernelGGL(k_g2p<false>, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), n, x, grid_v4, 4, x_next, v_next, C_next, (int32_t*)nullptr,
                     (int32_t*)nullptr, (int32_t*)nullptr, BinRuns{});
  return check_launch("dt_mpm_g2p4");
}

int dt_mpm_g2p4_adj(const dt_mpm_params* prm, int32_t n, const float* x, const float* grid_v4,
                    const float* gx_next, const float* gv_next, const float* gC_next, float* gx,
                    float* g_grid_v4, void* stream) {
  DT_REQUIRE(!prm || !prm->env_scale, "ragged environments need the binned path");
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && grid_v4 && gx_next && gv_next && gC_next && gx && g_grid_v4, "NULL pointer");
  hipLaunchKernelGGL(k_g2p_adj<1>, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), n, x, grid_v4, 4, gx_next, gv_next, gC_next, gx, g_grid_v4, 4);
  return check_launch("dt_mpm_g2p4_adj");
}

int dt_mpm_grid_op4_adj(const dt_mpm_params* prm, const float* grid4, const float* grid_f,
                        const float* g_grid_v4, float* g_grid4, float* g_grid_f, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(grid4 && g_grid_v4 && g_grid4, "NULL pointer");
  long long nn = grid_nodes(prm);
  hipLaunchKernelGGL(k_grid_op4_adj, dim3(node_blocks(*prm, nn)), dim3(kBlock), 0,
                     DT_STREAM(stream), to_dev_nodes(*prm), nn, (const float4*)grid4, grid_f,
                     (const float4*)g_grid_v4, (float4*)g_grid4, g_grid_f);
  return check_launch("dt_mpm_grid_op4_adj");
}

int dt_mpm_zero_nodes(const dt_mpm_params* prm, float* buf, int32_t width, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(buf && width > 0, "NULL buffer or width < 1");
  long long nn = grid_nodes(prm);
  hipLaunchKernelGGL(k_zero_nodes, dim3(node_blocks(*prm, nn)), dim3(kBlock), 0,
                     DT_STREAM(stream), to_dev_nodes(*prm), nn, buf, (int)width);
  return check_launch("dt_mpm_zero_nodes");
}

int dt_mpm_brick_nodes_copy(const dt_mpm_params* prm, int32_t n_arrays, void* const* dense,
                            void* const* compact, const int32_t* widths, int32_t capacity,
                            int32_t direction, int32_t* overflow, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(prm->active_bricks && prm->n_active_bricks,
             "dt_mpm_brick_nodes_copy needs the node set of prm (active_bricks)");
  DT_REQUIRE(n_arrays >= 1 && n_arrays <= kMaxNodeArrays && dense && compact && widths,
             "n_arrays must be in [1, 4] with its pointer and width arrays");
  DT_REQUIRE(capacity >= 0, "capacity must be >= 0");
  DT_REQUIRE(direction == 0 || direction == 1, "direction must be 0 (pack) or 1 (unpack)");
  NodeArrays A{};
  A.n = n_arrays;
  for (int a = 0; a < n_arrays; ++a) {
    DT_REQUIRE(dense[a] && compact[a] && widths[a] > 0, "NULL array or width < 1");
    const bool pack = direction == 0;
    A.src[a] = static_cast<const uint32_t*>(pack ? dense[a] : compact[a]);
    A.dst[a] = static_cast<uint32_t*>(pack ? compact[a] : dense[a]);
    A.width[a] = widths[a];
  }
  const MpmP P = to_dev_nodes(*prm);
  const dim3 grid(node_blocks(*prm, grid_nodes(prm)));
  if (direction == 0)
    hipLaunchKernelGGL(k_brick_nodes_copy<true>, grid, dim3(kBlock), 0, DT_STREAM(stream), P, A,
                       (int)capacity, overflow);
  else
    hipLaunchKernelGGL(k_brick_nodes_copy<false>, grid, dim3(kBlock), 0, DT_STREAM(stream), P, A,
                       (int)capacity, nullptr);
  return check_launch("dt_mpm_brick_nodes_copy");
}

int dt_mpm_p2g_fused_adj(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats,
                         int32_t n, const int32_t* mat_id, const float* x, const float* v,
                         const float* C, const float* F, const float* g_grid4,
                         const float* gF_next, float* gx, float* gv, float* gC, float* gF,
                         float* g_mats, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  if ((rc = check_mats(mats, n_mats))) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && v && C && F && g_grid4 && gF_next && gx && gv && gC && gF, "NULL pointer");
  launch_p2g_fused_adj(to_dev(*prm), DT_STREAM(stream), mats, n_mats, n, mat_id, x, v, C, F,
                       g_grid4, gF_next, gx, gv, gC, gF, g_mats);
  return check_launch("dt_mpm_p2g_fused_adj");
}

int dt_mpm_zero4_bricks(const dt_mpm_params* prm, const int32_t* flags, float* buf4,
                        void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(flags && buf4, "NULL pointer");
  int bgrid = (int)((n_bricks(prm) + 3) / 4);
  hipLaunchKernelGGL(k_zero4_bricks, dim3(bgrid), dim3(kBlock), 0, DT_STREAM(stream), to_dev(*prm),
                     flags, (float4*)buf4);
  return check_launch("dt_mpm_zero4_bricks");
}

int dt_mpm_grid_op4_bricks(const dt_mpm_params* prm, const int32_t* flags, const float* grid4,
                           float* grid_v4, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(flags && grid4 && grid_v4, "NULL pointer");
  int bgrid = (int)((n_bricks(prm) + 3) / 4);
  hipLaunchKernelGGL(k_grid_op4_bricks, dim3(bgrid), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), flags, (const float4*)grid4, (const float*)nullptr,
                     (float4*)grid_v4);
  return check_launch("dt_mpm_grid_op4_bricks");
}

int dt_mpm_grid_op4_adj_bricks(const dt_mpm_params* prm, const int32_t* flags, const float* grid4,
                               const float* g_grid_v4, float* g_grid4, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(flags && grid4 && g_grid_v4 && g_grid4, "NULL pointer");
  int bgrid = (int)((n_bricks(prm) + 3) / 4);
  hipLaunchKernelGGL(k_grid_op4_adj_bricks, dim3(bgrid), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), flags, (const float4*)grid4, (const float4*)g_grid_v4,
                     (float4*)g_grid4);
  return check_launch("dt_mpm_grid_op4_adj_bricks");
}

int64_t dt_mpm_brick_count(const dt_mpm_params* prm) {
  if (!prm || prm->n_grid < 3) return 0;
  return n_bricks(prm);
}

int dt_mpm_workspace_bytes(const dt_mpm_params* prm, int32_t n, size_t* bytes) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  DT_REQUIRE(bytes != nullptr, "NULL bytes");
  *bytes = bin_ws_layout(prm, n, nullptr, nullptr);
  return DT_OK;
}

int dt_mpm_workspace_history_bytes(const dt_mpm_params* prm, int32_t n, int32_t n_substeps,
                                   size_t* bytes) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(n >= 0 && n_substeps >= 0, "n and n_substeps must be >= 0");
  DT_REQUIRE(bytes != nullptr, "NULL bytes");
  *bytes = bin_history_slot_bytes(prm, n) * (size_t)n_substeps;
  return DT_OK;
}

int dt_mpm_bin(const dt_mpm_params* prm, int32_t n, const float* x, void* workspace,
               size_t workspace_bytes, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  DT_REQUIRE(n == 0 || x, "NULL x");
  BinWs ws{};
  if ((rc = workspace_views(prm, n, workspace, workspace_bytes, ws))) return rc;
  return launch_mark_bin(to_dev(*prm), prm, n, x, ws, DT_STREAM(stream));
}

int dt_mpm_active_bricks(const dt_mpm_params* prm, int32_t n, const void* workspace,
                         size_t workspace_bytes, const int32_t** active_bricks,
                         const int32_t** n_active_bricks) {
  int rc = validate_params(prm);
  if (rc) return rc;
  BinWs ws{};
  if ((rc = workspace_views(prm, n, const_cast<void*>(workspace), workspace_bytes, ws))) return rc;
  if (active_bricks) *active_bricks = ws.alist;
  if (n_active_bricks) *n_active_bricks = ws.n_alist;
  return DT_OK;
}

int dt_mpm_bin_perm(const dt_mpm_params* prm, int32_t n, const void* workspace,
                    size_t workspace_bytes, const int32_t** perm, const int32_t** brick_start) {
  int rc = validate_params(prm);
  if (rc) return rc;
  BinWs ws{};
  if ((rc = workspace_views(prm, n, const_cast<void*>(workspace), workspace_bytes, ws))) return rc;
  if (perm) *perm = ws.perm;
  if (brick_start) *brick_start = ws.start;
  return DT_OK;
}

int dt_mpm_bin_slot_bytes(const dt_mpm_params* prm, int32_t n, size_t* bytes) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  DT_REQUIRE(bytes, "NULL bytes");
  *bytes = bin_history_slot_bytes(prm, n);
  return DT_OK;
}

int dt_mpm_bin_copy(const dt_mpm_params* prm, int32_t n, void* workspace, size_t workspace_bytes,
                    void* slot, int32_t save, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  DT_REQUIRE(slot && ((uintptr_t)slot & 255u) == 0, "slot must be a 256-byte aligned buffer");
  BinWs ws{};
  if ((rc = workspace_views(prm, n, workspace, workspace_bytes, ws))) return rc;
  const BinWs sv = bin_view(ws, static_cast<char*>(slot), n);
  launch_bin_copy(save ? ws : sv, save ? sv : ws, n, DT_STREAM(stream));
  return check_launch("binning copy");
}

int dt_mpm_p2g_binned(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
                      const int32_t* mat_id, const float* x, const float* v, const float* C,
                      const float* F, float* F_next, void* workspace, size_t workspace_bytes,
                      void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  if ((rc = check_mats(mats, n_mats))) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  DT_REQUIRE(n == 0 || (x && v && C && F && F_next), "NULL pointer");
  BinWs ws{};
  if ((rc = workspace_views(prm, n, workspace, workspace_bytes, ws))) return rc;
  hipLaunchKernelGGL(k_p2g_binned, dim3(brick_blocks(ws)), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), mats, n, mat_id, x, v, C, F, F_next, (const int32_t*)ws.perm,
                     (const int32_t*)ws.start, (const int32_t*)ws.list,
                     (const int32_t*)ws.n_list, ws.tiles);
  return check_launch("dt_mpm_p2g_binned");
}

int dt_mpm_grid_op4_tiles(const dt_mpm_params* prm, int32_t n, const void* workspace,
                          size_t workspace_bytes, float* grid4, float* grid_v4, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(grid4 && grid_v4, "NULL grid pointer");
  BinWs ws{};
  if ((rc = workspace_views(prm, n, const_cast<void*>(workspace), workspace_bytes, ws))) return rc;
  hipLaunchKernelGGL(k_grid_op4_tiles, dim3(active_blocks(ws)), dim3(kBlock), 0,
                     DT_STREAM(stream), to_dev(*prm), (const int32_t*)ws.alist,
                     (const int32_t*)ws.n_alist, (const float4*)ws.tiles,
                     (const int32_t*)ws.start, (float4*)grid4, (float4*)grid_v4);
  return check_launch("dt_mpm_grid_op4_tiles");
}

int dt_mpm_tiles_sum(const dt_mpm_params* prm, int32_t n, const void* workspace,
                     size_t workspace_bytes, float* out4, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(out4, "NULL grid pointer");
  BinWs ws{};
  if ((rc = workspace_views(prm, n, const_cast<void*>(workspace), workspace_bytes, ws))) return rc;
  hipLaunchKernelGGL(k_tiles_sum, dim3(active_blocks(ws)), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), (const int32_t*)ws.alist, (const int32_t*)ws.n_alist,
                     (const float4*)ws.tiles, (const int32_t*)ws.start, (float4*)out4);
  return check_launch("dt_mpm_tiles_sum");
}

int dt_mpm_tiles_sum_zero(const dt_mpm_params* prm, int32_t n, const void* workspace,
                          size_t workspace_bytes, float* out4, float* zero_buf, int32_t width,
                          void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(out4 && zero_buf, "NULL grid pointer");
  DT_REQUIRE(width >= 1 && width <= 4, "width must be 1..4");
  BinWs ws{};
  if ((rc = workspace_views(prm, n, const_cast<void*>(workspace), workspace_bytes, ws))) return rc;
  hipLaunchKernelGGL(k_tiles_sum_zero, dim3(active_blocks(ws)), dim3(kBlock), 0,
                     DT_STREAM(stream), to_dev(*prm), (const int32_t*)ws.alist,
                     (const int32_t*)ws.n_alist, (const float4*)ws.tiles, (const int32_t*)ws.start,
                     (float4*)out4, zero_buf, width);
  return check_launch("dt_mpm_tiles_sum_zero");
}

int dt_mpm_g2p4_adj_binned(const dt_mpm_params* prm, int32_t n, const float* x,
                           const float* grid_v4, const float* gx_next, const float* gv_next,
                           const float* gC_next, float* gx, void* workspace,
                           size_t workspace_bytes, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  DT_REQUIRE(grid_v4, "NULL grid pointer");
  DT_REQUIRE(n == 0 || (x && gx_next && gv_next && gC_next && gx), "NULL pointer");
  BinWs ws{};
  if ((rc = workspace_views(prm, n, workspace, workspace_bytes, ws))) return rc;
  hipLaunchKernelGGL(k_g2p_adj_binned<false>, dim3(brick_blocks(ws)), dim3(kBlock), 0,
                     DT_STREAM(stream), to_dev(*prm), n, x, grid_v4, 4, gx_next, gv_next,
                     gC_next, gx, (const int32_t*)ws.perm, (const int32_t*)ws.start,
                     (const int32_t*)ws.list, (const int32_t*)ws.n_list, ws.tiles);
  return check_launch("dt_mpm_g2p4_adj_binned");
}

int dt_mpm_grid_op4_adj_tiles(const dt_mpm_params* prm, int32_t n, const void* workspace,
                              size_t workspace_bytes, const float* grid4, float* g_grid_v4,
                              float* g_grid4, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(grid4 && g_grid_v4 && g_grid4, "NULL grid pointer");
  BinWs ws{};
  if ((rc = workspace_views(prm, n, const_cast<void*>(workspace), workspace_bytes, ws))) return rc;
  hipLaunchKernelGGL(k_grid_op4_adj_tiles, dim3(active_blocks(ws)), dim3(kBlock), 0,
                     DT_STREAM(stream), to_dev(*prm), (const int32_t*)ws.alist,
                     (const int32_t*)ws.n_alist, (const float4*)grid4, (const float4*)ws.tiles,
                     (const int32_t*)ws.start, (float4*)g_grid_v4, (float4*)g_grid4);
  return check_launch("dt_mpm_grid_op4_adj_tiles");
}

int dt_mpm_grid_op4_adj_tiles_f(const dt_mpm_params* prm, int32_t n, const void* workspace,
                                size_t workspace_bytes, const float* grid4, const float* grid_f,
                                float* g_grid_v4, float* g_grid4, float* g_grid_f, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(grid4 && grid_f && g_grid4 && g_grid_f, "NULL grid pointer");
  BinWs ws{};
  if ((rc = workspace_views(prm, n, const_cast<void*>(workspace), workspace_bytes, ws))) return rc;
  hipLaunchKernelGGL(k_grid_op4_adj_tiles_f, dim3(active_blocks(ws)), dim3(kBlock), 0,
                     DT_STREAM(stream), to_dev(*prm), (const int32_t*)ws.alist,
                     (const int32_t*)ws.n_alist, (const float4*)grid4, grid_f,
                     (const float4*)ws.tiles, (const int32_t*)ws.start, (float4*)g_grid_v4,
                     (float4*)g_grid4, g_grid_f);
  return check_launch("dt_mpm_grid_op4_adj_tiles_f");
}

int dt_mpm_mark_bricks(const dt_mpm_params* prm, int32_t n, const float* x, int32_t* flags,
                       void* stream) {
  DT_REQUIRE(!prm || !prm->env_scale, "ragged environments need the binned path");
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  DT_REQUIRE(flags && (n == 0 || x), "NULL pointer");
  launch_mark(to_dev(*prm), prm, n, x, flags, DT_STREAM(stream));
  return check_launch("dt_mpm_mark_bricks");
}

int dt_mpm_cell_keys(const dt_mpm_params* prm, int32_t n, const float* x, int32_t* keys,
                     void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && keys, "NULL pointer");
  hipLaunchKernelGGL(k_cell_keys, dim3(nblocks(n)), dim3(kBlock), 0, DT_STREAM(stream),
                     to_dev(*prm), n, x, keys);
  return check_launch("dt_mpm_cell_keys");
}

int dt_gather_rows(int32_t n, int32_t width, const int32_t* perm, const float* src, float* dst,
                   void* stream) {
  DT_REQUIRE(n >= 0 && width > 0, "bad size");
  if (n == 0) return DT_OK;
  DT_REQUIRE(perm && src && dst && src != dst, "NULL or aliased pointer");
  long long total = (long long)n * width;
  hipLaunchKernelGGL(k_gather_rows, dim3(nblocks(total)), dim3(kBlock), 0, DT_STREAM(stream),
                     total, width, perm, src, dst);
  return check_launch("dt_gather_rows");
}

int dt_scatter_rows(int32_t n, int32_t width, const int32_t* perm, const float* src, float* dst,
                    void* stream) {
  DT_REQUIRE(n >= 0 && width > 0, "bad size");
  if (n == 0) return DT_OK;
  DT_REQUIRE(perm && src && dst && src != dst, "NULL or aliased pointer");
  long long total = (long long)n * width;
  hipLaunchKernelGGL(k_scatter_rows, dim3(nblocks(total)), dim3(kBlock), 0, DT_STREAM(stream),
                     total, width, perm, src, dst);
  return check_launch("dt_scatter_rows");
}

int dt_gather_rows_soa(int32_t n, int32_t width, const int32_t* perm, const float* src_rows,
                       float* dst_soa, void* stream) {
  DT_REQUIRE(n >= 0 && width > 0, "bad size");
  if (n == 0) return DT_OK;
  DT_REQUIRE(src_rows && dst_soa && src_rows != dst_soa, "NULL or aliased pointer");
  long long total = (long long)n * width;
  hipLaunchKernelGGL(k_gather_soa, dim3(nblocks(total)), dim3(kBlock), 0, DT_STREAM(stream),
                     (long long)n, width, perm, src_rows, dst_soa);
  return check_launch("dt_gather_rows_soa");
}

int dt_permute_soa(int32_t n, int32_t width, const int32_t* perm, int32_t scatter,
                   const float* src_soa, float* dst_soa, void* stream) {
  DT_REQUIRE(n >= 0 && width > 0, "bad size");
  if (n == 0) return DT_OK;
  DT_REQUIRE(perm && src_soa && dst_soa && src_soa != dst_soa, "NULL or aliased pointer");
  long long total = (long long)n * width;
  if (scatter)
    hipLaunchKernelGGL(k_permute_soa<true>, dim3(nblocks(total)), dim3(kBlock), 0,
                       DT_STREAM(stream), (long long)n, width, perm, src_soa, dst_soa);
  else
    hipLaunchKernelGGL(k_permute_soa<false>, dim3(nblocks(total)), dim3(kBlock), 0,
                       DT_STREAM(stream), (long long)n, width, perm, src_soa, dst_soa);
  return check_launch("dt_permute_soa");
}

int dt_scatter_rows_soa(int32_t n, int32_t width, const int32_t* perm, const float* src_soa,
                        float* dst_rows, void* stream) {
  DT_REQUIRE(n >= 0 && width > 0, "bad size");
  if (n == 0) return DT_OK;
  DT_REQUIRE(src_soa && dst_rows && src_soa != dst_rows, "NULL or aliased pointer");
  long long total = (long long)n * width;
  hipLaunchKernelGGL(k_scatter_soa, dim3(nblocks(total)), dim3(kBlock), 0, DT_STREAM(stream),
                     (long long)n, width, perm, src_soa, dst_rows);
  return check_launch("dt_scatter_rows_soa");
}

}  // extern "C"
