// Device-side math shared by the MPM, FEM and contact kernels (gfx950).
//
// Conventions
//   * 3x3 matrices are row-major float[9] (m[r*3+c]), matching the reference's
//     ti.Matrix.field to_numpy() layout [.., 3, 3].
//   * Functions whose results must match the oracle bit for bit (cell base,
//     closest-triangle distance, cell-centre queries) disable FP contraction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define DT_HD __device__ __forceinline__

namespace dt {

// ----------------------------------------------------------------------------
// small dense algebra
// ----------------------------------------------------------------------------
struct M3 { float a[9]; };
struct M3d { double a[9]; };

DT_HD M3 m3_zero() { M3 r; for (int i = 0; i < 9; ++i) r.a[i] = 0.f; return r; }
DT_HD M3 m3_eye() { M3 r = m3_zero(); r.a[0] = r.a[4] = r.a[8] = 1.f; return r; }
DT_HD M3 m3_mul(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.a[i * 3 + j] = x.a[i * 3 + 0] * y.a[0 * 3 + j] + x.a[i * 3 + 1] * y.a[1 * 3 + j] +
                       x.a[i * 3 + 2] * y.a[2 * 3 + j];
  return r;
}
// x * y^T
DT_HD M3 m3_mul_bt(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.a[i * 3 + j] = x.a[i * 3 + 0] * y.a[j * 3 + 0] + x.a[i * 3 + 1] * y.a[j * 3 + 1] +
                       x.a[i * 3 + 2] * y.a[j * 3 + 2];
  return r;
}
// x^T * y
DT_HD M3 m3_mul_at(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.a[i * 3 + j] = x.a[0 * 3 + i] * y.a[0 * 3 + j] + x.a[1 * 3 + i] * y.a[1 * 3 + j] +
                       x.a[2 * 3 + i] * y.a[2 * 3 + j];
  return r;
}
DT_HD M3 m3_t(const M3& x) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = x.a[j * 3 + i];
  return r;
}
DT_HD float m3_det(const M3& m) {
  return m.a[0] * (m.a[4] * m.a[8] - m.a[5] * m.a[7]) - m.a[1] * (m.a[3] * m.a[8] - m.a[5] * m.a[6]) +
         m.a[2] * (m.a[3] * m.a[7] - m.a[4] * m.a[6]);
}
// cofactor matrix: d det(m) / d m
DT_HD M3 m3_cof(const M3& m) {
  M3 c;
  c.a[0] = m.a[4] * m.a[8] - m.a[5] * m.a[7];
  c.a[1] = m.a[5] * m.a[6] - m.a[3] * m.a[8];
  c.a[2] = m.a[3] * m.a[7] - m.a[4] * m.a[6];
  c.a[3] = m.a[2] * m.a[7] - m.a[1] * m.a[8];
  c.a[4] = m.a[0] * m.a[8] - m.a[2] * m.a[6];
  c.a[5] = m.a[1] * m.a[6] - m.a[0] * m.a[7];
  c.a[6] = m.a[1] * m.a[5] - m.a[2] * m.a[4];
  c.a[7] = m.a[2] * m.a[3] - m.a[0] * m.a[5];
  c.a[8] = m.a[0] * m.a[4] - m.a[1] * m.a[3];
  return c;
}
DT_HD M3 m3_inv(const M3& m) {
  M3 c = m3_cof(m);
  float id = 1.f / m3_det(m);
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = c.a[j * 3 + i] * id;
  return r;
}

DT_HD M3d m3d_from(const M3& x) { M3d r; for (int i = 0; i < 9; ++i) r.a[i] = (double)x.a[i]; return r; }
DT_HD M3d m3d_mul(const M3d& x, const M3d& y) {
  M3d r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.a[i * 3 + j] = x.a[i * 3 + 0] * y.a[0 * 3 + j] + x.a[i * 3 + 1] * y.a[1 * 3 + j] +
                       x.a[i * 3 + 2] * y.a[2 * 3 + j];
  return r;
}
DT_HD M3d m3d_mul_at(const M3d& x, const M3d& y) {
  M3d r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.a[i * 3 + j] = x.a[0 * 3 + i] * y.a[0 * 3 + j] + x.a[1 * 3 + i] * y.a[1 * 3 + j] +
                       x.a[2 * 3 + i] * y.a[2 * 3 + j];
  return r;
}
DT_HD M3d m3d_mul_bt(const M3d& x, const M3d& y) {
  M3d r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.a[i * 3 + j] = x.a[i * 3 + 0] * y.a[j * 3 + 0] + x.a[i * 3 + 1] * y.a[j * 3 + 1] +
                       x.a[i * 3 + 2] * y.a[j * 3 + 2];
  return r;
}

// ----------------------------------------------------------------------------
// global memory helpers
// ----------------------------------------------------------------------------
DT_HD M3 ld_m3(const float* __restrict__ p) { M3 r; for (int i = 0; i < 9; ++i) r.a[i] = p[i]; return r; }
DT_HD void st_m3(float* __restrict__ p, const M3& m) { for (int i = 0; i < 9; ++i) p[i] = m.a[i]; }
DT_HD void add_m3(float* __restrict__ p, const M3& m) { for (int i = 0; i < 9; ++i) p[i] += m.a[i]; }

DT_HD void atomic_add_f(float* p, float v) { unsafeAtomicAdd(p, v); }

// Particle array layout: element (p, c) at p*sp + c*sc.  AoS (the reference's
// [n][k] layout): sp = k, sc = 1.  SoA ([k][n], coalesced per component):
// sp = 1, sc = n.
struct PL {
  long long sp, sc;
};
DT_HD PL pl_vec(int soa, long long n) { return soa ? PL{1, n} : PL{3, 1}; }
DT_HD PL pl_mat(int soa, long long n) { return soa ? PL{1, n} : PL{9, 1}; }
DT_HD void ld_v3(const float* __restrict__ b, long long p, PL L, float (&o)[3]) {
  for (int c = 0; c < 3; ++c) o[c] = b[p * L.sp + c * L.sc];
}
DT_HD void st_v3(float* __restrict__ b, long long p, PL L, const float (&o)[3]) {
  for (int c = 0; c < 3; ++c) b[p * L.sp + c * L.sc] = o[c];
}
DT_HD M3 ld_m3L(const float* __restrict__ b, long long p, PL L) {
  M3 r;
  for (int c = 0; c < 9; ++c) r.a[c] = b[p * L.sp + c * L.sc];
  return r;
}
DT_HD void st_m3L(float* __restrict__ b, long long p, PL L, const M3& m) {
  for (int c = 0; c < 9; ++c) b[p * L.sp + c * L.sc] = m.a[c];
}

// ----------------------------------------------------------------------------
// 3x3 SVD: U, V in SO(3), sigma_1 >= sigma_2 >= |sigma_3|, sign on sigma_3.
//
// Method: cyclic Jacobi eigen-decomposition of the symmetric S = A^T A with
// exact Jacobi rotations (up to 6 sweeps, stopping at convergence), column
// sort of B = A V by norm,
// then Givens QR of B; sigma = diag(R).  This is the structure of McAdams et
// al. 2011 ("Computing the SVD of 3x3 matrices with minimal branching") with
// exact rotations instead of the quaternion approximations.  S is kept as its
// six distinct entries and each rotation uses the closed-form update of the
// symmetric Schur decomposition (Golub & Van Loan, Matrix Computations,
// sec. 8.5.2): with t = tan(theta) zeroing s_pq,
//   s_pp -= t s_pq, s_qq += t s_pq, s_pq = 0,
//   (s_rp, s_rq) <- (c s_rp - s s_rq, s s_rp + c s_rq).
// Reciprocals and square roots use the hardware v_rcp / v_rsq / v_sqrt (about
// 1 ulp): a rotation only has to be orthogonal to working precision, which
// c = rsq(1 + t^2), s = t c guarantees.
// ----------------------------------------------------------------------------
DT_HD float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
DT_HD float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
DT_HD float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
DT_HD float rcp_r(float x) { return fast_rcp(x); }
DT_HD double rcp_r(double x) { return 1.0 / x; }

// rotation angle zeroing apq in the (p, q) plane of a symmetric matrix
DT_HD void jacobi_cs(float app, float aqq, float apq, float& c, float& s, float& t) {
  c = 1.f;
  s = 0.f;
  t = 0.f;
  if (fabsf(apq) > 1e-30f) {
    const float tau = (aqq - app) * fast_rcp(2.f * apq);
    t = copysignf(fast_rcp(fabsf(tau) + fast_sqrt(1.f + tau * tau)), tau);
    c = fast_rsq(1.f + t * t);
    s = t * c;
  }
}

DT_HD void rot_cols(float (&V)[3][3], int p, int q, float c, float s) {
  for (int k = 0; k < 3; ++k) {
    const float vkp = V[k][p], vkq = V[k][q];
    V[k][p] = c * vkp - s * vkq;
    V[k][q] = s * vkp + c * vkq;
  }
}

// One rotation on the symmetric (s_pp, s_qq, s_pq, s_rp, s_rq) with r the
// third index; V accumulates the rotations.
DT_HD void jacobi_step(float& spp, float& sqq, float& spq, float& srp, float& srq,
                       float (&V)[3][3], int p, int q) {
  float c, s, t;
  jacobi_cs(spp, sqq, spq, c, s, t);
  spp -= t * spq;
  sqq += t * spq;
  spq = 0.f;
  const float rp = srp, rq = srq;
  srp = c * rp - s * rq;
  srq = s * rp + c * rq;
  rot_cols(V, p, q, c, s);
}

DT_HD void swap_cols_neg(float (&B)[3][3], float (&V)[3][3], int i, int j) {
  // swap columns i, j of B and V and negate the new column j to keep det V = +1
  for (int k = 0; k < 3; ++k) {
    float t = B[k][i]; B[k][i] = B[k][j]; B[k][j] = -t;
    t = V[k][i]; V[k][i] = V[k][j]; V[k][j] = -t;
  }
}

// Givens QR step: zero B[r][c] against B[p][c] (rows p < r); accumulates U.
DT_HD void givens_qr(float (&B)[3][3], float (&U)[3][3], int p, int r, int c) {
  float a = B[p][c], b = B[r][c];
  float d = a * a + b * b;
  float cs = 1.f, sn = 0.f;
  if (d > 1e-38f) {
    float ir = fast_rsq(d);
    cs = a * ir;
    sn = b * ir;
  }
  // rows p, r of B <- G B with G = [[cs, sn], [-sn, cs]]
  for (int k = 0; k < 3; ++k) {
    float bp = B[p][k], br = B[r][k];
    B[p][k] = cs * bp + sn * br;
    B[r][k] = -sn * bp + cs * br;
  }
  // U <- U G^T (columns p, r)
  for (int k = 0; k < 3; ++k) {
    float up = U[k][p], ur = U[k][r];
    U[k][p] = cs * up + sn * ur;
    U[k][r] = -sn * up + cs * ur;
  }
}

DT_HD void svd3(const M3& A, M3& Uo, float (&sig)[3], M3& Vo) {
  float V[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
  // S = A^T A (six distinct entries)
  auto ata = [&](int i, int j) {
    return A.a[0 * 3 + i] * A.a[0 * 3 + j] + A.a[1 * 3 + i] * A.a[1 * 3 + j] +
           A.a[2 * 3 + i] * A.a[2 * 3 + j];
  };
  float s00 = ata(0, 0), s11 = ata(1, 1), s22 = ata(2, 2);
  float s01 = ata(0, 1), s02 = ata(0, 2), s12 = ata(1, 2);
  // at most 6 sweeps; stop once the off-diagonal part is below float32
  // resolution of the diagonal (|off|^2 <= 1e-14 |diag|^2): quadratic
  // convergence makes that 2-4 sweeps
#pragma unroll 1
  for (int sweep = 0; sweep < 6; ++sweep) {
    const float off = s01 * s01 + s02 * s02 + s12 * s12;
    const float dg = s00 * s00 + s11 * s11 + s22 * s22;
    if (off <= 1e-14f * dg) break;
    jacobi_step(s00, s11, s01, s02, s12, V, 0, 1);   // r = 2: (s_20, s_21)
    jacobi_step(s00, s22, s02, s01, s12, V, 0, 2);   // r = 1: (s_10, s_12)
    jacobi_step(s11, s22, s12, s01, s02, V, 1, 2);   // r = 0: (s_01, s_02)
  }
  // B = A V
  float B[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      B[i][j] = A.a[i * 3 + 0] * V[0][j] + A.a[i * 3 + 1] * V[1][j] + A.a[i * 3 + 2] * V[2][j];
  // sort columns by squared norm, descending
  float n0 = B[0][0] * B[0][0] + B[1][0] * B[1][0] + B[2][0] * B[2][0];
  float n1 = B[0][1] * B[0][1] + B[1][1] * B[1][1] + B[2][1] * B[2][1];
  float n2 = B[0][2] * B[0][2] + B[1][2] * B[1][2] + B[2][2] * B[2][2];
  if (n0 < n1) { swap_cols_neg(B, V, 0, 1); float t = n0; n0 = n1; n1 = t; }
  if (n0 < n2) { swap_cols_neg(B, V, 0, 2); float t = n0; n0 = n2; n2 = t; }
  if (n1 < n2) { swap_cols_neg(B, V, 1, 2); float t = n1; n1 = n2; n2 = t; }
  // QR of B
  float U[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
  givens_qr(B, U, 0, 1, 0);
  givens_qr(B, U, 0, 2, 0);
  givens_qr(B, U, 1, 2, 1);
  sig[0] = B[0][0];
  sig[1] = B[1][1];
  sig[2] = B[2][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      Uo.a[i * 3 + j] = U[i][j];
      Vo.a[i * 3 + j] = V[i][j];
    }
}

// ----------------------------------------------------------------------------
// MPM stencil.  base = trunc(fl(fl(x*inv_dx) - 0.5f)); fx = fl(x*inv_dx) - base
// ----------------------------------------------------------------------------
DT_HD int cell_base1(float x, float inv_dx) {
#pragma clang fp contract(off)
  float t = x * inv_dx;
  float u = t - 0.5f;
  return (int)u;
}

struct Stencil {
  int base[3];
  float fx[3];
  float w[3][3];   // w[offset][axis]
  float dw[3][3];  // d w / d fx
};

// x0: first x-node of the particle's environment (e n_grid when environments
// are stacked along x, else 0); the position x is local to the environment.
DT_HD void make_stencil(const float (&x)[3], float inv_dx, int n_grid, Stencil& s, int x0 = 0) {
#pragma clang fp contract(off)
  for (int a = 0; a < 3; ++a) {
    float t = x[a] * inv_dx;
    int b = (int)(t - 0.5f);
    // Clamp keeps the 3x3x3 stencil inside the grid.  The reference has no
    // bounds check; particles are kept away from the walls by the boundary
    // band, so the clamp only changes cases the reference leaves undefined.
    b = b < 0 ? 0 : (b > n_grid - 3 ? n_grid - 3 : b);
    s.base[a] = a == 0 ? b + x0 : b;
    float f = t - (float)b;
    s.fx[a] = f;
    float d0 = 1.5f - f, d1 = f - 1.f, d2 = f - 0.5f;
    s.w[0][a] = 0.5f * (d0 * d0);
    s.w[1][a] = 0.75f - d1 * d1;
    s.w[2][a] = 0.5f * (d2 * d2);
    s.dw[0][a] = -d0;
    s.dw[1][a] = -2.f * d1;
    s.dw[2][a] = d2;
  }
}

// cell key used for particle sorting: 4x4x4-cell bricks in row-major order,
// cells row-major inside a brick.  key = brick * 64 + local.
DT_HD int cell_key(const float (&x)[3], float inv_dx, int n_grid, int x0 = 0) {
  int b[3];
  for (int a = 0; a < 3; ++a) {
    int v = cell_base1(x[a], inv_dx);
    b[a] = v < 0 ? 0 : (v > n_grid - 1 ? n_grid - 1 : v);
  }
  b[0] += x0;
  int nb = (n_grid + 3) >> 2;
  int brick = ((b[0] >> 2) * nb + (b[1] >> 2)) * nb + (b[2] >> 2);
  int local = (((b[0] & 3) << 2) + (b[1] & 3)) * 4 + (b[2] & 3);
  return brick * 64 + local;
}

// Binning key of a particle: the cell of its (clamped) stencil base, ordered
// like cell_key (brick * 64 + local).  Uses the base expression of
// make_stencil so the key names exactly the cell whose stencil the particle
// scatters through.
DT_HD int bin_key(const float (&x)[3], float inv_dx, int n_grid, int x0 = 0) {
  int b[3];
  for (int a = 0; a < 3; ++a) {
    int v = cell_base1(x[a], inv_dx);
    b[a] = v < 0 ? 0 : (v > n_grid - 3 ? n_grid - 3 : v);
  }
  b[0] += x0;
  int nb = (n_grid + 3) >> 2;
  int brick = ((b[0] >> 2) * nb + (b[1] >> 2)) * nb + (b[2] >> 2);
  int local = (((b[0] & 3) << 2) + (b[1] & 3)) * 4 + (b[2] & 3);
  return brick * 64 + local;
}

// ----------------------------------------------------------------------------
// materials
// ----------------------------------------------------------------------------
enum { KIND_ELASTIC = 0, KIND_VON_MISES = 1, KIND_DRUCKER_PRAGER = 2 };

struct Mat {
  float mu, lam, yield_stress, dp_alpha;
  int kind;
};

// Return map + fixed-corotated Kirchhoff stress.
//   in : F_new, U, sig, V (pre-return-map SVD), material
//   out: tau, F_p (post-return-map F)
struct StressFwd {
  M3 Fp;
  M3 R;
  float J;
  float s[3];        // clamped sigma used by the return map
  float epsn[3];     // post-return-map Hencky strain (yielded cases)
  float eh[3];       // deviatoric strain
  float ehn;         // |eh|_eps
  float trv;         // tr(eps)
  float pa, pG;      // yielded: d_k = exp(epsn_k) = pG * s_k^pa (s = max(sigma, 0.05))
  int mode;          // 0 none, 1 von Mises yield, 2 DP shear, 3 DP tension
};

DT_HD void udv(const M3& U, const float (&d)[3], const M3& V, M3& out) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      out.a[i * 3 + j] = U.a[i * 3 + 0] * d[0] * V.a[j * 3 + 0] + U.a[i * 3 + 1] * d[1] * V.a[j * 3 + 1] +
                         U.a[i * 3 + 2] * d[2] * V.a[j * 3 + 2];
}

DT_HD void stress_fwd(const M3& Fn, const M3& U, const float (&sig)[3], const M3& V, const Mat& m,
                      StressFwd& st, M3& tau) {
  st.mode = 0;
  st.Fp = Fn;
  st.pa = 0.f;
  st.pG = 1.f;
  float J;
  if (m.kind != KIND_ELASTIC) {
    float eps[3];
    for (int i = 0; i < 3; ++i) {
      st.s[i] = fmaxf(sig[i], 0.05f);
      eps[i] = __logf(st.s[i]);
    }
    float tr = eps[0] + eps[1] + eps[2];
    st.trv = tr;
    for (int i = 0; i < 3; ++i) st.eh[i] = eps[i] - tr * (1.f / 3.f);
    st.ehn = sqrtf(st.eh[0] * st.eh[0] + st.eh[1] * st.eh[1] + st.eh[2] * st.eh[2] + 1e-8f);
    // eps' = mean(eps) + pa (eps - mean(eps)), so d_k = pG s_k^pa with
    // pG = exp((1 - pa) mean(eps))
    if (m.kind == KIND_VON_MISES) {
      float dg = st.ehn - m.yield_stress / (2.f * m.mu);
      if (dg > 0.f) {
        st.mode = 1;
        float r = dg * fast_rcp(st.ehn);
        for (int i = 0; i < 3; ++i) st.epsn[i] = eps[i] - r * st.eh[i];
        st.pa = 1.f - r;
      }
    } else {
      if (tr >= 0.f) {
        st.mode = 3;
        for (int i = 0; i < 3; ++i) st.epsn[i] = 0.f;
        st.pa = 0.f;
      } else {
        float dg = st.ehn + (3.f * m.lam + 2.f * m.mu) / (2.f * m.mu) * tr * m.dp_alpha;
        if (dg > 0.f) {
          st.mode = 2;
          float r = dg * fast_rcp(st.ehn);
          for (int i = 0; i < 3; ++i) st.epsn[i] = eps[i] - r * st.eh[i];
          st.pa = 1.f - r;
        }
      }
    }
    st.pG = st.mode == 3 ? 1.f : __expf((1.f - st.pa) * tr * (1.f / 3.f));
    st.R = m3_mul_bt(U, V);
    if (st.mode == 3) {
      // cone tip: F_p = U V^T = R and J = 1 exactly, so tau = 0 exactly (as
      // U diag(1) V^T and det of it in float32 it would be rounding noise of
      // size mu 1e-7, whose adjoint is noise too)
      st.Fp = st.R;
      J = 1.f;
    } else {
      if (st.mode) {
        float d[3] = {__expf(st.epsn[0]), __expf(st.epsn[1]), __expf(st.epsn[2])};
        udv(U, d, V, st.Fp);
      }
      J = m3_det(st.Fp);
    }
  } else {
    J = sig[0] * sig[1] * sig[2];
    st.R = m3_mul_bt(U, V);
  }
  // J is kept in a scalar until here: assigning st.J in the branches made the
  // compiler keep it in private memory
  st.J = J;
  // tau = 2 mu (Fp - R) Fp^T + lam J (J - 1) I
  M3 D;
  for (int i = 0; i < 9; ++i) D.a[i] = st.Fp.a[i] - st.R.a[i];
  tau = m3_mul_bt(D, st.Fp);
  float l = m.lam * J * (J - 1.f);
  for (int i = 0; i < 9; ++i) tau.a[i] *= 2.f * m.mu;
  tau.a[0] += l; tau.a[4] += l; tau.a[8] += l;
}

// Adjoint of stress_fwd + svd (reference svd_grad) + nothing else.
//   in : gtau, gFp_extra (adjoint of F_next = F_p), forward quantities
//   out: gFn (adjoint of F_new); gmu, glam (adjoints of the material's mu and
//        lam: the stress terms and, inside the yield surface's return maps,
//        the von Mises threshold tau_y / (2 mu) and the Drucker-Prager cone
//        slope (3 lam + 2 mu) / (2 mu) alpha)
// The U/V adjoints and the svd_grad ff-terms are evaluated in float64: the
// ff_ij = 1/(s_j^2 - s_i^2) factors reach 1e8 for near-equal singular values
// and the float32 products they multiply lose every digit there.
DT_HD double clampd(double a) { return a >= 0.0 ? fmax(a, 1e-8) : fmin(a, -1e-8); }

template <bool kExact>
DT_HD void stress_adj(const M3& Fn, const M3& U, const float (&sig)[3], const M3& V, const Mat& m,
                      const StressFwd& st, const M3& gtau, const M3& gFp_in, M3& gFn, float& gmu,
                      float& glam) {
  // Scalar type of the return-map and pair-coefficient terms: the exact form's
  // terms are well conditioned (no small denominators) and run in float32;
  // the reference form keeps float64 for its 1/(s_j^2 - s_i^2) factors.
  using R = typename std::conditional<kExact, float, double>::type;
  // tau adjoint
  M3 D;
  for (int i = 0; i < 9; ++i) D.a[i] = st.Fp.a[i] - st.R.a[i];
  M3 gFp = m3_mul(gtau, st.Fp);           // gtau Fp
  M3 t2 = m3_mul_at(gtau, D);             // gtau^T (Fp - R)
  M3 gR = gFp;
  // d tau / d mu = 2 (Fp - R) Fp^T: gtau : that = 2 (gtau Fp) : (Fp - R)
  float gm = 0.f;
  for (int i = 0; i < 9; ++i) gm += gFp.a[i] * D.a[i];
  for (int i = 0; i < 9; ++i) gFp.a[i] = 2.f * m.mu * (gFp.a[i] + t2.a[i]) + gFp_in.a[i];
  for (int i = 0; i < 9; ++i) gR.a[i] *= -2.f * m.mu;
  float trg = gtau.a[0] + gtau.a[4] + gtau.a[8];
  float gJ = m.lam * (2.f * st.J - 1.f) * trg;
  gmu = 2.f * gm;
  glam = trg * st.J * (st.J - 1.f);
  if (m.kind != KIND_ELASTIC && st.mode == 3) {
    // cone tip: tau = 0 whatever F, mu and lam (F_p = R, J = 1), so it has no
    // adjoint, and F_next = R takes its adjoint to F through the rotation
    // alone (the two paths' tau terms would cancel only to rounding)
    gR = gFp_in;
    gFp = m3_zero();
    gJ = 0.f;
    gmu = glam = 0.f;
  }

  R gsig[3] = {R(0), R(0), R(0)};
  // U/V adjoints of the return-map path (F_p = U diag(d) V^T), reference form
  // only; the rotation path R = U V^T is kept apart so the exact form can use
  // its closed form
  M3d gUp, gVp;
  if (!kExact)
    for (int i = 0; i < 9; ++i) gUp.a[i] = gVp.a[i] = 0.0;
  M3 N = m3_zero();   // U^T gF_p V of the return-map path
  bool pl = false;
  R d[3] = {R(1), R(1), R(1)};
  gFn = m3_zero();

  if (m.kind == KIND_ELASTIC) {
    gsig[0] += (R)gJ * sig[1] * sig[2];
    gsig[1] += (R)gJ * sig[0] * sig[2];
    gsig[2] += (R)gJ * sig[0] * sig[1];
    gFn = gFp;
  } else {
    M3 cof = m3_cof(st.Fp);
    for (int i = 0; i < 9; ++i) gFp.a[i] += gJ * cof.a[i];
    if (st.mode == 0) {
      gFn = gFp;
    } else {
      pl = true;
      for (int k = 0; k < 3; ++k) d[k] = (R)__expf(st.epsn[k]);
      N = m3_mul(m3_mul_at(U, gFp), V);
      if (!kExact) {
        M3d Ud = m3d_from(U), Vd = m3d_from(V), gFpd = m3d_from(gFp);
        M3d gFpV = m3d_mul(gFpd, Vd);      // gFp V
        M3d gFptU = m3d_mul_at(gFpd, Ud);  // gFp^T U
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            gUp.a[i * 3 + j] = gFpV.a[i * 3 + j] * d[j];
            gVp.a[i * 3 + j] = gFptU.a[i * 3 + j] * d[j];
          }
      }
      if (st.mode != 3) {
        R ge[3];   // adjoint of post-map eps
        for (int i = 0; i < 3; ++i) ge[i] = (R)N.a[i * 4] * d[i];
        const R eh[3] = {st.eh[0], st.eh[1], st.eh[2]};
        const R n = st.ehn;
        const R gdot = ge[0] * eh[0] + ge[1] * eh[1] + ge[2] * eh[2];
        R geps[3], geh[3], gtr = R(0);
        if (st.mode == 1) {
          // eps' = eps - eh + (c/n) eh, c = tau_y / (2 mu); dc/dmu = -c / mu
          const R c = (R)m.yield_stress * rcp_r(R(2) * (R)m.mu);
          const R rn = rcp_r(n);
          for (int i = 0; i < 3; ++i) {
            geps[i] = ge[i];
            geh[i] = (c * rn - R(1)) * ge[i] - c * gdot * (rn * rn * rn) * eh[i];
          }
          gmu += (float)(-gdot * rn * c * rcp_r((R)m.mu));
        } else {
          // eps' = eps - eh - r tr eh / n, r = (3 lam + 2 mu) / (2 mu) alpha;
          // dr/dlam = 3 alpha / (2 mu), dr/dmu = -3 alpha lam / (2 mu^2)
          const R r = (R(3) * (R)m.lam + R(2) * (R)m.mu) * rcp_r(R(2) * (R)m.mu) * (R)m.dp_alpha;
          const R tr = st.trv;
          const R rn = rcp_r(n);
          for (int i = 0; i < 3; ++i) {
            geps[i] = ge[i];
            geh[i] = -ge[i] - r * tr * (ge[i] * rn - gdot * eh[i] * (rn * rn * rn));
          }
          gtr = -r * gdot * rn;
          const R gr = -tr * gdot * rn;
          const R h = R(3) * (R)m.dp_alpha * rcp_r(R(2) * (R)m.mu);
          glam += (float)(gr * h);
          gmu += (float)(-gr * h * (R)m.lam * rcp_r((R)m.mu));
        }
        const R mean = (geh[0] + geh[1] + geh[2]) * (R(1) / R(3));
        for (int i = 0; i < 3; ++i) {
          geps[i] += geh[i] - mean + gtr;
          if (sig[i] > 0.05f) gsig[i] += geps[i] * rcp_r((R)st.s[i]);
        }
      }
    }
  }
  // SVD adjoint: G = U Mi V^T.
  //   exact form (matrix products in float32, the per-pair coefficients in
  //     float64, no small denominators): the rotation path as
  //     (M_ij - M_ji) / (s_i + s_j), M = U^T gR V; the return-map path
  //     (F_p = U diag(d) V^T, d_k = G s_k^a) as N_ij A_ij + N_ji B_ij with
  //     N = U^T gF_p V, A/B = +-(d_i + d_j) / (2 (s_i + s_j)) + Dd_ij / 2 and
  //     Dd_ij = (d_j - d_i) / (s_j - s_i), near s_i = s_j from a series;
  //   reference form (float64): svd_grad of gU = gR V + gUp, gV = gR^T U + gVp
  //     with the 1e-8 clamp on s_j^2 - s_i^2 (mpm_elastic.py:157-173).
  M3 Mi;
  if (kExact) {
    const M3 Mr = m3_mul(m3_mul_at(U, gR), V);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        if (i == j) {
          Mi.a[i * 3 + j] = (float)gsig[i];
          continue;
        }
        float den = sig[i] + sig[j];
        if (fabsf(den) < 1e-12f) den = den < 0.f ? -1e-12f : 1e-12f;
        const float rden = fast_rcp(den);
        float v = (Mr.a[i * 3 + j] - Mr.a[j * 3 + i]) * rden;
        if (pl) {
          // Dd = (d_j - d_i) / (s_j - s_i) with d_k = G s_k^a: near s_i = s_j
          // from the binomial series of ((1 + x)^a - 1) / x, x = del / s_i
          // (|x| < 0.1, six terms, truncation below 1e-7 relative), else as
          // the plain divided difference of the forward pass's d
          // (s_j - s_i is exact in float32 for s within a factor 2: Sterbenz)
          const float si = sig[i], sj = sig[j], del = sj - si, a = st.pa;
          float Dd = 0.f;
          const float rsi = fast_rcp(si);
          const float x = del * rsi;
          if (si > 0.05f && sj > 0.05f && fabsf(x) < 0.1f) {
            // a (1 + (a-1)/2 x (1 + (a-2)/3 x (1 + ... (a-5)/6 x))), inside out
            constexpr float kInv[7] = {0.f, 1.f, 1.f / 2.f, 1.f / 3.f, 1.f / 4.f, 1.f / 5.f,
                                       1.f / 6.f};
            float ser = 0.f;
#pragma unroll
            for (int k = 5; k >= 1; --k) ser = (a - k) * kInv[k + 1] * (1.f + x * ser);
            Dd = (float)d[i] * rsi * a * (1.f + x * ser);
          } else if (del != 0.f) {
            Dd = (float)(d[j] - d[i]) * fast_rcp(del);
          }
          const float hs = 0.5f * (float)(d[i] + d[j]) * rden;
          v += N.a[i * 3 + j] * (hs + 0.5f * Dd) + N.a[j * 3 + i] * (-hs + 0.5f * Dd);
        }
        Mi.a[i * 3 + j] = v;
      }
  } else {
    M3d Ud = m3d_from(U), Vd = m3d_from(V), gRd = m3d_from(gR);
    double s2[3] = {(double)sig[0] * sig[0], (double)sig[1] * sig[1], (double)sig[2] * sig[2]};
    M3d gU = gUp, gV = gVp;
    M3d a = m3d_mul(gRd, Vd), b = m3d_mul_at(gRd, Ud);
    for (int i = 0; i < 9; ++i) {
      gU.a[i] += a.a[i];
      gV.a[i] += b.a[i];
    }
    M3d Xu = m3d_mul_at(Ud, gU);   // U^T gU
    M3d Xv = m3d_mul_at(Vd, gV);   // V^T gV
    M3d Mid;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double v = 0.0;
        if (i != j) {
          double ff = 1.0 / clampd(s2[j] - s2[i]);
          double xu = Xu.a[i * 3 + j] - Xu.a[j * 3 + i];
          double xv = Xv.a[i * 3 + j] - Xv.a[j * 3 + i];
          v = ff * xu * (double)sig[j] + (double)sig[i] * ff * xv;
        } else {
          v = gsig[i];
        }
        Mid.a[i * 3 + j] = v;
      }
    M3d Gd = m3d_mul_bt(m3d_mul(Ud, Mid), Vd);
    for (int i = 0; i < 9; ++i) gFn.a[i] += (float)Gd.a[i];
    return;
  }
  const M3 G = m3_mul_bt(m3_mul(U, Mi), V);
  for (int i = 0; i < 9; ++i) gFn.a[i] += G.a[i];
}

template <bool kExact>
DT_HD void stress_adj(const M3& Fn, const M3& U, const float (&sig)[3], const M3& V, const Mat& m,
                      const StressFwd& st, const M3& gtau, const M3& gFp_in, M3& gFn) {
  float gmu, glam;
  stress_adj<kExact>(Fn, U, sig, V, m, st, gtau, gFp_in, gFn, gmu, glam);
}
}  // namespace dt
