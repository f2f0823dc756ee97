// CPU restatement of one MLS-MPM substep and its adjoint, C++17 / OpenMP,
// float32 (TEST AND BASELINE INFRASTRUCTURE ONLY: the checker's CPU timing
// leg, never on the product path).
//
// Follows the reference's loop structure
// (difftactile/object_model/mpm_elastic.py, mpm_plastic.py, multi_obj.py):
//   compute_new_F   mpm_elastic.py:134-137   F_new = (I + dt C) F
//   svd             mpm_elastic.py:139-142   F_new = U S V^T, U, V in SO(3)
//   p2g             mpm_elastic.py:175-208   per-particle scatter, atomics
//                   mpm_plastic.py:150-166   von Mises return map (Hencky strain)
//   grid_op         mpm_elastic.py:216-240   dense n^3 pass, one-sided walls
//   g2p             mpm_elastic.py:243-260   per-particle gather
// and the reverse of each in the opposite order (the reference takes them
// from Taichi's AD, with its hand-written svd_grad, mpm_elastic.py:144-173),
// plus the Drucker-Prager return map of this build's multi-material model
// (Klar et al. 2016, no reference implementation).  The per-particle
// arithmetic restates oracle/mpm.py (the torch oracle) with svd_exact off;
// tests/test_oracle_cpu_mpm.py checks one substep and its adjoint against it.
//
// Layout: particles [n][3] and [n][3][3] row-major; grid node (i, j, k) at
// (i n + j) n + k.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Prm {
  int32_t n_grid;
  float dx, inv_dx, dt, p_vol, p_mass, eps;
  float gravity[3];
  int32_t bound, sticky_floor;
};

struct Mat {
  float mu, lam, yield_stress, dp_alpha;
  int32_t kind;  // 0 fixed corotated, 1 von Mises, 2 Drucker-Prager
};

using M3 = float[9];

inline void mul(const float* A, const float* B, float* C) {  // C = A B
  float t[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      t[r * 3 + c] = A[r * 3] * B[c] + A[r * 3 + 1] * B[3 + c] + A[r * 3 + 2] * B[6 + c];
  std::memcpy(C, t, sizeof t);
}
inline void mul_bt(const float* A, const float* B, float* C) {  // C = A B^T
  float t[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      t[r * 3 + c] = A[r * 3] * B[c * 3] + A[r * 3 + 1] * B[c * 3 + 1] + A[r * 3 + 2] * B[c * 3 + 2];
  std::memcpy(C, t, sizeof t);
}
inline void mul_at(const float* A, const float* B, float* C) {  // C = A^T B
  float t[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      t[r * 3 + c] = A[r] * B[c] + A[3 + r] * B[3 + c] + A[6 + r] * B[6 + c];
  std::memcpy(C, t, sizeof t);
}
inline float det3(const float* a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}
// d det / d a (the cofactor matrix)
inline void cof3(const float* a, float* c) {
  c[0] = a[4] * a[8] - a[5] * a[7];
  c[1] = a[5] * a[6] - a[3] * a[8];
  c[2] = a[3] * a[7] - a[4] * a[6];
  c[3] = a[2] * a[7] - a[1] * a[8];
  c[4] = a[0] * a[8] - a[2] * a[6];
  c[5] = a[1] * a[6] - a[0] * a[7];
  c[6] = a[1] * a[5] - a[2] * a[4];
  c[7] = a[2] * a[3] - a[0] * a[5];
  c[8] = a[0] * a[4] - a[1] * a[3];
}
// U diag(d) V^T
inline void udv(const float* U, const float* d, const float* V, float* out) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      out[r * 3 + c] = U[r * 3] * d[0] * V[c * 3] + U[r * 3 + 1] * d[1] * V[c * 3 + 1] +
                       U[r * 3 + 2] * d[2] * V[c * 3 + 2];
}

// One-sided Jacobi SVD: F = U diag(s) V^T, s sorted by magnitude (s0, s1 >=
// 0, the sign on s2), U, V rotations.
void svd3(const float* F, float* U, float* s, float* V) {
  float A[9], W[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::memcpy(A, F, sizeof A);
  for (int sweep = 0; sweep < 8; ++sweep) {
    float off = 0.f;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        float al = 0.f, be = 0.f, ga = 0.f;
        for (int r = 0; r < 3; ++r) {
          al += A[r * 3 + p] * A[r * 3 + p];
          be += A[r * 3 + q] * A[r * 3 + q];
          ga += A[r * 3 + p] * A[r * 3 + q];
        }
        off = std::max(off, std::fabs(ga) / std::sqrt(std::max(al * be, 1e-38f)));
        if (std::fabs(ga) < 1e-30f) continue;
        const float z = (be - al) / (2.f * ga);
        const float t = (z >= 0.f ? 1.f : -1.f) / (std::fabs(z) + std::sqrt(1.f + z * z));
        const float c = 1.f / std::sqrt(1.f + t * t), sn = c * t;
        for (int r = 0; r < 3; ++r) {
          const float ap = A[r * 3 + p], aq = A[r * 3 + q];
          A[r * 3 + p] = c * ap - sn * aq;
          A[r * 3 + q] = sn * ap + c * aq;
          const float wp = W[r * 3 + p], wq = W[r * 3 + q];
          W[r * 3 + p] = c * wp - sn * wq;
          W[r * 3 + q] = sn * wp + c * wq;
        }
      }
    if (off < 1e-7f) break;
  }
  float nrm[3];
  for (int c = 0; c < 3; ++c)
    nrm[c] = std::sqrt(A[c] * A[c] + A[3 + c] * A[3 + c] + A[6 + c] * A[6 + c]);
  int o[3] = {0, 1, 2};
  std::sort(o, o + 3, [&](int a, int b) { return nrm[a] > nrm[b]; });
  float As[9], Vs[9];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) As[r * 3 + c] = A[r * 3 + o[c]], Vs[r * 3 + c] = W[r * 3 + o[c]];
  if (det3(Vs) < 0.f)
    for (int r = 0; r < 3; ++r) Vs[r * 3 + 2] = -Vs[r * 3 + 2], As[r * 3 + 2] = -As[r * 3 + 2];
  float u0[3], u1[3];
  const float s0 = nrm[o[0]];
  for (int r = 0; r < 3; ++r) u0[r] = s0 > 1e-30f ? As[r * 3] / s0 : (r == 0 ? 1.f : 0.f);
  // second column: the second Jacobi column, or any unit vector normal to u0
  float s1 = nrm[o[1]];
  if (s1 > 1e-30f) {
    for (int r = 0; r < 3; ++r) u1[r] = As[r * 3 + 1] / s1;
  } else {
    const float e[3] = {std::fabs(u0[0]) < 0.9f ? 1.f : 0.f, std::fabs(u0[0]) < 0.9f ? 0.f : 1.f, 0.f};
    const float d = e[0] * u0[0] + e[1] * u0[1] + e[2] * u0[2];
    float l = 0.f;
    for (int r = 0; r < 3; ++r) u1[r] = e[r] - d * u0[r], l += u1[r] * u1[r];
    l = std::sqrt(l);
    for (int r = 0; r < 3; ++r) u1[r] /= l;
    s1 = 0.f;
  }
  const float u2[3] = {u0[1] * u1[2] - u0[2] * u1[1], u0[2] * u1[0] - u0[0] * u1[2],
                       u0[0] * u1[1] - u0[1] * u1[0]};
  for (int r = 0; r < 3; ++r) U[r * 3] = u0[r], U[r * 3 + 1] = u1[r], U[r * 3 + 2] = u2[r];
  s[0] = s0;
  s[1] = s1;
  s[2] = u2[0] * As[2] + u2[1] * As[5] + u2[2] * As[8];
  std::memcpy(V, Vs, sizeof Vs);
}

// Hencky return map pieces shared by von Mises and Drucker-Prager (shear branch)
struct Ret {
  float sc[3], eps[3], tr, eh[3], ehn, dg, k, epsn[3], d[3];
  bool changed, tension;
};

inline void return_map(const Mat& m, const float* s, Ret& r) {
  for (int i = 0; i < 3; ++i) r.sc[i] = std::max(s[i], 0.05f), r.eps[i] = std::log(r.sc[i]);
  r.tr = r.eps[0] + r.eps[1] + r.eps[2];
  float q = 0.f;
  for (int i = 0; i < 3; ++i) r.eh[i] = r.eps[i] - r.tr / 3.f, q += r.eh[i] * r.eh[i];
  r.ehn = std::sqrt(q + 1e-8f);
  r.tension = false;
  if (m.kind == 1) {
    r.dg = r.ehn - m.yield_stress / (2.f * m.mu);
    r.changed = r.dg > 0.f;
  } else {
    r.dg = r.ehn + (3.f * m.lam + 2.f * m.mu) / (2.f * m.mu) * r.tr * m.dp_alpha;
    r.tension = r.tr >= 0.f;
    r.changed = r.tension || r.dg > 0.f;
  }
  r.k = r.dg / r.ehn;
  for (int i = 0; i < 3; ++i) {
    r.epsn[i] = r.tension ? 0.f : r.eps[i] - r.k * r.eh[i];
    r.d[i] = std::exp(r.epsn[i]);
  }
}

// tau and F_p of one particle (oracle particle_stress, svd_exact off)
inline void stress(const Mat& m, const float* Fn, const float* U, const float* s, const float* V,
                   float* tau, float* Fp, float* R, float& J, Ret& ret) {
  mul_bt(U, V, R);
  ret.changed = false;
  if (m.kind == 0) {
    std::memcpy(Fp, Fn, 9 * sizeof(float));
    J = s[0] * s[1] * s[2];
  } else {
    return_map(m, s, ret);
    if (ret.changed)
      udv(U, ret.d, V, Fp);
    else
      std::memcpy(Fp, Fn, 9 * sizeof(float));
    J = det3(Fp);
  }
  float D[9], T[9];
  for (int i = 0; i < 9; ++i) D[i] = Fp[i] - R[i];
  mul_bt(D, Fp, T);
  for (int i = 0; i < 9; ++i) tau[i] = 2.f * m.mu * T[i];
  const float iso = m.lam * J * (J - 1.f);
  tau[0] += iso, tau[4] += iso, tau[8] += iso;
}

inline void weights(const float* x, float inv_dx, int* base, float (&fx)[3], float (&w)[3][3],
                    float (&dw)[3][3]) {
  for (int a = 0; a < 3; ++a) {
    base[a] = (int)(x[a] * inv_dx - 0.5f);
    fx[a] = x[a] * inv_dx - (float)base[a];
    w[0][a] = 0.5f * (1.5f - fx[a]) * (1.5f - fx[a]);
    w[1][a] = 0.75f - (fx[a] - 1.f) * (fx[a] - 1.f);
    w[2][a] = 0.5f * (fx[a] - 0.5f) * (fx[a] - 0.5f);
    dw[0][a] = -(1.5f - fx[a]);
    dw[1][a] = -2.f * (fx[a] - 1.f);
    dw[2][a] = fx[a] - 0.5f;
  }
}

inline void atomic_add(float* p, float v) {
#pragma omp atomic
  *p += v;
}

constexpr int kPart = 30;  // per-particle floats kept by the forward: F_new, U, s, V

}  // namespace

extern "C" {

// floats of the forward's workspace (read by the adjoint) and of the adjoint's scratch
long long cpu_mpm_ws_floats(const Prm* P, int n) {
  const long long nn = (long long)P->n_grid * P->n_grid * P->n_grid;
  return (long long)kPart * n + 7 * nn;
}
long long cpu_mpm_gws_floats(const Prm* P, int n) {
  (void)n;
  const long long nn = (long long)P->n_grid * P->n_grid * P->n_grid;
  return 7 * nn;
}
int cpu_mpm_threads(void) { return omp_get_max_threads(); }

int cpu_mpm_substep(const Prm* P, const Mat* mats, int n_mats, int n, const int32_t* mat_id,
                    const float* x, const float* v, const float* C, const float* F, float* x1,
                    float* v1, float* C1, float* F1, float* ws) {
  const int ng = P->n_grid;
  const long long nn = (long long)ng * ng * ng;
  float* part = ws;
  float* gmv = ws + (long long)kPart * n;
  float* gm = gmv + 3 * nn;
  float* gvo = gm + nn;
  std::memset(gmv, 0, sizeof(float) * 4 * nn);
  const float coef = -P->dt * P->p_vol * 4.f * P->inv_dx * P->inv_dx;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    const Mat& m = mats[mat_id ? std::min(std::max(mat_id[p], 0), n_mats - 1) : 0];
    const float* Cp = C + 9ll * p;
    float* Fn = part + (long long)kPart * p;
    float *U = Fn + 9, *s = Fn + 18, *V = Fn + 21;
    float A[9];
    for (int i = 0; i < 9; ++i) A[i] = P->dt * Cp[i] + (i % 4 == 0 ? 1.f : 0.f);
    mul(A, F + 9ll * p, Fn);
    svd3(Fn, U, s, V);
    float tau[9], R[9], J;
    Ret ret;
    stress(m, Fn, U, s, V, tau, F1 + 9ll * p, R, J, ret);
    float aff[9];
    for (int i = 0; i < 9; ++i) aff[i] = coef * tau[i] + P->p_mass * Cp[i];
    int b[3];
    float fx[3], w[3][3], dw[3][3];
    weights(x + 3ll * p, P->inv_dx, b, fx, w, dw);
    const float* vp = v + 3ll * p;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          const float W = w[i][0] * w[j][1] * w[k][2];
          const float dp[3] = {((float)i - fx[0]) * P->dx, ((float)j - fx[1]) * P->dx,
                               ((float)k - fx[2]) * P->dx};
          const long long node = ((long long)(b[0] + i) * ng + (b[1] + j)) * ng + (b[2] + k);
          for (int r = 0; r < 3; ++r) {
            const float mv = W * (P->p_mass * vp[r] + aff[r * 3] * dp[0] + aff[r * 3 + 1] * dp[1] +
                                  aff[r * 3 + 2] * dp[2]);
            atomic_add(gmv + 3 * node + r, mv);
          }
          atomic_add(gm + node, W * P->p_mass);
        }
  }
#pragma omp parallel for schedule(static)
  for (long long node = 0; node < nn; ++node) {
    float* out = gvo + 3 * node;
    const float mm = gm[node];
    if (!(mm > P->eps)) {
      out[0] = out[1] = out[2] = 0.f;
      continue;
    }
    const int idx[3] = {(int)(node / ((long long)ng * ng)), (int)((node / ng) % ng),
                        (int)(node % ng)};
    const float inv = 1.f / (mm + P->eps);
    for (int c = 0; c < 3; ++c) {
      float u = inv * gmv[3 * node + c] + P->dt * P->gravity[c];
      if ((idx[c] < P->bound && u < 0.f) || (idx[c] > ng - P->bound && u > 0.f)) u = 0.f;
      if (P->sticky_floor && idx[1] < P->bound) u = 0.f;
      out[c] = u;
    }
  }
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    int b[3];
    float fx[3], w[3][3], dw[3][3];
    weights(x + 3ll * p, P->inv_dx, b, fx, w, dw);
    float nv[3] = {0, 0, 0}, nC[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          const float W = w[i][0] * w[j][1] * w[k][2];
          const float dp[3] = {(float)i - fx[0], (float)j - fx[1], (float)k - fx[2]};
          const long long node = ((long long)(b[0] + i) * ng + (b[1] + j)) * ng + (b[2] + k);
          const float* g = gvo + 3 * node;
          for (int r = 0; r < 3; ++r) {
            nv[r] += W * g[r];
            for (int c = 0; c < 3; ++c) nC[r * 3 + c] += 4.f * P->inv_dx * W * g[r] * dp[c];
          }
        }
    for (int r = 0; r < 3; ++r) {
      v1[3ll * p + r] = nv[r];
      x1[3ll * p + r] = x[3ll * p + r] + P->dt * nv[r];
    }
    std::memcpy(C1 + 9ll * p, nC, sizeof nC);
  }
  return 0;
}

int cpu_mpm_substep_adj(const Prm* P, const Mat* mats, int n_mats, int n, const int32_t* mat_id,
                        const float* x, const float* v, const float* C, const float* F,
                        const float* ws, const float* gx1, const float* gv1, const float* gC1,
                        const float* gF1, float* gx, float* gv, float* gC, float* gF,
                        float* g_mats, float* gws) {
  const int ng = P->n_grid;
  const long long nn = (long long)ng * ng * ng;
  const float* part = ws;
  const float* gmv_f = ws + (long long)kPart * n;
  const float* gm_f = gmv_f + 3 * nn;
  const float* gvo = gm_f + nn;
  float* g_gvo = gws;
  float* g_mv = gws + 3 * nn;
  float* g_m = g_mv + 3 * nn;
  std::memset(g_gvo, 0, sizeof(float) * 3 * nn);
  // g2p
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    int b[3];
    float fx[3], w[3][3], dw[3][3];
    weights(x + 3ll * p, P->inv_dx, b, fx, w, dw);
    float gvt[3];
    for (int r = 0; r < 3; ++r) gvt[r] = gv1[3ll * p + r] + P->dt * gx1[3ll * p + r];
    const float* gc = gC1 + 9ll * p;
    float gfx[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          const int o[3] = {i, j, k};
          const float W = w[i][0] * w[j][1] * w[k][2];
          const float dp[3] = {(float)i - fx[0], (float)j - fx[1], (float)k - fx[2]};
          const long long node = ((long long)(b[0] + i) * ng + (b[1] + j)) * ng + (b[2] + k);
          const float* g = gvo + 3 * node;
          float dLdW = 0.f, gdp[3] = {0, 0, 0};
          for (int r = 0; r < 3; ++r) {
            float cd = 0.f;
            for (int c = 0; c < 3; ++c) {
              cd += gc[r * 3 + c] * dp[c];
              gdp[c] += 4.f * P->inv_dx * W * gc[r * 3 + c] * g[r];
            }
            atomic_add(g_gvo + 3 * node + r, W * gvt[r] + 4.f * P->inv_dx * W * cd);
            dLdW += gvt[r] * g[r] + 4.f * P->inv_dx * g[r] * cd;
          }
          for (int a = 0; a < 3; ++a) {
            float dW = dw[o[a]][a];
            for (int q = 0; q < 3; ++q)
              if (q != a) dW *= w[o[q]][q];
            gfx[a] += dLdW * dW - gdp[a];
          }
        }
    for (int a = 0; a < 3; ++a) gx[3ll * p + a] = gx1[3ll * p + a] + P->inv_dx * gfx[a];
  }
  // grid_op
#pragma omp parallel for schedule(static)
  for (long long node = 0; node < nn; ++node) {
    const float mm = gm_f[node];
    float* o = g_mv + 3 * node;
    o[0] = o[1] = o[2] = 0.f;
    g_m[node] = 0.f;
    if (!(mm > P->eps)) continue;
    const int idx[3] = {(int)(node / ((long long)ng * ng)), (int)((node / ng) % ng),
                        (int)(node % ng)};
    const float inv = 1.f / (mm + P->eps);
    float gmm = 0.f;
    for (int c = 0; c < 3; ++c) {
      const float u = inv * gmv_f[3 * node + c] + P->dt * P->gravity[c];
      const bool zero = (idx[c] < P->bound && u < 0.f) || (idx[c] > ng - P->bound && u > 0.f) ||
                        (P->sticky_floor && idx[1] < P->bound);
      if (zero) continue;
      const float g = g_gvo[3 * node + c];
      o[c] = inv * g;
      gmm -= inv * inv * gmv_f[3 * node + c] * g;
    }
    g_m[node] = gmm;
  }
  // p2g, stress, svd_grad, compute_new_F
  const float coef = -P->dt * P->p_vol * 4.f * P->inv_dx * P->inv_dx;
  const int nt = omp_get_max_threads();
  std::vector<float> gmat_t((size_t)nt * 2 * n_mats, 0.f);
#pragma omp parallel
  {
    float* gmat = gmat_t.data() + (size_t)omp_get_thread_num() * 2 * n_mats;
#pragma omp for schedule(static)
    for (int p = 0; p < n; ++p) {
      const int mi = mat_id ? std::min(std::max(mat_id[p], 0), n_mats - 1) : 0;
      const Mat& m = mats[mi];
      const float* Cp = C + 9ll * p;
      const float* Fn = part + (long long)kPart * p;
      const float *U = Fn + 9, *s = Fn + 18, *V = Fn + 21;
      float tau[9], Fp[9], R[9], J;
      Ret ret;
      stress(m, Fn, U, s, V, tau, Fp, R, J, ret);
      float aff[9];
      for (int i = 0; i < 9; ++i) aff[i] = coef * tau[i] + P->p_mass * Cp[i];
      int b[3];
      float fx[3], w[3][3], dw[3][3];
      weights(x + 3ll * p, P->inv_dx, b, fx, w, dw);
      const float* vp = v + 3ll * p;
      float gvp[3] = {0, 0, 0}, gaff[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gfx[3] = {0, 0, 0};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            const int o[3] = {i, j, k};
            const float W = w[i][0] * w[j][1] * w[k][2];
            const float dp[3] = {((float)i - fx[0]) * P->dx, ((float)j - fx[1]) * P->dx,
                                 ((float)k - fx[2]) * P->dx};
            const long long node = ((long long)(b[0] + i) * ng + (b[1] + j)) * ng + (b[2] + k);
            const float* g = g_mv + 3 * node;
            float dLdW = P->p_mass * g_m[node], gdp[3] = {0, 0, 0};
            for (int r = 0; r < 3; ++r) {
              const float ad = aff[r * 3] * dp[0] + aff[r * 3 + 1] * dp[1] + aff[r * 3 + 2] * dp[2];
              dLdW += g[r] * (P->p_mass * vp[r] + ad);
              gvp[r] += W * P->p_mass * g[r];
              for (int c = 0; c < 3; ++c) {
                gaff[r * 3 + c] += W * g[r] * dp[c];
                gdp[c] += W * aff[r * 3 + c] * g[r];
              }
            }
            for (int a = 0; a < 3; ++a) {
              float dW = dw[o[a]][a];
              for (int q = 0; q < 3; ++q)
                if (q != a) dW *= w[o[q]][q];
              gfx[a] += dLdW * dW - P->dx * gdp[a];
            }
          }
      for (int a = 0; a < 3; ++a) {
        gx[3ll * p + a] += P->inv_dx * gfx[a];
        gv[3ll * p + a] = gvp[a];
      }
      float gCp[9], gtau[9];
      for (int i = 0; i < 9; ++i) gCp[i] = P->p_mass * gaff[i], gtau[i] = coef * gaff[i];
      // tau = 2 mu (Fp - R) Fp^T + lam J (J - 1) I;  F_next = Fp
      float D[9], T1[9], T2[9], gFp[9], gR[9];
      for (int i = 0; i < 9; ++i) D[i] = Fp[i] - R[i];
      mul(gtau, Fp, T1);     // gtau Fp
      mul_at(gtau, D, T2);   // gtau^T (Fp - R)
      for (int i = 0; i < 9; ++i) {
        gFp[i] = 2.f * m.mu * (T1[i] + T2[i]) + gF1[9ll * p + i];
        gR[i] = -2.f * m.mu * T1[i];
      }
      const float trg = gtau[0] + gtau[4] + gtau[8];
      const float gJ = m.lam * (2.f * J - 1.f) * trg;
      float DF[9];
      mul_bt(D, Fp, DF);
      float gmu = 0.f;
      for (int i = 0; i < 9; ++i) gmu += 2.f * gtau[i] * DF[i];
      float glam = J * (J - 1.f) * trg;
      float gU[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gV[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      float gs[3] = {0, 0, 0}, gFn[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      if (m.kind == 0) {
        gs[0] += gJ * s[1] * s[2];
        gs[1] += gJ * s[0] * s[2];
        gs[2] += gJ * s[0] * s[1];
        for (int i = 0; i < 9; ++i) gFn[i] += gFp[i];
      } else {
        float cf[9];
        cof3(Fp, cf);
        for (int i = 0; i < 9; ++i) gFp[i] += gJ * cf[i];
        if (!ret.changed) {
          for (int i = 0; i < 9; ++i) gFn[i] += gFp[i];
        } else {
          // Fp = U diag(d) V^T
          float N[9], t[9];
          mul_at(U, gFp, t);
          mul(t, V, N);  // U^T gFp V
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
              float a = 0.f, bb = 0.f;
              for (int q = 0; q < 3; ++q) {
                a += gFp[r * 3 + q] * V[q * 3 + c];   // (gFp V)_rc
                bb += gFp[q * 3 + r] * U[q * 3 + c];  // (gFp^T U)_rc
              }
              gU[r * 3 + c] += a * ret.d[c];
              gV[r * 3 + c] += bb * ret.d[c];
            }
          if (!ret.tension) {
            float ge[3];
            for (int i = 0; i < 3; ++i) ge[i] = N[i * 4] * ret.d[i];  // d(exp)/d eps_new
            // eps_new = eps - k eh, k = dg / ehn
            float geps[3], geh[3], gk = 0.f;
            for (int i = 0; i < 3; ++i) {
              geps[i] = ge[i];
              gk -= ge[i] * ret.eh[i];
              geh[i] = -ret.k * ge[i];
            }
            const float gdg = gk / ret.ehn;
            float gehn = -gk * ret.dg / (ret.ehn * ret.ehn);
            float gtr = 0.f;
            if (m.kind == 1) {
              gehn += gdg;
              gmu += gdg * m.yield_stress / (2.f * m.mu * m.mu);
            } else {
              const float c = (3.f * m.lam + 2.f * m.mu) / (2.f * m.mu);
              gehn += gdg;
              gtr += gdg * c * m.dp_alpha;
              const float gc = gdg * m.dp_alpha * ret.tr;
              gmu += gc * (-3.f * m.lam / (2.f * m.mu * m.mu));
              glam += gc * (3.f / (2.f * m.mu));
            }
            for (int i = 0; i < 3; ++i) geh[i] += gehn * ret.eh[i] / ret.ehn;
            const float sh = (geh[0] + geh[1] + geh[2]) / 3.f;
            for (int i = 0; i < 3; ++i) geps[i] += geh[i] - sh + gtr;
            for (int i = 0; i < 3; ++i)
              if (s[i] > 0.05f) gs[i] += geps[i] / ret.sc[i];
          }
        }
      }
      // R = U V^T
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          float a = 0.f, bb = 0.f;
          for (int q = 0; q < 3; ++q) {
            a += gR[r * 3 + q] * V[q * 3 + c];
            bb += gR[q * 3 + r] * U[q * 3 + c];
          }
          gU[r * 3 + c] += a;
          gV[r * 3 + c] += bb;
        }
      // svd_grad (mpm_elastic.py:157-173)
      {
        float ff[9];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            if (i == j) {
              ff[i * 3 + j] = 0.f;
              continue;
            }
            float d = s[j] * s[j] - s[i] * s[i];
            d = d >= 0.f ? std::max(d, 1e-8f) : std::min(d, -1e-8f);
            ff[i * 3 + j] = 1.f / d;
          }
        float A1[9], A2[9], K[9], t[9], out[9];
        mul_at(U, gU, A1);
        mul_at(gU, U, A2);
        for (int i = 0; i < 9; ++i) K[i] = ff[i] * (A1[i] - A2[i]);
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) t[r * 3 + c] = K[r * 3 + c] * s[c];  // K S
        mul(U, t, out);
        mul_bt(out, V, out);  // U (K S) V^T
        for (int i = 0; i < 9; ++i) gFn[i] += out[i];
        mul_at(V, gV, A1);
        mul_at(gV, V, A2);
        for (int i = 0; i < 9; ++i) K[i] = ff[i] * (A1[i] - A2[i]);
        mul_bt(K, V, t);  // K V^T
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) t[r * 3 + c] *= s[r];  // S (K V^T)
        mul(U, t, out);
        for (int i = 0; i < 9; ++i) gFn[i] += out[i];
        udv(U, gs, V, out);
        for (int i = 0; i < 9; ++i) gFn[i] += out[i];
      }
      // F_new = (I + dt C) F
      const float* Fq = F + 9ll * p;
      float A[9], t[9];
      mul_bt(gFn, Fq, t);
      for (int i = 0; i < 9; ++i) gC[9ll * p + i] = gCp[i] + P->dt * t[i];
      for (int i = 0; i < 9; ++i) A[i] = P->dt * Cp[i] + (i % 4 == 0 ? 1.f : 0.f);
      mul_at(A, gFn, gF + 9ll * p);
      gmat[2 * mi] += gmu;
      gmat[2 * mi + 1] += glam;
    }
  }
  if (g_mats)
    for (int t = 0; t < nt; ++t)
      for (int i = 0; i < 2 * n_mats; ++i) g_mats[i] += gmat_t[(size_t)t * 2 * n_mats + i];
  return 0;
}

}  // extern "C"
