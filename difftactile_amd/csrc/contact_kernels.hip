// Grid penalty contact between the MPM object and an FEM pad (gfx950).
//
// Reference (difftactile/):
//   find_closest            sensor_model/fem_sensor.py:343-359
//   find_sdf                sensor_model/fem_sensor.py:361-394
//   update_contact_force    sensor_model/fem_sensor.py:401-406, object_model/mpm_elastic.py:130-132
//   calculate_contact_force tasks/grasp_plastic.py:204-216 (grasp form) and
//                           tasks/cable_straightening.py:211-223 (cable form:
//                           no max(., 0) on n.v_rel, shear norm without eps)
//   check_collision         tasks/grasp_plastic.py:269-276
//   collision               tasks/grasp_plastic.py:279-299
//
// The closest-triangle search is the reference's brute force (first minimum
// of the centroid distance wins), evaluated with one fixed float32 rounding
// rule so it is reproducible bit for bit; chunks of unoccupied nodes are
// skipped whole.
#include "dt_common.h"
#include "dt_device.h"

namespace dt {

struct ContactP {
  float kn, kd, kt, mu_f, norm_eps, sdf_offset;
  int unclamped_vn;   // 1: the cable form (no max(., 0) on the normal velocity)
};

// Lower bound on the squared centroid distance from the expanded form,
// lb = (1 - 4e-6)(|c|^2 + |q|^2) - 2 c.q: three FMAs per triangle.  With
// S = |c|^2 + |q|^2 and u = 2^-24, its rounding error is below 13u S (about
// 13 roundings of terms bounded by 2S), while the float32 squared distance the
// search compares is at least |c - q|^2 - 12u S (six relative roundings of a
// value at most 2S).  So lb < |c - q|^2 - (4e-6 - 25u) S < that distance, and a
// triangle with lb >= the node's best squared distance cannot win.
constexpr float kBoundShrink = 1.0f - 4e-6f;

// centroid c = fl32(1/3) * ((p1 + p2) + p3), no contraction
__global__ __launch_bounds__(kBlock) void k_centroids(int nt, const int32_t* __restrict__ tris,
                                                      const float* __restrict__ pos,
                                                      float4* __restrict__ cen) {
#pragma clang fp contract(off)
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
  const float third = 1.0f / 3.0f;
  float x = third * ((pos[3ll * a] + pos[3ll * b]) + pos[3ll * c]);
  float y = third * ((pos[3ll * a + 1] + pos[3ll * b + 1]) + pos[3ll * c + 1]);
  float z = third * ((pos[3ll * a + 2] + pos[3ll * b + 2]) + pos[3ll * c + 2]);
  // |c|^2 shrunk by 4e-6: part of the search's lower bound on a squared distance
  cen[t] = make_float4(x, y, z, kBoundShrink * ((x * x + y * y) + z * z));
}

__device__ __forceinline__ float node_centre(int i, float dx) {
#pragma clang fp contract(off)
  return ((float)i + 0.5f) * dx;
}

// squared centroid distance plus the 1e-10 of the reference's norm: its
// square root is the distance the search compares
__device__ __forceinline__ float centroid_dist2(const float4 c, float qx, float qy, float qz) {
#pragma clang fp contract(off)
  float d0 = c.x - qx, d1 = c.y - qy, d2 = c.z - qz;
  float dd = (d0 * d0 + d1 * d1) + d2 * d2;
  return dd + 1e-10f;
}

// check_collision for one pad: contact_idx[node*stride + col] = closest triangle
// for occupied nodes (grid_m[node*m_stride] > eps); other nodes are untouched.
// tri_start (environments stacked along x): the node of environment e = i /
// n_grid searches only triangles [tri_start[e], tri_start[e+1]) (its own pad)
// and positions are local to the environment; NULL: all nt triangles.
//
// The reference scans triangles in index order from (100, -1) and keeps a
// distance only when it is strictly smaller: the result is the smallest index
// among those at the minimum distance (-1 if none is below 100).  That is the
// lexicographic minimum of (distance, index), which no visiting order
// changes, so the search here visits triangles in whatever order skips most
// of them:
//
// * a block takes one chunk of 64 nodes at a time (an active brick, or 64
//   consecutive nodes); the chunk's triangle range is cut into runs of
//   kTriBatch triangles and groups of kTriGroup runs, its eight wavefronts
//   take the groups in turn (wave w: groups w, w + 8, ...), reading centroids
//   at wave-uniform indices (scalar loads), and the waves' minima are
//   combined at the end;
// * a node may start from a hinted triangle (its result of the previous
//   substep, say: any index in its range is a valid start), which makes its
//   best distance tight from the first run on;
// * with bounds (the bounding box of each group's and each run's
//   centroids), a group, and then a run, whose box lies beyond the best
//   distance of all 64 nodes is skipped without loading what it holds (boxes
//   rather than spheres: a pad's surface is flat on the scale of a run, and a
//   flat box stays tight across it, so a node deep under a face visits only
//   the runs around its foot point);
// * inside a visited run, sets of four centroids whose distance lower bounds
//   (centroid .w) all exceed the best are skipped.
//
// Skips use bounds with relative margins far above float32 rounding (stated
// at each test), so a skipped triangle's float32 distance is strictly larger
// than the best one: it could neither win nor tie.
constexpr int kClosestThreads = 512;
constexpr int kClosestWaves = kClosestThreads / 64;
constexpr int kTriBatch = DT_CONTACT_TRI_RUN;
constexpr int kTriGroup = DT_CONTACT_TRI_GROUP;

// best = lexicographic minimum of (distance, index) so far, best_s its squared
// distance, thr_s = best_s (1 + 1e-6): every s whose square root (correctly
// rounded, so monotone) is at most best lies below thr_s.
struct Closest {
  float best, best_s, thr_s;
  int best_k;
};

template <bool kPerLane>
__device__ __forceinline__ void closest_step(const float4 c, int t, int lo, int hi, float qx,
                                             float qy, float qz, Closest& r) {
  const float sq = centroid_dist2(c, qx, qy, qz);
  if (sq <= r.thr_s && (!kPerLane || (t >= lo && t < hi))) {
    const float d = sqrtf(sq);
    if (d < r.best || (d == r.best && t < r.best_k)) {
      r.best = d;
      r.best_s = sq;
      r.best_k = t;
      r.thr_s = sq * (1.0f + 1e-6f);
    }
  }
}

// kTriBatch centroids from t (wave-uniform), loaded eight at a time (scalar
// loads issued together); groups of four whose lower bounds lb (kBoundShrink,
// qb = (1 - 4e-6)|q|^2) all exceed thr_s are skipped (s > lb > thr_s)
template <bool kPerLane>
__device__ __forceinline__ void closest_batch(const float4* __restrict__ cen, int t, int lo, int hi,
                                              float qx, float qy, float qz, float qb,
                                              Closest& r) {
  constexpr int kG = 4, kL = 8;
  const float mx = -2.f * qx, my = -2.f * qy, mz = -2.f * qz;
#pragma unroll
  for (int h = 0; h < kTriBatch; h += kL) {
    float4 c[kL];
#pragma unroll
    for (int u = 0; u < kL; ++u) c[u] = cen[t + h + u];
#pragma unroll
    for (int g = 0; g < kL; g += kG) {
      float lb = 3.4e38f;
#pragma unroll
      for (int u = g; u < g + kG; ++u)
        lb = fminf(lb, fmaf(c[u].x, mx, fmaf(c[u].y, my, fmaf(c[u].z, mz, c[u].w + qb))));
      if (lb <= r.thr_s) {
#pragma unroll
        for (int u = g; u < g + kG; ++u)
          closest_step<kPerLane>(c[u], t + h + u, lo, hi, qx, qy, qz, r);
      }
    }
  }
}

// lexicographic (distance, index) order; index -1 (nothing found) is larger
// than any found triangle
__device__ __forceinline__ bool closest_before(float d, int k, float bd, int bk) {
  if (k < 0) return false;
  if (bk < 0) return true;
  return d < bd || (d == bd && k < bk);
}

__device__ __forceinline__ int wave_min_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// Whether some lane that takes triangles of [t0, t1) (its own range [lo, hi))
// may find one at or below its best inside the box b[0] (lower corner, .w
// unused) .. b[1] (upper corner).  The box is out of a lane's reach when the
// float32 squared distance dd from q to the box exceeds T^2, T = best
// (1 + 1e-5): each of dd's per-axis gaps is one correctly rounded difference
// and dd adds three squares (a few 2^-24 relative in all), so every centroid
// c inside is then more than best (1 + 0.99e-5) from q, and its float32
// distance (six roundings more) is still above best: it could neither win
// nor tie.
__device__ __forceinline__ bool box_reachable(const float4* __restrict__ b, int t0, int t1, int lo,
                                              int hi, float qx, float qy, float qz, float best) {
  const float4 l = b[0], u = b[1];
  const float g0 = fmaxf(fmaxf(l.x - qx, qx - u.x), 0.f);
  const float g1 = fmaxf(fmaxf(l.y - qy, qy - u.y), 0.f);
  const float g2 = fmaxf(fmaxf(l.z - qz, qz - u.z), 0.f);
  const float dd = g0 * g0 + g1 * g1 + g2 * g2;
  const float reach = best * (1.0f + 1e-5f);
  return __ballot(t1 > t0 && lo < t1 && hi > t0 && !(dd > reach * reach)) != 0ull;
}

// Wave w's part of one chunk's search over the triangles [t_lo, t_hi): the
// groups of runs overlapping it, w, w + kClosestWaves, ... of them.  bounds
// (kBounds): the boxes of dt_contact_tri_batches, two float4 each, n_runs run
// boxes then the group boxes.  mixed: some lane's range [lo, hi) is not [t_lo, t_hi)
// (each triangle is then checked against the lane's own range).
template <bool kBounds>
__device__ __forceinline__ void closest_search(const float4* __restrict__ cen,
                                               const float4* __restrict__ bounds, int n_runs,
                                               int w, int t_lo, int t_hi, int lo, int hi,
                                               bool mixed, float qx, float qy, float qz,
                                               Closest& r) {
  const float qb = kBoundShrink * ((qx * qx + qy * qy) + qz * qz);
  const int b_lo = t_lo / kTriBatch;
  const int b_hi = t_hi > t_lo ? (t_hi + kTriBatch - 1) / kTriBatch : b_lo;
  const int g_lo = b_lo / kTriGroup;
  const int g_hi = b_hi > b_lo ? (b_hi + kTriGroup - 1) / kTriGroup : g_lo;
  for (int g = g_lo + w; g < g_hi; g += kClosestWaves) {
    const int r0 = max(g * kTriGroup, b_lo), r1 = min(g * kTriGroup + kTriGroup, b_hi);
    if (kBounds &&
        !box_reachable(bounds + 2 * (n_runs + g), max(r0 * kTriBatch, t_lo),
                       min(r1 * kTriBatch, t_hi), lo, hi, qx, qy, qz, r.best))
      continue;
    for (int b = r0; b < r1; ++b) {
      const int t0 = max(b * kTriBatch, t_lo), t1 = min(b * kTriBatch + kTriBatch, t_hi);
      if (kBounds && !box_reachable(bounds + 2 * b, t0, t1, lo, hi, qx, qy, qz, r.best)) continue;
      if (t1 - t0 == kTriBatch) {
        if (mixed)
          closest_batch<true>(cen, t0, lo, hi, qx, qy, qz, qb, r);
        else
          closest_batch<false>(cen, t0, lo, hi, qx, qy, qz, qb, r);
      } else {
        for (int t = t0; t < t1; ++t) closest_step<true>(cen[t], t, lo, hi, qx, qy, qz, r);
      }
    }
  }
}

template <bool kBounds>
__global__ __launch_bounds__(kClosestThreads) void k_closest(MpmP P, long long nn,
                                                    const float* __restrict__ gm, int m_stride,
                                                    int nt, const int32_t* __restrict__ tri_start,
                                                    const float4* __restrict__ cen,
                                                    const float4* __restrict__ bounds,
                                                    const int32_t* hint, int32_t* idx,
                                                    int stride, int col) {
  __shared__ float s_best[kClosestWaves][64];
  __shared__ int s_k[kClosestWaves][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n = P.n_grid;
  const long long n_chunks = P.alist ? (long long)*P.n_alist : (nn + 63) / 64;
  for (long long ch = blockIdx.x; ch < n_chunks; ch += gridDim.x) {
    // every wave of the block holds the same 64 nodes, so the tests below
    // agree across the block and the barriers are uniform
    // node coordinates from the brick and lane (or the node index)
    int i, j, k;
    long long nd;
    if (P.alist) {
      const int nb = (n + 3) >> 2, b = P.alist[ch];
      const int bz = b % nb, by = (b / nb) % nb, bx = b / (nb * nb);
      i = bx * 4 + (lane >> 4), j = by * 4 + ((lane >> 2) & 3), k = bz * 4 + (lane & 3);
      nd = (i < P.nx && j < n && k < n) ? ((long long)i * n + j) * n + k : -1;
    } else {
      nd = ch * 64 + lane;
      if (nd >= nn) nd = -1;
      const long long nn2 = (long long)n * n;
      i = (int)(nd / nn2), j = (int)((nd / n) % n), k = (int)(nd % n);
    }
    const bool occ = nd >= 0 && gm[nd * m_stride] > P.eps;
    if (__ballot(occ) == 0ull) continue;
    float qx = 0.f, qy = 0.f, qz = 0.f;
    int lo = 0, hi = nt;
    if (nd >= 0) {
      if (tri_start) {
        const int e = i / n;
        lo = tri_start[e];
        hi = tri_start[e + 1];
      }
      const float dxe = at_xnode(P, i).dx;   // the node's environment
      qx = node_centre(i % n, dxe);
      qy = node_centre(j, dxe);
      qz = node_centre(k, dxe);
    }
    // the chunk searches the union of its occupied nodes' triangle ranges; a
    // node outside the common range checks each triangle against its own
    int t_lo = 0, t_hi = nt;
    bool mixed = false;
    if (tri_start) {
      t_lo = __builtin_amdgcn_readfirstlane(wave_min_i(occ ? lo : 0x7fffffff));
      t_hi = __builtin_amdgcn_readfirstlane(wave_max_i(occ ? hi : 0));
      mixed = __ballot(occ && (lo != t_lo || hi != t_hi)) != 0ull;
    }
    if (!occ) lo = hi = 0;   // idle lanes take no triangle
    Closest r{100.0f, 1e4f, 1e4f * (1.0f + 1e-6f), -1};
    // seed from the hint (hint and idx may be the same array: every read of
    // a node's entry precedes the barrier before its write)
    if (hint && occ) {
      const int h = hint[nd * stride + col];
      if (h >= lo && h < hi) closest_step<false>(cen[h], h, lo, hi, qx, qy, qz, r);
    }
    closest_search<kBounds>(cen, bounds, (nt + kTriBatch - 1) / kTriBatch, w, t_lo, t_hi, lo, hi,
                            mixed, qx, qy, qz, r);
    s_best[w][lane] = r.best;
    s_k[w][lane] = r.best_k;
    __syncthreads();
    if (w == 0 && occ) {
      float bd = 100.0f;
      int bk = -1;
      for (int q = 0; q < kClosestWaves; ++q)
        if (closest_before(s_best[q][lane], s_k[q][lane], bd, bk)) {
          bd = s_best[q][lane];
          bk = s_k[q][lane];
        }
      idx[nd * stride + col] = bk;
    }
    __syncthreads();
  }
}

// Bounding box of each run of kTriBatch centroids and of each group of
// kTriGroup runs (exact minima and maxima), two float4 each: the lower corner
// (.w = 0) then the upper.  One thread per run; the threads of a group are
// kTriGroup consecutive lanes of a wavefront, and threads past the last run
// take an empty run.
__global__ __launch_bounds__(kBlock) void k_tri_batches(int nt, const float4* __restrict__ cen,
                                                        float4* __restrict__ bounds) {
  static_assert(64 % kTriGroup == 0 && (kTriGroup & (kTriGroup - 1)) == 0, "group of lanes");
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int n_runs = (nt + kTriBatch - 1) / kTriBatch;
  const int t0 = min(b * kTriBatch, nt), t1 = min(t0 + kTriBatch, nt);
  float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  for (int t = t0; t < t1; ++t) {
    const float4 c = cen[t];
    lo[0] = fminf(lo[0], c.x), hi[0] = fmaxf(hi[0], c.x);
    lo[1] = fminf(lo[1], c.y), hi[1] = fmaxf(hi[1], c.y);
    lo[2] = fminf(lo[2], c.z), hi[2] = fmaxf(hi[2], c.z);
  }
  float glo[3] = {lo[0], lo[1], lo[2]}, ghi[3] = {hi[0], hi[1], hi[2]};
  for (int o = 1; o < kTriGroup; o <<= 1)
    for (int a = 0; a < 3; ++a) {
      glo[a] = fminf(glo[a], __shfl_xor(glo[a], o));
      ghi[a] = fmaxf(ghi[a], __shfl_xor(ghi[a], o));
    }
  if (b >= n_runs) return;
  bounds[2 * b] = make_float4(lo[0], lo[1], lo[2], 0.f);
  bounds[2 * b + 1] = make_float4(hi[0], hi[1], hi[2], 0.f);
  if (b % kTriGroup == 0) {
    const int g = n_runs + b / kTriGroup;
    bounds[2 * g] = make_float4(glo[0], glo[1], glo[2], 0.f);
    bounds[2 * g + 1] = make_float4(ghi[0], ghi[1], ghi[2], 0.f);
  }
}

// persistent grid of the closest-triangle search: one block per chunk up to a cap
static unsigned closest_blocks(const dt_mpm_params& p, long long nn) {
  long long c;
  if (p.active_bricks && p.n_active_bricks) {
    const long long nb = (p.n_grid + 3) >> 2;
    const long long nbx = ((long long)(p.n_env > 1 ? p.n_env : 1) * p.n_grid + 3) >> 2;
    c = nbx * nb * nb;
  } else {
    c = (nn + 63) / 64;
  }
  if (c < 1) c = 1;
  return (unsigned)(c < 8192 ? c : 8192);
}

// ---------------------------------------------------------------------------
// per-node contact force (forward) and its adjoint
// ---------------------------------------------------------------------------
struct ContactFwd {
  float p1[3], e1[3], e2[3], cr[3], s, sgn, n[3], sdf;
  float r[3];     // -rel_v, the relative velocity argument of the force
  float vnr, vn, fn[3], fnn, vt[3], vtn, mag, f[3];
  bool flag, shear, kt_branch;
};

__device__ __forceinline__ void contact_fwd(const ContactP& C, const float (&q)[3],
                                            const float (&qv)[3], const float* __restrict__ pos,
                                            const float* __restrict__ vel, const int (&tv)[3],
                                            const float (&out_dir)[3], ContactFwd& w) {
  float p1[3], p2[3], p3[3], v1[3], v2[3], v3[3];
  for (int a = 0; a < 3; ++a) {
    p1[a] = pos[3ll * tv[0] + a];
    p2[a] = pos[3ll * tv[1] + a];
    p3[a] = pos[3ll * tv[2] + a];
    v1[a] = vel[3ll * tv[0] + a];
    v2[a] = vel[3ll * tv[1] + a];
    v3[a] = vel[3ll * tv[2] + a];
    w.p1[a] = p1[a];
    w.e1[a] = p2[a] - p1[a];
    w.e2[a] = p3[a] - p1[a];
  }
  w.cr[0] = w.e1[1] * w.e2[2] - w.e1[2] * w.e2[1];
  w.cr[1] = w.e1[2] * w.e2[0] - w.e1[0] * w.e2[2];
  w.cr[2] = w.e1[0] * w.e2[1] - w.e1[1] * w.e2[0];
  w.s = sqrtf(w.cr[0] * w.cr[0] + w.cr[1] * w.cr[1] + w.cr[2] * w.cr[2] + 1e-10f);
  float nrm[3] = {w.cr[0] / w.s, w.cr[1] / w.s, w.cr[2] / w.s};
  float dn = nrm[0] * out_dir[0] + nrm[1] * out_dir[1] + nrm[2] * out_dir[2];
  w.sgn = dn > 0.f ? 1.f : (dn < 0.f ? -1.f : 0.f);
  for (int a = 0; a < 3; ++a) w.n[a] = w.sgn * nrm[a];
  float ds[3] = {q[0] - p1[0], q[1] - p1[1], q[2] - p1[2]};
  float sdf = ds[0] * w.n[0] + ds[1] * w.n[1] + ds[2] * w.n[2];
  float proj[3] = {q[0] - sdf * w.n[0], q[1] - sdf * w.n[1], q[2] - sdf * w.n[2]};
  // barycentric on v0 = p3 - p1, v1 = p2 - p1
  float vv2[3] = {proj[0] - p1[0], proj[1] - p1[1], proj[2] - p1[2]};
  float dot00 = w.e2[0] * w.e2[0] + w.e2[1] * w.e2[1] + w.e2[2] * w.e2[2];
  float dot01 = w.e2[0] * w.e1[0] + w.e2[1] * w.e1[1] + w.e2[2] * w.e1[2];
  float dot02 = w.e2[0] * vv2[0] + w.e2[1] * vv2[1] + w.e2[2] * vv2[2];
  float dot11 = w.e1[0] * w.e1[0] + w.e1[1] * w.e1[1] + w.e1[2] * w.e1[2];
  float dot12 = w.e1[0] * vv2[0] + w.e1[1] * vv2[1] + w.e1[2] * vv2[2];
  float inv_d = 1.f / (dot00 * dot11 - dot01 * dot01);
  float u = (dot11 * dot02 - dot01 * dot12) * inv_d;
  float v = (dot00 * dot12 - dot01 * dot02) * inv_d;
  w.sdf = sdf - C.sdf_offset;
  for (int a = 0; a < 3; ++a) {
    float rel = qv[a] - (1.f / 3.f) * (v1[a] + v2[a] + v3[a]);
    w.r[a] = -rel;
  }
  w.flag = (w.sdf < 0.f) && (u >= 0.f) && (v >= 0.f) && (u + v <= 1.f);
  // calculate_contact_force(sdf, n, r): grasp form, or the cable form when
  // C.unclamped_vn (with C.norm_eps = 0)
  w.vnr = w.n[0] * w.r[0] + w.n[1] * w.r[1] + w.n[2] * w.r[2];
  w.vn = C.unclamped_vn ? w.vnr : fmaxf(w.vnr, 0.f);
  float kk = -(C.kn + C.kd * w.vn) * w.sdf;
  for (int a = 0; a < 3; ++a) w.fn[a] = kk * w.n[a];
  for (int a = 0; a < 3; ++a) w.vt[a] = w.r[a] - w.vnr * w.n[a];
  w.vtn = sqrtf(w.vt[0] * w.vt[0] + w.vt[1] * w.vt[1] + w.vt[2] * w.vt[2] + C.norm_eps);
  w.fnn = sqrtf(w.fn[0] * w.fn[0] + w.fn[1] * w.fn[1] + w.fn[2] * w.fn[2]);
  w.shear = w.vtn > 1e-4f;
  float a1 = C.kt * w.vtn, a2 = C.mu_f * w.fnn;
  w.kt_branch = a1 < a2;
  w.mag = fminf(a1, a2);
  for (int a = 0; a < 3; ++a) w.f[a] = w.fn[a] + (w.shear ? w.vt[a] / w.vtn * w.mag : 0.f);
}

// collision for one pad.  grid: momentum mv (stride smv), mass m (stride sm),
// impulse grid_f [n^3,3] (+= f*dt); pad external force ext [nv,3] (+= -f/3).
__global__ __launch_bounds__(kBlock) void k_collision(MpmP P, ContactP C, long long nn,
                                                      const float* __restrict__ mv, int smv,
                                                      const float* __restrict__ m, int sm,
                                                      const int32_t* __restrict__ idx, int stride,
                                                      int col, const int32_t* __restrict__ tris,
                                                      const float* __restrict__ pos,
                                                      const float* __restrict__ vel, float o0,
                                                      float o1, float o2,
                                                      float* __restrict__ gridf,
                                                      float* __restrict__ ext) {
  node_loop(P, nn, [&](long long nd) {
    float mm = m[nd * sm];
    if (!(mm > P.eps)) return;
    int k = idx[nd * stride + col];
    if (k < 0) return;
    const int n = P.n_grid;
    int i = (int)(nd / ((long long)n * n)), j = (int)((nd / n) % n), kk = (int)(nd % n);
    const float dxe = at_xnode(P, i).dx;
    float q[3] = {node_centre(i % n, dxe), node_centre(j, dxe), node_centre(kk, dxe)};
    float inv = 1.f / (mm + P.eps);
    float qv[3] = {mv[nd * smv] * inv, mv[nd * smv + 1] * inv, mv[nd * smv + 2] * inv};
    int tv[3] = {tris[3 * k], tris[3 * k + 1], tris[3 * k + 2]};
    float od[3] = {o0, o1, o2};
    ContactFwd w;
    contact_fwd(C, q, qv, pos, vel, tv, od, w);
    if (!w.flag) return;
    for (int a = 0; a < 3; ++a) {
      gridf[3 * nd + a] += w.f[a] * P.dt;
      float e = -w.f[a] / 3.f;
      atomic_add_f(ext + 3ll * tv[0] + a, e);
      atomic_add_f(ext + 3ll * tv[1] + a, e);
      atomic_add_f(ext + 3ll * tv[2] + a, e);
    }
  });
}

// Adjoint of one contact force given gf = dL/df: the adjoints of the pad's
// three vertex velocities (gvv, the same for the three) and positions (gp[c]
// for corner c), and of the query point q and its velocity qv.
__device__ __forceinline__ void contact_adj_terms(const ContactP& C, const float (&q)[3],
                                                  const ContactFwd& w, const float (&gf)[3],
                                                  float (&gvv)[3], float (&gp)[3][3],
                                                  float (&gq)[3], float (&gqv)[3]) {
  float gfn[3] = {gf[0], gf[1], gf[2]};
  float gvt[3] = {0.f, 0.f, 0.f};
  if (w.shear) {
    // ft = vt * (mag / vtn)
    float s = gf[0] * w.vt[0] + gf[1] * w.vt[1] + gf[2] * w.vt[2];
    float ratio = w.mag / w.vtn;
    for (int a = 0; a < 3; ++a) gvt[a] += ratio * gf[a];
    float gmag = s / w.vtn;
    float gvtn = -s * w.mag / (w.vtn * w.vtn);
    if (w.kt_branch) {
      gvtn += C.kt * gmag;
    } else if (w.fnn > 0.f) {
      float gfnn = C.mu_f * gmag;
      for (int a = 0; a < 3; ++a) gfn[a] += gfnn * w.fn[a] / w.fnn;
    }
    for (int a = 0; a < 3; ++a) gvt[a] += gvtn * w.vt[a] / w.vtn;
  }
  // vt = r - vnr n
  float gr[3], gn[3], gvnr = 0.f;
  for (int a = 0; a < 3; ++a) {
    gr[a] = gvt[a];
    gvnr -= gvt[a] * w.n[a];
    gn[a] = -w.vnr * gvt[a];
  }
  // fn = -(kn + kd vn) sdf n
  float kk2 = -(C.kn + C.kd * w.vn) * w.sdf;
  float gfn_n = gfn[0] * w.n[0] + gfn[1] * w.n[1] + gfn[2] * w.n[2];
  for (int a = 0; a < 3; ++a) gn[a] += kk2 * gfn[a];
  float gsdf = -(C.kn + C.kd * w.vn) * gfn_n;
  float gvn = -C.kd * w.sdf * gfn_n;
  if (C.unclamped_vn || w.vnr > 0.f) gvnr += gvn;
  // vnr = n . r
  for (int a = 0; a < 3; ++a) {
    gn[a] += gvnr * w.r[a];
    gr[a] += gvnr * w.n[a];
  }
  // r = -qv + (1/3) sum vel
  for (int a = 0; a < 3; ++a) {
    gqv[a] = -gr[a];
    gvv[a] = gr[a] / 3.f;
  }
  // sdf = (q - p1) . n - offset
  float gp1[3], ge1[3], ge2[3];
  for (int a = 0; a < 3; ++a) {
    gn[a] += gsdf * (q[a] - w.p1[a]);
    gp1[a] = -gsdf * w.n[a];
    gq[a] = gsdf * w.n[a];
  }
  // n = sgn * cr / s
  float gnr[3] = {w.sgn * gn[0], w.sgn * gn[1], w.sgn * gn[2]};
  float dotc = gnr[0] * w.cr[0] + gnr[1] * w.cr[1] + gnr[2] * w.cr[2];
  float s3 = w.s * w.s * w.s;
  float gcr[3];
  for (int a = 0; a < 3; ++a) gcr[a] = gnr[a] / w.s - w.cr[a] * dotc / s3;
  // cr = e1 x e2: g_e1 = e2 x g_cr, g_e2 = g_cr x e1
  ge1[0] = w.e2[1] * gcr[2] - w.e2[2] * gcr[1];
  ge1[1] = w.e2[2] * gcr[0] - w.e2[0] * gcr[2];
  ge1[2] = w.e2[0] * gcr[1] - w.e2[1] * gcr[0];
  ge2[0] = gcr[1] * w.e1[2] - gcr[2] * w.e1[1];
  ge2[1] = gcr[2] * w.e1[0] - gcr[0] * w.e1[2];
  ge2[2] = gcr[0] * w.e1[1] - gcr[1] * w.e1[0];
  for (int a = 0; a < 3; ++a) {
    gp[0][a] = gp1[a] - ge1[a] - ge2[a];
    gp[1][a] = ge1[a];
    gp[2][a] = ge2[a];
  }
}

// contact_adj_terms with the vertex adjoints added to g_pos / g_vel by float
// atomics (order not fixed)
__device__ __forceinline__ void contact_adj_core(const ContactP& C, const float (&q)[3],
                                                 const ContactFwd& w, const float (&gf)[3],
                                                 const int (&tv)[3], float* __restrict__ g_pos,
                                                 float* __restrict__ g_vel, float (&gq)[3],
                                                 float (&gqv)[3]) {
  float gvv[3], gp[3][3];
  contact_adj_terms(C, q, w, gf, gvv, gp, gq, gqv);
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 3; ++a) {
      atomic_add_f(g_vel + 3ll * tv[c] + a, gvv[a]);
      atomic_add_f(g_pos + 3ll * tv[c] + a, gp[c][a]);
    }
}

// Adjoint of k_collision.  Inputs: g_gridf (adjoint of grid_f), g_ext (adjoint
// of the pad's external force).  Accumulates into g_mv, g_m (grid, strided),
// g_pos, g_vel (pad, atomics).
__global__ __launch_bounds__(kBlock) void k_collision_adj(
    MpmP P, ContactP C, long long nn, const float* __restrict__ mv, int smv,
    const float* __restrict__ m, int sm, const int32_t* __restrict__ idx, int stride, int col,
    const int32_t* __restrict__ tris, const float* __restrict__ pos, const float* __restrict__ vel,
    float o0, float o1, float o2, const float* __restrict__ g_gridf,
    const float* __restrict__ g_ext, float* __restrict__ g_mv, int sgmv, float* __restrict__ g_m,
    int sgm, float* __restrict__ g_pos, float* __restrict__ g_vel) {
  node_loop(P, nn, [&](long long nd) {
    float mm = m[nd * sm];
    if (!(mm > P.eps)) return;
    int k = idx[nd * stride + col];
    if (k < 0) return;
    const int n = P.n_grid;
    int i = (int)(nd / ((long long)n * n)), j = (int)((nd / n) % n), kk = (int)(nd % n);
    const float dxe = at_xnode(P, i).dx;
    float q[3] = {node_centre(i % n, dxe), node_centre(j, dxe), node_centre(kk, dxe)};
    float inv = 1.f / (mm + P.eps);
    float mvv[3] = {mv[nd * smv], mv[nd * smv + 1], mv[nd * smv + 2]};
    float qv[3] = {mvv[0] * inv, mvv[1] * inv, mvv[2] * inv};
    int tv[3] = {tris[3 * k], tris[3 * k + 1], tris[3 * k + 2]};
    float od[3] = {o0, o1, o2};
    ContactFwd w;
    contact_fwd(C, q, qv, pos, vel, tv, od, w);
    if (!w.flag) return;
    // adjoint of f: grid_f += f dt; ext[t] += -f/3
    float gf[3];
    for (int a = 0; a < 3; ++a)
      gf[a] = g_gridf[3 * nd + a] * P.dt -
              (g_ext[3ll * tv[0] + a] + g_ext[3ll * tv[1] + a] + g_ext[3ll * tv[2] + a]) / 3.f;
    float gq[3], gqv[3];
    contact_adj_core(C, q, w, gf, tv, g_pos, g_vel, gq, gqv);
    // qv = mv / (m + eps); q is the fixed cell centre
    float gmm = 0.f;
    for (int a = 0; a < 3; ++a) {
      atomic_add_f(g_mv + nd * sgmv + a, gqv[a] * inv);
      gmm -= gqv[a] * mvv[a] * inv * inv;
    }
    atomic_add_f(g_m + nd * sgm, gmm);
  });
}

// ---------------------------------------------------------------------------
// exact (order-free) sums of the pad vertex terms
// ---------------------------------------------------------------------------
// The pads' vertex terms (forward -f/3, adjoint d/d vel and d/d pos) are added
// as int64 fixed-point numbers, whose sums do not depend on the order the
// atomics run in: pass 1 takes the largest term magnitude (atomicMax on its
// float bits, itself order-free), pass 2 adds every term scaled by 2^(40 - e)
// (e: that magnitude's binary exponent, so a term stays below 2^41 and 2^21
// terms on one vertex below 2^62), pass 3 converts each vertex's sums back to
// float (one rounding) and clears them.  The quantum is 2^-40 of the largest
// term.  scale_state: per scale, the pass-1 maximum (cleared by pass 3) and
// the copy pass 2 leaves for pass 3.
__device__ __forceinline__ int fx_exp(uint32_t bits) {
  int e = (int)((bits >> 23) & 255u) - 127;
  return e < -80 ? -80 : (e > 80 ? 80 : e);
}
__device__ __forceinline__ long long fx_q(float v, int e) {
  return __float2ll_rn(v * ldexpf(1.f, 40 - e));
}
__device__ __forceinline__ void fx_add(long long* acc, float v, int e) {
  atomicAdd(reinterpret_cast<unsigned long long*>(acc), (unsigned long long)fx_q(v, e));
}
__device__ __forceinline__ float fx_float(long long a, int e) {
  return (float)a * ldexpf(1.f, e - 40);
}
__device__ __forceinline__ void fx_max(uint32_t* m, float v) {
  atomicMax(m, __float_as_uint(fabsf(v)));
}

// the grid contact of one node (collision's body): flag, force and triangle
__device__ __forceinline__ bool node_contact(const MpmP& P, const ContactP& C, long long nd,
                                             const float* __restrict__ mv, int smv,
                                             const float* __restrict__ m, int sm,
                                             const int32_t* __restrict__ idx, int stride,
                                             int col, const int32_t* __restrict__ tris,
                                             const float* __restrict__ pos,
                                             const float* __restrict__ vel, const float (&od)[3],
                                             ContactFwd& w, int (&tv)[3], float (&q)[3],
                                             float& inv, float (&mvv)[3]) {
  const float mm = m[nd * sm];
  if (!(mm > P.eps)) return false;
  const int k = idx[nd * stride + col];
  if (k < 0) return false;
  const int n = P.n_grid;
  const int i = (int)(nd / ((long long)n * n)), j = (int)((nd / n) % n), kk = (int)(nd % n);
  const float dxe = at_xnode(P, i).dx;
  q[0] = node_centre(i % n, dxe);
  q[1] = node_centre(j, dxe);
  q[2] = node_centre(kk, dxe);
  inv = 1.f / (mm + P.eps);
  for (int a = 0; a < 3; ++a) mvv[a] = mv[nd * smv + a];
  const float qv[3] = {mvv[0] * inv, mvv[1] * inv, mvv[2] * inv};
  tv[0] = tris[3 * k];
  tv[1] = tris[3 * k + 1];
  tv[2] = tris[3 * k + 2];
  contact_fwd(C, q, qv, pos, vel, tv, od, w);
  return w.flag;
}

// forward pass 1: grid_f += f dt; the largest |f_a| / 3
__global__ __launch_bounds__(kBlock) void k_collision_fx1(
    MpmP P, ContactP C, long long nn, const float* __restrict__ mv, int smv,
    const float* __restrict__ m, int sm, const int32_t* __restrict__ idx, int stride, int col,
    const int32_t* __restrict__ tris, const float* __restrict__ pos, const float* __restrict__ vel,
    float o0, float o1, float o2, float* __restrict__ gridf, uint32_t* __restrict__ scale) {
  const float od[3] = {o0, o1, o2};
  node_loop(P, nn, [&](long long nd) {
    ContactFwd w;
    int tv[3];
    float q[3], inv, mvv[3];
    if (!node_contact(P, C, nd, mv, smv, m, sm, idx, stride, col, tris, pos, vel, od, w, tv, q,
                      inv, mvv))
      return;
    float mx = 0.f;
    for (int a = 0; a < 3; ++a) {
      gridf[3 * nd + a] += w.f[a] * P.dt;
      mx = fmaxf(mx, fabsf(w.f[a]));
    }
    fx_max(scale, mx / 3.f);
  });
}

// forward pass 2: the -f/3 terms into the fixed-point vertex sums
__global__ __launch_bounds__(kBlock) void k_collision_fx2(
    MpmP P, ContactP C, long long nn, const float* __restrict__ mv, int smv,
    const float* __restrict__ m, int sm, const int32_t* __restrict__ idx, int stride, int col,
    const int32_t* __restrict__ tris, const float* __restrict__ pos, const float* __restrict__ vel,
    float o0, float o1, float o2, long long* __restrict__ acc, uint32_t* __restrict__ scale) {
  const float od[3] = {o0, o1, o2};
  const uint32_t bits = scale[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) scale[1] = bits;
  const int e = fx_exp(bits);
  node_loop(P, nn, [&](long long nd) {
    ContactFwd w;
    int tv[3];
    float q[3], inv, mvv[3];
    if (!node_contact(P, C, nd, mv, smv, m, sm, idx, stride, col, tris, pos, vel, od, w, tv, q,
                      inv, mvv))
      return;
    for (int c = 0; c < 3; ++c)
      for (int a = 0; a < 3; ++a) fx_add(acc + 3ll * tv[c] + a, -w.f[a] / 3.f, e);
  });
}

// pass 3: out[v] += sums, sums cleared, the pass-1 maxima cleared
__global__ __launch_bounds__(kBlock) void k_fx_out(int nv, int n_arrays,
                                                   long long* __restrict__ acc,
                                                   uint32_t* __restrict__ scale,
                                                   float* __restrict__ out0,
                                                   float* __restrict__ out1) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = 3ll * nv;
  if (i < per * n_arrays) {
    const int s = (int)(i / per);
    const long long r = i - s * per;
    const int e = fx_exp(scale[n_arrays + s]);
    float* out = s == 0 ? out0 : out1;
    out[r] += fx_float(acc[i], e);
    acc[i] = 0;
  }
  if (i < n_arrays) scale[i] = 0u;
}

// adjoint pass 1: the largest vertex terms (velocity, position)
__global__ __launch_bounds__(kBlock) void k_collision_adj_fx1(
    MpmP P, ContactP C, long long nn, const float* __restrict__ mv, int smv,
    const float* __restrict__ m, int sm, const int32_t* __restrict__ idx, int stride, int col,
    const int32_t* __restrict__ tris, const float* __restrict__ pos, const float* __restrict__ vel,
    float o0, float o1, float o2, const float* __restrict__ g_gridf,
    const float* __restrict__ g_ext, uint32_t* __restrict__ scale) {
  const float od[3] = {o0, o1, o2};
  node_loop(P, nn, [&](long long nd) {
    ContactFwd w;
    int tv[3];
    float q[3], inv, mvv[3];
    if (!node_contact(P, C, nd, mv, smv, m, sm, idx, stride, col, tris, pos, vel, od, w, tv, q,
                      inv, mvv))
      return;
    float gf[3];
    for (int a = 0; a < 3; ++a)
      gf[a] = g_gridf[3 * nd + a] * P.dt -
              (g_ext[3ll * tv[0] + a] + g_ext[3ll * tv[1] + a] + g_ext[3ll * tv[2] + a]) / 3.f;
    float gvv[3], gp[3][3], gq[3], gqv[3];
    contact_adj_terms(C, q, w, gf, gvv, gp, gq, gqv);
    float mv_ = 0.f, mp = 0.f;
    for (int a = 0; a < 3; ++a) {
      mv_ = fmaxf(mv_, fabsf(gvv[a]));
      mp = fmaxf(mp, fmaxf(fabsf(gp[0][a]), fmaxf(fabsf(gp[1][a]), fabsf(gp[2][a]))));
    }
    fx_max(scale, mv_);
    fx_max(scale + 1, mp);
  });
}

// adjoint pass 2: the grid adjoints (node-owned) and the vertex terms into the
// fixed-point sums acc[0 .. 3 nv) (velocity) and acc[3 nv .. 6 nv) (position)
__global__ __launch_bounds__(kBlock) void k_collision_adj_fx2(
    MpmP P, ContactP C, long long nn, const float* __restrict__ mv, int smv,
    const float* __restrict__ m, int sm, const int32_t* __restrict__ idx, int stride, int col,
    const int32_t* __restrict__ tris, const float* __restrict__ pos, const float* __restrict__ vel,
    float o0, float o1, float o2, const float* __restrict__ g_gridf,
    const float* __restrict__ g_ext, float* __restrict__ g_mv, int sgmv, float* __restrict__ g_m,
    int sgm, int nv, long long* __restrict__ acc, uint32_t* __restrict__ scale) {
  const float od[3] = {o0, o1, o2};
  const uint32_t bv = scale[0], bp = scale[1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scale[2] = bv;
    scale[3] = bp;
  }
  const int ev = fx_exp(bv), ep = fx_exp(bp);
  node_loop(P, nn, [&](long long nd) {
    ContactFwd w;
    int tv[3];
    float q[3], inv, mvv[3];
    if (!node_contact(P, C, nd, mv, smv, m, sm, idx, stride, col, tris, pos, vel, od, w, tv, q,
                      inv, mvv))
      return;
    float gf[3];
    for (int a = 0; a < 3; ++a)
      gf[a] = g_gridf[3 * nd + a] * P.dt -
              (g_ext[3ll * tv[0] + a] + g_ext[3ll * tv[1] + a] + g_ext[3ll * tv[2] + a]) / 3.f;
    float gvv[3], gp[3][3], gq[3], gqv[3];
    contact_adj_terms(C, q, w, gf, gvv, gp, gq, gqv);
    // qv = mv / (m + eps): one node per thread, so these adds are ordered
    float gmm = 0.f;
    for (int a = 0; a < 3; ++a) {
      g_mv[nd * sgmv + a] += gqv[a] * inv;
      gmm -= gqv[a] * mvv[a] * inv * inv;
    }
    g_m[nd * sgm] += gmm;
    for (int c = 0; c < 3; ++c)
      for (int a = 0; a < 3; ++a) {
        fx_add(acc + 3ll * tv[c] + a, gvv[a], ev);
        fx_add(acc + 3ll * nv + 3ll * tv[c] + a, gp[c][a], ep);
      }
  });
}

// ---------------------------------------------------------------------------
// particle-wise contact of a rope with one pad (cable_straightening.py:287-319)
// ---------------------------------------------------------------------------
// find_closest with the particle position as the query (strict < from
// (100, -1) in index order: the lexicographic minimum of (distance, index)),
// searched as k_closest searches nodes: a block takes 64 particles at a time,
// its eight wavefronts take the groups of triangle runs in turn (with
// kBounds, groups and runs whose boxes lie beyond every lane's best are
// skipped) and their
// minima are combined in LDS.  hint as for k_closest (it may be idx itself:
// every read of a particle's entry precedes the barrier before its write).
template <bool kBounds>
__global__ __launch_bounds__(kClosestThreads) void k_point_closest(
    int n, const float* __restrict__ x, int nt, const float4* __restrict__ cen,
    const float4* __restrict__ bounds, const int32_t* hint, int32_t* idx, int stride, int col) {
  __shared__ float s_best[kClosestWaves][64];
  __shared__ int s_k[kClosestWaves][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n_chunks = (n + 63) / 64;
  const int n_runs = (nt + kTriBatch - 1) / kTriBatch;
  for (int ch = blockIdx.x; ch < n_chunks; ch += gridDim.x) {
    const int p = ch * 64 + lane;
    const bool live = p < n;
    const float qx = live ? x[3ll * p] : 0.f, qy = live ? x[3ll * p + 1] : 0.f,
                qz = live ? x[3ll * p + 2] : 0.f;
    Closest r{100.0f, 1e4f, 1e4f * (1.0f + 1e-6f), -1};
    if (hint && live) {
      const int h = hint[(long long)p * stride + col];
      if (h >= 0 && h < nt) closest_step<false>(cen[h], h, 0, nt, qx, qy, qz, r);
    }
    // lanes past the end take no triangle
    closest_search<kBounds>(cen, bounds, n_runs, w, 0, nt, 0, live ? nt : 0, false, qx, qy, qz,
                            r);
    s_best[w][lane] = r.best;
    s_k[w][lane] = r.best_k;
    __syncthreads();
    if (w == 0 && live) {
      float bd = 100.0f;
      int bk = -1;
      for (int q = 0; q < kClosestWaves; ++q)
        if (closest_before(s_best[q][lane], s_k[q][lane], bd, bk)) {
          bd = s_best[q][lane];
          bk = s_k[q][lane];
        }
      idx[(long long)p * stride + col] = bk;
    }
    __syncthreads();
  }
}

// rope impulse += f dt, pad external force += -f/3 per triangle vertex,
// flag[p] = 1 where the particle touches this pad (C.sdf_offset: the radius)
__global__ __launch_bounds__(kBlock) void k_rope_collision(
    ContactP C, float dt, int n, const float* __restrict__ x, const float* __restrict__ v,
    const int32_t* __restrict__ idx, int stride, int col, const int32_t* __restrict__ tris,
    const float* __restrict__ pos, const float* __restrict__ vel, float o0, float o1, float o2,
    float* __restrict__ impulse, float* __restrict__ ext, int32_t* __restrict__ flag) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int k = idx[(long long)p * stride + col];
  if (k < 0) return;
  const float q[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
  const float qv[3] = {v[3 * p], v[3 * p + 1], v[3 * p + 2]};
  const int tv[3] = {tris[3 * k], tris[3 * k + 1], tris[3 * k + 2]};
  const float od[3] = {o0, o1, o2};
  ContactFwd w;
  contact_fwd(C, q, qv, pos, vel, tv, od, w);
  if (!w.flag) return;
  for (int a = 0; a < 3; ++a) {
    impulse[3 * p + a] += w.f[a] * dt;
    const float e = -w.f[a] / 3.f;
    atomic_add_f(ext + 3ll * tv[0] + a, e);
    atomic_add_f(ext + 3ll * tv[1] + a, e);
    atomic_add_f(ext + 3ll * tv[2] + a, e);
  }
  if (flag) flag[p] = 1;
}

// adjoint: g_x, g_v (rope) and g_pos, g_vel (pad) accumulate
__global__ __launch_bounds__(kBlock) void k_rope_collision_adj(
    ContactP C, float dt, int n, const float* __restrict__ x, const float* __restrict__ v,
    const int32_t* __restrict__ idx, int stride, int col, const int32_t* __restrict__ tris,
    const float* __restrict__ pos, const float* __restrict__ vel, float o0, float o1, float o2,
    const float* __restrict__ g_impulse, const float* __restrict__ g_ext,
    float* __restrict__ g_x, float* __restrict__ g_v, float* __restrict__ g_pos,
    float* __restrict__ g_vel) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int k = idx[(long long)p * stride + col];
  if (k < 0) return;
  const float q[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
  const float qv[3] = {v[3 * p], v[3 * p + 1], v[3 * p + 2]};
  const int tv[3] = {tris[3 * k], tris[3 * k + 1], tris[3 * k + 2]};
  const float od[3] = {o0, o1, o2};
  ContactFwd w;
  contact_fwd(C, q, qv, pos, vel, tv, od, w);
  if (!w.flag) return;
  float gf[3];
  for (int a = 0; a < 3; ++a)
    gf[a] = g_impulse[3 * p + a] * dt -
            (g_ext[3ll * tv[0] + a] + g_ext[3ll * tv[1] + a] + g_ext[3ll * tv[2] + a]) / 3.f;
  float gq[3], gqv[3];
  contact_adj_core(C, q, w, gf, tv, g_pos, g_vel, gq, gqv);
  for (int a = 0; a < 3; ++a) {
    g_x[3 * p + a] += gq[a];
    g_v[3 * p + a] += gqv[a];
  }
}

// the rope contact with the pad-vertex sums exact (the fixed-point passes
// above): pass 1 the largest term, pass 2 the impulse on the rope (one
// particle per thread) and the fixed-point pad terms
__device__ __forceinline__ bool rope_contact_at(const ContactP& C, int p,
                                                const float* __restrict__ x,
                                                const float* __restrict__ v,
                                                const int32_t* __restrict__ idx, int stride,
                                                int col, const int32_t* __restrict__ tris,
                                                const float* __restrict__ pos,
                                                const float* __restrict__ vel,
                                                const float (&od)[3], float (&q)[3], int (&tv)[3],
                                                ContactFwd& w) {
  const int k = idx[(long long)p * stride + col];
  if (k < 0) return false;
  for (int a = 0; a < 3; ++a) q[a] = x[3 * p + a];
  const float qv[3] = {v[3 * p], v[3 * p + 1], v[3 * p + 2]};
  tv[0] = tris[3 * k];
  tv[1] = tris[3 * k + 1];
  tv[2] = tris[3 * k + 2];
  contact_fwd(C, q, qv, pos, vel, tv, od, w);
  return w.flag;
}

template <int kPass>
__global__ __launch_bounds__(kBlock) void k_rope_collision_fx(
    ContactP C, float dt, int n, const float* __restrict__ x, const float* __restrict__ v,
    const int32_t* __restrict__ idx, int stride, int col, const int32_t* __restrict__ tris,
    const float* __restrict__ pos, const float* __restrict__ vel, float o0, float o1, float o2,
    float* __restrict__ impulse, int32_t* __restrict__ flag, long long* __restrict__ acc,
    uint32_t* __restrict__ scale) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const float od[3] = {o0, o1, o2};
  const uint32_t bits = kPass == 2 ? scale[0] : 0u;
  if (kPass == 2 && p == 0) scale[1] = bits;
  if (p >= n) return;
  float q[3];
  int tv[3];
  ContactFwd w;
  if (!rope_contact_at(C, p, x, v, idx, stride, col, tris, pos, vel, od, q, tv, w)) return;
  if (kPass == 1) {
    fx_max(scale, fmaxf(fabsf(w.f[0]), fmaxf(fabsf(w.f[1]), fabsf(w.f[2]))) / 3.f);
    return;
  }
  const int e = fx_exp(bits);
  for (int a = 0; a < 3; ++a) {
    impulse[3 * p + a] += w.f[a] * dt;
    for (int c = 0; c < 3; ++c) fx_add(acc + 3ll * tv[c] + a, -w.f[a] / 3.f, e);
  }
  if (flag) flag[p] = 1;
}

template <int kPass>
__global__ __launch_bounds__(kBlock) void k_rope_collision_adj_fx(
    ContactP C, float dt, int n, const float* __restrict__ x, const float* __restrict__ v,
    const int32_t* __restrict__ idx, int stride, int col, const int32_t* __restrict__ tris,
    const float* __restrict__ pos, const float* __restrict__ vel, float o0, float o1, float o2,
    const float* __restrict__ g_impulse, const float* __restrict__ g_ext, float* __restrict__ g_x,
    float* __restrict__ g_v, int nv, long long* __restrict__ acc, uint32_t* __restrict__ scale) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const float od[3] = {o0, o1, o2};
  uint32_t bv = 0u, bp = 0u;
  if (kPass == 2) {
    bv = scale[0];
    bp = scale[1];
    if (p == 0) {
      scale[2] = bv;
      scale[3] = bp;
    }
  }
  if (p >= n) return;
  float q[3];
  int tv[3];
  ContactFwd w;
  if (!rope_contact_at(C, p, x, v, idx, stride, col, tris, pos, vel, od, q, tv, w)) return;
  float gf[3];
  for (int a = 0; a < 3; ++a)
    gf[a] = g_impulse[3 * p + a] * dt -
            (g_ext[3ll * tv[0] + a] + g_ext[3ll * tv[1] + a] + g_ext[3ll * tv[2] + a]) / 3.f;
  float gvv[3], gp[3][3], gq[3], gqv[3];
  contact_adj_terms(C, q, w, gf, gvv, gp, gq, gqv);
  if (kPass == 1) {
    float mv_ = 0.f, mp = 0.f;
    for (int a = 0; a < 3; ++a) {
      mv_ = fmaxf(mv_, fabsf(gvv[a]));
      mp = fmaxf(mp, fmaxf(fabsf(gp[0][a]), fmaxf(fabsf(gp[1][a]), fabsf(gp[2][a]))));
    }
    fx_max(scale, mv_);
    fx_max(scale + 1, mp);
    return;
  }
  const int ev = fx_exp(bv), ep = fx_exp(bp);
  for (int a = 0; a < 3; ++a) {
    g_x[3 * p + a] += gq[a];
    g_v[3 * p + a] += gqv[a];
  }
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 3; ++a) {
      fx_add(acc + 3ll * tv[c] + a, gvv[a], ev);
      fx_add(acc + 3ll * nv + 3ll * tv[c] + a, gp[c][a], ep);
    }
}

static ContactP contact_dev(const dt_contact_params* c) {
  ContactP d;
  d.kn = c->kn;
  d.kd = c->kd;
  d.kt = c->kt;
  d.mu_f = c->mu_f;
  d.norm_eps = c->norm_eps;
  d.sdf_offset = c->sdf_offset;
  d.unclamped_vn = c->unclamped_vn ? 1 : 0;
  return d;
}

}  // namespace dt

using namespace dt;

extern "C" {

int dt_contact_centroids(int32_t n_tris, const int32_t* tris, const float* pos, float* centroids4,
                         void* stream) {
  DT_REQUIRE(n_tris >= 0, "n_tris must be >= 0");
  if (n_tris == 0) return DT_OK;
  DT_REQUIRE(tris && pos && centroids4, "NULL pointer");
  hipLaunchKernelGGL(k_centroids, dim3(nblocks(n_tris)), dim3(kBlock), 0, (hipStream_t)stream,
                     n_tris, tris, pos, (float4*)centroids4);
  return check_launch("dt_contact_centroids");
}

static int closest_launch(const dt_mpm_params* prm, const float* grid_m, int32_t m_stride,
                          const int32_t* tri_start, int32_t n_tris, const float* centroids4,
                          const float* bounds4, const int32_t* hint, int32_t* contact_idx,
                          int32_t idx_stride, int32_t idx_col, void* stream, const char* name) {
  DT_REQUIRE(n_tris >= 0 && m_stride > 0 && idx_stride > 0 && idx_col >= 0 &&
                 idx_col < idx_stride,
             "bad size/stride");
  DT_REQUIRE(grid_m && contact_idx && (n_tris == 0 || centroids4), "NULL pointer");
  long long nn = (long long)(prm->n_env > 1 ? prm->n_env : 1) * prm->n_grid * prm->n_grid *
                 prm->n_grid;
  if (bounds4 && n_tris > 0)
    hipLaunchKernelGGL(k_closest<true>, dim3(closest_blocks(*prm, nn)), dim3(kClosestThreads), 0,
                       (hipStream_t)stream, to_dev_nodes(*prm), nn, grid_m, m_stride, n_tris,
                       tri_start, (const float4*)centroids4, (const float4*)bounds4, hint,
                       contact_idx, idx_stride, idx_col);
  else
    hipLaunchKernelGGL(k_closest<false>, dim3(closest_blocks(*prm, nn)), dim3(kClosestThreads), 0,
                       (hipStream_t)stream, to_dev_nodes(*prm), nn, grid_m, m_stride, n_tris,
                       tri_start, (const float4*)centroids4, (const float4*)nullptr, hint,
                       contact_idx, idx_stride, idx_col);
  return check_launch(name);
}

int dt_contact_closest(const dt_mpm_params* prm, const float* grid_m, int32_t m_stride,
                       int32_t n_tris, const float* centroids4, int32_t* contact_idx,
                       int32_t idx_stride, int32_t idx_col, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(prm->n_env <= 1, "stacked environments: use dt_contact_closest_env");
  return closest_launch(prm, grid_m, m_stride, nullptr, n_tris, centroids4, nullptr, nullptr,
                        contact_idx, idx_stride, idx_col, stream, "dt_contact_closest");
}

int dt_contact_closest_env(const dt_mpm_params* prm, const float* grid_m, int32_t m_stride,
                           const int32_t* tri_start, int32_t n_tris, const float* centroids4,
                           int32_t* contact_idx, int32_t idx_stride, int32_t idx_col,
                           void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(tri_start, "NULL tri_start");
  return closest_launch(prm, grid_m, m_stride, tri_start, n_tris, centroids4, nullptr, nullptr,
                        contact_idx, idx_stride, idx_col, stream, "dt_contact_closest_env");
}

int dt_contact_tri_batches(int32_t n_tris, const float* centroids4, float* batch_bounds4,
                           void* stream) {
  DT_REQUIRE(n_tris >= 0, "n_tris must be >= 0");
  if (n_tris == 0) return DT_OK;
  DT_REQUIRE(centroids4 && batch_bounds4, "NULL pointer");
  const int nb = (n_tris + kTriBatch - 1) / kTriBatch;
  hipLaunchKernelGGL(k_tri_batches, dim3(nblocks(nb)), dim3(kBlock), 0, (hipStream_t)stream,
                     n_tris, (const float4*)centroids4, (float4*)batch_bounds4);
  return check_launch("dt_contact_tri_batches");
}

int dt_contact_closest_batched(const dt_mpm_params* prm, const float* grid_m, int32_t m_stride,
                               const int32_t* tri_start, int32_t n_tris, const float* centroids4,
                               const float* batch_bounds4, const int32_t* hint,
                               int32_t* contact_idx, int32_t idx_stride, int32_t idx_col,
                               void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(prm->n_env <= 1 || tri_start, "stacked environments need tri_start");
  DT_REQUIRE(n_tris == 0 || batch_bounds4, "NULL batch bounds");
  return closest_launch(prm, grid_m, m_stride, tri_start, n_tris, centroids4, batch_bounds4, hint,
                        contact_idx, idx_stride, idx_col, stream, "dt_contact_closest_batched");
}

int dt_contact_collision(const dt_mpm_params* prm, const dt_contact_params* cp,
                         const float* grid_mv, int32_t mv_stride, const float* grid_m,
                         int32_t m_stride, const int32_t* contact_idx, int32_t idx_stride,
                         int32_t idx_col, int32_t n_tris, const int32_t* tris, const float* pos,
                         const float* vel, const float* out_dir, float* grid_f, float* ext_force,
                         void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(cp && out_dir, "NULL contact params / out_dir");
  DT_REQUIRE(mv_stride >= 3 && m_stride > 0 && idx_stride > 0 && idx_col >= 0 &&
                 idx_col < idx_stride && n_tris >= 0,
             "bad size/stride");
  DT_REQUIRE(grid_mv && grid_m && contact_idx && tris && pos && vel && grid_f && ext_force,
             "NULL pointer");
  long long nn = (long long)(prm->n_env > 1 ? prm->n_env : 1) * prm->n_grid * prm->n_grid *
                 prm->n_grid;
  hipLaunchKernelGGL(k_collision, dim3(node_blocks(*prm, nn)), dim3(kBlock), 0,
                     (hipStream_t)stream, to_dev_nodes(*prm), contact_dev(cp), nn, grid_mv, mv_stride, grid_m, m_stride,
                     contact_idx, idx_stride, idx_col, tris, pos, vel, out_dir[0], out_dir[1],
                     out_dir[2], grid_f, ext_force);
  return check_launch("dt_contact_collision");
}

int dt_contact_collision_adj(const dt_mpm_params* prm, const dt_contact_params* cp,
                             const float* grid_mv, int32_t mv_stride, const float* grid_m,
                             int32_t m_stride, const int32_t* contact_idx, int32_t idx_stride,
                             int32_t idx_col, int32_t n_tris, const int32_t* tris,
                             const float* pos, const float* vel, const float* out_dir,
                             const float* g_grid_f, const float* g_ext_force, float* g_grid_mv,
                             int32_t gmv_stride, float* g_grid_m, int32_t gm_stride, float* g_pos,
                             float* g_vel, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(cp && out_dir, "NULL contact params / out_dir");
  DT_REQUIRE(mv_stride >= 3 && m_stride > 0 && gmv_stride >= 3 && gm_stride > 0 &&
                 idx_stride > 0 && idx_col >= 0 && idx_col < idx_stride && n_tris >= 0,
             "bad size/stride");
  DT_REQUIRE(grid_mv && grid_m && contact_idx && tris && pos && vel && g_grid_f && g_ext_force &&
                 g_grid_mv && g_grid_m && g_pos && g_vel,
             "NULL pointer");
  long long nn = (long long)(prm->n_env > 1 ? prm->n_env : 1) * prm->n_grid * prm->n_grid *
                 prm->n_grid;
  hipLaunchKernelGGL(k_collision_adj, dim3(node_blocks(*prm, nn)), dim3(kBlock), 0,
                     (hipStream_t)stream, to_dev_nodes(*prm), contact_dev(cp), nn, grid_mv, mv_stride, grid_m, m_stride,
                     contact_idx, idx_stride, idx_col, tris, pos, vel, out_dir[0], out_dir[1],
                     out_dir[2], g_grid_f, g_ext_force, g_grid_mv, gmv_stride, g_grid_m,
                     gm_stride, g_pos, g_vel);
  return check_launch("dt_contact_collision_adj");
}

int dt_contact_collision_fx(const dt_mpm_params* prm, const dt_contact_params* cp,
                            const float* grid_mv, int32_t mv_stride, const float* grid_m,
                            int32_t m_stride, const int32_t* contact_idx, int32_t idx_stride,
                            int32_t idx_col, int32_t n_tris, const int32_t* tris,
                            const float* pos, const float* vel, const float* out_dir,
                            float* grid_f, int32_t n_verts, float* ext_force, int64_t* acc,
                            uint32_t* scale_state, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(cp && out_dir, "NULL contact params / out_dir");
  DT_REQUIRE(mv_stride >= 3 && m_stride > 0 && idx_stride > 0 && idx_col >= 0 &&
                 idx_col < idx_stride && n_tris >= 0 && n_verts >= 0,
             "bad size/stride");
  DT_REQUIRE(grid_mv && grid_m && contact_idx && tris && pos && vel && grid_f && ext_force &&
                 acc && scale_state,
             "NULL pointer");
  long long nn = (long long)(prm->n_env > 1 ? prm->n_env : 1) * prm->n_grid * prm->n_grid *
                 prm->n_grid;
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = (unsigned)node_blocks(*prm, nn);
  const MpmP P = to_dev_nodes(*prm);
  const ContactP C = contact_dev(cp);
  hipLaunchKernelGGL(k_collision_fx1, dim3(nb), dim3(kBlock), 0, st, P, C, nn, grid_mv, mv_stride,
                     grid_m, m_stride, contact_idx, idx_stride, idx_col, tris, pos, vel,
                     out_dir[0], out_dir[1], out_dir[2], grid_f, scale_state);
  hipLaunchKernelGGL(k_collision_fx2, dim3(nb), dim3(kBlock), 0, st, P, C, nn, grid_mv, mv_stride,
                     grid_m, m_stride, contact_idx, idx_stride, idx_col, tris, pos, vel,
                     out_dir[0], out_dir[1], out_dir[2], (long long*)acc, scale_state);
  hipLaunchKernelGGL(k_fx_out, dim3(nblocks(3ll * (n_verts > 0 ? n_verts : 1))), dim3(kBlock), 0,
                     st, (int)n_verts, 1, (long long*)acc, scale_state, ext_force,
                     (float*)nullptr);
  return check_launch("dt_contact_collision_fx");
}

int dt_contact_collision_adj_fx(const dt_mpm_params* prm, const dt_contact_params* cp,
                                const float* grid_mv, int32_t mv_stride, const float* grid_m,
                                int32_t m_stride, const int32_t* contact_idx,
                                int32_t idx_stride, int32_t idx_col, int32_t n_tris,
                                const int32_t* tris, const float* pos, const float* vel,
                                const float* out_dir, const float* g_grid_f,
                                const float* g_ext_force, float* g_grid_mv, int32_t gmv_stride,
                                float* g_grid_m, int32_t gm_stride, int32_t n_verts,
                                float* g_pos, float* g_vel, int64_t* acc,
                                uint32_t* scale_state, void* stream) {
  int rc = validate_params(prm);
  if (rc) return rc;
  DT_REQUIRE(cp && out_dir, "NULL contact params / out_dir");
  DT_REQUIRE(mv_stride >= 3 && m_stride > 0 && gmv_stride >= 3 && gm_stride > 0 &&
                 idx_stride > 0 && idx_col >= 0 && idx_col < idx_stride && n_tris >= 0 &&
                 n_verts >= 0,
             "bad size/stride");
  DT_REQUIRE(grid_mv && grid_m && contact_idx && tris && pos && vel && g_grid_f && g_ext_force &&
                 g_grid_mv && g_grid_m && g_pos && g_vel && acc && scale_state,
             "NULL pointer");
  long long nn = (long long)(prm->n_env > 1 ? prm->n_env : 1) * prm->n_grid * prm->n_grid *
                 prm->n_grid;
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = (unsigned)node_blocks(*prm, nn);
  const MpmP P = to_dev_nodes(*prm);
  const ContactP C = contact_dev(cp);
  hipLaunchKernelGGL(k_collision_adj_fx1, dim3(nb), dim3(kBlock), 0, st, P, C, nn, grid_mv,
                     mv_stride, grid_m, m_stride, contact_idx, idx_stride, idx_col, tris, pos,
                     vel, out_dir[0], out_dir[1], out_dir[2], g_grid_f, g_ext_force, scale_state);
  hipLaunchKernelGGL(k_collision_adj_fx2, dim3(nb), dim3(kBlock), 0, st, P, C, nn, grid_mv,
                     mv_stride, grid_m, m_stride, contact_idx, idx_stride, idx_col, tris, pos,
                     vel, out_dir[0], out_dir[1], out_dir[2], g_grid_f, g_ext_force, g_grid_mv,
                     gmv_stride, g_grid_m, gm_stride, (int)n_verts, (long long*)acc, scale_state);
  hipLaunchKernelGGL(k_fx_out, dim3(nblocks(6ll * (n_verts > 0 ? n_verts : 1))), dim3(kBlock), 0,
                     st, (int)n_verts, 2, (long long*)acc, scale_state, g_vel, g_pos);
  return check_launch("dt_contact_collision_adj_fx");
}

int dt_rope_collision_fx(const dt_contact_params* cp, float dt, int32_t n, const float* x,
                         const float* v, const int32_t* contact_idx, int32_t idx_stride,
                         int32_t idx_col, int32_t n_tris, const int32_t* tris, const float* pos,
                         const float* vel, const float* out_dir, float* impulse, int32_t n_verts,
                         float* ext_force, int32_t* contact_flag, int64_t* acc,
                         uint32_t* scale_state, void* stream) {
  DT_REQUIRE(cp && out_dir, "NULL contact params / out_dir");
  DT_REQUIRE(n >= 0 && n_tris >= 0 && n_verts >= 0 && idx_stride > 0 && idx_col >= 0 &&
                 idx_col < idx_stride,
             "bad size/stride");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && v && contact_idx && tris && pos && vel && impulse && ext_force && acc &&
                 scale_state,
             "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  const ContactP C = contact_dev(cp);
  hipLaunchKernelGGL(k_rope_collision_fx<1>, dim3(nblocks(n)), dim3(kBlock), 0, st, C, dt, (int)n,
                     x, v, contact_idx, idx_stride, idx_col, tris, pos, vel, out_dir[0],
                     out_dir[1], out_dir[2], impulse, contact_flag, (long long*)acc, scale_state);
  hipLaunchKernelGGL(k_rope_collision_fx<2>, dim3(nblocks(n)), dim3(kBlock), 0, st, C, dt, (int)n,
                     x, v, contact_idx, idx_stride, idx_col, tris, pos, vel, out_dir[0],
                     out_dir[1], out_dir[2], impulse, contact_flag, (long long*)acc, scale_state);
  hipLaunchKernelGGL(k_fx_out, dim3(nblocks(3ll * (n_verts > 0 ? n_verts : 1))), dim3(kBlock), 0,
                     st, (int)n_verts, 1, (long long*)acc, scale_state, ext_force,
                     (float*)nullptr);
  return check_launch("dt_rope_collision_fx");
}

int dt_rope_collision_adj_fx(const dt_contact_params* cp, float dt, int32_t n, const float* x,
                             const float* v, const int32_t* contact_idx, int32_t idx_stride,
                             int32_t idx_col, int32_t n_tris, const int32_t* tris,
                             const float* pos, const float* vel, const float* out_dir,
                             const float* g_impulse, const float* g_ext_force, float* g_x,
                             float* g_v, int32_t n_verts, float* g_pos, float* g_vel,
                             int64_t* acc, uint32_t* scale_state, void* stream) {
  DT_REQUIRE(cp && out_dir, "NULL contact params / out_dir");
  DT_REQUIRE(n >= 0 && n_tris >= 0 && n_verts >= 0 && idx_stride > 0 && idx_col >= 0 &&
                 idx_col < idx_stride,
             "bad size/stride");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && v && contact_idx && tris && pos && vel && g_impulse && g_ext_force && g_x &&
                 g_v && g_pos && g_vel && acc && scale_state,
             "NULL pointer");
  hipStream_t st = (hipStream_t)stream;
  const ContactP C = contact_dev(cp);
  hipLaunchKernelGGL(k_rope_collision_adj_fx<1>, dim3(nblocks(n)), dim3(kBlock), 0, st, C, dt,
                     (int)n, x, v, contact_idx, idx_stride, idx_col, tris, pos, vel, out_dir[0],
                     out_dir[1], out_dir[2], g_impulse, g_ext_force, g_x, g_v, (int)n_verts,
                     (long long*)acc, scale_state);
  hipLaunchKernelGGL(k_rope_collision_adj_fx<2>, dim3(nblocks(n)), dim3(kBlock), 0, st, C, dt,
                     (int)n, x, v, contact_idx, idx_stride, idx_col, tris, pos, vel, out_dir[0],
                     out_dir[1], out_dir[2], g_impulse, g_ext_force, g_x, g_v, (int)n_verts,
                     (long long*)acc, scale_state);
  hipLaunchKernelGGL(k_fx_out, dim3(nblocks(6ll * (n_verts > 0 ? n_verts : 1))), dim3(kBlock), 0,
                     st, (int)n_verts, 2, (long long*)acc, scale_state, g_vel, g_pos);
  return check_launch("dt_rope_collision_adj_fx");
}

static int point_closest_launch(int32_t n, const float* x, int32_t n_tris,
                                const float* centroids4, const float* bounds4,
                                const int32_t* hint, int32_t* contact_idx, int32_t idx_stride,
                                int32_t idx_col, void* stream, const char* name) {
  DT_REQUIRE(n >= 0 && n_tris >= 0 && idx_stride > 0 && idx_col >= 0 && idx_col < idx_stride,
             "bad size/stride");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && contact_idx && (n_tris == 0 || centroids4), "NULL pointer");
  const int blocks = (int)std::min<long long>((n + 63ll) / 64, 4096);
  if (bounds4 && n_tris > 0)
    hipLaunchKernelGGL(k_point_closest<true>, dim3(blocks), dim3(kClosestThreads), 0,
                       (hipStream_t)stream, (int)n, x, (int)n_tris, (const float4*)centroids4,
                       (const float4*)bounds4, hint, contact_idx, idx_stride, idx_col);
  else
    hipLaunchKernelGGL(k_point_closest<false>, dim3(blocks), dim3(kClosestThreads), 0,
                       (hipStream_t)stream, (int)n, x, (int)n_tris, (const float4*)centroids4,
                       (const float4*)nullptr, hint, contact_idx, idx_stride, idx_col);
  return check_launch(name);
}

int dt_rope_closest(int32_t n, const float* x, int32_t n_tris, const float* centroids4,
                    int32_t* contact_idx, int32_t idx_stride, int32_t idx_col, void* stream) {
  return point_closest_launch(n, x, n_tris, centroids4, nullptr, nullptr, contact_idx, idx_stride,
                              idx_col, stream, "dt_rope_closest");
}

int dt_rope_closest_batched(int32_t n, const float* x, int32_t n_tris, const float* centroids4,
                            const float* batch_bounds4, const int32_t* hint,
                            int32_t* contact_idx, int32_t idx_stride, int32_t idx_col,
                            void* stream) {
  DT_REQUIRE(n_tris == 0 || batch_bounds4, "NULL batch bounds");
  return point_closest_launch(n, x, n_tris, centroids4, batch_bounds4, hint, contact_idx,
                              idx_stride, idx_col, stream, "dt_rope_closest_batched");
}

int dt_rope_collision(const dt_contact_params* cp, float dt, int32_t n, const float* x,
                      const float* v, const int32_t* contact_idx, int32_t idx_stride,
                      int32_t idx_col, int32_t n_tris, const int32_t* tris, const float* pos,
                      const float* vel, const float* out_dir, float* impulse, float* ext_force,
                      int32_t* contact_flag, void* stream) {
  DT_REQUIRE(cp && out_dir, "NULL contact params / out_dir");
  DT_REQUIRE(n >= 0 && n_tris >= 0 && idx_stride > 0 && idx_col >= 0 && idx_col < idx_stride,
             "bad size/stride");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && v && contact_idx && tris && pos && vel && impulse && ext_force,
             "NULL pointer");
  hipLaunchKernelGGL(k_rope_collision, dim3(nblocks(n)), dim3(kBlock), 0, (hipStream_t)stream,
                     contact_dev(cp), dt, (int)n, x, v, contact_idx, idx_stride, idx_col, tris, pos,
                     vel, out_dir[0], out_dir[1], out_dir[2], impulse, ext_force, contact_flag);
  return check_launch("dt_rope_collision");
}

int dt_rope_collision_adj(const dt_contact_params* cp, float dt, int32_t n, const float* x,
                          const float* v, const int32_t* contact_idx, int32_t idx_stride,
                          int32_t idx_col, int32_t n_tris, const int32_t* tris, const float* pos,
                          const float* vel, const float* out_dir, const float* g_impulse,
                          const float* g_ext_force, float* g_x, float* g_v, float* g_pos,
                          float* g_vel, void* stream) {
  DT_REQUIRE(cp && out_dir, "NULL contact params / out_dir");
  DT_REQUIRE(n >= 0 && n_tris >= 0 && idx_stride > 0 && idx_col >= 0 && idx_col < idx_stride,
             "bad size/stride");
  if (n == 0) return DT_OK;
  DT_REQUIRE(x && v && contact_idx && tris && pos && vel && g_impulse && g_ext_force && g_x &&
                 g_v && g_pos && g_vel,
             "NULL pointer");
  hipLaunchKernelGGL(k_rope_collision_adj, dim3(nblocks(n)), dim3(kBlock), 0,
                     (hipStream_t)stream, contact_dev(cp), dt, (int)n, x, v, contact_idx,
                     idx_stride, idx_col, tris, pos, vel, out_dir[0], out_dir[1], out_dir[2],
                     g_impulse, g_ext_force, g_x, g_v, g_pos, g_vel);
  return check_launch("dt_rope_collision_adj");
}

}  // extern "C"
