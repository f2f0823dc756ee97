/*
 * difftactile.h -- C-ABI of libdifftactile_hip.so, the MI355X (gfx950) HIP
 * library behind the differentiable DiffTactile substep.
 *
 * Every entry point is `extern "C"`, takes plain device pointers, sizes and
 * plain-C parameter structs, enqueues its work on `stream` (a hipStream_t;
 * NULL = the default stream) without synchronising or allocating, and returns
 * 0 (DT_OK) or a negative status.  dt_last_error() returns a thread-local
 * message for the last failure on the calling thread.
 *
 * Array layouts are those of the reference's Taichi fields (to_numpy()):
 *   particle vectors  float[n][3]       (x_0, v_0 at one frame)
 *   particle matrices float[n][3][3]    row-major (C_0, F_0, F_new, U/S/V_svd)
 *   grid vectors      float[n^3][3]     node (i,j,k) at (i*n + j)*n + k
 *   grid scalars      float[n^3], int32[n^3]
 * Adjoint ("_adj") entry points ACCUMULATE (+=) into their gradient outputs,
 * as Taichi's kernel.grad does, unless stated otherwise.
 *
 * Reference interfaces replaced (paths relative to the DiffTactile repo):
 *   each dt_mpm_* entry point names the @ti.kernel it replaces.
 */
#ifndef DIFFTACTILE_H
#define DIFFTACTILE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DT_OK 0
#define DT_EINVAL (-1)
#define DT_ELAUNCH (-2)
#define DT_ENOMEM (-3)

/* Material kinds of the per-particle material table. */
#define DT_MAT_ELASTIC 0        /* fixed corotated, J = det(Sigma)   mpm_elastic.py:183-185 */
#define DT_MAT_VON_MISES 1      /* von Mises return map, J = det(F)  mpm_plastic.py:150-166 */
#define DT_MAT_DRUCKER_PRAGER 2 /* sand return map (no reference implementation) */

/* Scalar parameters of one MPM object (mpm_elastic.py:19-66). */
typedef struct dt_mpm_params {
  int32_t n_grid;       /* grid nodes per axis (reference: 64, mpm_elastic.py:35)   */
  int32_t bound;        /* wall band width (reference: 3, mpm_elastic.py:34)        */
  int32_t sticky_floor; /* 1: zero every velocity component at j < bound (no ref.)  */
  int32_t clamp_grad;   /* 1: grid_op_adj clips d/d grid_m to +-grad_clamp
                           (clamp_grid, grasp_plastic.py:191-194)                   */
  float dx, inv_dx, dt, p_vol, p_mass, eps; /* eps: mass threshold, 1e-5           */
  float gravity[3];
  float grad_clamp;     /* 1000 in the reference                                    */
  int32_t soa;          /* particle arrays: 0 = [n][k] (reference layout), 1 = [k][n].
                           Honoured by dt_mpm_g2p*, dt_mpm_p2g_fused*, dt_mpm_substep_*,
                           dt_mpm_advance, dt_mpm_run_* and dt_mpm_cell_keys; the other
                           per-kernel entry points require 0.                       */
  int32_t svd_exact;    /* fused adjoints (dt_mpm_p2g_fused_adj and the loops built on
                           it): 0 = svd_grad as the reference writes it (1e-8 clamp on
                           s_j^2 - s_i^2, mpm_elastic.py:149-155); 1 = exact adjoint
                           (rotation path (M_ij - M_ji)/(s_i + s_j), no clamp). The
                           per-kernel dt_mpm_svd_adj is always the reference form.     */
  int32_t n_env;        /* independent environments stacked along x (0 or 1: one).
                           The grid is (n_env n_grid) x n_grid x n_grid nodes;
                           environment e owns x-nodes [e n_grid, (e+1) n_grid).
                           Particle positions stay local to their environment;
                           env_id[p] names particle p's environment, whose stencil
                           (clamped as for one environment) is shifted by e n_grid
                           nodes along x; walls apply per environment, so
                           environments never exchange momentum.  Contact takes
                           one pad per environment (dt_contact_closest_env).       */
  const int32_t* env_id; /* device array [n] of environment ids (NULL: all 0); it
                           follows the particle order of the arrays passed with it */
  const int32_t* active_bricks;   /* optional node set of the grid-node passes
                           (dt_mpm_grid_op4(_adj), dt_mpm_zero_nodes and the contact
                           passes): device list of 4x4x4-node bricks and, */
  const int32_t* n_active_bricks; /* device count of it.  NULL: every node.  The
                           active-brick list of a binned workspace
                           (dt_mpm_active_bricks) covers every node a stencil of
                           its particles touches, so passes restricted to it leave
                           the other nodes unwritten and unread.                 */
  const int32_t* env_ngrid; /* ragged environments (NULL: every environment uses
                           n_grid, dx, p_mass, p_vol above): device [n_env] grid
                           size n_e <= n_grid of each environment; it occupies
                           nodes [0, n_e)^3 of its n_grid^3 slot, its walls act at
                           n_e - bound.  Honoured by the binned (sparse) path:
                           dt_mpm_bin / dt_mpm_cell_keys, dt_mpm_p2g_binned,
                           dt_mpm_tiles_sum*, dt_mpm_grid_op4*, dt_mpm_g2p4*,
                           dt_mpm_p2g_fused_adj, dt_mpm_run_fwd / _adj and the
                           contact passes; the dense entry points reject it.   */
  const float* env_scale; /* device [n_env][4]: dx, inv_dx, p_mass, p_vol of each
                           environment (with env_ngrid).                          */
} dt_mpm_params;

/* One row of the material table (multi_obj.py:65-74 generalised). */
typedef struct dt_material {
  float mu, lam;        /* Lame parameters                                          */
  float yield_stress;   /* von Mises yield (mpm_plastic.py:69,73)                   */
  float dp_alpha;       /* Drucker-Prager sqrt(2/3) 2 sin(phi) / (3 - sin(phi))     */
  int32_t kind;         /* DT_MAT_*                                                 */
  int32_t pad[3];
} dt_material;

/* Material adjoints.  The adjoint entry points that take g_mats (a device
   array of n_mats x 2 floats, or NULL to skip) ACCUMULATE d/d mu and d/d lam
   of each material row into g_mats[2 id] and g_mats[2 id + 1], as Taichi's
   p2g.grad accumulates mu_0.grad / lambda_0.grad (mpm_elastic.py:68-69,
   multi_obj.py:67-68): the stress terms plus, inside an active return map,
   the von Mises threshold tau_y / (2 mu) (mpm_plastic.py:157-161) and the
   Drucker-Prager cone slope (3 lam + 2 mu) / (2 mu) alpha.  dt_mpm_run_adj
   with a workspace sums them in a fixed order (per-block partials in the
   workspace, reduced at the end of the sweep), so they repeat bit for bit;
   the single-substep entry points add per-block sums with float atomics
   (order not fixed). */

/* FEM pad material and integration constants (fem_sensor.py:35-49). */
typedef struct dt_fem_params {
  float mu, lam;        /* Lame parameters of the stable Neo-Hookean model          */
  float dt, mass;       /* substep and per-vertex mass (vol * rho)                  */
} dt_fem_params;

/* Penalty contact constants (grasp_plastic.py:57-64, 96). */
typedef struct dt_contact_params {
  float kn, kd, kt, mu_f; /* normal stiffness, normal damping, shear stiffness, friction */
  float norm_eps;         /* eps of the shear-velocity norm: 1e-11 in the grasp form
                             (grasp_plastic.py:212), 0 in the cable form
                             (cable_straightening.py:219, .norm() without eps)       */
  float sdf_offset;       /* find_sdf offset (0 for the grid contact)                 */
  int32_t unclamped_vn;   /* 0: v_n = max(n.v_rel, 0) (grasp form, grasp_plastic.py:209);
                             1: v_n = n.v_rel (cable form, cable_straightening.py:216)  */
} dt_contact_params;

const char* dt_last_error(void);
int dt_version(void);

/* ---- stage timing (no reference counterpart: measurement support) ----
   dt_timing_enable(1) clears the records and makes every stage launch of
   dt_mpm_run_fwd / dt_mpm_run_adj / dt_mpm_advance record a HIP event pair on
   its stream; dt_timing_read synchronizes on those events and returns the
   summed duration and the launch count of one stage since the last enable. */
#define DT_STAGE_BIN 0          /* binning: counts, scan, flags, permutation */
#define DT_STAGE_P2G 1          /* brick-owned fused P2G */
#define DT_STAGE_GRID_OP 2      /* tile sums + grid_op */
#define DT_STAGE_G2P 3
#define DT_STAGE_G2P_ADJ 4      /* brick-owned G2P adjoint */
#define DT_STAGE_GRID_OP_ADJ 5
#define DT_STAGE_P2G_ADJ 6      /* fused P2G + stress adjoint */
int dt_timing_enable(int32_t on);
int dt_timing_read(int32_t stage, double* total_ms, int64_t* launches);

/* ---- graph replay (no reference counterpart: launch-overhead removal) ----
   dt_mpm_run_fwd / dt_mpm_run_adj / dt_mpm_advance capture their launch
   sequence into a HIP graph the first time they see a set of arguments and
   replay it with one hipGraphLaunch afterwards (256 most recent argument sets
   kept per process).  On by default; off while stage timing is on. */
int dt_graphs_enable(int32_t on);
int dt_graph_stats(int64_t* hits, int64_t* misses);

/* Graphs of a caller's own launch sequence (no reference counterpart):
   dt_graph_replay launches the graph cached under key (n_key words; the
   caller puts every pointer and value its sequence depends on in it) and
   sets *hit, or sets *hit = 0.  On a miss the caller brackets its launches,
   issued on a capture stream of its own (not the legacy default stream),
   with dt_graph_begin (captures that stream; *capturing = 0 when graphs are
   off: then run the launches eagerly and skip dt_graph_end) and
   dt_graph_end (instantiates, caches under key, launches the graph once on
   launch_stream).  The sequence may contain kernels only: no allocation,
   memset, copy or host synchronisation between the two calls. */
int dt_graph_replay(void* stream, const uint64_t* key, int32_t n_key, int32_t* hit);
int dt_graph_begin(void* stream, int32_t* capturing);
int dt_graph_end(void* stream, void* launch_stream, const uint64_t* key, int32_t n_key);
/* ends a capture begun by dt_graph_begin and discards it (the caller's
   launches failed) */
int dt_graph_abort(void* stream);

/* ---- launch shapes (no reference counterpart: the autotuner's knobs) ----
   Process-wide grid sizes of the persistent launches; 0 keeps the default.
   brick_blocks: blocks of the brick-owned scatter / gather passes (default
   2048; each block strides over the occupied bricks); node_blocks: blocks of
   the active-brick node passes (default 1024, 4 bricks per block and
   iteration); sort_blocks: blocks of the per-brick segment sort (default
   2048).  Results do not depend on them, except the order of the
   material-adjoint partial sums (per block of particles, not of bricks: they
   do not change either).  Graph keys include the setting. */
typedef struct dt_launch_config {
  int32_t brick_blocks;
  int32_t node_blocks;
  int32_t sort_blocks;
} dt_launch_config;
int dt_set_launch_config(const dt_launch_config* cfg);
int dt_get_launch_config(dt_launch_config* cfg);

/* ---- MPM object, one entry point per reference kernel (frame f views) ---- */

/* init_object, mpm_elastic.py:107-114: x = trans_h [p; 1] (trans16: HOST
   row-major 4x4), v = ball_vel3 (HOST), F = I, for the n rest particles p. */
int dt_mpm_init_object(int32_t n, const float* particles, const float* trans16,
                       const float* ball_vel3, float* x, float* v, float* F, void* stream);
/* compute_new_F, mpm_elastic.py:134-137: F_new = (I + dt C) F */
int dt_mpm_compute_new_F(int32_t n, const float* C, const float* F, float* F_new, float dt,
                         void* stream);
/* compute_new_F.grad: gC += dt gF_new F^T; gF += (I + dt C)^T gF_new */
int dt_mpm_compute_new_F_adj(int32_t n, const float* C, const float* F, const float* gF_new,
                             float* gC, float* gF, float dt, void* stream);

/* svd, mpm_elastic.py:139-142 (ti.svd): U, V in SO(3); S diagonal (3x3, sign on S[2][2]) */
int dt_mpm_svd(int32_t n, const float* F_new, float* U, float* S, float* V, void* stream);
/* svd_grad, mpm_elastic.py:144-173: gF_new += U[S_bar + ff(U^T U_bar - U_bar^T U)S
   + S ff(V^T V_bar - V_bar^T V)]V^T, ff terms evaluated in float64 */
int dt_mpm_svd_adj(int32_t n, const float* U, const float* S, const float* V, const float* gU,
                   const float* gS, const float* gV, float* gF_new, void* stream);

/* p2g, mpm_elastic.py:175-208 / mpm_plastic.py:221-256 / multi_obj.py:227-252.
   Scatters into grid_mv/grid_m (accumulating, like the reference's atomics)
   and writes F_next (= F_0[f+1]).  mat_id may be NULL (all material 0). */
int dt_mpm_p2g(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
               const int32_t* mat_id, const float* x, const float* v, const float* C,
               const float* F_new, const float* U, const float* S, const float* V, float* grid_mv,
               float* grid_m, float* F_next, void* stream);
/* p2g.grad: reads d/d grid_mv, d/d grid_m, d/d F_next; accumulates into
   gx, gv, gC, gF_new, gU, gS, gV (the svd factors' adjoints, consumed by
   dt_mpm_svd_adj as the reference's svd_grad consumes U/S/V_svd.grad) and
   g_mats (material adjoints, may be NULL). */
int dt_mpm_p2g_adj(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
                   const int32_t* mat_id, const float* x, const float* v, const float* C,
                   const float* F_new, const float* U, const float* S, const float* V,
                   const float* g_grid_mv, const float* g_grid_m, const float* gF_next, float* gx,
                   float* gv, float* gC, float* gF_new, float* gU, float* gS, float* gV,
                   float* g_mats, void* stream);

/* check_grid_occupy, mpm_elastic.py:210-214: occupy = 1 where grid_m > eps
   (only ever sets 1; the caller clears, as the reference's reset does) */
int dt_mpm_grid_occupy(const dt_mpm_params* prm, const float* grid_m, int32_t* occupy,
                       void* stream);
/* grid_op, mpm_elastic.py:216-240.  grid_f may be NULL (no contact impulse).
   occupy may be NULL (occupancy computed as grid_m > eps).  Writes every
   occupied node of grid_v_out; unoccupied nodes are left untouched. */
int dt_mpm_grid_op(const dt_mpm_params* prm, const float* grid_mv, const float* grid_m,
                   const float* grid_f, const int32_t* occupy, float* grid_v_out, void* stream);
/* grid_op.grad: accumulates into g_grid_mv, g_grid_m (clipped when
   prm->clamp_grad), g_grid_f (may be NULL). */
int dt_mpm_grid_op_adj(const dt_mpm_params* prm, const float* grid_mv, const float* grid_m,
                       const float* grid_f, const int32_t* occupy, const float* g_grid_v_out,
                       float* g_grid_mv, float* g_grid_m, float* g_grid_f, void* stream);

/* g2p, mpm_elastic.py:243-260: writes x_next, v_next, C_next (frame f+1) */
int dt_mpm_g2p(const dt_mpm_params* prm, int32_t n, const float* x, const float* grid_v_out,
               float* x_next, float* v_next, float* C_next, void* stream);
/* g2p.grad: reads gx_next, gv_next, gC_next; accumulates gx and g_grid_v_out */
int dt_mpm_g2p_adj(const dt_mpm_params* prm, int32_t n, const float* x, const float* grid_v_out,
                   const float* gx_next, const float* gv_next, const float* gC_next, float* gx,
                   float* g_grid_v_out, void* stream);

/* ---- fused substep (Contact.update / update_grad for one object without
   pads, grasp_plastic.py:151-178): compute_new_F + svd + p2g fused into one
   particle pass, grid_op, g2p.  Grid buffers are packed float4 per node
   (mv.x, mv.y, mv.z, m) and float4 (v.x, v.y, v.z, unused), n^3 nodes each. */

/* One forward substep: zeroes grid4, scatters, grid update, gather.
   Writes grid4 (momentum, mass), grid_v4 and the next state. */
int dt_mpm_substep_fwd(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats,
                       int32_t n, const int32_t* mat_id, const float* x, const float* v,
                       const float* C, const float* F, float* grid4, float* grid_v4,
                       float* x_next, float* v_next, float* C_next, float* F_next, void* stream);
/* Adjoint of dt_mpm_substep_fwd given its grid4 / grid_v4 outputs: reads the
   next state's adjoints and OVERWRITES gx, gv, gC, gF with the adjoints of
   the state (x, v, C, F).  g_grid4, g_grid_v4: n^3 float4 scratch.  g_mats:
   material adjoints, accumulated (may be NULL). */
int dt_mpm_substep_adj(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats,
                       int32_t n, const int32_t* mat_id, const float* x, const float* v,
                       const float* C, const float* F, const float* grid4, const float* grid_v4,
                       const float* gx_next, const float* gv_next, const float* gC_next,
                       const float* gF_next, float* g_grid4, float* g_grid_v4, float* gx,
                       float* gv, float* gC, float* gF, float* g_mats, void* stream);

/* Stages of the fused substep, for callers that couple other work between
   them (pads and contact run between p2g and grid_op):
   zero4 -> p2g_fused -> grid_op4 -> g2p4, and in reverse
   g2p4_adj -> grid_op4_adj -> p2g_fused_adj.  grid_f / g_grid_f ([n^3][3],
   the contact impulse f*dt, mpm_elastic.py:130-132) may be NULL. */
int dt_mpm_zero4(int64_t n_nodes, float* buf4, void* stream);
/* compute_new_F + svd + return map + p2g over (x, v, C, F); accumulates into grid4 */
int dt_mpm_p2g_fused(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
                     const int32_t* mat_id, const float* x, const float* v, const float* C,
                     const float* F, float* grid4, float* F_next, void* stream);
/* grid_op on packed grids; writes every node of grid_v4 (0 where unoccupied) */
int dt_mpm_grid_op4(const dt_mpm_params* prm, const float* grid4, const float* grid_f,
                    float* grid_v4, void* stream);
/* g2p from a packed velocity grid */
int dt_mpm_g2p4(const dt_mpm_params* prm, int32_t n, const float* x, const float* grid_v4,
                float* x_next, float* v_next, float* C_next, void* stream);
/* g2p adjoint: OVERWRITES gx with its partial; accumulates into g_grid_v4
   (which the caller zeroes first) */
int dt_mpm_g2p4_adj(const dt_mpm_params* prm, int32_t n, const float* x, const float* grid_v4,
                    const float* gx_next, const float* gv_next, const float* gC_next, float* gx,
                    float* g_grid_v4, void* stream);
/* grid_op adjoint: OVERWRITES g_grid4 (and g_grid_f when non-NULL) */
int dt_mpm_grid_op4_adj(const dt_mpm_params* prm, const float* grid4, const float* grid_f,
                        const float* g_grid_v4, float* g_grid4, float* g_grid_f, void* stream);
/* adjoint of dt_mpm_p2g_fused: gx holds the g2p partial on entry and the full
   adjoint on return; OVERWRITES gv, gC, gF; accumulates g_mats (may be NULL) */
int dt_mpm_p2g_fused_adj(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats,
                         int32_t n, const int32_t* mat_id, const float* x, const float* v,
                         const float* C, const float* F, const float* g_grid4,
                         const float* gF_next, float* gx, float* gv, float* gC, float* gF,
                         float* g_mats, void* stream);

/* Native trajectory loops over a frame-major state buffer: frame k of x is
   at x + k*n*3 (likewise v, C (n*9), F (n*9)).  dt_mpm_run_fwd advances
   frames [0, n_substeps] and, when grids4 / grids_v4 are non-NULL, keeps
   each substep's grids (n^3 float4 per substep) for dt_mpm_run_adj;
   otherwise grids4 / grids_v4 must be passed as single-substep scratch in
   scratch4 / scratch_v4.  workspace (dt_mpm_workspace_bytes(prm, n) bytes,
   256-byte aligned) selects the sparse path: each substep bins the particles
   by cell, visits only the 4x4x4-node bricks some stencil touches, and
   scatters with the brick-owned tile passes; grid nodes outside those bricks
   are neither written nor read.  NULL keeps the dense passes over all n^3
   nodes.  On the sparse path grids_v4 may be NULL with grids4 kept: the
   substep's velocity grid then lives in scratch_v4 and dt_mpm_run_adj (given
   grids_v4 NULL too) remakes each brick's velocities from grids4 by the same
   node update, bit for bit, so a kept trajectory holds 16 B per node and
   substep instead of 32. */
int dt_mpm_run_fwd(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
                   const int32_t* mat_id, float* x, float* v, float* C, float* F,
                   int32_t n_substeps, float* grids4, float* grids_v4, float* scratch4,
                   float* scratch_v4, void* workspace, size_t workspace_bytes, void* stream);
/* Reverse sweep over frames [0, n_substeps] stored by dt_mpm_run_fwd with
   grids kept.  Seeds: gx..gF hold the adjoint of frame n_substeps on entry and
   the adjoint of frame 0 on return; gx2..gF2 are same-size scratch.  g_mats
   accumulates the material adjoints over the sweep (may be NULL).  grids_v4
   NULL (sparse path only): the velocities come from grids4, as kept by a
   dt_mpm_run_fwd with grids_v4 NULL. */
int dt_mpm_run_adj(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
                   const int32_t* mat_id, const float* x, const float* v, const float* C,
                   const float* F, int32_t n_substeps, const float* grids4, const float* grids_v4,
                   float* g_grid4, float* g_grid_v4, float* gx, float* gv, float* gC, float* gF,
                   float* gx2, float* gv2, float* gC2, float* gF2, void* workspace,
                   size_t workspace_bytes, float* g_mats, void* stream);

/* Advances the state (x, v, C, F) by n_substeps in place, ping-ponging with
   the same-size scratch state (x2, v2, C2, F2); grids are single-substep
   scratch (n^3 float4 each).  Keeps no history (forward-only / checkpoint
   segments). */
int dt_mpm_advance(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
                   const int32_t* mat_id, float* x, float* v, float* C, float* F, float* x2,
                   float* v2, float* C2, float* F2, int32_t n_substeps, float* scratch4,
                   float* scratch_v4, void* workspace, size_t workspace_bytes, void* stream);

/* No counterpart in the reference: dt_mpm_advance for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc).  Same result as dt_mpm_advance on the
   sparse path, with the ping-pong state, the two grids and the binning
   workspace in a static scratch array the library holds on every device it
   runs on (64 MiB each, until the process ends).  The caller passes only the
   state, advanced in place.  Rejects (DT_EINVAL) a particle count and grid
   whose scratch exceeds 64 MiB: 96 B per particle, 32 B per grid node, plus
   the workspace (dt_mpm_workspace_bytes); a 64^3 grid holds about 400k
   particles.  Calls no hipMalloc / hipFree and captures no graph.  Every call
   on one device shares that device's scratch: calls that overlap on one GPU
   (different streams, threads) race. */
int dt_mpm_advance_static(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats,
                          int32_t n, const int32_t* mat_id, float* x, float* v, float* C,
                          float* F, int32_t n_substeps, void* stream);

/* ---- sparse path: workspace, binning and brick-owned scatter passes ----
   (no reference counterpart: the reference scatters with per-particle
   atomics into a dense grid, mpm_elastic.py:175-208; these stages give the
   same grids) */

/* Bytes of device workspace the sparse path needs for n particles:
   per-brick counts and starts, occupied- and active-brick lists,
   per-particle key / rank / permutation, one 6x6x6-node float4 tile per
   brick, scan scratch. */
int dt_mpm_workspace_bytes(const dt_mpm_params* prm, int32_t n, size_t* bytes);
/* Extra bytes, after dt_mpm_workspace_bytes, for the binning history of an
   n_substeps trajectory.  A workspace with that room makes a grid-keeping
   dt_mpm_run_fwd store each substep's binning, and a following dt_mpm_run_adj
   on the same frames (same x pointer, n and n_substeps, frames unmodified in
   between) reuses it instead of binning again. */
int dt_mpm_workspace_history_bytes(const dt_mpm_params* prm, int32_t n, int32_t n_substeps,
                                   size_t* bytes);
/* Bins the particles of x by the 4x4x4-cell brick of their stencil base
   (bricks numbered as in dt_mpm_cell_keys, bases clamped to the stencil's
   [0, n_grid-3] range) and marks the active bricks. */
int dt_mpm_bin(const dt_mpm_params* prm, int32_t n, const float* x, void* workspace,
               size_t workspace_bytes, void* stream);
/* Device pointers of the last dt_mpm_bin's active-brick list and its count
   (for dt_mpm_params.active_bricks / n_active_bricks). */
int dt_mpm_active_bricks(const dt_mpm_params* prm, int32_t n, const void* workspace,
                         size_t workspace_bytes, const int32_t** active_bricks,
                         const int32_t** n_active_bricks);
/* Zeroes width floats per node over the node set of prm (all nodes, or the
   active bricks). */
int dt_mpm_zero_nodes(const dt_mpm_params* prm, float* buf, int32_t width, void* stream);
/* Copies, for n_arrays (1..4) node arrays at once, widths[a] 32-bit words per
   node between the dense array dense[a] and the compact array compact[a]
   holding the nodes of prm's active bricks (64 per brick, in the order of the
   active-brick list; prm must carry the node set): direction 0 packs
   dense -> compact, 1 unpacks compact -> dense (other dense nodes are left as
   they are).  dense, compact and widths are host arrays of device pointers
   and widths.  Only the first capacity bricks are copied; a pack of more
   bricks sets *overflow (a device int32, NULL: not reported) to 1.  No
   counterpart in the reference, which keeps every substep's dense grids of a
   window in its [sub_steps, n_grid^3] fields (mpm_elastic.py:82-87): the
   coupled rollout keeps only the active bricks' nodes of each kept substep
   when whole grids do not fit its HBM budget. */
int dt_mpm_brick_nodes_copy(const dt_mpm_params* prm, int32_t n_arrays, void* const* dense,
                            void* const* compact, const int32_t* widths, int32_t capacity,
                            int32_t direction, int32_t* overflow, void* stream);
/* Device pointers into the workspace: perm (n) and brick_start
   (dt_mpm_brick_count + 1) of the last dt_mpm_bin; the particles of brick b are
   perm[brick_start[b] .. brick_start[b+1]) in no particular order. */
int dt_mpm_bin_perm(const dt_mpm_params* prm, int32_t n, const void* workspace,
                    size_t workspace_bytes, const int32_t** perm, const int32_t** brick_start);
/* Bytes of one saved binning for n particles (perm, brick starts, occupied-
   and active-brick lists and their counts); slots are 256-byte aligned. */
int dt_mpm_bin_slot_bytes(const dt_mpm_params* prm, int32_t n, size_t* bytes);
/* save = 1: copies the last dt_mpm_bin of the workspace into slot;
   save = 0: copies a saved binning from slot back into the workspace, after
   which the binned passes and dt_mpm_active_bricks see it as if dt_mpm_bin
   had run again on the same x (the adjoint of a substep reuses its forward
   binning this way). */
int dt_mpm_bin_copy(const dt_mpm_params* prm, int32_t n, void* workspace, size_t workspace_bytes,
                    void* slot, int32_t save, void* stream);
/* Brick-owned fused p2g over the binned particles: writes F_next and the
   per-brick tiles in the workspace (after dt_mpm_bin on the same x). */
int dt_mpm_p2g_binned(const dt_mpm_params* prm, const dt_material* mats, int32_t n_mats, int32_t n,
                      const int32_t* mat_id, const float* x, const float* v, const float* C,
                      const float* F, float* F_next, void* workspace, size_t workspace_bytes,
                      void* stream);
/* Sums the p2g tiles into grid4 (mv, m) on the active bricks and applies
   grid_op there (grid_v4). */
int dt_mpm_grid_op4_tiles(const dt_mpm_params* prm, int32_t n, const void* workspace,
                          size_t workspace_bytes, float* grid4, float* grid_v4, void* stream);
/* Writes the tile sums of the last brick pass (p2g: (mv, m); g2p adjoint:
   (d/d v, 0)) into out4 on the active bricks; other nodes are untouched. */
int dt_mpm_tiles_sum(const dt_mpm_params* prm, int32_t n, const void* workspace,
                     size_t workspace_bytes, float* out4, void* stream);
/* dt_mpm_tiles_sum that also zeroes width (1..4) floats per node of zero_buf
   on the same nodes (the coupled substep clears its contact force grid
   this way). */
int dt_mpm_tiles_sum_zero(const dt_mpm_params* prm, int32_t n, const void* workspace,
                          size_t workspace_bytes, float* out4, float* zero_buf, int32_t width,
                          void* stream);
/* Brick-owned g2p adjoint over the binned particles: OVERWRITES gx with its
   partial and writes the grid-velocity adjoint tiles. */
int dt_mpm_g2p4_adj_binned(const dt_mpm_params* prm, int32_t n, const float* x,
                           const float* grid_v4, const float* gx_next, const float* gv_next,
                           const float* gC_next, float* gx, void* workspace,
                           size_t workspace_bytes, void* stream);
/* Sums the adjoint tiles into g_grid_v4 on the active bricks and applies the
   grid_op adjoint there (g_grid4). */
int dt_mpm_grid_op4_adj_tiles(const dt_mpm_params* prm, int32_t n, const void* workspace,
                              size_t workspace_bytes, const float* grid4, float* g_grid_v4,
                              float* g_grid4, void* stream);
/* The same with the contact force grid of dt_mpm_grid_op4_adj (grid_f, 3
   floats per node): writes g_grid4 and g_grid_f on the active bricks, and the
   summed d/d grid_v_out into g_grid_v4 unless it is NULL.  The coupled
   substep's adjoint (grasp_plastic.py:167-178) uses it after
   dt_mpm_g2p4_adj_binned. */
int dt_mpm_grid_op4_adj_tiles_f(const dt_mpm_params* prm, int32_t n, const void* workspace,
                                size_t workspace_bytes, const float* grid4, const float* grid_f,
                                float* g_grid_v4, float* g_grid4, float* g_grid_f, void* stream);

/* Active-brick passes on explicit flags. */
int64_t dt_mpm_brick_count(const dt_mpm_params* prm);
/* flags[b] = 1 for every brick touched by a stencil of the particles at x
   (flags cleared first). */
int dt_mpm_mark_bricks(const dt_mpm_params* prm, int32_t n, const float* x, int32_t* flags,
                       void* stream);
/* brick-restricted forms of dt_mpm_zero4 / dt_mpm_grid_op4 / dt_mpm_grid_op4_adj
   (no contact impulse) */
int dt_mpm_zero4_bricks(const dt_mpm_params* prm, const int32_t* flags, float* buf4,
                        void* stream);
int dt_mpm_grid_op4_bricks(const dt_mpm_params* prm, const int32_t* flags, const float* grid4,
                           float* grid_v4, void* stream);
int dt_mpm_grid_op4_adj_bricks(const dt_mpm_params* prm, const int32_t* flags, const float* grid4,
                               const float* g_grid_v4, float* g_grid4, void* stream);

/* ---- particle ordering ---- */

/* Cell key of each particle's stencil base (4x4x4-cell bricks, row-major
   bricks, row-major cells inside a brick): key = brick*64 + local. */
int dt_mpm_cell_keys(const dt_mpm_params* prm, int32_t n, const float* x, int32_t* keys,
                     void* stream);
/* Stable LSD radix sort of (key, index) pairs: writes keys_out (sorted) and
   perm (perm[i] = original index of the i-th particle in sorted order).
   tmp: at least dt_sort_tmp_bytes(n) bytes of device scratch. */
int64_t dt_sort_tmp_bytes(int32_t n);
int dt_sort_by_key(int32_t n, int32_t key_bits, const int32_t* keys, int32_t* keys_out,
                   int32_t* perm, void* tmp, int64_t tmp_bytes, void* stream);
/* dst[i, :] = src[perm[i], :] for rows of `width` floats. */
int dt_gather_rows(int32_t n, int32_t width, const int32_t* perm, const float* src, float* dst,
                   void* stream);
/* dst[perm[i], :] = src[i, :]. */
int dt_scatter_rows(int32_t n, int32_t width, const int32_t* perm, const float* src, float* dst,
                    void* stream);
/* [n][width] -> [width][n]: dst[c*n + i] = src[perm[i]*width + c] (perm NULL = identity). */
int dt_gather_rows_soa(int32_t n, int32_t width, const int32_t* perm, const float* src_rows,
                       float* dst_soa, void* stream);
/* [width][n] -> [width][n] reordered (no reference counterpart: the re-sort of
   a trajectory's particles at its checkpoints): scatter 0: dst[c*n + i] =
   src[c*n + perm[i]]; scatter 1: dst[c*n + perm[i]] = src[c*n + i] (the
   inverse).  Any 4-byte element (int32 ids as their bit patterns). */
int dt_permute_soa(int32_t n, int32_t width, const int32_t* perm, int32_t scatter,
                   const float* src_soa, float* dst_soa, void* stream);
/* [width][n] -> [n][width]: dst[perm[i]*width + c] = src[c*n + i] (perm NULL = identity). */
int dt_scatter_rows_soa(int32_t n, int32_t width, const int32_t* perm, const float* src_soa,
                        float* dst_rows, void* stream);


/* ---- FEM pad (sensor_model/fem_sensor.py, gripper_fem.py) ----
   pos/vel/vpos/ext_f/cv: float[n_verts][3] at one frame; f2v int32[n_tets][4];
   B float[n_tets][3][3]; layer_id int32[n_verts] (0 = fixed base layer). */

/* init_pos, fem_sensor.py:334-338: B = [a-d, b-d, c-d]^-1 from world positions */
int dt_fem_init_B(int32_t n_tets, const int32_t* f2v, const float* pos0, float* B, void* stream);
/* update, fem_sensor.py:408-440: vel[v] += dt * force / mass (stable Neo-Hookean,
   current volume |det D|/6), in place on the frame's velocities */
int dt_fem_update(const dt_fem_params* prm, int32_t n_tets, const int32_t* f2v, const float* B,
                  const float* pos, float* vel, void* stream);
/* update.grad: reads g_vel (adjoint of vel after update); accumulates g_pos */
int dt_fem_update_adj(const dt_fem_params* prm, int32_t n_tets, const int32_t* f2v,
                      const float* B, const float* pos, const float* g_vel, float* g_pos,
                      void* stream);
/* Atomic-free forms of dt_fem_update / dt_fem_update_adj (same results up to
   float summation order): each tet writes its four vertex contributions to
   tet_buf (16 * 4 * n_tets bytes of device scratch), then each vertex adds the
   entries of its incident (tet, corner) pairs.  vert_start (n_verts + 1) and
   vert_tets (4 * n_tets entries, 4 * tet + corner, grouped by vertex) are the
   mesh's vertex-to-tet CSR, built once per mesh. */
int dt_fem_update_csr(const dt_fem_params* prm, int32_t n_tets, int32_t n_verts,
                      const int32_t* f2v, const float* B, const float* pos, float* vel,
                      const int32_t* vert_start, const int32_t* vert_tets, float* tet_buf,
                      void* stream);
int dt_fem_update_adj_csr(const dt_fem_params* prm, int32_t n_tets, int32_t n_verts,
                          const int32_t* f2v, const float* B, const float* pos,
                          const float* g_vel, float* g_pos, const int32_t* vert_start,
                          const int32_t* vert_tets, float* tet_buf, void* stream);
/* update2, fem_sensor.py:443-457 (cv_next NULL) / gripper_fem.py:403-418
   (cv_next = control_vel[f+1]): writes frame f+1 */
int dt_fem_update2(const dt_fem_params* prm, int32_t n_verts, const int32_t* layer_id,
                   const float* pos, const float* vel, const float* vpos, const float* ext_f,
                   const float* cv, const float* cv_next, float* pos_next, float* vel_next,
                   float* vpos_next, void* stream);
/* update2.grad: accumulates into g_pos, g_vel, g_vpos, g_ext_f and (when
   non-NULL) g_cv, g_cv_next */
int dt_fem_update2_adj(const dt_fem_params* prm, int32_t n_verts, const int32_t* layer_id,
                       const float* g_pos_next, const float* g_vel_next, const float* g_vpos_next,
                       float* g_pos, float* g_vel, float* g_vpos, float* g_ext_f, float* g_cv,
                       float* g_cv_next, void* stream);
/* clamp_grid, grasp_plastic.py:191-202: a[i] = clip(a[i], -limit, limit) */
int dt_clamp(int64_t n, float* a, float limit, void* stream);

/* ---- grid penalty contact (tasks/grasp_plastic.py:204-299) ----
   The grid is given as momentum (stride mv_stride floats per node) and mass
   (stride m_stride): (grid_v_in, grid_m) -> strides 3 and 1; packed grid4 ->
   grid4 and grid4 + 3 with strides 4 and 4.  contact_idx is int32 with
   idx_stride entries per node (one column per pad), -1 = none.
   out_dir is a HOST pointer to 3 floats (the pad's outward direction). */

/* triangle centroids (1/3)((p1 + p2) + p3), float4 per triangle: xyz the
   centroid, w = (1 - 4e-6)|c|^2, a bound the closest-triangle search uses to
   skip triangles (any w between 0 and that value gives the same indices) */
int dt_contact_centroids(int32_t n_tris, const int32_t* tris, const float* pos, float* centroids4,
                         void* stream);
/* check_collision + find_closest (fem_sensor.py:343-359) for one pad: for each
   node with grid_m > eps, the first triangle minimising
   sqrt(|c - q|^2 + 1e-10) over q = cell centre ((i+.5)dx, (j+.5)dx, (k+.5)dx);
   -1 when none is below 100.  Bit-exact float32 rule. */
int dt_contact_closest(const dt_mpm_params* prm, const float* grid_m, int32_t m_stride,
                       int32_t n_tris, const float* centroids4, int32_t* contact_idx,
                       int32_t idx_stride, int32_t idx_col, void* stream);
/* The same search over stacked environments (prm->n_env, one pad per
   environment, the pads' triangles concatenated): a node of environment
   e = i / n_grid searches triangles [tri_start[e], tri_start[e+1]) only and
   its cell centre is local to the environment.  tri_start: device int32
   [n_env + 1].  The collision entry points take the concatenated pads as
   they are (their positions local to each environment). */
int dt_contact_closest_env(const dt_mpm_params* prm, const float* grid_m, int32_t m_stride,
                           const int32_t* tri_start, int32_t n_tris, const float* centroids4,
                           int32_t* contact_idx, int32_t idx_stride, int32_t idx_col,
                           void* stream);
/* Accelerated form of the same search (same indices, bit for bit): the
   bounding boxes of each run of DT_CONTACT_TRI_RUN centroids and of each
   group of DT_CONTACT_TRI_GROUP runs let the search skip groups and runs
   beyond a node's best distance.  dt_contact_tri_batches writes the boxes
   from the centroids of dt_contact_centroids, two float4 each (lower corner,
   upper corner; .w = 0), DT_CONTACT_TRI_BOUNDS(n_tris) float4 in all: the
   ceil(n_tris / DT_CONTACT_TRI_RUN) runs' boxes first, then the groups'.  tri_start as in dt_contact_closest_env, or
   NULL for a single environment.  hint (optional, same layout as contact_idx, may be
   contact_idx itself): a triangle per node to start from, such as the
   previous substep's result; entries outside the node's range are ignored,
   and any start gives the same result.  Replaces fem_sensor.py:343-359 like
   dt_contact_closest. */
#define DT_CONTACT_TRI_RUN 8
#define DT_CONTACT_TRI_GROUP 8
#define DT_CONTACT_TRI_BOUNDS(n_tris)                                      \
  (2 * (((n_tris) + DT_CONTACT_TRI_RUN - 1) / DT_CONTACT_TRI_RUN +         \
        ((n_tris) + DT_CONTACT_TRI_RUN * DT_CONTACT_TRI_GROUP - 1) /       \
            (DT_CONTACT_TRI_RUN * DT_CONTACT_TRI_GROUP)))
int dt_contact_tri_batches(int32_t n_tris, const float* centroids4, float* batch_bounds4,
                           void* stream);
int dt_contact_closest_batched(const dt_mpm_params* prm, const float* grid_m, int32_t m_stride,
                               const int32_t* tri_start, int32_t n_tris, const float* centroids4,
                               const float* batch_bounds4, const int32_t* hint,
                               int32_t* contact_idx, int32_t idx_stride, int32_t idx_col,
                               void* stream);
/* collision (one pad): find_sdf + calculate_contact_force; grid_f += f*dt,
   ext_force[tri verts] += -f/3 */
int dt_contact_collision(const dt_mpm_params* prm, const dt_contact_params* cp,
                         const float* grid_mv, int32_t mv_stride, const float* grid_m,
                         int32_t m_stride, const int32_t* contact_idx, int32_t idx_stride,
                         int32_t idx_col, int32_t n_tris, const int32_t* tris, const float* pos,
                         const float* vel, const float* out_dir, float* grid_f, float* ext_force,
                         void* stream);
/* collision.grad: reads g_grid_f, g_ext_force; accumulates into g_grid_mv,
   g_grid_m (strided like the grid) and g_pos, g_vel of the pad */
int dt_contact_collision_adj(const dt_mpm_params* prm, const dt_contact_params* cp,
                             const float* grid_mv, int32_t mv_stride, const float* grid_m,
                             int32_t m_stride, const int32_t* contact_idx, int32_t idx_stride,
                             int32_t idx_col, int32_t n_tris, const int32_t* tris,
                             const float* pos, const float* vel, const float* out_dir,
                             const float* g_grid_f, const float* g_ext_force, float* g_grid_mv,
                             int32_t gmv_stride, float* g_grid_m, int32_t gm_stride, float* g_pos,
                             float* g_vel, void* stream);
/* The same two kernels with the pad-vertex sums exact and order-free, so they
   repeat bit for bit (no reference counterpart; collision and collision.grad,
   grasp_plastic.py:279-299): the vertex terms are added as int64 fixed-point
   numbers at 2^-40 of the largest term (three launches: largest term, sums,
   conversion).  acc: n_verts*3 (forward) or n_verts*6 (adjoint) int64,
   scale_state: 2 (forward) or 4 (adjoint) uint32; both zero before the first
   call and left zero by every call (one pair per pad, not shared by calls
   that may overlap).  ext_force, g_pos, g_vel accumulate (+=). */
int dt_contact_collision_fx(const dt_mpm_params* prm, const dt_contact_params* cp,
                            const float* grid_mv, int32_t mv_stride, const float* grid_m,
                            int32_t m_stride, const int32_t* contact_idx, int32_t idx_stride,
                            int32_t idx_col, int32_t n_tris, const int32_t* tris,
                            const float* pos, const float* vel, const float* out_dir,
                            float* grid_f, int32_t n_verts, float* ext_force, int64_t* acc,
                            uint32_t* scale_state, void* stream);
int dt_contact_collision_adj_fx(const dt_mpm_params* prm, const dt_contact_params* cp,
                                const float* grid_mv, int32_t mv_stride, const float* grid_m,
                                int32_t m_stride, const int32_t* contact_idx,
                                int32_t idx_stride, int32_t idx_col, int32_t n_tris,
                                const int32_t* tris, const float* pos, const float* vel,
                                const float* out_dir, const float* g_grid_f,
                                const float* g_ext_force, float* g_grid_mv, int32_t gmv_stride,
                                float* g_grid_m, int32_t gm_stride, int32_t n_verts,
                                float* g_pos, float* g_vel, int64_t* acc,
                                uint32_t* scale_state, void* stream);


/* ---- losses on the measured path ---- */

/* Reductions here are float32 in a fixed order (per-block partials in a
   caller scratch of DT_REDUCE_BLOCKS x terms floats, then one block over
   them), so every result repeats bit for bit. */
/* compute_obj_pos, tasks/grasp_elastic.py:240-243: out3 = mean_p x[p] (device
   float[3]); scratch 3 DT_REDUCE_BLOCKS floats; soa: x layout. */
int dt_loss_com(int32_t n, int32_t soa, const float* x, float* scratch, float* out3, void* stream);
/* adjoint of the mean position: gx[p] += g_com3 / n (g_com3: device float[3]) */
int dt_loss_com_adj(int32_t n, int32_t soa, const float* g_com3, float* gx, void* stream);
/* compute_contact_force, tasks/grasp_plastic.py:218-230: out3 = 1/6 sum over
   the pad's contact triangles of the external force at their three vertices
   (device float[3]); scratch 3 DT_REDUCE_BLOCKS floats. */
int dt_contact_force_sum(int32_t n_tris, const int32_t* tris, const float* ext, float* scratch,
                         float* out3, void* stream);
/* adjoint: g_ext[v] += (triangles holding v) / 6 * g3 (g3: device float[3]) */
int dt_contact_force_sum_adj(int32_t n_tris, const int32_t* tris, const float* g3, float* g_ext,
                             void* stream);
/* compute_force_loss, tasks/grasp_plastic.py:232-238, over n_pads (1..8) pads:
   p_k = inv_rot_k cf_k (inv_rot9: HOST row-major 3x3 per pad, the sensors'
   inv_rot; cf: device n_pads x 3); loss = beta sum_k |p_k.x| / |mu p_k.y| when
   every |p_k.x| > 1e-3, else 0.  out (device, 1 + 3 n_pads floats): out[0] =
   loss, out[1 + 3k + c] = d loss / d cf_k[c]. */
int dt_force_loss(int32_t n_pads, const float* cf, const float* inv_rot9, float mu, float beta,
                  float* out, void* stream);
/* compute_marker_motion, tasks/grasp_plastic.py:240-246: out2 = mean over the
   markers of pred - virt (device float[2]); scratch 2 DT_REDUCE_BLOCKS floats.
   The reference's loop divides both sensors' sums by sensor 1's marker count;
   here each pad's by its own (the same for pads of one mesh). */
int dt_marker_offset(int32_t n_markers, const float* pred, const float* virt, float* scratch,
                     float* out2, void* stream);
/* adjoint: g_pred[m] += g2 / n_markers; g_virt[m] -= g2 / n_markers (g_virt may
   be NULL; g2: device float[2]) */
int dt_marker_offset_adj(int32_t n_markers, const float* g2, float* g_pred, float* g_virt,
                         void* stream);
/* extract_markers, sensor_model/fem_sensor.py:161-187 with project_3d_2d,
   fisheye_model.py:13-30: surface vertices (surface_id) of pos / vpos mapped
   by inv_trans16 (HOST row-major 4x4), reordered (x, z, y), projected with
   fisheye5 = HOST {f, m, cx, cy, z_offset} into surf_loc / surf_vloc
   [n_surface][2]; markers interpolated with k_closest inverse-distance weights
   into pred / virt [n_markers][2]. */
int dt_markers_fwd(int32_t n_surface, const int32_t* surface_id, const float* pos,
                   const float* vpos, const float* inv_trans16, const float* fisheye5,
                   int32_t n_markers, int32_t k_closest, const int32_t* interp_idx,
                   const float* interp_w, float* surf_loc, float* surf_vloc, float* pred,
                   float* virt, void* stream);
/* adjoint of pred w.r.t. pos: g_surf_loc (scratch [n_surface][2]) is written,
   g_pos accumulates.  The interpolation's adjoint gathers per surface point:
   surf_entry[surf_start[s] .. surf_start[s+1]) lists m k_closest + j for every
   interp_idx[m][j] == s (surface ids distinct). */
int dt_markers_adj(int32_t n_surface, const int32_t* surface_id, const float* pos,
                   const float* inv_trans16, const float* fisheye5, int32_t n_markers,
                   int32_t k_closest, const int32_t* surf_start, const int32_t* surf_entry,
                   const float* interp_w, const float* g_pred, float* g_surf_loc, float* g_pos,
                   void* stream);

/* ---- parallel-jaw kinematics and pad control velocities ----
   (sensor_model/gripper_kinematics.py).  A trajectory's frames are six
   row-major 4x4 matrices each (DT_JAW_FRAME_FLOATS floats): ee_local, ee,
   finger1_local, finger2_local, finger1_ee, finger2_ee (the fields of
   Gripper, gripper_kinematics.py:82-93).  Controls are seven floats per step:
   d_pos xyz, d_ori xyz (degrees per unit time), d_gripper (the d_pos /
   d_ori / d_gripper fields, 104-107). */
#define DT_JAW_FRAME_FLOATS 96
typedef struct dt_jaw_params {
  float ee_world[16];    /* end effector -> world (init_params, 125-127)      */
  float finger_t[2][16]; /* end effector -> finger 1 / 2 (139-158)           */
  float dt;              /* time step of the chain (Gripper.dt)               */
} dt_jaw_params;
/* Frames 1..n_steps from frame 0 (filled by the caller, as init_params does)
   and controls[n_steps][7]: pos_transformation (186-192) then grip_kinematic
   (169-183) per step. */
int dt_jaw_chain(const dt_jaw_params* prm, int32_t n_steps, const float* controls, float* frames,
                 void* stream);
/* Adjoint of dt_jaw_chain (kinematic_grad, 219-229, for the chain part):
   consumes g_frames[0..n_steps] (accumulated into, frame 0 last) and
   accumulates g_controls[n_steps][7]. */
int dt_jaw_chain_adj(const dt_jaw_params* prm, int32_t n_steps, const float* controls,
                     const float* frames, float* g_frames, float* g_controls, void* stream);
/* fem_sensor{1,2}_set_control_vel (194-216) for steps 0..n_steps-1:
   control_vel[f][i] = ((E[f+1] x_i) - (E[f] x_i))_{0:3} / dt with E the pose of
   finger 1 or 2 in frames[0..n_steps] and x_i = (init_x[i], 1). */
int dt_jaw_control_vel(int32_t n_steps, int32_t n_verts, int32_t finger, const float* frames,
                       const float* init_x, float dt, float* control_vel, void* stream);
/* Adjoint: g_frames' finger-pose entries of frames 0..n_steps accumulate. */
int dt_jaw_control_vel_adj(int32_t n_steps, int32_t n_verts, int32_t finger, const float* init_x,
                           float dt, const float* g_control_vel, float* g_frames, void* stream);

/* ---- dome pad pose control (sensor_model/fem_sensor.py:196-266) ----
   A pose is DT_DOME_POSE_FLOATS floats of device memory holding the
   FEMDomeSensor fields below (row-major).  d6 (device) = (d_pos xyz, d_ori xyz
   in degrees per unit time); step_time = dt (sub_steps - 1), the span one
   call covers.  The step transform is eul2mat (196-210): R = Rz Ry Rx of
   d_ori step_time, t = d_pos step_time. */
#define DT_DOME_TRANS_WORLD 0    /* [16] pad -> world at init (fixed) */
#define DT_DOME_TRANS_LOCAL 16   /* [16] accumulated local steps      */
#define DT_DOME_TRANS_H 32       /* [16] trans_world trans_local      */
#define DT_DOME_INV_TRANS_H 48   /* [16]                              */
#define DT_DOME_DTRANS_H 64      /* [16] W T W^-1 of the last step    */
#define DT_DOME_ROT_WORLD 80     /* [9]                               */
#define DT_DOME_ROT_LOCAL 89     /* [9]                               */
#define DT_DOME_ROT_H 98         /* [9]                               */
#define DT_DOME_INV_ROT 107      /* [9]                               */
#define DT_DOME_POSE_FLOATS 116
/* mode 0: set_pose_control (229-245): dtrans_h = W T W^-1, trans_local =
   T trans_local, trans_h = W trans_local, rot_local = R rot_local, rot_h =
   rot_world rot_local; mode 1: set_trans_h (217-227): trans_h = T trans_h,
   rot_h = R rot_h; mode 2: set_pose_control_bp (248-256): dtrans_h only.
   Every mode then sets inv_trans_h and inv_rot (general inverses).
   pose_out may be pose_in. */
int dt_dome_pose_control(int32_t mode, const float* d6, float step_time, const float* pose_in,
                         float* pose_out, void* stream);
/* adjoint (set_pose_control.grad): g_pose_out holds d/d(every field of the
   output pose); g_pose_in receives the gradient of the input pose (the
   fields the mode overwrites get 0, the ones it reads their gradient; it may
   be g_pose_out), g_d6 accumulates d/d(d_pos, d_ori). */
int dt_dome_pose_control_adj(int32_t mode, const float* d6, float step_time, const float* pose_in,
                             const float* g_pose_out, float* g_pose_in, float* g_d6,
                             void* stream);
/* set_control_vel (258-266): control_vel[i] = ((dtrans_h [vpos_i; 1]) -
   vpos_i)_{0:3} / step_time. */
int dt_dome_control_vel(int32_t n_verts, const float* vpos, const float* pose, float step_time,
                        float* control_vel, void* stream);
/* adjoint: g_pose's dtrans_h accumulates sum_i g_cv_i [vpos_i; 1]^T / step_time
   (fixed order; scratch 12 DT_REDUCE_BLOCKS floats); g_vpos (may be NULL)
   accumulates (dtrans_h^T g_cv_i - g_cv_i) / step_time. */
int dt_dome_control_vel_adj(int32_t n_verts, const float* vpos, const float* pose,
                            const float* g_control_vel, float step_time, float* scratch,
                            float* g_pose, float* g_vpos, void* stream);

/* ---- grid-mass deformation losses (tasks/deform_loss.py) ---- */
/* Mass-only p2g (compute_grid_m_kernel, mpm_plastic.py:168-178): grid_m
   [n^3] += sum_p w(node) p_mass over x [n][3]; one environment. */
int dt_mpm_grid_mass(const dt_mpm_params* prm, int32_t n, const float* x, float* grid_m,
                     void* stream);
/* Its adjoint w.r.t. x (base held fixed): gx [n][3] accumulates. */
int dt_mpm_grid_mass_adj(const dt_mpm_params* prm, int32_t n, const float* x,
                         const float* g_grid_m, float* gx, void* stream);
/* update_target (deform_loss.py:61-86): n_sweeps sweeps (the reference runs
   2 n_grid) of the nearest-point distance propagation from the nodes whose
   target density exceeds 1e-4; sdf [n^3] and nearest [n^3][3] receive the
   result, sdf_tmp / nearest_tmp are scratch of the same sizes.  Nodes at
   i dx; distance sqrt(|d|^2 + 1e-8); 1000 where no point has arrived. */
int dt_deform_target_sdf(int32_t n_grid, float dx, const float* density, int32_t n_sweeps,
                         float* sdf, float* nearest, float* sdf_tmp, float* nearest_tmp,
                         void* stream);
/* Sum of block partials: reductions use a caller scratch of
   DT_REDUCE_BLOCKS x (number of terms) floats and a fixed order. */
#define DT_REDUCE_BLOCKS 256
/* out2 = (sum |m - t|, sum sdf m) (compute_density_loss_kernel,
   compute_sdf_loss_kernel, deform_loss.py:99-107); scratch 2 DT_REDUCE_BLOCKS. */
int dt_deform_loss(int64_t n_nodes, const float* grid_m, const float* target_density,
                   const float* target_sdf, float* scratch, float* out2, void* stream);
/* g_grid_m += g_loss (w_density sign(m - t) + w_sdf sdf), sign(0) = 0;
   g_loss is a device scalar (sum_up_loss_kernel.grad, deform_loss.py:108-111). */
int dt_deform_loss_adj(int64_t n_nodes, const float* grid_m, const float* target_density,
                       const float* target_sdf, float w_density, float w_sdf, const float* g_loss,
                       float* g_grid_m, void* stream);
/* out5 = (max m, max t, sum m t, sum m, sum t), the terms of iou
   (deform_loss.py:139-153: I = sum m t / max m / max t, U = sum m / max m +
   sum t / max t, iou = I / (U - I)); scratch 5 DT_REDUCE_BLOCKS. */
int dt_grid_iou_terms(int64_t n_nodes, const float* grid_m, const float* target_density,
                      float* scratch, float* out5, void* stream);

/* ---- Kabsch rotation angle of the masked particles ----
   (compute_COM / compute_H / compute_H_svd / compute_R, multi_obj.py:309-340,
   and one term of compute_angle, box_open.py:400-406).  y = x + dt v_next;
   c0 = sum_masked x / n, c1 = sum_masked y / n (all n particles, as the
   reference divides); H = sum_masked (x - c0)(y - c1)^T = U diag(s) V^T;
   R = V U^T; angle = atan2(R_10, R_11) in state[47].  scratch: 9
   DT_REDUCE_BLOCKS floats; state: DT_KABSCH_STATE floats, also read by the
   adjoint.  The adjoint differentiates R exactly through the polar rotation
   of H (the reference's H_svd_grad, multi_obj.py:209-224, clamps 1/(s_j^2 -
   s_i^2) instead); gx and gv_next accumulate, g_angle is a device scalar. */
#define DT_KABSCH_STATE 64
int dt_kabsch_angle(int32_t n, const float* x, const float* v_next, float dt, const int32_t* mask,
                    float* scratch, float* state, void* stream);
int dt_kabsch_angle_adj(int32_t n, const float* x, const float* v_next, float dt,
                        const int32_t* mask, float* state, const float* g_angle, float* gx,
                        float* gv_next, void* stream);

/* ---- position-based rope and its particle-wise pad contact ----
   (object_model/PBD_rope.py, tasks/cable_straightening.py) */
typedef struct dt_rope_params {
  float dt, p_mass, table_height, floor_friction; /* PBD_rope.py:17-28,71      */
  float gravity[3];
  float rest_length, bending_length;              /* 34-42                      */
  float stretch_compliance, bending_compliance;   /* 49-52                      */
  float stretch_relaxation, bending_relaxation;
} dt_rope_params;
/* update (PBD_rope.py:111-180) for n particles [n][3]: ext_impulse is the
   contact impulse ext_f; particle 0 goes to fixed_pos (init_pos[0]); work:
   12 n floats of scratch. */
int dt_rope_step(const dt_rope_params* prm, int32_t n, const float* fixed_pos, const float* pos,
                 const float* vel, const float* ext_impulse, float* work, float* pos_next,
                 float* vel_next, void* stream);
/* Adjoint (update.grad): g_pos, g_vel, g_ext_impulse accumulate. */
int dt_rope_step_adj(const dt_rope_params* prm, int32_t n, const float* fixed_pos,
                     const float* pos, const float* vel, const float* ext_impulse,
                     const float* g_pos_next, const float* g_vel_next, float* work, float* g_pos,
                     float* g_vel, float* g_ext_impulse, void* stream);
/* check_collision for rope particles (cable_straightening.py:287-292): the
   closest triangle to each particle position, as dt_contact_closest. */
int dt_rope_closest(int32_t n, const float* x, int32_t n_tris, const float* centroids4,
                    int32_t* contact_idx, int32_t idx_stride, int32_t idx_col, void* stream);
/* dt_rope_closest with the boxes of dt_contact_tri_batches (written for
   these centroids) and an optional hint per particle, as
   dt_contact_closest_batched: same result, most runs skipped.  hint may be
   contact_idx itself (its previous contents) or NULL. */
int dt_rope_closest_batched(int32_t n, const float* x, int32_t n_tris, const float* centroids4,
                            const float* batch_bounds4, const int32_t* hint,
                            int32_t* contact_idx, int32_t idx_stride, int32_t idx_col,
                            void* stream);
/* collision for rope particles (294-319) with one pad: cp->sdf_offset is the
   particle radius; impulse [n][3] += f dt, ext_force (pad) += -f/3 per
   triangle vertex, contact_flag[p] = 1 where the particle touches (NULL:
   not written). */
int dt_rope_collision(const dt_contact_params* cp, float dt, int32_t n, const float* x,
                      const float* v, const int32_t* contact_idx, int32_t idx_stride,
                      int32_t idx_col, int32_t n_tris, const int32_t* tris, const float* pos,
                      const float* vel, const float* out_dir, float* impulse, float* ext_force,
                      int32_t* contact_flag, void* stream);
/* Adjoint: g_x, g_v (rope) and g_pos, g_vel (pad) accumulate. */
int dt_rope_collision_adj(const dt_contact_params* cp, float dt, int32_t n, const float* x,
                          const float* v, const int32_t* contact_idx, int32_t idx_stride,
                          int32_t idx_col, int32_t n_tris, const int32_t* tris, const float* pos,
                          const float* vel, const float* out_dir, const float* g_impulse,
                          const float* g_ext_force, float* g_x, float* g_v, float* g_pos,
                          float* g_vel, void* stream);
/* The same two kernels with the pad-vertex sums exact and order-free (the
   int64 fixed-point passes of dt_contact_collision_fx; no reference
   counterpart), so they repeat bit for bit; same acc / scale_state rules
   (n_verts*3 or *6 int64, 2 or 4 uint32, zero before and after).  The
   impulse, g_x, g_v (per particle) and ext_force, g_pos, g_vel accumulate. */
int dt_rope_collision_fx(const dt_contact_params* cp, float dt, int32_t n, const float* x,
                         const float* v, const int32_t* contact_idx, int32_t idx_stride,
                         int32_t idx_col, int32_t n_tris, const int32_t* tris, const float* pos,
                         const float* vel, const float* out_dir, float* impulse, int32_t n_verts,
                         float* ext_force, int32_t* contact_flag, int64_t* acc,
                         uint32_t* scale_state, void* stream);
int dt_rope_collision_adj_fx(const dt_contact_params* cp, float dt, int32_t n, const float* x,
                             const float* v, const int32_t* contact_idx, int32_t idx_stride,
                             int32_t idx_col, int32_t n_tris, const int32_t* tris,
                             const float* pos, const float* vel, const float* out_dir,
                             const float* g_impulse, const float* g_ext_force, float* g_x,
                             float* g_v, int32_t n_verts, float* g_pos, float* g_vel,
                             int64_t* acc, uint32_t* scale_state, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DIFFTACTILE_H */
