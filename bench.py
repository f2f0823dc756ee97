"""Benchmark: fwd+bwd particle-substeps/s of the differentiable MPM substep.

Workload (one GPU): the multi-material sphere of 1M particles on a 128^3 grid
(rigid elastic core E=1e7 inside a Drucker-Prager shell), float32.  One step
is a 1000-substep trajectory forward plus its backward, checkpointed every
50 substeps: segments whose frames and grids fit the resident budget (half
of the 288 GB of HBM) are kept from the forward pass, the others are
recomputed from their checkpoints in the backward pass.  Every step starts from the
same deformed state (the undeformed sphere advanced 200 substeps), and the
inputs are resident in HBM before the timed region.

Prints ONE JSON line (rank 0).  Also reports, for the dominant kernel, the
achieved algorithmic bandwidth against MI355X's 8 TB/s (HIP events on the
stream the kernel runs on), a CPU baseline (the oracle's C++/OpenMP
restatement of the kernels over 96 substeps of the same trajectory, all host
threads) and the relative errors of the GPU gradient against FP64 central
differences of the oracle for 16 seeded parameters of a scaled sphere.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--n-grid", type=int, default=128)
    ap.add_argument("--segment", type=int, default=50)
    ap.add_argument("--substeps", type=int, default=1000,
                    help="sphere workload: substeps of the trajectory one step runs fwd+bwd")
    ap.add_argument("--predeform", type=int, default=200,
                    help="sphere workload: substeps advanced before timing (the start state)")
    ap.add_argument("--resident-frac", type=float, default=0.5,
                    help="sphere workload: fraction of HBM for resident trajectory segments")
    ap.add_argument("--no-resort", action="store_true",
                    help="sphere workload: sort the particles once per trajectory, not at "
                         "every checkpoint")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--no-tune", action="store_true",
                    help="sphere workload: use the cached (else default) launch shapes")
    ap.add_argument("--breakdown", action="store_true",
                    help="also time each stage in isolation on the initial state")
    ap.add_argument("--cpu-substeps", type=int, default=96,
                    help="CPU baseline: substeps of the trajectory timed fwd+bwd on the host")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-grad-check", action="store_true")
    ap.add_argument("--no-population", action="store_true",
                    help="sphere workload: skip the batched-population line (C4 with pads)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain 256 MiB device-to-device copy (the reference "
                         "kernel of known traffic for rocprofv3 FETCH_SIZE / WRITE_SIZE runs)")
    ap.add_argument("--workload", default="sphere",
                    choices=["sphere", "pad_press", "population", "gripper"],
                    help="sphere: 1M-particle multi-material sphere (default, the metric's "
                         "workload); pad_press: slab pad pressing a von Mises block; "
                         "population: 64 environments, batched vs a per-environment loop; "
                         "gripper: jaw, two pads, a cable and a 2M-particle block")
    ap.add_argument("--pop-substeps", type=int, default=300,
                    help="population workload: substeps per fwd+bwd step")
    ap.add_argument("--pad-recompute", action="store_true",
                    help="pad_press: keep checkpoints only (the backward pass recomputes "
                         "every segment) instead of the resident trajectory")
    ap.add_argument("--pop-grouped", action="store_true",
                    help="population workload: one stacked batch per grid size instead of one "
                         "ragged batch")
    ap.add_argument("--pop-pads", action="store_true",
                    help="population workload: one FEM pad (10,800 tets) per environment "
                         "pressing its object")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the trajectory loops eagerly instead of replaying their "
                         "captured HIP graphs")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_stage(torch, fn, reps):
    """Average duration (ms) of `fn` (one kernel launch) over reps, HIP events on
    the current stream, which is the stream the library launches on."""
    fn()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


def kernel_breakdown(torch, sim, state, mat_id, reps):
    """Per-stage event timings of one fwd+bwd substep of the sparse path on
    resident inputs (the stages dt_mpm_run_fwd / dt_mpm_run_adj launch)."""
    from difftactile_amd import _lib
    from difftactile_amd.sim import workspace_bytes
    prm = sim.c_prm
    n = sim.n
    nn = sim.prm.n_nodes
    dev = state[0].device
    x, v, C, F = state
    g4 = torch.zeros(nn, 4, device=dev)
    gv4 = torch.zeros(nn, 4, device=dev)
    x1, v1 = torch.empty_like(x), torch.empty_like(v)
    C1, F1 = torch.empty_like(C), torch.empty_like(F)
    nbytes = workspace_bytes(prm, n)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    mats = sim.mats
    nm = len(sim.materials)
    st = _lib.stream()
    gen = torch.Generator(device=dev).manual_seed(0)
    gnext = [torch.randn(t.shape, device=dev, generator=gen) for t in state]
    gg4 = torch.zeros(nn, 4, device=dev)
    ggv4 = torch.zeros(nn, 4, device=dev)
    gout = [torch.empty_like(t) for t in state]

    def binning():
        _lib.call("dt_mpm_bin", prm, n, x, ws, nbytes, st)

    def p2g():
        _lib.call("dt_mpm_p2g_binned", prm, mats, nm, n, mat_id, x, v, C, F, F1, ws, nbytes, st)

    def grid_op():
        _lib.call("dt_mpm_grid_op4_tiles", prm, n, ws, nbytes, g4, gv4, st)

    def g2p():
        _lib.call("dt_mpm_g2p4", prm, n, x, gv4, x1, v1, C1, st)

    def g2p_adj():
        _lib.call("dt_mpm_g2p4_adj_binned", prm, n, x, gv4, gnext[0], gnext[1], gnext[2],
                  gout[0], ws, nbytes, st)

    def grid_op_adj():
        _lib.call("dt_mpm_grid_op4_adj_tiles", prm, n, ws, nbytes, g4, ggv4, gg4, st)

    def p2g_adj():
        _lib.call("dt_mpm_p2g_fused_adj", prm, mats, nm, n, mat_id, x, v, C, F, gg4,
                  gnext[3], gout[0], gout[1], gout[2], gout[3], None, st)

    # order matters for valid inputs: each stage consumes the previous one's output
    stages = [("bin", binning), ("p2g_binned", p2g), ("grid_op4_tiles", grid_op), ("g2p4", g2p),
              ("g2p4_adj_binned", g2p_adj), ("grid_op4_adj_tiles", grid_op_adj),
              ("p2g_fused_adj", p2g_adj)]
    for _, fn in stages:
        fn()
    out = {}
    for name, fn in stages:
        out[name] = time_stage(torch, fn, reps)
    active = int((g4[:, 3] > sim.prm.eps).sum().item())
    return out, active


STAGES = {"bin": 0, "p2g_binned": 1, "grid_op4_tiles": 2, "g2p4": 3, "g2p4_adj_binned": 4,
          "grid_op4_adj_tiles": 5, "p2g_fused_adj": 6}


def live_stage_times(_lib):
    """Average duration (ms) and launch count of every stage over the launches
    since dt_timing_enable(1): HIP event pairs recorded by the library around
    each launch on its stream."""
    import ctypes
    out = {}
    for name, sid in STAGES.items():
        t, c = ctypes.c_double(0.0), ctypes.c_int64(0)
        _lib.check(_lib.lib().dt_timing_read(sid, ctypes.byref(t), ctypes.byref(c)))
        if c.value:
            out[name] = (t.value / c.value, int(c.value))
    return out


def graph_stats(_lib):
    import ctypes
    h, m = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(_lib.lib().dt_graph_stats(ctypes.byref(h), ctypes.byref(m)))
    return int(h.value), int(m.value)


def active_nodes(torch, sim, st):
    """Grid nodes with mass above the occupancy threshold after the first P2G
    of the workload (binned P2G on the sorted state)."""
    from difftactile_amd import _lib
    from difftactile_amd.sim import workspace_bytes
    n, nn = sim.n, sim.prm.n_nodes
    _, perm = sim.sort_order(st[0].contiguous())
    ss = [sim._gather(perm, st[0], 3), sim._gather(perm, st[1], 3),
          sim._gather(perm, st[2].contiguous(), 9), sim._gather(perm, st[3], 9)]
    mid = sim._gather_i(perm, sim.mat_id)
    nbytes = workspace_bytes(sim.c_prm, n)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=st[0].device)
    g4 = torch.zeros(nn, 4, device=st[0].device)
    F1 = torch.empty_like(ss[3])
    s = _lib.stream()
    _lib.call("dt_mpm_bin", sim.c_prm, n, ss[0], ws, nbytes, s)
    _lib.call("dt_mpm_p2g_binned", sim.c_prm, sim.mats, len(sim.materials), n, mid, *ss, F1, ws,
              nbytes, s)
    _lib.call("dt_mpm_tiles_sum", sim.c_prm, n, ws, nbytes, g4, s)
    return int((g4[:, 3] > sim.prm.eps).sum().item())


def algorithmic_bytes(n, active, has_mat):
    """Algorithmic HBM bytes per launch of each fused stage (float32).

    Per particle: x 12, v 12, C 36, F 36 (+4 material id).  Per active node:
    packed (mv, m) 16 B or (v, pad) 16 B.
    """
    mid = 4 if has_mat else 0
    return {
        # read x (mark + key), write the permutation
        "bin": n * 16,
        # read x v C F (+id), write F_next; produce (mv, m) per active node
        "p2g_binned": n * (96 + mid + 36) + active * 16,
        # read (mv, m), write (v) per active node
        "grid_op4_tiles": active * 32,
        # read x, write x v C; gather v per active node
        "g2p4": n * (12 + 60) + active * 16,
        # read x, gx' gv' gC' (60), write gx; read the forward's grid velocity
        # (remade from the kept (mv, m)) and produce g_v per active node
        "g2p4_adj_binned": n * (12 + 60 + 12) + active * 32,
        # read (mv, m), g_v; write (g_mv, g_m) per active node
        "grid_op4_adj_tiles": active * 48,
        # read x v C F (+id), gx partial, gF'; write gx gv gC gF; gather (g_mv, g_m)
        "p2g_fused_adj": n * (96 + mid + 12 + 36 + 96) + active * 16,
    }


def cpu_baseline(prm, mats, mid, state, n_substeps):
    """The oracle's C++/OpenMP restatement (oracle/cpu_mpm.cpp: the reference's
    loop structure -- per-particle scatter with atomics, dense n^3 grid pass,
    per-particle gather -- and its adjoint in reverse order, float32) on the
    whole workload: an n_substeps trajectory fwd, then bwd, from the timed
    start state, on every host thread."""
    from oracle.cpu_mpm import CpuMPM
    x, v, C, F = state
    n = len(x)
    cpu = CpuMPM(prm, mats, mid)
    t0 = time.perf_counter()
    frames, wss = [[x, v, C, F]], []
    for _ in range(n_substeps):
        out, ws = cpu.substep(*frames[-1])
        frames.append(out)
        wss.append(ws)
    g = [np.ones_like(a) for a in frames[-1]]
    gm_tot = np.zeros((len(mats), 2), np.float32)
    for k in reversed(range(n_substeps)):
        g, gm = cpu.substep_adj(*frames[k], wss[k], g)
        gm_tot += gm
    dt_s = time.perf_counter() - t0
    model = "unknown CPU"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": n * n_substeps / dt_s, "unit": "particle-substeps/s",
            "cores": cpu.threads(), "kind": "port",
            "sample": f"CPU restatement of the reference's kernels (C++17/OpenMP float32, not "
                      f"Taichi; the reference's svd_grad with its 1e-8 clamp): fwd then bwd of "
                      f"{n_substeps} substeps of the same {n}-particle {prm.n_grid}^3 trajectory "
                      f"from the same start state, {cpu.threads()} threads on {model}; "
                      f"{dt_s:.1f} s"}


GC_SCENE = dict(n_particles=20000, n_grid=64)
GC_STEPS = 100
GC_EPS = 1e-4
GC_NV = 12          # velocity directions; then mu, lam of the core and of the shell


def _gc_problem():
    """The scaled sphere of the gradient check and its 16 seeded parameters:
    theta_k (k < 12) moves v0 along a seeded direction d_k; theta_12..15 scale
    (mu_core, lam_core, mu_shell, lam_shell) by (1 + theta).  Loss: a seeded
    linear functional of the final positions."""
    from difftactile_amd import scenes
    prm, mats, x, v, mid = scenes.config3_multi_material(**GC_SCENE)
    rng = np.random.default_rng(11)
    dirs = rng.normal(size=(GC_NV,) + x.shape)
    w = rng.normal(size=x.shape)
    return prm, mats, x, v, mid, dirs, w


def _gc_init():
    import torch
    torch.set_num_threads(1)


def _gc_fd_loss(job):
    """Oracle (FP64) loss at theta_k = sign * eps (one pool job)."""
    import torch
    from oracle import mpm as om
    k, sign = job
    prm, mats, x, v, mid, dirs, w = _gc_problem()
    op = om.MPMParams(n_grid=prm.n_grid, rho=prm.rho, dt=prm.dt, gravity=prm.gravity,
                      svd_exact=prm.svd_exact)
    oms = [om.Material(m.E, m.nu, m.kind, m.yield_stress) for m in mats]
    for o, m in zip(oms, mats):
        o.dp_alpha = m.dp_alpha
    mu, lam, kind, yld, alpha = om.material_arrays(oms, mid, torch.float64)
    vv = v.astype(np.float64)
    if k < GC_NV:
        vv = vv + sign * GC_EPS * dirs[k]
    else:
        sel = torch.as_tensor(mid == (k - GC_NV) // 2)
        scale = 1.0 + sign * GC_EPS
        if (k - GC_NV) % 2 == 0:
            mu = torch.where(sel, mu * scale, mu)
        else:
            lam = torch.where(sel, lam * scale, lam)
    n = len(x)
    st = (torch.tensor(x, dtype=torch.float64), torch.tensor(vv),
          torch.zeros(n, 3, 3, dtype=torch.float64),
          torch.eye(3, dtype=torch.float64).repeat(n, 1, 1))
    with torch.no_grad():
        out = om.rollout(op, st, (mu, lam, kind, yld, alpha), GC_STEPS)
    return float((out[0] * torch.tensor(w)).sum())


def _gc_oracle_ad(torch, dtype):
    """The oracle's reverse mode (the exact derivative of its arithmetic) of
    the check's 16 parameters, in dtype, on the host's cores."""
    from torch.utils.checkpoint import checkpoint
    from oracle import mpm as om
    prm, mats, x, v, mid, dirs, w = _gc_problem()
    op = om.MPMParams(n_grid=prm.n_grid, rho=prm.rho, dt=prm.dt, gravity=prm.gravity,
                      svd_exact=prm.svd_exact)
    oms = [om.Material(m.E, m.nu, m.kind, m.yield_stress) for m in mats]
    for o, m in zip(oms, mats):
        o.dp_alpha = m.dp_alpha
    n = len(x)
    mu_t = torch.tensor([m.mu for m in oms], dtype=dtype, requires_grad=True)
    lam_t = torch.tensor([m.lam for m in oms], dtype=dtype, requires_grad=True)
    _, _, kind, yld, alpha = om.material_arrays(oms, mid, dtype)
    idx = torch.as_tensor(mid).long()
    v0 = torch.tensor(v).to(dtype).requires_grad_(True)

    def seg(x_, v_, C_, F_, mu_, lam_):
        s = (x_, v_, C_, F_)
        for _ in range(10):
            s = om.substep(op, s, (mu_[idx], lam_[idx], kind, yld, alpha))
        return s

    s = (torch.tensor(x).to(dtype), v0, torch.zeros(n, 3, 3, dtype=dtype),
         torch.eye(3, dtype=dtype).repeat(n, 1, 1))
    for _ in range(GC_STEPS // 10):
        s = checkpoint(seg, *s, mu_t, lam_t, use_reentrant=False)
    loss = (s[0] * torch.tensor(w).to(dtype)).sum()
    gv, gmu, glam = torch.autograd.grad(loss, [v0, mu_t, lam_t])
    ad = [float((gv.double().numpy() * dirs[k]).sum()) for k in range(GC_NV)]
    for c in range(2):
        ad += [oms[c].mu * float(gmu[c]), oms[c].lam * float(glam[c])]
    return ad


def grad_check(torch, pool):
    """16 directional derivatives of a trajectory loss: the GPU adjoint
    (float32, both 50-substep segments recomputed from checkpoints) against
    FP64 central differences (eps 1e-4) of the oracle, computed on the host's
    cores by the worker pool."""
    from difftactile_amd.sim import MPMSim
    prm, mats, x, v, mid, dirs, w = _gc_problem()
    n = len(x)
    jobs = [(k, s) for k in range(GC_NV + 4) for s in (1, -1)]
    pending = pool.map_async(_gc_fd_loss, jobs)
    sim = MPMSim(prm, mats, n, mat_id=mid, checkpoint_every=50, resident_bytes=0,
                 material_grad=True)
    st = [torch.tensor(x).cuda(), torch.tensor(v).cuda().requires_grad_(True),
          torch.zeros(n, 3, 3).cuda(), torch.eye(3).repeat(n, 1, 1).cuda()]
    out = sim.rollout(*st, GC_STEPS)
    loss = (out[0] * torch.tensor(w, dtype=torch.float32).cuda()).sum()
    gv = torch.autograd.grad(loss, st[1])[0].double().cpu().numpy()
    gm = sim.material_grad.double().cpu().numpy()
    ad = [float((gv * dirs[k]).sum()) for k in range(GC_NV)]
    for c in range(2):
        ad += [mats[c].mu * gm[c, 0], mats[c].lam * gm[c, 1]]
    fl = pending.get()
    fd = [(fl[2 * k] - fl[2 * k + 1]) / (2 * GC_EPS) for k in range(GC_NV + 4)]
    threads = torch.get_num_threads()
    torch.set_num_threads(min(16, os.cpu_count() or 1))
    rm64 = _gc_oracle_ad(torch, torch.float64)
    rm32 = _gc_oracle_ad(torch, torch.float32)
    torch.set_num_threads(threads)
    rel = lambda a, b: [abs(p - q) / max(abs(q), 1e-30) for p, q in zip(a, b)]
    r_fd, r_rm, fd_rm, f32 = rel(ad, fd), rel(ad, rm64), rel(fd, rm64), rel(rm32, rm64)
    return {"rel_err_max": max(r_fd), "rel_err_median": float(np.median(r_fd)),
            "rel_err": r_fd,
            "rel_err_vs_fp64_reverse_mode": r_rm,
            "rel_err_vs_fp64_reverse_mode_median": float(np.median(r_rm)),
            "fd_vs_fp64_reverse_mode": fd_rm,
            "fp32_oracle_vs_fp64_reverse_mode": f32,
            "gpu": ad, "fd_fp64": fd, "reverse_mode_fp64": rm64,
            "note": "rel_err: GPU vs FP64 central differences; fd_vs_fp64_reverse_mode: the "
                    "oracle's own FD vs its exact FP64 derivative (non-zero where Drucker-"
                    "Prager return-map kinks lie within eps of the trajectory); "
                    "fp32_oracle_vs_fp64_reverse_mode: the oracle's reverse mode in float32 "
                    "(the float32 floor of each parameter)",
            "case": f"{n} particles, {prm.n_grid}^3, {GC_STEPS} substeps (two 50-substep "
                    "segments, both recomputed), core + Drucker-Prager shell; 16 seeded "
                    "parameters: 12 directions of v0, mu and lam of each material; loss a "
                    "seeded linear functional of the final positions; FP64 central "
                    "differences of the oracle, eps 1e-4"}


def bench_pad_press(args, torch):
    """Pad pressing a block: coupled fwd+bwd over the full press (500 substeps,
    segments of 50); loss = summed squared marker displacement of the pad at
    the end; gradient w.r.t. the pad velocity sequence.  The segments the
    forward pass keeps within the object's resident budget (a quarter of HBM)
    are not recomputed by the backward pass; --pad-recompute keeps checkpoints
    only."""
    import numpy as np
    from difftactile_amd import scenes
    from difftactile_amd.coupled import CoupledSim, Pad
    from difftactile_amd.losses import MarkerModel, marker_displacement
    from difftactile_amd.sim import MPMSim
    c = scenes.config2_pad_press()
    prm, n = c["prm"], len(c["x"])
    T = c["n_substeps"]
    nodes, f2v, tris, layer, od = c["mesh"]
    sim = MPMSim(prm, c["mats"], n, checkpoint_every=50, sort=True,
                 resident_bytes=0 if args.pad_recompute else None)
    pad = Pad(nodes.astype(np.float32), f2v, tris, layer, od, c["pad_mu"], c["pad_lam"],
              c["pad_mass"], prm.dt)
    # segment graphs: this scene's substeps are short enough that launching
    # them one by one shows (148 -> 115 ms per step); on the gripper and the
    # population the kernels are long and eager launches are 3-5 % faster
    cs = CoupledSim(sim, [pad], **c["contact"], graphs=True)
    surf_ids = np.where(layer == layer.max())[0]
    mm = MarkerModel(nodes, surf_ids, c["inv_T"])
    dev = torch.device("cuda")
    x0 = torch.tensor(c["x"], device=dev)
    v0 = torch.tensor(c["v"], device=dev)
    C0 = torch.zeros(n, 3, 3, device=dev)
    F0 = torch.eye(3, device=dev).repeat(n, 1, 1)
    pos0 = torch.tensor(nodes, dtype=torch.float32, device=dev)
    cv = torch.tensor(c["control_vel"], device=dev)

    def step():
        cvl = cv.clone().requires_grad_(True)
        pst = (pos0.clone(), cv[0].clone(), pos0.clone())
        out = cs.rollout(x0, v0, C0, F0, [pst], [cvl], T)
        pos, vel, vpos = out[4][0]
        loss = marker_displacement(mm, pos, vpos)
        g = torch.autograd.grad(loss, cvl)[0]
        # contact evidence: how far the press pushed the block down
        dy = float((out[0][:, 1] - x0[:, 1]).min().detach())
        return loss, g, dy

    log(f"[bench] pad_press: {n} particles, {len(f2v)} tets, {T} substeps")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, g, dy = step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gv = g.sum(1)          # d loss / d (pad velocity of frame f), summed over vertices
    res = {"metric": "fwd+bwd particle-substeps/s (pad press)",
           "value": n * T * args.steps / wall, "unit": "particle-substeps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
           "substeps_per_s": T * args.steps / wall, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"slab pad ({len(f2v)} tets) pressing a von Mises block "
                                  f"({n} particles, {prm.n_grid}^3), {T} substeps fwd+bwd, "
                                  "loss = summed marker displacement"},
           "loss": float(loss.detach()), "grad_finite": bool(torch.isfinite(g).all()),
           "grad_dL_dvy_first_last": [float(gv[0, 1]), float(gv[-1, 1])],
           "block_min_dy": dy,
           "resident_segments": f"{cs.last_resident[0]} of {cs.last_resident[1]}",
           "resident_gb": round(cs.last_resident[2] / 1e9, 1),
           "segment_graphs": {"captures": cs.graph_misses, "replays": cs.graph_hits}}
    print(json.dumps(res), flush=True)


def bench_population(args, torch):
    print(json.dumps(population(args, torch)), flush=True)


def population(args, torch):
    """Batched population: 64 environments (5k-200k particles, 64^3 or 128^3
    grids) advanced fwd+bwd for --pop-substeps substeps, once batched (one
    ragged EnvBatch over both grid sizes, or with --pop-grouped one per grid
    size) and once as a loop over per-environment simulations.  Both recompute 10-substep segments from checkpoints (no
    resident segments: the stacked 128^3 grids of 300 substeps exceed HBM).
    Loss per environment: its final centre of mass along a seeded direction;
    the gradient is taken with respect to its initial velocities."""
    from difftactile_amd import scenes
    from difftactile_amd.batch import CoupledEnvBatch, EnvBatch
    from difftactile_amd.coupled import CoupledSim, Pad
    from difftactile_amd.sim import MPMSim
    pads = args.pop_pads
    envs = scenes.config4_population(pads=pads)
    T, seg = args.pop_substeps, 10
    contact = dict(kn=1e4, kd=269.44, kt=154.78, mu_f=53.85)
    dev = torch.device("cuda")
    n_tot = sum(len(d["x"]) for d in envs)
    rng = np.random.default_rng(5)
    w = [torch.tensor(rng.normal(size=3), dtype=torch.float32, device=dev) for _ in envs]
    xs = [torch.tensor(d["x"], device=dev) for d in envs]
    vs = [torch.tensor(d["v"], device=dev) for d in envs]

    def state(e):
        n = len(envs[e]["x"])
        return [xs[e], vs[e].clone().requires_grad_(True), torch.zeros(n, 3, 3, device=dev),
                torch.eye(3, device=dev).repeat(n, 1, 1)]

    def pad_state(e):
        nodes = torch.tensor(envs[e]["pad_mesh"][0], device=dev)
        cv = torch.tensor(envs[e]["pad_cv"], device=dev)
        return (nodes.clone(), cv.clone(), nodes.clone()), cv.expand(T + 1, -1, -1).contiguous()

    groups = {}
    for e, d in enumerate(envs):
        groups.setdefault(d["prm"].n_grid, []).append(e)
    batches, sims = [], []
    # ragged (default): all environments in one launch sequence, each keeping
    # its own grid inside a slot of the largest; grouped: one per grid size
    layout = "grouped" if args.pop_grouped else "ragged"
    batch_groups = list(groups.values()) if args.pop_grouped else [list(range(len(envs)))]
    for ids in batch_groups:
        spec = [(envs[e]["mats"], len(envs[e]["x"]), envs[e]["mat_id"], envs[e]["prm"])
                for e in ids]
        kw = dict(checkpoint_every=seg, sort=True, resident_bytes=0)
        if pads:
            d0 = envs[ids[0]]
            b = CoupledEnvBatch(d0["prm"], spec, [envs[e]["pad_mesh"] for e in ids],
                                dict(out_dir=d0["pad_out"], **d0["pad"]), contact, **kw)
        else:
            b = EnvBatch(envs[ids[0]]["prm"], spec, **kw)
        batches.append((b, ids))
    for d in envs:
        sim = MPMSim(d["prm"], d["mats"], len(d["x"]), mat_id=d["mat_id"], checkpoint_every=seg,
                     sort=True, resident_bytes=0)
        if pads:
            nodes, f2v, tris, layer = d["pad_mesh"]
            pad = Pad(nodes, f2v, tris, layer, d["pad_out"], d["pad"]["mu"], d["pad"]["lam"],
                      d["pad"]["mass"], d["prm"].dt)
            sim = CoupledSim(sim, [pad], **contact)
        sims.append(sim)

    def run_batched():
        g = [None] * len(envs)
        xs_out = [None] * len(envs)
        for batch, ids in batches:
            sts = [state(e) for e in ids]
            if pads:
                ps = [pad_state(e) for e in ids]
                res = batch.rollout(sts, [p[0] for p in ps], [p[1] for p in ps], T)
                out = [r[0] for r in res]
            else:
                out = batch.rollout(sts, T)
            loss = sum((o[0].mean(0) * w[e]).sum() for o, e in zip(out, ids))
            for e, o in zip(ids, out):
                xs_out[e] = o[0].detach()
            for e, ge in zip(ids, torch.autograd.grad(loss, [st[1] for st in sts])):
                g[e] = ge
        return g, xs_out

    def run_loop():
        g, xs_out = [], []
        for e, sim in enumerate(sims):
            st = state(e)
            if pads:
                p0, cv = pad_state(e)
                out = sim.rollout(*st, [p0], [cv], T)
            else:
                out = sim.rollout(*st, T)
            xs_out.append(out[0].detach())
            g.append(torch.autograd.grad((out[0].mean(0) * w[e]).sum(), st[1])[0])
        return g, xs_out

    def timed(fn, reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, out

    log(f"[bench] population: {len(envs)} environments, {n_tot} particles, "
        f"{', '.join(f'{len(ids)} on {n}^3' for n, ids in groups.items())}, {T} substeps")
    g_b0, _ = run_batched()
    g_l0, _ = run_loop()
    reps = max(1, min(args.steps, 2))
    t_b, (g_b, x_b) = timed(run_batched, reps)
    log(f"[bench] batched {t_b * 1e3:.1f} ms per fwd+bwd step")
    t_l, (g_l, x_l) = timed(run_loop, reps)
    log(f"[bench] per-environment loop {t_l * 1e3:.1f} ms per fwd+bwd step")
    num = sum(float(((a - b) ** 2).sum()) for a, b in zip(g_b, g_l))
    den = sum(float((b ** 2).sum()) for b in g_l)
    rel = lambda a, b: float((a - b).norm() / max(float(b.norm()), 1e-30))
    x_rel = [rel(a, b) for a, b in zip(x_b, x_l)]
    g_rel = sorted(rel(a, b) for a, b in zip(g_b, g_l))
    # the same loop twice: float atomics order sums differently from run to run
    g_self = sorted(rel(a, b) for a, b in zip(g_l0, g_l))
    log(f"[bench] batched vs loop: final x rel diff max {max(x_rel):.2e}; per-environment "
        f"gradient rel diff median {g_rel[len(g_rel) // 2]:.2e}, max {g_rel[-1]:.2e}; "
        f"gradient norms {min(float(t.norm()) for t in g_l):.3e}..{max(float(t.norm()) for t in g_l):.3e}; "
        f"loop vs itself: median {g_self[len(g_self) // 2]:.2e}, max {g_self[-1]:.2e}")
    res = {"metric": "fwd+bwd env-substeps/s (64-environment population"
                     + (", one FEM pad each)" if pads else ")"),
           "value": len(envs) * T / t_b, "unit": "env-substeps/s", "n_gpus": 1, "steps": reps,
           "warmup": 1, "ms_per_step": t_b * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "particle_substeps_per_s": n_tot * T / t_b,
           "loop": {"ms_per_step": t_l * 1e3, "value": len(envs) * T / t_l},
           "speedup_vs_loop": t_l / t_b,
           "grad_rel_diff_batched_vs_loop": (num / max(den, 1e-30)) ** 0.5,
           "state_rel_diff_batched_vs_loop_max": max(x_rel),
           "grad_rel_diff_per_env_median": g_rel[len(g_rel) // 2],
           "grad_rel_diff_loop_vs_itself_median": g_self[len(g_self) // 2],
           "config": {"workload": f"{len(envs)} environments, {n_tot} particles "
                                  "(log-uniform 5e3-2e5 per environment), 64^3 below 3e4 "
                                  f"particles else 128^3, {T} substeps fwd+bwd, 10-substep "
                                  "segments recomputed; batched = "
                                  + ("one launch sequence over all environments (ragged grids)"
                                     if layout == "ragged" else
                                     "one stacked launch sequence per grid size")
                                  + ", loop = one simulation per environment"
                                  + ("; each environment with a 10,800-tet FEM pad pressing "
                                     "its object (penalty contact)" if pads else ""),
                      "groups": {f"{n}^3": len(ids) for n, ids in groups.items()},
                      "layout": layout}}
    return res


def bench_gripper(args, torch):
    """Full gripper scene: the jaw's controls drive two 45,630-tet pads that
    grip a 2M-particle elastoplastic block on a 256^3 grid, with a
    2,001-particle cable touching the left pad; fwd+bwd over a 1000-substep
    trajectory checkpointed every 50 substeps (segments beyond the resident
    budget recomputed); loss = squared distance of the block's centre of mass
    from a target plus the reference's marker loss (squared mean marker offset
    of each pad); gradient w.r.t. the jaw controls (through the kinematic
    chain)."""
    import numpy as np
    from difftactile_amd import scenes
    from difftactile_amd.coupled import CoupledSim, Pad
    from difftactile_amd.losses import MarkerModel, com_loss, marker_motion_loss
    from difftactile_amd.object_model.pbd_rope import PBDRope
    from difftactile_amd.sensor_model.gripper_kinematics import Gripper
    from difftactile_amd.sim import MPMSim
    c = scenes.config5_gripper()
    prm, n = c["prm"], len(c["x"])
    T = c["n_substeps"]
    dev = torch.device("cuda")
    total_mem = torch.cuda.get_device_properties(dev).total_memory
    sim = MPMSim(prm, c["mats"], n, mat_id=c["mat_id"], checkpoint_every=50, sort=True,
                 resident_bytes=int(args.resident_frac * total_mem))
    pads = [Pad(nodes, f2v, tris, layer, od, c["pad_mu"], c["pad_lam"], c["pad_mass"], prm.dt)
            for nodes, f2v, tris, layer, od in c["pads"]]
    r = c["rope"]
    rope = PBDRope(dt=prm.dt, p_rad=r["p_rad"], p_rho=r["p_rho"], n_particles=r["n"],
                   rest_length=r["rest_length"])
    rope.stretch_compliance = rope.bending_compliance = r["compliance"]
    rp0, rv0 = rope.init(r["start"], r["ori"])
    cs = CoupledSim(sim, pads, **c["contact"], rope=rope, rope_contact=c["rope_contact"])
    g = c["gripper"]
    grip = Gripper(dt=g["dt"])
    grip.init_params(g["rot"], g["trans"], g["grip_width"])
    init_x, mms = [], []
    for k, (nodes, f2v, tris, layer, od) in enumerate(c["pads"]):
        Pk = grip.finger_pose(k + 1).astype(np.float64)
        nh = np.concatenate([nodes.astype(np.float64), np.ones((len(nodes), 1))], 1)
        init_x.append(torch.tensor((nh @ np.linalg.inv(Pk).T)[:, :3].astype(np.float32),
                                   device=dev))
        mms.append(MarkerModel(nodes, np.where(layer == layer.max())[0], c["inv_T"][k]))
    x0 = torch.tensor(c["x"], device=dev)
    v0 = torch.tensor(c["v"], device=dev)
    C0 = torch.zeros(n, 3, 3, device=dev)
    F0 = torch.eye(3, device=dev).repeat(n, 1, 1)
    pos0 = [torch.tensor(p[0], device=dev) for p in c["pads"]]
    ctrl0 = torch.tensor(c["controls"], device=dev)
    tgt = torch.tensor(c["target_com"], dtype=torch.float32, device=dev)

    def step():
        ctrl = ctrl0.clone().requires_grad_(True)
        cv1, cv2 = grip.control_velocities(ctrl, *init_x)
        pst = [(p.clone(), cvk[0].detach().clone(), p.clone()) for p, cvk in zip(pos0, (cv1, cv2))]
        out = cs.rollout(x0, v0, C0, F0, pst, [cv1, cv2], T, rope_state=(rp0.clone(), rv0.clone()))
        loss = ((com_loss(out[0]) - tgt) ** 2).sum() + \
            marker_motion_loss(mms, [(pos, vpos) for pos, _, vpos in out[4]])
        gc = torch.autograd.grad(loss, ctrl)[0]
        return loss, gc, out

    log(f"[bench] gripper: {n} particles, {prm.n_grid}^3, 2 x {pads[0].n_tets} tets, "
        f"{rope.n_particles} cable particles, {T} substeps")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, gc, out = step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gap = [float(out[4][0][0][:, 0].detach().max() - float(c["pads"][0][0][:, 0].max())),
           float(out[4][1][0][:, 0].detach().min() - float(c["pads"][1][0][:, 0].min()))]
    # the cable's velocity is (p' - p) / dt of float32 positions: a substep
    # moves a particle only when it displaces it by half a position ulp
    cable_move = float((out[5][0].detach() - rp0).norm(dim=1).max())
    res = {"metric": "fwd+bwd particle-substeps/s (gripper scene)",
           "value": n * T * args.steps / wall, "unit": "particle-substeps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
           "substeps_per_s": T * args.steps / wall, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"parallel jaw with two pads ({pads[0].n_tets} tets each), "
                                  f"a {rope.n_particles}-particle cable and an elastoplastic "
                                  f"block ({n} particles, {prm.n_grid}^3), {T} substeps fwd+bwd "
                                  "checkpointed every 50, loss = centre-of-mass error + mean "
                                  "marker offset, gradient w.r.t. the jaw controls",
                      "resident_segments": f"{cs.last_resident[0]} of {cs.last_resident[1]}",
                      "resident_gb": round(cs.last_resident[2] / 1e9, 1)},
           "loss": float(loss.detach()), "grad_finite": bool(torch.isfinite(gc).all()),
           "grad_abs_max": float(gc.abs().max()), "pad_face_travel": gap,
           "cable_max_displacement": cable_move,
           "segment_graphs": {"captures": cs.graph_misses, "replays": cs.graph_hits}}
    print(json.dumps(res), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU runs are not "
              "part of this build", file=sys.stderr)
        sys.exit(2)
    pool = None
    if args.workload == "sphere" and not args.no_grad_check:
        # host workers for the FP64 finite differences, forked before anything
        # touches the GPU (they run the CPU oracle only)
        import multiprocessing as mp
        pool = mp.get_context("fork").Pool(min(16, os.cpu_count() or 1), initializer=_gc_init)
    import torch
    if args.no_graphs:
        from difftactile_amd import _lib
        _lib.call("dt_graphs_enable", 0)
    if args.workload == "pad_press":
        bench_pad_press(args, torch)
        return
    if args.workload == "population":
        bench_population(args, torch)
        return
    if args.workload == "gripper":
        bench_gripper(args, torch)
        return
    from difftactile_amd import scenes
    from difftactile_amd.sim import MPMSim, _segments

    dev = torch.device("cuda:0")
    prm, mats, x, v, mid = scenes.config3_multi_material(n_particles=args.particles,
                                                         n_grid=args.n_grid)
    n = len(x)
    total_mem = torch.cuda.get_device_properties(dev).total_memory
    sim = MPMSim(prm, mats, n, mat_id=mid, device=dev, checkpoint_every=args.segment, sort=True,
                 resident_bytes=int(args.resident_frac * total_mem), resort=not args.no_resort)
    st = [torch.tensor(x, device=dev), torch.tensor(v, device=dev),
          torch.zeros(n, 3, 3, device=dev), torch.tensor(scenes.identity_F(n), device=dev)]
    log(f"[bench] {n} particles, {prm.n_grid}^3: advancing {args.predeform} substeps to the "
        "start state")
    st = list(sim.step_no_grad(*st, args.predeform))
    # launch shapes: tuned on this workload (4-substep fwd+bwd rollouts from the start state)
    from difftactile_amd import autotune
    if args.no_tune:
        launch, tune_ms = autotune.apply_cached(prm, n) or autotune.get_config(), None
    else:
        log("[bench] tuning launch shapes")
        launch, tune_ms = autotune.tune(sim, st, n_substeps=4, reps=3)
    w = torch.randn(n, 3, device=dev, generator=torch.Generator(device=dev).manual_seed(1))
    T = args.substeps
    # which segments the forward pass keeps (the rule of sim._Rollout.forward)
    used, kept, segs = 0, 0, _segments(T, args.segment)
    for _, k in segs:
        if used + sim.segment_bytes(k) <= sim.resident_bytes:
            used += sim.segment_bytes(k)
            kept += 1
    recomputed_substeps = sum(k for _, k in segs[kept:])

    def step():
        leaves = [t.detach().requires_grad_(True) for t in st]
        out = sim.rollout(*leaves, T)
        loss = (out[0] * w).sum() + out[1].sum()
        return torch.autograd.grad(loss, leaves[1])[0]

    log(f"[bench] trajectory {T} substeps, segments of {args.segment}: {kept} of {len(segs)} "
        f"resident ({used / 1e9:.1f} GB), {recomputed_substeps} substeps recomputed; "
        f"warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} steps")
    from difftactile_amd import _lib
    lib = _lib.lib()
    h0, m0 = graph_stats(_lib)
    t0 = time.perf_counter()
    for i in range(args.steps):
        g = step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    h1, m1 = graph_stats(_lib)
    # stage durations: one more step of the same workload with every stage launch
    # bracketed by HIP events on its stream (kept out of the timed region: the
    # event records add gaps between launches, not time to the kernels)
    _lib.check(lib.dt_timing_enable(1))
    step()
    torch.cuda.synchronize()
    times = live_stage_times(_lib)
    _lib.check(lib.dt_timing_enable(0))
    finite = bool(torch.isfinite(g).all().item())
    ms = wall * 1e3 / args.steps
    value = n * T * args.steps / wall
    log(f"[bench] {ms:.2f} ms/step, {value:.3e} particle-substeps/s, grad finite={finite}")

    active = active_nodes(torch, sim, st)
    ab = algorithmic_bytes(n, active, True)
    kern = {k: {"ms": round(t, 5), "launches": c, "GBps": round(ab[k] / (t * 1e-3) / 1e9, 1)}
            for k, (t, c) in times.items()}
    dom = max(times, key=lambda k: times[k][0])
    achieved = ab[dom] / (times[dom][0] * 1e-3) / 1e9
    log(f"[bench] stages (average over the timed launches): {json.dumps(kern)}")
    isolated = None
    if args.breakdown:
        _, perm = sim.sort_order(st[0].contiguous())
        sorted_state = [sim._gather(perm, st[0], 3), sim._gather(perm, st[1], 3),
                        sim._gather(perm, st[2].contiguous(), 9), sim._gather(perm, st[3], 9)]
        mat_sorted = sim._gather_i(perm, sim.mat_id)
        iso, _ = kernel_breakdown(torch, sim, sorted_state, mat_sorted, args.kernel_reps)
        isolated = {k: round(v, 5) for k, v in iso.items()}
        log(f"[bench] isolated stages on the start state (ms): {json.dumps(isolated)}")
    if args.calib_copy:
        src = torch.ones(64 * 2 ** 20, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(dom)
    # whole-substep algorithmic bytes: 480 B/particle + 128 B/active node (+8 B/particle ids);
    # recompute traffic is not counted
    b_sub = 480 * n + 128 * active + 8 * n
    result = {
        "metric": "fwd+bwd particle-substeps/s at 1M particles",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"multi-material sphere (elastic core E=1e7 + Drucker-Prager "
                               f"shell), {n} particles, {prm.n_grid}^3 grid; one step = a "
                               f"{T}-substep trajectory fwd+bwd from a state deformed by "
                               f"{args.predeform} substeps, checkpointed every {args.segment} "
                               f"substeps; {kept} of {len(segs)} segments kept in HBM "
                               f"({used / 1e9:.1f} GB, {args.resident_frac:g} of HBM), the "
                               f"others recomputed from their checkpoints in the backward pass",
                   "particles": n, "n_grid": prm.n_grid, "substeps_per_step": T,
                   "checkpoint_every": args.segment, "segments": len(segs),
                   "segments_resident": kept, "substeps_recomputed": recomputed_substeps,
                   "recompute_share": recomputed_substeps / T,
                   "predeform_substeps": args.predeform, "active_nodes": active,
                   "launch_shapes": launch, "tuned_4substep_ms": tune_ms},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "substep_roofline_frac": b_sub * (value / n) / (HBM_PEAK_GBS * 1e9),
        "kernels": kern,
        "graph_replays": {"hits": h1 - h0, "captures": m1 - m0},
        "grad_finite": finite,
    }
    if isolated is not None:
        result["kernels_isolated_ms"] = isolated
    if pool is not None:
        log("[bench] gradient check: 16 parameters vs FP64 central differences")
        result["grad_check"] = grad_check(torch, pool)
        pool.close()
        pool.join()
    if not args.no_population:
        # the batched population (C4 with pads) beside the headline: the
        # sphere's buffers go first
        log("[bench] population (64 environments with pads, batched vs loop)")
        del step
        sim.release_buffers()
        del sim
        torch.cuda.empty_cache()
        pargs = argparse.Namespace(**vars(args))
        pargs.pop_pads, pargs.steps = True, 1
        pop = population(pargs, torch)
        result["population"] = {k: pop[k] for k in (
            "metric", "value", "unit", "ms_per_step", "loop", "speedup_vs_loop",
            "state_rel_diff_batched_vs_loop_max", "grad_rel_diff_per_env_median")}
        result["population"]["workload"] = pop["config"]["workload"]
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline")
        host = [t.detach().cpu().numpy() for t in st]
        result["cpu_baseline"] = cpu_baseline(prm, mats, mid, host, args.cpu_substeps)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
