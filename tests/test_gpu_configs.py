"""The benchmark configurations on the GPU.

* C2 (slab pad pressing a von Mises block) scaled so the FP64 oracle covers
  it: state, marker loss and d loss / d (pad velocity sequence) against the
  oracle's reverse mode over the same composition, with the adjoint clamp on;
  and at full size (200k particles, 128^3, 40 x 40 x 6 pad) the properties
  that hold at any size: Newton's third law per substep (sum grid_f / dt +
  sum ext = 0), mass conservation of the scatter, a finite non-zero gradient.
* C4 (64 environments of uneven size, one FEM pad each): the batched form
  against the per-environment loop for every environment, and the smallest
  environment against the oracle.
"""
import numpy as np
import pytest
import torch

from difftactile_amd import _lib, scenes
from difftactile_amd.coupled import CoupledSim, Pad
from difftactile_amd.losses import MarkerModel, marker_displacement
from difftactile_amd.sim import MPMSim, workspace_bytes
from oracle import coupled as ocp
from oracle import fem as of
from oracle import markers as ok
from oracle import mpm as om
from tests.mpm_cases import oracle_mats, oracle_params, rel_l2

pytestmark = pytest.mark.gpu
F64 = torch.float64
P = _lib.ptr


def _gpu_c2(c, checkpoint_every=4, resident_bytes=None):
    prm, n = c["prm"], len(c["x"])
    nodes, f2v, tris, layer, od = c["mesh"]
    sim = MPMSim(prm, c["mats"], n, checkpoint_every=checkpoint_every, sort=True,
                 resident_bytes=resident_bytes)
    pad = Pad(nodes.astype(np.float32), f2v, tris, layer, od, c["pad_mu"], c["pad_lam"],
              c["pad_mass"], prm.dt)
    cs = CoupledSim(sim, [pad], **c["contact"])
    mm = MarkerModel(nodes, np.where(layer == layer.max())[0], c["inv_T"])
    return sim, pad, cs, mm


def _oracle_pad(pos, vel, vpos, cv, cv1, f2v, B, layer, tris, od, c, dt):
    return dict(pos=pos, vel=vel, vpos=vpos, cv=cv, cv_next=cv1, f2v=torch.tensor(f2v).long(),
                B=B, layer=torch.tensor(layer), tris=tris, out_dir=torch.tensor(od, dtype=F64),
                mu=c["pad_mu"], lam=c["pad_lam"], mass=c["pad_mass"], dt=dt)


def test_c2_scaled_pad_press_matches_the_oracle(gpu):
    """C2 with its materials and constants (von Mises yield 50, kn = 1e4, the
    grasp task's kd / kt / friction, clamp on, the press rate of 2 mm per 500
    substeps, the marker loss) on 6,000 particles, a 32^3 grid and a 12 x 12 x 3
    pad, 6 substeps.

    Gradients w.r.t. the pad velocity sequence and the initial positions are
    compared on a random linear functional of the final pad and object
    positions (1e-3).  The marker loss after 6 substeps is a few float32 ulps
    of the marker pixel coordinates (~440 px, markers move ~1e-4 px), so any
    float32 evaluation of it, the reference's included, is rounding-bound:
    it is held to three times the float32 oracle's distance from float64."""
    steps = 6
    c = scenes.config2_pad_press(n_particles=6000, n_grid=32, n_substeps=steps,
                                 pad_verts=(12, 12, 3), press_substeps=500)
    prm, n = c["prm"], len(c["x"])
    nodes, f2v, tris, layer, od = c["mesh"]
    sim, pad, cs, mm = _gpu_c2(c)
    dev = torch.device("cuda")
    rng = np.random.default_rng(17)
    wx = rng.normal(size=(n, 3))
    wp = rng.normal(size=(len(nodes), 3))
    x0 = torch.tensor(c["x"], device=dev).requires_grad_(True)
    v0 = torch.tensor(c["v"], device=dev)
    C0 = torch.zeros(n, 3, 3, device=dev)
    F0 = torch.eye(3, device=dev).repeat(n, 1, 1)
    pos0 = torch.tensor(nodes, dtype=torch.float32, device=dev)
    cv = torch.tensor(c["control_vel"], device=dev).requires_grad_(True)
    out = cs.rollout(x0, v0, C0, F0, [(pos0.clone(), cv[0].detach().clone(), pos0.clone())],
                     [cv], steps)
    pos, vel, vpos = out[4][0]
    loss_m = marker_displacement(mm, pos, vpos)
    loss_l = (out[0] * torch.tensor(wx, dtype=torch.float32, device=dev)).sum() + \
        (pos * torch.tensor(wp, dtype=torch.float32, device=dev)).sum()
    g_cv, g_x = torch.autograd.grad(loss_l, [cv, x0])
    sid = torch.as_tensor(np.where(layer == layer.max())[0]).long()

    def oracle(dtype):
        op = oracle_params(prm)
        mats = oracle_mats(c["mats"], np.zeros(n, int), dtype)
        so = [torch.tensor(c["x"]).to(dtype).requires_grad_(True), torch.tensor(c["v"]).to(dtype),
              torch.zeros(n, 3, 3, dtype=dtype), torch.eye(3, dtype=dtype).repeat(n, 1, 1)]
        cvo = torch.tensor(c["control_vel"]).to(dtype).requires_grad_(True)
        B = pad.B.detach().cpu().to(dtype)
        p32 = torch.tensor(nodes.astype(np.float32)).to(dtype)
        state, ps = tuple(so), (p32, cvo[0].detach().clone(), p32.clone())
        contacts = 0
        for f in range(steps):
            pd = _oracle_pad(*ps, cvo[f], cvo[f + 1], f2v, B, layer, tris, od, c, prm.dt)
            pd["out_dir"] = pd["out_dir"].to(dtype)
            info = {}
            state, po, _, _ = ocp.coupled_substep(op, state, mats, [pd],
                                                  tuple(c["contact"].values()), clamp=True,
                                                  info=info)
            contacts = max(contacts, int((info["grid_f"].abs().sum(1) > 0).sum()))
            ps = po[0]
        pr, vr = ok.extract_markers(ps[0], ps[2], sid, torch.tensor(c["inv_T"]).to(dtype),
                                    mm.idx.long().cpu(), mm.w.cpu().to(dtype))
        return state, ps, so, cvo, ((pr - vr) ** 2).sum(), contacts

    state, ps, so, cvo, lm64, contacts = oracle(F64)
    assert contacts > 50, "the press must make contact"
    for got, want in zip(out[:4], state):
        assert rel_l2(got.detach().cpu(), want.detach()) < 1e-4
    for got, want in zip(out[4][0], ps):
        assert rel_l2(got.detach().cpu(), want.detach()) < 1e-4
    lo = (state[0] * torch.tensor(wx)).sum() + (ps[0] * torch.tensor(wp)).sum()
    go_cv, go_x = torch.autograd.grad(lo, [cvo, so[0]])
    assert float(go_cv.abs().max()) > 0
    assert rel_l2(g_cv.cpu(), go_cv) < 1e-3
    assert rel_l2(g_x.cpu(), go_x) < 1e-3
    lm32 = oracle(torch.float32)[4]
    tol = 3 * abs(float(lm32.detach()) - float(lm64.detach()))
    lm = float(lm64.detach())
    assert abs(float(loss_m.detach()) - lm) <= max(tol, 1e-4 * lm)


def _sorted_state(sim, x, v):
    n = sim.n
    dev = sim.device
    x, v = torch.tensor(x, device=dev), torch.tensor(v, device=dev)
    C = torch.zeros(n, 3, 3, device=dev)
    F = torch.eye(3, device=dev).repeat(n, 1, 1)
    perm = sim.sort_order(x)[1]
    st = [sim._gather(perm, t, w) for t, w in zip((x, v, C, F), (3, 3, 9, 9))]
    sim._set_env(None)
    return st


def test_c2_full_size_contact_balance_mass_and_gradient(gpu):
    """Full C2: every substep's contact impulses balance the pad's forces,
    the binned scatter conserves mass, and a 10-substep marker-loss gradient
    w.r.t. the pad velocities is finite and non-zero."""
    c = scenes.config2_pad_press()
    prm, n = c["prm"], len(c["x"])
    nodes, f2v, tris, layer, od = c["mesh"]
    sim, pad, cs, mm = _gpu_c2(c, checkpoint_every=5)
    dev = sim.device
    T = 8
    st = _sorted_state(sim, c["x"], c["v"])
    pos0 = torch.tensor(nodes, dtype=torch.float32, device=dev)
    cv = torch.tensor(c["control_vel"][:T + 1], device=dev)
    pst = [[pos0.clone(), cv[0].clone(), pos0.clone()]]
    frames, pframes, stores = cs._run_keep(st, pst, [cv], 0, T, None)
    torch.cuda.synchronize()
    nn = prm.n_nodes
    ws = torch.empty(workspace_bytes(sim.c_prm, n), dtype=torch.uint8, device=dev)
    touched = 0
    for j, store in enumerate(stores):
        f = store["gf"]
        # the kept grid is written on the active bricks only: take its nodes
        g4 = torch.zeros(nn, 4, device=dev)
        fr = frames[j]
        F1 = torch.empty_like(fr[3])
        s = _lib.stream()
        _lib.call("dt_mpm_bin", sim.c_prm, n, P(fr[0]), P(ws, torch.uint8), ws.numel(), s)
        _lib.call("dt_mpm_p2g_binned", sim.c_prm, sim.mats, 1, n, None, *[P(t) for t in fr],
                  P(F1), P(ws, torch.uint8), ws.numel(), s)
        _lib.call("dt_mpm_tiles_sum", sim.c_prm, n, P(ws, torch.uint8), ws.numel(), P(g4), s)
        m_tot = float(g4[:, 3].double().sum())
        assert abs(m_tot - n * prm.p_mass) <= 1e-5 * n * prm.p_mass, j
        # contact nodes: the occupied nodes of the stored grid with an impulse
        occ = g4[:, 3] > prm.eps
        fo = f[occ].double()
        ext = store["ext"][0].double()
        touched = max(touched, int((fo.abs().sum(1) > 0).sum()))
        bal = fo.sum(0) / prm.dt + ext.sum(0)
        mag = fo.abs().sum() / prm.dt + ext.abs().sum()
        assert float(bal.abs().max()) <= 1e-5 * float(mag) + 1e-12, (j, bal, mag)
    assert touched > 1000, "the press must make contact"
    # gradient over a short press
    x0 = torch.tensor(c["x"], device=dev)
    v0 = torch.tensor(c["v"], device=dev)
    C0 = torch.zeros(n, 3, 3, device=dev)
    F0 = torch.eye(3, device=dev).repeat(n, 1, 1)
    cvl = torch.tensor(c["control_vel"][:11], device=dev).requires_grad_(True)
    out = cs.rollout(x0, v0, C0, F0, [(pos0.clone(), cvl[0].detach().clone(), pos0.clone())],
                     [cvl], 10)
    loss = marker_displacement(mm, out[4][0][0], out[4][0][2])
    g = torch.autograd.grad(loss, cvl)[0]
    assert torch.isfinite(g).all() and float(g.abs().max()) > 0


def test_c4_population_batched_matches_the_loop_and_the_oracle(gpu):
    """C4 at full size (64 environments, 5k-200k particles, 64^3 / 128^3
    grids, a 10,800-tet pad each): one stacked batch per grid size, and all
    64 in one ragged batch, against one simulation per environment, state and
    d loss / d v0 per environment over 3 substeps (the ragged batch to 1e-6 /
    1e-5); the smallest environment also against the FP64 oracle."""
    from difftactile_amd.batch import CoupledEnvBatch
    envs = scenes.config4_population(pads=True)
    assert len(envs) == 64
    T = 3
    contact = dict(kn=1e4, kd=269.44, kt=154.78, mu_f=53.85)
    dev = torch.device("cuda")
    rng = np.random.default_rng(5)
    w = [torch.tensor(rng.normal(size=3), dtype=torch.float32, device=dev) for _ in envs]

    def state(e):
        d = envs[e]
        n = len(d["x"])
        return [torch.tensor(d["x"], device=dev),
                torch.tensor(d["v"], device=dev).requires_grad_(True),
                torch.zeros(n, 3, 3, device=dev), torch.eye(3, device=dev).repeat(n, 1, 1)]

    def pad_state(e):
        nodes = torch.tensor(envs[e]["pad_mesh"][0], device=dev)
        cv = torch.tensor(envs[e]["pad_cv"], device=dev)
        return (nodes.clone(), cv.clone(), nodes.clone()), cv.expand(T + 1, -1, -1).contiguous()

    groups = {}
    for e, d in enumerate(envs):
        groups.setdefault(d["prm"].n_grid, []).append(e)
    got_x, got_g = [None] * 64, [None] * 64
    for ids in groups.values():
        d0 = envs[ids[0]]
        spec = [(envs[e]["mats"], len(envs[e]["x"]), envs[e]["mat_id"]) for e in ids]
        b = CoupledEnvBatch(d0["prm"], spec, [envs[e]["pad_mesh"] for e in ids],
                            dict(out_dir=d0["pad_out"], **d0["pad"]), contact,
                            checkpoint_every=2, sort=True)
        sts = [state(e) for e in ids]
        ps = [pad_state(e) for e in ids]
        res = b.rollout(sts, [p[0] for p in ps], [p[1] for p in ps], T)
        loss = sum((r[0][0].mean(0) * w[e]).sum() for r, e in zip(res, ids))
        for e, r, ge in zip(ids, res, torch.autograd.grad(loss, [s[1] for s in sts])):
            got_x[e], got_g[e] = r[0][0].detach(), ge
    # all 64 in one ragged batch: one launch sequence over both grid sizes
    d0 = envs[0]
    rb = CoupledEnvBatch(d0["prm"], [(d["mats"], len(d["x"]), d["mat_id"], d["prm"]) for d in envs],
                         [d["pad_mesh"] for d in envs], dict(out_dir=d0["pad_out"], **d0["pad"]),
                         contact, checkpoint_every=2, sort=True)
    assert rb.batch.ragged
    sts = [state(e) for e in range(64)]
    ps = [pad_state(e) for e in range(64)]
    res = rb.rollout(sts, [p[0] for p in ps], [p[1] for p in ps], T)
    loss = sum((r[0][0].mean(0) * w[e]).sum() for e, r in enumerate(res))
    rag_g = torch.autograd.grad(loss, [s[1] for s in sts])
    rag_x = [r[0][0].detach() for r in res]
    del rb, res
    small = int(np.argmin([len(d["x"]) for d in envs]))
    for e, d in enumerate(envs):
        nodes, f2v, tris, layer = d["pad_mesh"]
        sim = MPMSim(d["prm"], d["mats"], len(d["x"]), mat_id=d["mat_id"], checkpoint_every=2,
                     sort=True)
        pad = Pad(nodes, f2v, tris, layer, d["pad_out"], d["pad"]["mu"], d["pad"]["lam"],
                  d["pad"]["mass"], d["prm"].dt)
        cs = CoupledSim(sim, [pad], **contact)
        st = state(e)
        p0, cvs = pad_state(e)
        out = cs.rollout(*st, [p0], [cvs], T)
        g = torch.autograd.grad((out[0].mean(0) * w[e]).sum(), st[1])[0]
        assert rel_l2(got_x[e], out[0].detach()) < 1e-5, e
        assert rel_l2(got_g[e], g) < 1e-4, e
        # the pads' vertex sums are exact fixed point at the scale of the
        # largest term of the whole call, so a batch rounds them at its own
        # scale: the object sums match bit for bit, the contact sums to rounding
        assert rel_l2(rag_x[e], out[0].detach()) < 1e-6, e
        assert rel_l2(rag_g[e], g) < 1e-5, e
        if e != small:
            continue
        # the smallest environment against the oracle
        n = len(d["x"])
        op = oracle_params(d["prm"])
        mats = oracle_mats(d["mats"], d["mat_id"], F64)
        so = [torch.tensor(d["x"], dtype=F64),
              torch.tensor(d["v"], dtype=F64, requires_grad=True),
              torch.zeros(n, 3, 3, dtype=F64), torch.eye(3, dtype=F64).repeat(n, 1, 1)]
        B = pad.B.detach().cpu().to(F64)
        cvo = torch.tensor(d["pad_cv"], dtype=F64)
        pp = torch.tensor(nodes, dtype=F64)
        state_o, ps = tuple(so), (pp, cvo.clone(), pp.clone())
        cc = dict(pad_mu=d["pad"]["mu"], pad_lam=d["pad"]["lam"], pad_mass=d["pad"]["mass"])
        for f in range(T):
            pd = _oracle_pad(*ps, cvo, cvo, f2v, B, layer, tris, d["pad_out"], cc, d["prm"].dt)
            state_o, po, _, _ = ocp.coupled_substep(op, state_o, mats, [pd],
                                                    tuple(contact.values()), clamp=True)
            ps = po[0]
        assert rel_l2(out[0].detach().cpu(), state_o[0].detach()) < 1e-5
        lo = (state_o[0].mean(0) * w[e].cpu().to(F64)).sum()
        go = torch.autograd.grad(lo, so[1])[0]
        assert rel_l2(g.cpu(), go) < 1e-3
