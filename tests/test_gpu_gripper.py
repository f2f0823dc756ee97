"""The gripper scene on the GPU: two FEM pads on the parallel jaw, a
position-based cable and an elastoplastic object in one differentiable
rollout (coupled.CoupledSim with a cable).

* Scaled (6,000 particles, 32^3, 12 x 12 x 3 pads, a 101-particle cable, 6
  substeps): object, pad and cable states and the last substep's pad
  external forces against the FP64 oracle composition
  (oracle.gripper_scene); the gradients of a random linear functional of the
  final object, pad and cable positions and of the grasp task's pose + force
  loss (centre-of-mass error, sliding ratio of the pads' contact forces) with
  respect to the jaw controls (through Gripper.control_velocities), the
  initial object positions and the initial cable velocities against the
  oracle's reverse mode, with the adjoint clamp on.
* Full size (2M particles, 256^3, 40 x 40 x 6 pads, a 2,001-particle cable):
  over a few substeps, the contact impulses on the object and the cable
  balance the pads' contact forces, the scatter conserves mass, and the
  gradient with respect to the controls is finite and non-zero.
"""
import numpy as np
import pytest
import torch

from difftactile_amd import _lib, scenes
from difftactile_amd.coupled import CoupledSim, Pad
from difftactile_amd.losses import com_loss, contact_force, force_loss
from difftactile_amd.object_model.pbd_rope import PBDRope
from difftactile_amd.sensor_model.gripper_kinematics import Gripper
from difftactile_amd.sim import MPMSim, workspace_bytes
from oracle import gripper_scene as ogs
from oracle import kinematics as ok
from oracle import markers as ok_m
from tests.gripper_cases import oracle_B, oracle_pads, oracle_rope_params, small_scene
from tests.mpm_cases import oracle_mats, oracle_params, rel_l2

pytestmark = pytest.mark.gpu
F64 = torch.float64
P = _lib.ptr


def _gpu_scene(c, checkpoint_every=3):
    prm, n = c["prm"], len(c["x"])
    sim = MPMSim(prm, c["mats"], n, mat_id=c["mat_id"], checkpoint_every=checkpoint_every,
                 sort=True)
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
    init_x = []
    for k, (nodes, *_) in enumerate(c["pads"]):
        Pk = grip.finger_pose(k + 1).astype(np.float64)
        nh = np.concatenate([nodes.astype(np.float64), np.ones((len(nodes), 1))], 1)
        init_x.append((nh @ np.linalg.inv(Pk).T)[:, :3].astype(np.float32))
    return sim, pads, rope, cs, grip, init_x, rp0, rv0


def test_c5_scaled_gripper_matches_the_oracle(gpu):
    steps = 6
    c = small_scene(steps)
    prm, n = c["prm"], len(c["x"])
    sim, pads, rope, cs, grip, init_x, rp0, rv0 = _gpu_scene(c)
    dev = torch.device("cuda")
    nr = rope.n_particles
    rng = np.random.default_rng(23)
    wx = rng.normal(size=(n, 3))
    wp = [rng.normal(size=(pd.n_verts, 3)) for pd in pads]
    wr = rng.normal(size=(nr, 3))
    rv0 = torch.tensor(rng.normal(0, 0.5, size=(nr, 3)).astype(np.float32), device=dev)
    ctrl = torch.tensor(c["controls"], device=dev).requires_grad_(True)
    cv1, cv2 = grip.control_velocities(ctrl, *[torch.tensor(x, device=dev) for x in init_x])
    # the jaw closes each pad along its face normal at the closing speed
    for cvk, (*_, od) in zip((cv1, cv2), c["pads"]):
        want = 10.0 * torch.tensor(od, dtype=torch.float32, device=dev)
        assert float((cvk.detach() - want).abs().max()) < 0.05
    x0 = torch.tensor(c["x"], device=dev).requires_grad_(True)
    v0 = torch.tensor(c["v"], device=dev)
    C0 = torch.zeros(n, 3, 3, device=dev)
    F0 = torch.eye(3, device=dev).repeat(n, 1, 1)
    rvl = rv0.clone().requires_grad_(True)
    pst = []
    for k, pd in enumerate(pads):
        pos0 = torch.tensor(c["pads"][k][0], device=dev)
        pst.append((pos0.clone(), (cv1, cv2)[k][0].detach().clone(), pos0.clone()))
    out = cs.rollout(x0, v0, C0, F0, pst, [cv1, cv2], steps, rope_state=(rp0.clone(), rvl))
    objo, pado, ropeo = out[:4], out[4], out[5]
    f32 = lambda a: torch.tensor(a, dtype=torch.float32, device=dev)
    loss = (objo[0] * f32(wx)).sum() + sum((p[0] * f32(w)).sum() for p, w in zip(pado, wp)) + \
        (ropeo[0] * f32(wr)).sum()
    g_ctrl, g_x, g_rv, g_cv1, g_cv2 = torch.autograd.grad(loss, [ctrl, x0, rvl, cv1, cv2])
    # the grasp task's pose and force terms: centre-of-mass error and the
    # sliding ratio of the last substep's pad contact forces (a second rollout:
    # a rollout's backward runs once)
    cv1, cv2 = grip.control_velocities(ctrl, *[torch.tensor(x, device=dev) for x in init_x])
    outf = cs.rollout(x0, v0, C0, F0, pst, [cv1, cv2], steps, rope_state=(rp0.clone(), rv0),
                      return_ext=True)
    exto = outf[6]
    tgt = torch.tensor(c["target_com"], dtype=torch.float32, device=dev)
    mu_f = c["contact"]["mu_f"]
    tris_d = [torch.tensor(np.asarray(p[2]), dtype=torch.int32, device=dev) for p in c["pads"]]
    loss_f = ((com_loss(outf[0]) - tgt) ** 2).sum() + \
        force_loss([contact_force(t, e) for t, e in zip(tris_d, exto)], c["pad_inv_rot"], mu_f)
    gf_ctrl, gf_x = torch.autograd.grad(loss_f, [ctrl, x0])

    def oracle(dtype):
        """The oracle's rollout from the same float32 inputs (control
        velocities included) in dtype."""
        op = oracle_params(prm)
        mats = oracle_mats(c["mats"], c["mat_id"], dtype)
        so = [torch.tensor(c["x"]).to(dtype).requires_grad_(True), torch.tensor(c["v"]).to(dtype),
              torch.zeros(n, 3, 3, dtype=dtype), torch.eye(3, dtype=dtype).repeat(n, 1, 1)]
        cvo = [t.detach().cpu().to(dtype).requires_grad_(True) for t in (cv1, cv2)]
        B = oracle_B(c, dtype)
        rp = oracle_rope_params(c)
        rpos = rp0.detach().cpu().to(dtype)
        rvo = rv0.detach().cpu().to(dtype).requires_grad_(True)
        fixed = rpos[0].clone()
        state = tuple(so)
        pos = [torch.tensor(p[0]).to(dtype) for p in c["pads"]]
        ps = [(pos[k], cvo[k][0].detach().clone(), pos[k].clone()) for k in range(2)]
        rs = (rpos, rvo)
        g_contacts = r_contacts = 0
        info = {}
        for f in range(steps):
            pd = oracle_pads(c, [p[0] for p in ps], [p[1] for p in ps], [p[2] for p in ps], cvo, f,
                             B, dtype)
            info = {}
            state, po, rs = ogs.gripper_substep(op, state, mats, pd, (rs[0], rs[1], fixed), rp,
                                                tuple(c["contact"].values()),
                                                tuple(c["rope_contact"].values()), clamp=True,
                                                info=info)
            g_contacts = max(g_contacts, int((info["grid_f"].abs().sum(1) > 0).sum()))
            r_contacts = max(r_contacts, int((info["impulse"].abs().sum(1) > 0).sum()))
            ps = [tuple(p) for p in po]
        return state, ps, rs, (so, cvo, rvo), (g_contacts, r_contacts), info["ext"]

    state, ps, rs, (so, cvo, rvo), (g_contacts, r_contacts), ext64 = oracle(F64)
    assert g_contacts > 20 and r_contacts > 5, (g_contacts, r_contacts)
    for got, want in zip(objo, state):
        assert rel_l2(got.detach().cpu(), want.detach()) < 1e-4
    for k in range(2):
        for got, want in zip(pado[k], ps[k]):
            assert rel_l2(got.detach().cpu(), want.detach()) < 1e-4
    assert rel_l2(ropeo[0].detach().cpu(), rs[0].detach()) < 1e-4
    # the cable's velocity is (p' - p) / dt of float32 positions: at dt = 1e-5
    # and |p| ~ 5 one position ulp is 0.05 of velocity, so any float32
    # evaluation of it is rounding-bound; it is held to three times the
    # float32 oracle's distance from float64
    o32 = oracle(torch.float32)
    rs32, ext32 = o32[2], o32[5]
    tol = 3 * rel_l2(rs32[1].detach().double(), rs[1].detach())
    assert rel_l2(ropeo[1].detach().cpu(), rs[1].detach()) <= max(tol, 1e-4), tol
    lo = (state[0] * torch.tensor(wx)).sum() + sum((p[0] * torch.tensor(w)).sum()
                                                   for p, w in zip(ps, wp)) + \
        (rs[0] * torch.tensor(wr)).sum()
    go_cv1, go_cv2, go_x, go_rv = torch.autograd.grad(lo, [cvo[0], cvo[1], so[0], rvo],
                                                      retain_graph=True)
    for got, want in ((g_cv1, go_cv1), (g_cv2, go_cv2), (g_x, go_x), (g_rv, go_rv)):
        assert float(want.abs().max()) > 0
        assert rel_l2(got.cpu().double(), want) < 1e-3
    # controls: the oracle's jaw chain applied to the oracle's d loss / d velocities
    g = c["gripper"]
    co = torch.tensor(c["controls"], dtype=F64, requires_grad=True)
    W, f1t, f2t, frame0 = ok.init_frames(g["rot"], g["trans"], g["grip_width"])
    fo = ok.chain(W, f1t, f2t, frame0, co, prm.dt)
    cvs = [ok.control_vel(fo, k + 1, torch.tensor(init_x[k], dtype=F64), prm.dt) for k in range(2)]
    want = torch.autograd.grad(cvs, co, grad_outputs=[go_cv1, go_cv2], retain_graph=True)[0]
    assert float(want.abs().max()) > 0
    assert rel_l2(g_ctrl.cpu().double(), want) < 1e-3
    # pose + force loss.  The cable's pad force carries its rounding-bound
    # velocity (kd n.v), and the force loss is the ratio of the pads' small
    # tangential force sum to the normal one: both are held to three times
    # the float32 oracle's distance from float64
    for got, want_e, e32 in zip(exto, ext64, ext32):
        tol = 3 * rel_l2(e32.detach().double(), want_e.detach())
        assert rel_l2(got.detach().cpu(), want_e.detach()) <= max(tol, 1e-4), tol

    def pose_force(st0, ext):
        cf = [ok_m.contact_force(e, p[2]) for e, p in zip(ext, c["pads"])]
        return ((st0.mean(0) - torch.tensor(c["target_com"]).to(st0.dtype)) ** 2).sum() + \
            ok_m.force_loss(cf, c["pad_inv_rot"], mu_f)

    lof = pose_force(state[0], ext64)
    lof32 = pose_force(o32[0][0], ext32)
    tol = 3 * abs(float(lof32.detach()) - float(lof.detach()))
    assert abs(float(loss_f.detach()) - float(lof.detach())) <= max(tol, 1e-4 * abs(float(lof)))
    gof_cv1, gof_cv2, gof_x = torch.autograd.grad(lof, [cvo[0], cvo[1], so[0]])
    wantf = torch.autograd.grad(cvs, co, grad_outputs=[gof_cv1, gof_cv2])[0]
    assert float(wantf.abs().max()) > 0
    assert rel_l2(gf_ctrl.cpu().double(), wantf) < 1e-3
    assert rel_l2(gf_x.cpu().double(), gof_x) < 1e-3


def test_c5_full_size_contact_balance_mass_and_gradient(gpu):
    c = scenes.config5_gripper()
    prm, n = c["prm"], len(c["x"])
    sim, pads, rope, cs, grip, init_x, rp0, rv0 = _gpu_scene(c, checkpoint_every=4)
    dev = sim.device
    T = 4
    ctrl = torch.tensor(c["controls"][:T + 1], device=dev)
    with torch.no_grad():
        cvs = grip.control_velocities(ctrl, *[torch.tensor(x, device=dev) for x in init_x])
    st = [torch.tensor(c["x"], device=dev), torch.tensor(c["v"], device=dev),
          torch.zeros(n, 3, 3, device=dev), torch.eye(3, device=dev).repeat(n, 1, 1)]
    perm = sim.sort_order(st[0])[1]
    st = [sim._gather(perm, t, w) for t, w in zip(st, (3, 3, 9, 9))]
    mid = sim._gather_i(perm, sim.mat_id)
    sim._set_env(None)
    pst = []
    for k, pd in enumerate(pads):
        pos0 = torch.tensor(c["pads"][k][0], device=dev)
        pst.append([pos0.clone(), cvs[k][0].clone(), pos0.clone()])
    pst.append([rp0.clone(), rv0.clone()])
    frames, pframes, stores = cs._run_keep(st, pst, list(cvs), 0, T, mid)
    torch.cuda.synchronize()
    ws = torch.empty(workspace_bytes(sim.c_prm, n), dtype=torch.uint8, device=dev)
    g_touch = r_touch = 0
    for j, store in enumerate(stores):
        g4 = torch.zeros(prm.n_nodes, 4, device=dev)
        fr = frames[j]
        F1 = torch.empty_like(fr[3])
        s = _lib.stream()
        _lib.call("dt_mpm_bin", sim.c_prm, n, P(fr[0]), P(ws, torch.uint8), ws.numel(), s)
        _lib.call("dt_mpm_p2g_binned", sim.c_prm, sim.mats, 2, n, P(mid, torch.int32),
                  *[P(t) for t in fr], P(F1), P(ws, torch.uint8), ws.numel(), s)
        _lib.call("dt_mpm_tiles_sum", sim.c_prm, n, P(ws, torch.uint8), ws.numel(), P(g4), s)
        m_tot = float(g4[:, 3].double().sum())
        assert abs(m_tot - n * prm.p_mass) <= 1e-5 * n * prm.p_mass, j
        occ = g4[:, 3] > prm.eps
        fo = store["gf"][occ].double()
        imp = store["rimp"].double()
        ext = sum(e.double().sum(0) for e in store["ext"])
        # contact nodes on each pad's side of the block
        hit = occ & (store["gf"].abs().sum(1) > 0)
        i_node = torch.nonzero(hit).squeeze(1) // (prm.n_grid * prm.n_grid)
        g_touch = max(g_touch, min(int((i_node < prm.n_grid // 2).sum()),
                                   int((i_node >= prm.n_grid // 2).sum())))
        r_touch = max(r_touch, int((imp.abs().sum(1) > 0).sum()))
        bal = fo.sum(0) / prm.dt + imp.sum(0) / prm.dt + ext
        mag = fo.abs().sum() / prm.dt + imp.abs().sum() / prm.dt + \
            sum(e.double().abs().sum() for e in store["ext"])
        assert float(bal.abs().max()) <= 1e-5 * float(mag) + 1e-12, (j, bal, mag)
        for t in pframes[j + 1][-1]:
            assert torch.isfinite(t).all()
    assert g_touch > 300, "both pads must touch the block"
    assert r_touch > 20, "the cable must touch a pad"
    del frames, pframes, stores
    # gradient of a cable-and-object loss w.r.t. the controls
    ctrl = ctrl.clone().requires_grad_(True)
    cv1, cv2 = grip.control_velocities(ctrl, *[torch.tensor(x, device=dev) for x in init_x])
    st0 = [torch.tensor(c["x"], device=dev), torch.tensor(c["v"], device=dev),
           torch.zeros(n, 3, 3, device=dev), torch.eye(3, device=dev).repeat(n, 1, 1)]
    pst = [(p[0].clone(), cvk[0].detach().clone(), p[0].clone())
           for p, cvk in zip(pst[:2], (cv1, cv2))]
    out = cs.rollout(*st0, pst, [cv1, cv2], T, rope_state=(rp0.clone(), rv0.clone()))
    tgt = torch.tensor(c["target_com"], dtype=torch.float32, device=dev)
    loss = ((out[0].mean(0) - tgt) ** 2).sum() + (out[5][0][:, 0] ** 2).mean()
    g = torch.autograd.grad(loss, ctrl)[0]
    assert torch.isfinite(g).all() and float(g.abs().max()) > 0
