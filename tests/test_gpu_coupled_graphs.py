"""Graph capture of the coupled rollout (CoupledSim(graphs=True)): each
checkpoint segment's forward and adjoint launch sequences are captured once
and replayed by a steady caller's later rollouts; every replay gives the same
bits as the eager launches, with kept, recomputed and compact-store segments,
without gradients, and with a cable and two pads."""
import numpy as np
import pytest
import torch

from difftactile_amd import scenes
from difftactile_amd.coupled import CoupledSim, Pad
from difftactile_amd.object_model.pbd_rope import PBDRope
from difftactile_amd.sim import MPMSim
from tests.gripper_cases import small_scene

pytestmark = pytest.mark.gpu


def _c2(graphs, mode="resident"):
    c = scenes.config2_pad_press(n_particles=6000, n_grid=32, n_substeps=12,
                                 pad_verts=(12, 12, 3), press_substeps=500)
    prm, n = c["prm"], len(c["x"])
    nodes, f2v, tris, layer, od = c["mesh"]
    sim = MPMSim(prm, c["mats"], n, checkpoint_every=4, sort=True,
                 resident_bytes=0 if mode == "recompute" else None)
    pad = Pad(nodes.astype(np.float32), f2v, tris, layer, od, c["pad_mu"], c["pad_lam"],
              c["pad_mass"], prm.dt)
    cs = CoupledSim(sim, [pad], **c["contact"], graphs=graphs)
    if mode == "compact":
        cs.compact_stores = True
    if mode == "partial":
        # the first two of three segments kept (whole-grid stores)
        cs.compact_stores = False
        sim.resident_bytes = 2 * 5 * cs.substep_bytes()
    return c, cs


def _run(c, cs, T=12, grad=True):
    dev = torch.device("cuda")
    n = len(c["x"])
    nodes = c["mesh"][0]
    x0 = torch.tensor(c["x"], device=dev).requires_grad_(grad)
    v0 = torch.tensor(c["v"], device=dev)
    C0 = torch.zeros(n, 3, 3, device=dev)
    F0 = torch.eye(3, device=dev).repeat(n, 1, 1)
    pos0 = torch.tensor(nodes, dtype=torch.float32, device=dev)
    cv = torch.tensor(c["control_vel"][:T + 1], device=dev).requires_grad_(grad)
    with torch.set_grad_enabled(grad):
        out = cs.rollout(x0, v0, C0, F0, [(pos0.clone(), cv[0].detach().clone(), pos0.clone())],
                         [cv], T, return_ext=True)
    res = [out[0].detach().clone(), out[4][0][0].detach().clone(), out[5][0].detach().clone()]
    if grad:
        w = torch.randn(out[0].shape, device=dev,
                        generator=torch.Generator("cuda").manual_seed(2))
        loss = (out[0] * w).sum() + out[4][0][0].sum() + out[5][0].sum()
        res += list(torch.autograd.grad(loss, [x0, cv]))
    torch.cuda.synchronize()
    return res


@pytest.mark.parametrize("mode", ["resident", "recompute", "compact", "partial"])
def test_captured_segments_repeat_the_eager_bits(gpu, mode):
    """Three segments of four substeps: the first rollout captures each
    segment's forward (and, for segments not kept, its recompute) and its
    adjoint; the second and third replay them, with eager work and
    synchronisation between the replays.  partial: two segments kept, the
    third recomputed (the pooled buffers come back in the same order, so the
    second rollout replays too)."""
    c, eager = _c2(False, mode)
    ref = _run(c, eager)
    c, cs = _c2(True, mode)
    runs = []
    for r in range(3):
        runs.append(_run(c, cs))
        # the program's own eager work between replays
        torch.cuda.synchronize()
        _ = (torch.randn(4096, device="cuda") * 2).sum().item()
    for got in runs:
        for a, b in zip(ref, got):
            assert torch.equal(a, b)
    per = {"resident": 6, "recompute": 9, "compact": 6, "partial": 7}[mode]
    assert cs.last_resident[0] == {"resident": 3, "recompute": 0, "compact": 3, "partial": 2}[mode]
    assert cs.graph_misses == per, (cs.graph_misses, cs.graph_hits)
    assert cs.graph_hits == 2 * per, (cs.graph_misses, cs.graph_hits)
    assert cs.graphs
    assert float(ref[-1].abs().max()) > 0


def test_captured_segments_without_gradients(gpu):
    c, eager = _c2(False)
    ref = _run(c, eager, grad=False)
    c, cs = _c2(True)
    for _ in range(3):
        got = _run(c, cs, grad=False)
        for a, b in zip(ref, got):
            assert torch.equal(a, b)
    assert cs.graph_misses == 3 and cs.graph_hits == 6


def _gripper(graphs):
    c = small_scene(6)
    prm, n = c["prm"], len(c["x"])
    sim = MPMSim(prm, c["mats"], n, mat_id=c["mat_id"], checkpoint_every=3, sort=True)
    pads = [Pad(nodes, f2v, tris, layer, od, c["pad_mu"], c["pad_lam"], c["pad_mass"], prm.dt)
            for nodes, f2v, tris, layer, od in c["pads"]]
    r = c["rope"]
    rope = PBDRope(dt=prm.dt, p_rad=r["p_rad"], p_rho=r["p_rho"], n_particles=r["n"],
                   rest_length=r["rest_length"])
    rope.stretch_compliance = rope.bending_compliance = r["compliance"]
    rp0, _ = rope.init(r["start"], r["ori"])
    cs = CoupledSim(sim, pads, **c["contact"], rope=rope, rope_contact=c["rope_contact"],
                    graphs=graphs)
    return c, cs, rp0


def _run_gripper(c, cs, rp0):
    dev = torch.device("cuda")
    n = len(c["x"])
    rng = np.random.default_rng(5)
    nr = rp0.shape[0]
    rv0 = torch.tensor(rng.normal(0, 0.5, size=(nr, 3)).astype(np.float32),
                       device=dev).requires_grad_(True)
    x0 = torch.tensor(c["x"], device=dev).requires_grad_(True)
    pst, cvs = [], []
    for k, (nodes, *_, od) in enumerate(c["pads"]):
        cv = (10.0 * torch.tensor(od, dtype=torch.float32, device=dev)).expand(
            7, len(nodes), 3).contiguous().requires_grad_(True)
        pos0 = torch.tensor(nodes, device=dev)
        pst.append((pos0.clone(), cv[0].detach().clone(), pos0.clone()))
        cvs.append(cv)
    out = cs.rollout(x0, torch.tensor(c["v"], device=dev), torch.zeros(n, 3, 3, device=dev),
                     torch.eye(3, device=dev).repeat(n, 1, 1), pst, cvs, 6,
                     rope_state=(rp0.clone(), rv0), return_ext=True)
    w = torch.randn(n, 3, device=dev, generator=torch.Generator("cuda").manual_seed(3))
    loss = (out[0] * w).sum() + sum(p[0].sum() for p in out[4]) + out[5][0].sum() + \
        sum(e.sum() for e in out[6])
    g = torch.autograd.grad(loss, [x0, rv0] + cvs)
    torch.cuda.synchronize()
    return [out[0].detach().clone(), out[5][0].detach().clone()] + \
        [p[0].detach().clone() for p in out[4]] + [e.detach().clone() for e in out[6]] + list(g)


def test_captured_gripper_segments_repeat_the_eager_bits(gpu):
    """Two pads and a cable, two segments of three substeps."""
    c, eager, rp0 = _gripper(False)
    ref = _run_gripper(c, eager, rp0)
    c, cs, rp0 = _gripper(True)
    for _ in range(3):
        got = _run_gripper(c, cs, rp0)
        for a, b in zip(ref, got):
            assert torch.equal(a, b)
    assert cs.graph_misses == 4 and cs.graph_hits == 8
    assert all(float(t.abs().max()) > 0 for t in ref[-3:])
