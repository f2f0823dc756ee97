"""Repeatability: the sparse path gives bit-identical states and gradients
when a rollout is run twice on the same inputs (the binning permutation is
the stable counting sort by brick and the brick passes stage each chunk in a
fixed order, so no float sum depends on the order atomics ran in), and the
re-sort of the particles at every checkpoint changes results only at float
rounding."""
import numpy as np
import pytest
import torch

from difftactile_amd import scenes
from difftactile_amd.sim import MPMSim
from tests.mpm_cases import rel_l2

pytestmark = pytest.mark.gpu


def _run(sim, x, v, w, T):
    n = len(x)
    st = [torch.tensor(x).cuda(), torch.tensor(v).cuda().requires_grad_(True),
          torch.zeros(n, 3, 3).cuda().requires_grad_(True),
          torch.eye(3).repeat(n, 1, 1).cuda().requires_grad_(True)]
    out = sim.rollout(*st, T)
    loss = (out[0] * w).sum() + (out[1] * w).sum()
    g = torch.autograd.grad(loss, st[1:])
    torch.cuda.synchronize()
    return [t.detach().clone() for t in out], [t.clone() for t in g]


@pytest.mark.parametrize("resident", [True, False])
def test_object_rollout_repeats_bit_for_bit(gpu, resident):
    prm, mats, x, v, mid = scenes.config3_multi_material(n_particles=60000, n_grid=64)
    n = len(x)
    w = torch.randn(n, 3, generator=torch.Generator().manual_seed(3)).cuda()
    T = 60
    runs = []
    for _ in range(2):
        sim = MPMSim(prm, mats, n, mat_id=mid, checkpoint_every=25,
                     resident_bytes=None if resident else 0, material_grad=True)
        out, g = _run(sim, x, v, w, T)
        runs.append((out, g, sim.material_grad.clone()))
    for a, b in zip(runs[0][0] + runs[0][1], runs[1][0] + runs[1][1]):
        assert torch.equal(a, b)
    # material adjoints: reduced in a fixed order too
    assert torch.equal(runs[0][2], runs[1][2])
    # a simulation reused for further rollouts (its pooled buffers and the
    # graphs replayed on them) repeats them too, after a rollout of other inputs
    sim = MPMSim(prm, mats, n, mat_id=mid, checkpoint_every=25,
                 resident_bytes=None if resident else 0, material_grad=True)
    _run(sim, x[::-1].copy(), -v, w, T)
    sim.zero_material_grad()
    out, g = _run(sim, x, v, w, T)
    for a, b in zip(out + g, runs[0][0] + runs[0][1]):
        assert torch.equal(a, b)
    assert torch.equal(sim.material_grad, runs[0][2])


def test_resorting_at_checkpoints_changes_only_rounding(gpu):
    prm, mats, x, v, mid = scenes.config3_multi_material(n_particles=60000, n_grid=64)
    n = len(x)
    w = torch.randn(n, 3, generator=torch.Generator().manual_seed(4)).cuda()
    res = []
    for resort in (True, False):
        sim = MPMSim(prm, mats, n, mat_id=mid, checkpoint_every=20, resort=resort)
        res.append(_run(sim, x, v, w, 60))
    # the orders of the float sums differ: positions and velocities to 1e-6;
    # C (a velocity gradient, 4 inv_dx times grid-velocity differences) and F
    # after 60 substeps to 1e-4; the gradients to the rollout bar of 1e-3
    tol = (1e-6, 1e-6, 1e-4, 1e-4)
    for a, b, t in zip(res[0][0], res[1][0], tol):
        assert rel_l2(a.cpu(), b.cpu()) < t
    for a, b in zip(res[0][1], res[1][1]):
        assert rel_l2(a.cpu(), b.cpu()) < 1e-3


def test_coupled_rollout_repeats_bit_for_bit(gpu):
    """The pad press (object, FEM pad, contact, clamp) twice: the pad-vertex
    contact sums are exact fixed-point sums and the FEM vertex forces are
    gathered per vertex, so states and the gradients w.r.t. the pad velocity
    sequence and the initial positions repeat bit for bit."""
    from difftactile_amd.coupled import CoupledSim, Pad
    c = scenes.config2_pad_press(n_particles=6000, n_grid=32, n_substeps=20,
                                 pad_verts=(12, 12, 3), press_substeps=100)
    prm, n = c["prm"], len(c["x"])
    nodes, f2v, tris, layer, od = c["mesh"]
    runs = []
    for _ in range(2):
        sim = MPMSim(prm, c["mats"], n, checkpoint_every=8)
        pad = Pad(nodes.astype(np.float32), f2v, tris, layer, od, c["pad_mu"], c["pad_lam"],
                  c["pad_mass"], prm.dt)
        cs = CoupledSim(sim, [pad], **c["contact"])
        x0 = torch.tensor(c["x"]).cuda().requires_grad_(True)
        cv = torch.tensor(c["control_vel"]).cuda().requires_grad_(True)
        pos0 = torch.tensor(nodes, dtype=torch.float32).cuda()
        out = cs.rollout(x0, torch.tensor(c["v"]).cuda(), torch.zeros(n, 3, 3).cuda(),
                         torch.eye(3).repeat(n, 1, 1).cuda(),
                         [(pos0.clone(), cv[0].detach().clone(), pos0.clone())], [cv], 20)
        w = torch.randn(n, 3, generator=torch.Generator().manual_seed(8)).cuda()
        wp = torch.randn(len(nodes), 3, generator=torch.Generator().manual_seed(9)).cuda()
        loss = (out[0] * w).sum() + (out[4][0][0] * wp).sum() + (out[4][0][1] * wp).sum()
        g = torch.autograd.grad(loss, [x0, cv])
        torch.cuda.synchronize()
        runs.append([t.detach().clone() for t in out[:4] + tuple(out[4][0])] +
                    [t.clone() for t in g])
    assert float(runs[0][-1].abs().max()) > 0
    for a, b in zip(runs[0], runs[1]):
        assert torch.equal(a, b)


@pytest.mark.parametrize("mode", ["dense", "overflow", "recompute_overflow"])
def test_compact_stores_match_dense_stores(gpu, mode):
    """Kept coupled substeps store only their active bricks' nodes (compact
    stores, used when whole-grid stores do not fit the budget): the pad press's states
    and gradients equal those of whole-grid stores bit for bit.  overflow: the
    compact stores get room for one brick, so every kept segment overflows and
    the backward pass recomputes it (its recompute overflows too and runs on
    whole-grid stores); recompute_overflow: nothing is kept, the recomputed
    segments overflow and run again on whole-grid stores."""
    from difftactile_amd.coupled import CoupledSim, Pad
    c = scenes.config2_pad_press(n_particles=6000, n_grid=32, n_substeps=20,
                                 pad_verts=(12, 12, 3), press_substeps=100)
    prm, n = c["prm"], len(c["x"])
    nodes, f2v, tris, layer, od = c["mesh"]

    def run(compact, tiny_cap=False, resident=True):
        sim = MPMSim(prm, c["mats"], n, checkpoint_every=8,
                     resident_bytes=None if resident else 0)
        pad = Pad(nodes.astype(np.float32), f2v, tris, layer, od, c["pad_mu"], c["pad_lam"],
                  c["pad_mass"], prm.dt)
        cs = CoupledSim(sim, [pad], **c["contact"])
        cs.compact_stores = compact
        if tiny_cap:
            cs.store_capacity = lambda x: 1
        x0 = torch.tensor(c["x"]).cuda().requires_grad_(True)
        cv = torch.tensor(c["control_vel"]).cuda().requires_grad_(True)
        pos0 = torch.tensor(nodes, dtype=torch.float32).cuda()
        out = cs.rollout(x0, torch.tensor(c["v"]).cuda(), torch.zeros(n, 3, 3).cuda(),
                         torch.eye(3).repeat(n, 1, 1).cuda(),
                         [(pos0.clone(), cv[0].detach().clone(), pos0.clone())], [cv], 20)
        kept = cs.last_resident[0]
        w = torch.randn(n, 3, generator=torch.Generator().manual_seed(8)).cuda()
        wp = torch.randn(len(nodes), 3, generator=torch.Generator().manual_seed(9)).cuda()
        loss = (out[0] * w).sum() + (out[4][0][0] * wp).sum() + (out[4][0][1] * wp).sum()
        g = torch.autograd.grad(loss, [x0, cv])
        torch.cuda.synchronize()
        return kept, [t.detach().clone() for t in out[:4] + tuple(out[4][0])] + \
            [t.clone() for t in g]

    kept_d, ref = run(False)
    assert kept_d == 3 and float(ref[-1].abs().max()) > 0
    if mode == "dense":
        kept_c, got = run(True)
        assert kept_c == 3
    elif mode == "overflow":
        kept_c, got = run(True, tiny_cap=True)
        assert kept_c == 0
    else:
        kept_c, got = run(True, tiny_cap=True, resident=False)
        assert kept_c == 0
    for a, b in zip(ref, got):
        assert torch.equal(a, b)
