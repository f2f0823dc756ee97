"""Batched populations (EnvBatch): environments stacked along x in one MPMSim
advance exactly as the same environments simulated one at a time."""
import numpy as np
import pytest
import torch

from difftactile_amd.batch import EnvBatch
from difftactile_amd.sim import MPMSim
from tests.mpm_cases import blob, default_prm, materials, rel_l2

pytestmark = pytest.mark.gpu


def _envs():
    """Three environments: one pushed into its +x wall, one multi-material, one
    pushed into its -x wall (the walls act per environment)."""
    out = []
    x, v, C, F = blob(1500, seed=41, centre=0.8, spread=0.12)
    v[:, 0] += 3.0
    out.append(("elastic", x, v, 0.1 * C, F, None))
    x, v, C, F = blob(2500, seed=42)
    mid = np.random.default_rng(42).integers(0, 3, 2500).astype(np.int32)
    out.append(("multi", x, v, 0.1 * C, F, mid))
    x, v, C, F = blob(800, seed=43, centre=0.2, spread=0.1)
    v[:, 0] -= 3.0
    out.append(("von_mises", x, v, 0.1 * C, F, None))
    return out


def _seeds(states, base):
    g = torch.Generator().manual_seed(base)
    return [[torch.randn(t.shape, generator=g).cuda() for t in st] for st in states]


@pytest.mark.parametrize("sparse", [True, False])
def test_batched_environments_match_separate_simulations(gpu, sparse):
    prm = default_prm(32)
    envs = _envs()
    steps, every = 12, 5
    # both sides sum in float32 in orders fixed by atomics (binning ranks, run
    # scatter), and 12 substeps of wall impacts and plastic flow amplify those
    # roundings: state to 1e-4, gradients to 1e-3 (the rollout-vs-oracle bar)
    batch = EnvBatch(prm, [(materials(k), len(x), mid) for k, x, _, _, _, mid in envs],
                     checkpoint_every=every, sort=True, sparse_grid=sparse, material_grad=True)
    states = [[torch.tensor(a).cuda().requires_grad_(True) for a in (x, v, C, F)]
              for _, x, v, C, F, _ in envs]
    out = batch.rollout(states, steps)
    seeds = _seeds(out, 7)
    loss = sum((o * s).sum() for oe, se in zip(out, seeds) for o, s in zip(oe, se))
    grads = torch.autograd.grad(loss, [t for st in states for t in st])
    gmat = batch.material_grad()
    for e, (kind, x, v, C, F, mid) in enumerate(envs):
        sim = MPMSim(prm, materials(kind), len(x), mat_id=mid, checkpoint_every=every, sort=True,
                     sparse_grid=sparse, material_grad=True)
        st = [torch.tensor(a).cuda().requires_grad_(True) for a in (x, v, C, F)]
        ref = sim.rollout(*st, steps)
        for got, want in zip(out[e], ref):
            assert rel_l2(got.detach(), want.detach()) < 1e-4, (e, "state")
        loss_e = sum((o * s).sum() for o, s in zip(ref, seeds[e]))
        g_ref = torch.autograd.grad(loss_e, st)
        for k, (got, want) in enumerate(zip(grads[4 * e:4 * e + 4], g_ref)):
            err = rel_l2(got, want)
            assert err < 1e-3, (e, "xvCF"[k], err)
        err = rel_l2(gmat[e], sim.material_grad)
        assert err < 1e-3, (e, "mu/lam", err)


def test_environment_ids_are_validated(gpu):
    prm = default_prm(32)
    prm.n_env = 2
    with pytest.raises(ValueError, match="env_id"):
        MPMSim(prm, materials("elastic"), 10)
    with pytest.raises(ValueError, match="out of range"):
        MPMSim(prm, materials("elastic"), 10, env_id=np.full(10, 2, np.int32))


@pytest.mark.parametrize("sparse", [True, False])
def test_batched_pads_match_separate_coupled_simulations(gpu, sparse):
    """Environments with one FEM pad each (pads concatenated, each grid node
    contacting only its own environment's pad) against the same environments
    as separate coupled simulations: states, pads and gradients, including
    d/d(pad velocity sequence)."""
    from difftactile_amd.batch import CoupledEnvBatch
    from difftactile_amd.coupled import CoupledSim, Pad
    from difftactile_amd.params import MPMParams, Material
    from difftactile_amd.sensor_model.meshes import slab_mesh
    prm = MPMParams(n_grid=64, rho=1000.0, dt=2e-5)
    steps, every = 5, 3
    mu, lam, mass = 1e5 / 2 / 1.45, 1e5 * 0.45 / 1.45 / 0.1, 2e-4
    contact = dict(kn=1e3, kd=10.0, kt=50.0, mu_f=0.5)
    h = 1.0 / 128
    envs = []
    for e, (n, origin) in enumerate([(4000, (0.4, 0.52, 0.4)), (2500, (0.35, 0.5, 0.42)),
                                     (3000, (0.42, 0.53, 0.38))]):
        nodes, f2v, surf, layer, od = slab_mesh(24 - 4 * e, 24, 4, h=h, origin=origin)
        rng = np.random.default_rng(50 + e)
        pos = (nodes + 2e-4 * rng.normal(size=nodes.shape)).astype(np.float32)
        lo = np.asarray(origin) + np.array([0.0, -0.08, 0.0])
        x = rng.uniform(lo, lo + np.array([0.17, 0.11, 0.17]), size=(n, 3)).astype(np.float32)
        v = rng.normal(0, 0.2, size=(n, 3)).astype(np.float32)
        cv = np.zeros((steps + 1, len(nodes), 3), np.float32)
        cv[:, :, 1] = -0.5 + 0.05 * rng.normal(size=(steps + 1, 1))
        envs.append(dict(mesh=(nodes, f2v, surf, layer), od=od, pos=pos, x=x, v=v, cv=cv))

    def leaves(d):
        n = len(d["x"])
        st = [torch.tensor(d["x"]).cuda().requires_grad_(True),
              torch.tensor(d["v"]).cuda().requires_grad_(True),
              torch.zeros(n, 3, 3, device="cuda", requires_grad=True),
              torch.eye(3, device="cuda").repeat(n, 1, 1).requires_grad_(True)]
        pst = (torch.tensor(d["pos"]).cuda().requires_grad_(True),
               torch.tensor(d["cv"][0]).cuda().requires_grad_(True), torch.tensor(d["pos"]).cuda())
        return st, pst, torch.tensor(d["cv"]).cuda().requires_grad_(True)

    batch = CoupledEnvBatch(prm, [([Material(5e4, 0.3)], len(d["x"]), None) for d in envs],
                            [d["mesh"] for d in envs],
                            dict(out_dir=envs[0]["od"], mu=mu, lam=lam, mass=mass), contact,
                            checkpoint_every=every, sort=True, sparse_grid=sparse)
    lv = [leaves(d) for d in envs]
    out = batch.rollout([l[0] for l in lv], [l[1] for l in lv], [l[2] for l in lv], steps)
    g = torch.Generator().manual_seed(3)
    seeds = [[torch.randn(t.shape, generator=g).cuda() for t in list(o) + list(p)]
             for o, p in out]
    loss = sum((t * s).sum() for (o, p), se in zip(out, seeds)
               for t, s in zip(list(o) + list(p), se))
    flat = [t for st, pst, cvt in lv for t in st + [pst[0], pst[1], cvt]]
    grads = torch.autograd.grad(loss, flat)
    touched = 0
    for e, d in enumerate(envs):
        sim = MPMSim(prm, [Material(5e4, 0.3)], len(d["x"]), checkpoint_every=every, sort=True,
                     sparse_grid=sparse)
        nodes, f2v, surf, layer = d["mesh"]
        pad = Pad(nodes.astype(np.float32), f2v, surf, layer, d["od"], mu, lam, mass, prm.dt)
        cs = CoupledSim(sim, [pad], **contact)
        st, pst, cvt = leaves(d)
        ref = cs.rollout(*st, [pst], [cvt], steps)
        ref_t = list(ref[:4]) + list(ref[4][0])
        got_t = list(out[e][0]) + list(out[e][1])
        for k, (a, b) in enumerate(zip(got_t, ref_t)):
            assert rel_l2(a.detach(), b.detach()) < 1e-4, (e, k)
        touched += int((ref[4][0][1] - pst[1]).abs().max() > 0)
        loss_e = sum((t * s).sum() for t, s in zip(ref_t, seeds[e]))
        g_ref = torch.autograd.grad(loss_e, st + [pst[0], pst[1], cvt])
        for k, (a, b) in enumerate(zip(grads[7 * e:7 * e + 7], g_ref)):
            err = rel_l2(a, b)
            assert err < 1e-3, (e, k, err)
    assert touched == len(envs), "every pad must move"


def test_ragged_environments_match_separate_simulations(gpu):
    """Environments of 32^3, 48^3 and 64^3 grids in one ragged batch (one
    launch sequence, each environment keeping its grid, dx, particle mass and
    walls inside a 64^3 slot) against each one simulated alone on its own
    grid: states and d loss / d (x0, v0, C0, F0) bit for bit (the brick-owned
    sums see the same particles in the same order)."""
    from difftactile_amd.params import MPMParams
    grids = (32, 48, 64)
    prms = [MPMParams(n_grid=g, rho=1000.0, dt=2e-5) for g in grids]
    envs = _envs()
    steps, every = 8, 3
    batch = EnvBatch(prms[0], [(materials(k), len(x), mid, q) for (k, x, _, _, _, mid), q in
                               zip(envs, prms)], checkpoint_every=every, sort=True)
    assert batch.ragged and batch.prm.n_grid == 64
    states = [[torch.tensor(a).cuda().requires_grad_(True) for a in (x, v, C, F)]
              for _, x, v, C, F, _ in envs]
    out = batch.rollout(states, steps)
    seeds = _seeds(out, 11)
    loss = sum((o * s).sum() for oe, se in zip(out, seeds) for o, s in zip(oe, se))
    grads = torch.autograd.grad(loss, [t for st in states for t in st])
    for e, (kind, x, v, C, F, mid) in enumerate(envs):
        sim = MPMSim(prms[e], materials(kind), len(x), mat_id=mid, checkpoint_every=every,
                     sort=True)
        st = [torch.tensor(a).cuda().requires_grad_(True) for a in (x, v, C, F)]
        ref = sim.rollout(*st, steps)
        assert float((ref[0] - st[0]).abs().max()) > 0
        for got, want in zip(out[e], ref):
            assert torch.equal(got.detach(), want.detach()), (e, rel_l2(got.detach(), want.detach()))
        g_ref = torch.autograd.grad(sum((o * s).sum() for o, s in zip(ref, seeds[e])), st)
        for got, want in zip(grads[4 * e: 4 * e + 4], g_ref):
            assert torch.equal(got, want), (e, rel_l2(got, want))


def test_ragged_batches_reject_the_dense_path(gpu):
    from difftactile_amd.params import MPMParams
    prms = [MPMParams(n_grid=g, rho=1000.0, dt=2e-5) for g in (32, 64)]
    with pytest.raises(ValueError):
        EnvBatch(prms[0], [(materials("elastic"), 10, None, q) for q in prms], sparse_grid=False)
    bad = MPMParams(n_grid=64, rho=1000.0, dt=1e-5)
    with pytest.raises(ValueError):
        EnvBatch(prms[0], [(materials("elastic"), 10, None, prms[0]),
                           (materials("elastic"), 10, None, bad)])
