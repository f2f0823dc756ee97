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
