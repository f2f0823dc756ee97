"""Launch shapes on the GPU: a rollout's states and gradients do not depend
on the persistent-grid sizes, and the tuner returns a candidate setting."""
import pytest
import torch

from difftactile_amd import autotune, scenes
from difftactile_amd.sim import MPMSim

pytestmark = pytest.mark.gpu


def _case():
    prm, mats, x, v, mid = scenes.config3_multi_material(n_particles=20000, n_grid=64)
    sim = MPMSim(prm, mats, len(x), mat_id=mid, checkpoint_every=3)
    dev = torch.device("cuda")
    n = len(x)
    st = [torch.tensor(x, device=dev), torch.tensor(v, device=dev),
          torch.zeros(n, 3, 3, device=dev), torch.eye(3, device=dev).repeat(n, 1, 1)]
    return sim, st


def _run(sim, st, T=7):
    leaves = [t.detach().clone().requires_grad_(True) for t in st]
    out = sim.rollout(*leaves, T)
    w = torch.randn(out[0].shape, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    g = torch.autograd.grad((out[0] * w).sum() + out[1].sum(), leaves[:2])
    torch.cuda.synchronize()
    return [o.detach() for o in out], list(g)


def test_results_do_not_depend_on_the_launch_shapes(gpu):
    sim, st = _case()
    try:
        autotune.set_config({})
        ref = _run(sim, st)
        for cfg in ({"brick_blocks": 257, "node_blocks": 33, "sort_blocks": 100},
                    {"brick_blocks": 4096, "node_blocks": 4096, "sort_blocks": 4096}):
            autotune.set_config(cfg)
            assert autotune.get_config() == cfg
            got = _run(sim, st)
            for a, b in zip(ref[0] + ref[1], got[0] + got[1]):
                assert torch.equal(a, b), cfg
    finally:
        autotune.set_config({})
    assert autotune.get_config() == autotune.DEFAULT


def test_tune_returns_a_candidate_and_leaves_it_set(gpu):
    sim, st = _case()
    cands = {"brick_blocks": [512, 2048], "node_blocks": [256, 1024], "sort_blocks": [2048]}
    try:
        cfg, ms = autotune.tune(sim, st, n_substeps=2, reps=1, save=False, candidates=cands)
        assert all(cfg[k] in cands[k] for k in cands) and ms > 0
        assert autotune.get_config() == cfg
    finally:
        autotune.set_config({})
