"""The batched task-protocol shim (envs.PadPressEnvs): a population stepped
with per-environment actions reproduces each environment stepped alone."""
import numpy as np
import pytest
import torch

from difftactile_amd import scenes
from difftactile_amd.envs import PadPressEnvs

pytestmark = pytest.mark.gpu
CONTACT = dict(kn=1e4, kd=269.44, kt=154.78, mu_f=53.85)


def _envs():
    envs = scenes.config4_population(n_env=8, seed=3, pads=True)
    small = [d for d in envs if d["prm"].n_grid == 64][:3]
    assert len(small) == 3
    return small


def test_population_steps_like_its_environments_alone(gpu):
    envs = _envs()
    pop = PadPressEnvs(envs, CONTACT, sub_steps=8)
    rng = np.random.default_rng(0)
    acts = [rng.normal(0, 0.05, size=(3, 3)).astype(np.float32) for _ in range(3)]
    obs_b, loss_b = [], []
    for a in acts:
        obs, rew, done, bad = pop.step(a)
        assert not bad.any() and np.isfinite(rew).all()
        obs_b.append(obs)
        loss_b.append(-rew)
    assert pop.loss[None].shape == (3,)
    for e in range(3):
        one = PadPressEnvs([envs[e]], CONTACT, sub_steps=8)
        for k, a in enumerate(acts):
            obs, rew, _, _ = one.step(a[e])
            assert np.allclose(obs[0], obs_b[k][e], rtol=1e-4, atol=1e-6), (e, k)
            assert abs(-rew[0] - loss_b[k][e]) <= 1e-4 * max(abs(loss_b[k][e]), 1e-12), (e, k)
    # reset returns to the initial state
    pop.reset()
    obs0, _, _, _ = pop.step(acts[0])
    assert np.allclose(obs0, obs_b[0], rtol=1e-5, atol=1e-7)
