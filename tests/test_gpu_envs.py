"""The batched task-protocol shim (envs.PadPressEnvs): a population stepped
with per-environment actions reproduces each environment stepped alone, and
one environment stepped through the shim reproduces the oracle."""
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


def test_shim_steps_match_the_oracle(gpu):
    """The shim's protocol against the FP64 oracle: one environment (the
    smallest of the population, 64^3) stepped twice (10 substeps each) under
    two actions equals the oracle's coupled substeps with the pad's control
    velocity plus the action (the reference's d_pos per outer step): object
    positions, pad state, the loss (summed squared vertex displacement from
    the virtual pad) and the observations.

    The displacement field behind the loss is 2e-5 .. 8e-5 against vertex
    positions near 0.6, so float32 positions carry it to about 2e-3 relative
    (the oracle itself run in float32 differs from float64 by 1.2e-3 .. 2.5e-3
    in the loss on these two steps); loss and displacements are held to 1e-2."""
    from oracle import coupled as ocp
    from tests.mpm_cases import oracle_mats, oracle_params, rel_l2
    F64 = torch.float64
    envs = scenes.config4_population(n_env=8, seed=3, pads=True)
    d = envs[int(np.argmin([len(e["x"]) for e in envs]))]
    T = 10
    acts = [np.array([0.5, -2.0, 0.3], np.float32), np.array([-1.0, 1.0, 0.5], np.float32)]
    pop = PadPressEnvs([d], CONTACT, sub_steps=T)
    res = []
    for a in acts:
        r = pop.step(a)
        res.append(r + ([t.cpu() for t in pop.pads[0]], pop.states[0][0].cpu()))
    n = len(d["x"])
    nodes, f2v, tris, layer = d["pad_mesh"]
    op = oracle_params(d["prm"])
    mats = oracle_mats(d["mats"], d["mat_id"], F64)
    state = (torch.tensor(d["x"], dtype=F64), torch.tensor(d["v"], dtype=F64),
             torch.zeros(n, 3, 3, dtype=F64), torch.eye(3, dtype=F64).repeat(n, 1, 1))
    B = pop.batch.pad.B.detach().cpu().to(F64)
    base = torch.tensor(d["pad_cv"], dtype=F64)
    pp = torch.tensor(nodes, dtype=F64)
    ps = (pp, base.clone(), pp.clone())
    for a, (obs, rew, done, bad, pads, x) in zip(acts, res):
        cv = base + torch.tensor(a, dtype=F64)
        for _ in range(T):
            pd = dict(pos=ps[0], vel=ps[1], vpos=ps[2], cv=cv, cv_next=cv,
                      f2v=torch.tensor(f2v).long(), B=B, layer=torch.tensor(layer), tris=tris,
                      out_dir=torch.tensor(d["pad_out"], dtype=F64), mu=d["pad"]["mu"],
                      lam=d["pad"]["lam"], mass=d["pad"]["mass"], dt=d["prm"].dt)
            state, po, _, _ = ocp.coupled_substep(op, state, mats, [pd], tuple(CONTACT.values()),
                                                  clamp=True)
            ps = po[0]
        disp = ps[0] - ps[2]
        loss = float((disp ** 2).sum())
        assert not bad.any() and not done.any()
        assert abs(-float(rew[0]) - loss) <= 1e-2 * loss, (float(rew[0]), loss)
        assert rel_l2(pads[0] - pads[2], disp) < 1e-2
        assert rel_l2(x, state[0]) < 1e-5
        assert rel_l2(pads[0], ps[0]) < 1e-5 and rel_l2(pads[2], ps[2]) < 1e-5
        assert rel_l2(pads[1], ps[1]) < 1e-3
        com, dmean, vmean = obs[0][:3], obs[0][3:6], obs[0][6:]
        assert np.allclose(com, state[0].mean(0).numpy(), rtol=1e-5, atol=0)
        dm = disp.mean(0).numpy()
        assert np.all(np.abs(dmean - dm) <= 1e-2 * disp.abs().mean().item()), (dmean, dm)
        assert np.allclose(vmean, ps[1].mean(0).numpy(), rtol=1e-3, atol=1e-6)
