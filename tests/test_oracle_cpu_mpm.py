"""Pins the oracle's C++/OpenMP restatement (oracle/cpu_mpm.cpp, the CPU
baseline's code) against the torch oracle: one substep forward and its
adjoint on the elastic, von Mises and Drucker-Prager materials, including
the material (mu, lam) gradients."""
import numpy as np
import pytest
import torch

from oracle import mpm as om
from oracle.cpu_mpm import CpuMPM
from tests.mpm_cases import (blob, default_prm, materials, oracle_mats_leaf, oracle_params,
                             rel_l2)

F64 = torch.float64


@pytest.mark.parametrize("kind", ["elastic", "von_mises", "drucker_prager", "multi"])
def test_cpu_restatement_matches_the_torch_oracle(kind):
    prm = default_prm(32)
    mats = materials(kind)
    n = 3000
    x, v, C, F = blob(n, seed=5, fpert=0.15, cscale=4.0)
    mid = np.random.default_rng(1).integers(0, len(mats), n).astype(np.int32)
    cpu = CpuMPM(prm, mats, mid)
    (x1, v1, C1, F1), ws = cpu.substep(x, v, C, F)
    op = oracle_params(prm)
    omats, mu_t, lam_t = oracle_mats_leaf(mats, mid, F64)
    st = [torch.tensor(a, dtype=F64, requires_grad=True) for a in (x, v, C, F)]
    ref = om.substep(op, tuple(st), omats)
    for got, want in zip((x1, v1, C1, F1), ref):
        assert rel_l2(got, want.detach()) < 1e-5
    rng = np.random.default_rng(2)
    gn = [rng.normal(size=a.shape).astype(np.float32) for a in (x, v, C, F)]
    (gx, gv, gC, gF), gm = cpu.substep_adj(x, v, C, F, ws, gn)
    loss = sum((r * torch.tensor(g, dtype=F64)).sum() for r, g in zip(ref, gn))
    want = torch.autograd.grad(loss, st + [mu_t, lam_t])
    for name, got, w in zip("xvCF", (gx, gv, gC, gF), want):
        assert rel_l2(got, w) < 1e-4, name
    assert rel_l2(gm[:, 0], want[4]) < 1e-4
    assert rel_l2(gm[:, 1], want[5]) < 1e-4
