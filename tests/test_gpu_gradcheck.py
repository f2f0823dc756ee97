"""GPU adjoint vs FP64 central differences of the oracle on the elastic-cube workload.

Elastic cube: 8000 particles (20^3 lattice at dx/2, +-0.25 dx jitter), 64^3
grid, fixed corotated E=5e4 nu=0.3 rho=1, gravity -9.8, sticky floor, 100
substeps; loss = w . (final centre of mass); gradient w.r.t. the initial
velocities (computed as w . mean displacement, which differs from w . COM by a
constant).  With rho = 1 the reference's 1e-5 mass regulariser damps grid
velocities by m / (m + 1e-5) ~ 0.75 per substep, so the gradient is small
(~1e-9) but well resolved.  The specified dt = 1e-4 violates the CFL bound
(dt*sqrt(E/rho)/dx = 1.43; the FP64 oracle itself goes unstable), so the check
runs at dt = 1e-5.  Tolerance: 1e-3 relative (BASELINE target) for directional
derivatives against FD with eps = 1e-4, and for the full gradient against the
oracle's FP64 reverse mode.
"""
import numpy as np
import pytest
import torch

from difftactile_amd import scenes
from difftactile_amd.sim import MPMSim
from oracle import mpm as om
from tests.mpm_cases import oracle_mats, oracle_params, rel_l2

pytestmark = pytest.mark.gpu
F64 = torch.float64


def test_elastic_cube_com_gradient_vs_fp64_fd(gpu):
    prm, mats, x, v, mid = scenes.config1_elastic_cube(seed=0)
    prm.dt = 1e-5
    n = len(x)
    steps = 100
    w = torch.tensor([0.3, 1.0, -0.5])
    sim = MPMSim(prm, mats, n, mat_id=mid, checkpoint_every=25, sort=True)
    st = [torch.tensor(x).cuda(), torch.tensor(v).cuda().requires_grad_(True),
          torch.zeros(n, 3, 3).cuda(), torch.eye(3).repeat(n, 1, 1).cuda()]
    out = sim.rollout(*st, steps)
    # w . mean displacement: the same gradient as w . COM, without the O(1)
    # constant that would cancel FD digits
    loss = ((out[0] - st[0]).mean(0) * w.cuda()).sum()
    g = torch.autograd.grad(loss, st[1])[0].double().cpu()

    op = oracle_params(prm)
    omats = oracle_mats(mats, mid, F64)

    def L(v0):
        s = (torch.tensor(x, dtype=F64), v0, torch.zeros(n, 3, 3, dtype=F64),
             torch.eye(3, dtype=F64).repeat(n, 1, 1))
        o = om.rollout(op, s, omats, steps)
        return ((o[0] - s[0]).mean(0) * w.double()).sum()

    v0 = torch.tensor(v, dtype=F64, requires_grad=True)
    g_ref = torch.autograd.grad(L(v0), v0)[0]
    assert rel_l2(g, g_ref) < 1e-3
    rng = np.random.default_rng(1)
    eps = 1e-4
    with torch.no_grad():
        for _ in range(2):
            d = torch.tensor(rng.normal(size=(n, 3)), dtype=F64)
            d /= d.norm()
            fd = (L(v0.detach() + eps * d) - L(v0.detach() - eps * d)) / (2 * eps)
            gd = (g * d).sum()
            assert abs(gd - fd) <= 1e-3 * abs(fd), (gd.item(), fd.item())
