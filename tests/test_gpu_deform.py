"""GPU parity of the deformation losses (mass-only scatter, target distance
field, density / SDF losses, IoU) against the oracle.

Tolerances: grid mass and losses within 1e-6 relative of FP64 on the same
float32 inputs (float32 sums); the distance field bit-exact against the
oracle's float32 sweep (same neighbour order and rounding); position
gradients within 1e-4 relative L2 of the oracle's FP64 reverse mode.
"""
import numpy as np
import pytest
import torch

from difftactile_amd import _lib
from difftactile_amd.deform_loss import DeformLoss
from difftactile_amd.params import MPMParams
from oracle import deform as od
from oracle import mpm as om
from tests.mpm_cases import rel_l2

pytestmark = pytest.mark.gpu
D64 = torch.float64


def _scene(n=24, seed=0):
    prm = MPMParams(n_grid=n, rho=1000.0)
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.3, 0.6, size=(4000, 3)).astype(np.float32)
    # target: a ball of density, offset from the particles
    ii = np.arange(n)
    I = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3) * prm.dx
    t = np.where(np.linalg.norm(I - [0.5, 0.42, 0.47], axis=1) < 0.17,
                 rng.uniform(0.5, 1.5, len(I)) * 2e-3, 0.0).astype(np.float32)
    return prm, x, t


def _oprm(prm):
    return om.MPMParams(n_grid=prm.n_grid, rho=prm.rho)


def test_grid_mass_and_its_adjoint_match_the_oracle(gpu):
    prm, x, _ = _scene()
    dl = DeformLoss(prm)
    m = dl.compute_grid_mass(torch.tensor(x).cuda()).cpu().double()
    xo = torch.tensor(x, dtype=D64, requires_grad=True)
    mo = od.grid_mass(_oprm(prm), xo)
    assert rel_l2(m, mo.detach()) < 1e-6
    w = torch.tensor(np.random.default_rng(1).normal(size=m.numel()), dtype=torch.float32)
    gx = torch.zeros(len(x), 3, device="cuda")
    _lib.call("dt_mpm_grid_mass_adj", prm.c_struct(), len(x), torch.tensor(x).cuda(), w.cuda(),
              gx, _lib.stream())
    want = torch.autograd.grad((mo * w.double()).sum(), xo)[0]
    assert rel_l2(gx.cpu().double(), want) < 1e-5


@pytest.mark.parametrize("sweeps", [7, 48])
def test_target_distance_field_is_bit_exact(gpu, sweeps):
    prm, _, t = _scene()
    dl = DeformLoss(prm)
    dl.target_density.copy_(torch.tensor(t).cuda())
    dl.update_target(sweeps)
    want_sdf, want_pt = od.target_sdf_np(t, prm.n_grid, prm.dx, sweeps)
    assert np.array_equal(dl.target_sdf.cpu().numpy(), want_sdf)
    reached = want_sdf < od.SDF_INF
    assert np.array_equal(dl.nearest_point.cpu().numpy()[reached], want_pt[reached])
    assert reached.sum() > 1000


def test_losses_iou_and_position_gradient_match_the_oracle(gpu):
    prm, x, t = _scene(seed=2)
    dl = DeformLoss(prm)
    dl.initialize(t)
    assert abs(dl.target_iou - od.iou(t, t)) < 1e-5
    dl.set_weights(sdf=0.7, density=1.3)
    xg = torch.tensor(x).cuda().requires_grad_(True)
    loss = dl.loss(xg)
    gx = torch.autograd.grad(loss, xg)[0].cpu().double()
    sdf = dl.target_sdf.cpu().double()
    tt = torch.tensor(t, dtype=D64)
    xo = torch.tensor(x, dtype=D64, requires_grad=True)
    mo = od.grid_mass(_oprm(prm), xo)
    dens, sl = od.deform_losses(mo, tt, sdf)
    lo = 1.3 * dens + 0.7 * sl
    assert abs(loss.item() - lo.item()) <= 1e-6 * abs(lo.item())
    want = torch.autograd.grad(lo, xo)[0]
    assert rel_l2(gx, want) < 1e-4
    m = dl.grid_mass.cpu().numpy()
    assert abs(dl.iou() - od.iou(m, t)) < 1e-5


def test_kabsch_angle_and_its_adjoint_match_the_oracle(gpu):
    from difftactile_amd.losses import kabsch_angle
    rng = np.random.default_rng(9)
    n, dt = 5000, 1e-3
    x = rng.uniform(0.3, 0.7, size=(n, 3))
    th = np.radians(1.5)
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    y = (x - 0.5) @ Rz.T + 0.5 + 1e-4 * rng.normal(size=x.shape)
    v = (y - x) / dt
    mask = rng.random(n) < 0.6
    x32, v32 = x.astype(np.float32), v.astype(np.float32)
    xg = torch.tensor(x32).cuda().requires_grad_(True)
    vg = torch.tensor(v32).cuda().requires_grad_(True)
    a = kabsch_angle(xg, vg, torch.tensor(mask).cuda(), dt)
    gx, gv = torch.autograd.grad(a, [xg, vg])
    xo = torch.tensor(x32, dtype=D64, requires_grad=True)
    vo = torch.tensor(v32, dtype=D64, requires_grad=True)
    ao = od.kabsch_angle(xo, vo, mask, dt)
    gxo, gvo = torch.autograd.grad(ao, [xo, vo])
    # float32 sums of 3000 centred products: the angle to 2e-5 rad
    assert abs(a.item() - ao.item()) < 2e-5
    assert rel_l2(gx.cpu().double(), gxo) < 1e-3
    assert rel_l2(gv.cpu().double(), gvo) < 1e-3
