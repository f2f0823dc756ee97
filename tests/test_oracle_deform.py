"""Pins the deformation-loss oracle (CPU)."""
import numpy as np
import torch

from oracle import deform as od
from oracle import mpm as om

D64 = torch.float64


def _prm(n=16):
    return om.MPMParams(n_grid=n, rho=1000.0)


def test_grid_mass_conserves_mass_and_first_moment():
    prm = _prm()
    rng = np.random.default_rng(0)
    x = torch.tensor(rng.uniform(0.2, 0.8, size=(500, 3)), dtype=D64)
    m = od.grid_mass(prm, x)
    n = prm.n_grid
    assert abs(m.sum().item() - 500 * prm.p_mass) < 1e-12 * 500
    ii = torch.arange(n, dtype=D64) * prm.dx
    node = torch.stack(torch.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
    first = (m[:, None] * node).sum(0)
    assert torch.allclose(first, prm.p_mass * x.sum(0), rtol=1e-12, atol=1e-15)


def test_grid_mass_gradient_matches_central_differences():
    prm = _prm()
    rng = np.random.default_rng(1)
    x0 = torch.tensor(rng.uniform(0.3, 0.7, size=(50, 3)), dtype=D64)
    wgt = torch.tensor(rng.normal(size=prm.n_grid ** 3), dtype=D64)
    x = x0.clone().requires_grad_(True)
    g = torch.autograd.grad((od.grid_mass(prm, x) * wgt).sum(), x)[0]
    h = 1e-7
    for p, a in [(0, 0), (7, 1), (33, 2), (49, 0)]:
        e = torch.zeros_like(x0)
        e[p, a] = h
        fd = ((od.grid_mass(prm, x0 + e) * wgt).sum() - (od.grid_mass(prm, x0 - e) * wgt).sum()) / (2 * h)
        assert abs(fd.item() - g[p, a].item()) < 1e-6 * max(1.0, abs(fd.item()))


def test_target_sdf_of_one_point_is_the_euclidean_distance():
    n, dx = 12, 0.1
    dens = np.zeros((n, n, n), np.float32)
    dens[4, 6, 5] = 1.0
    sdf, pt = od.target_sdf_np(dens, n, dx, 2 * n)
    ii = np.arange(n)
    I = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
    want = np.sqrt(((I - [4, 6, 5]) ** 2).sum(1) * dx * dx + 1e-8)
    want[np.all(I == [4, 6, 5], axis=1)] = 0.0
    assert np.allclose(sdf, want, rtol=1e-6, atol=1e-6)
    assert np.allclose(pt, np.float32([4, 6, 5]) * np.float32(dx), atol=1e-7)


def test_target_sdf_of_a_plane_is_the_distance_to_it():
    n, dx = 10, 0.1
    dens = np.zeros((n, n, n), np.float32)
    dens[:, 3, :] = 1.0
    sdf, _ = od.target_sdf_np(dens, n, dx, 2 * n)
    j = np.arange(n)[None, :, None] * np.ones((n, n, n))
    want = np.sqrt((np.abs(j - 3) * dx) ** 2 + 1e-8).reshape(-1)
    want[(j == 3).reshape(-1)] = 0.0
    assert np.allclose(sdf, want, rtol=1e-6, atol=1e-6)


def test_vectorised_sweep_matches_the_scalar_loop_bit_for_bit():
    n, dx = 6, 0.125
    rng = np.random.default_rng(2)
    dens = (rng.random((n, n, n)) < 0.06).astype(np.float32)
    dens[1, 1, 1] = 1.0
    a, pa = od.target_sdf_np(dens, n, dx, 5)
    b, pb = od.target_sdf_scalar(dens, n, dx, 5)
    assert np.array_equal(a, b) and np.array_equal(pa, pb)


def test_losses_and_iou_closed_forms():
    rng = np.random.default_rng(3)
    t = torch.tensor(rng.random(1000), dtype=D64)
    sdf = torch.tensor(rng.random(1000), dtype=D64)
    dl, _ = od.deform_losses(t.clone(), t, sdf)
    assert dl.item() == 0.0
    sdf_on = torch.where(t > 0.5, torch.zeros_like(sdf), sdf)
    m_on = torch.where(t > 0.5, t, torch.zeros_like(t))
    assert od.deform_losses(m_on, t, sdf_on)[1].item() == 0.0
    tb = (t > 0.3).double().numpy()
    assert abs(od.iou(tb, tb) - 1.0) < 1e-12
    assert abs(od.iou(2.5 * tb, tb) - 1.0) < 1e-12     # each side normalised by its max
    a = np.zeros((8, 8, 8))
    b = np.zeros((8, 8, 8))
    a[:4] = 1.0
    b[2:6] = 1.0
    inter, union = 2 * 64, 4 * 64 + 4 * 64 - 2 * 64
    assert abs(od.iou(a.reshape(-1), b.reshape(-1)) - inter / union) < 1e-12


def _turn(theta_deg, n=300, seed=4):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.3, 0.7, size=(n, 3))
    th = np.radians(theta_deg)
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    return x, Rz


def test_kabsch_angle_of_a_rigid_turn():
    dt = 1e-2
    x, Rz = _turn(3.0)
    # every particle masked, turned about the origin and translated
    y = x @ Rz.T + [0.01, -0.02, 0.005]
    v = (y - x) / dt
    a = od.kabsch_angle(torch.tensor(x), torch.tensor(v), np.ones(len(x), bool), dt)
    assert abs(a.item() - np.radians(3.0)) < 1e-10
    # a partial mask and no translation: H = A R^T with A SPD
    mask = np.arange(len(x)) % 3 == 0
    y = x @ Rz.T
    v = (y - x) / dt
    a = od.kabsch_angle(torch.tensor(x), torch.tensor(v), mask, dt)
    assert abs(a.item() - np.radians(3.0)) < 1e-10


def test_kabsch_angle_gradient_matches_central_differences():
    dt = 1e-2
    x, Rz = _turn(2.0, n=60, seed=5)
    rng = np.random.default_rng(6)
    y = x @ Rz.T + 0.003 * rng.normal(size=x.shape)
    v = (y - x) / dt
    mask = rng.random(len(x)) < 0.7
    xt = torch.tensor(x, requires_grad=True)
    vt = torch.tensor(v, requires_grad=True)
    gx, gv = torch.autograd.grad(od.kabsch_angle(xt, vt, mask, dt), [xt, vt])
    h = 1e-6
    for (p, a) in [(0, 0), (5, 1), (17, 2), (40, 0)]:
        for arr, g, is_x in ((x, gx, True), (v, gv, False)):
            e = np.zeros_like(arr)
            e[p, a] = h
            args_p = (torch.tensor(x + e), torch.tensor(v)) if is_x else (torch.tensor(x), torch.tensor(v + e))
            args_m = (torch.tensor(x - e), torch.tensor(v)) if is_x else (torch.tensor(x), torch.tensor(v - e))
            fd = (od.kabsch_angle(*args_p, mask, dt) - od.kabsch_angle(*args_m, mask, dt)) / (2 * h)
            assert abs(fd.item() - g[p, a].item()) < 1e-6 * max(1.0, abs(fd.item())), (p, a, is_x)
