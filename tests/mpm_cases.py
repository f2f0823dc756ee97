"""Seeded MPM inputs shared by the oracle and GPU parity tests."""
import numpy as np
import torch

from difftactile_amd.params import DRUCKER_PRAGER, ELASTIC, VON_MISES, Material, MPMParams
from oracle import mpm as om


def blob(n, n_grid=32, seed=0, spread=0.18, vel=0.3, cscale=2.0, fpert=0.05, centre=0.5):
    """n particles in a ball, random velocity, small random C, F = I + noise."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = spread * rng.random(n) ** (1 / 3)
    x = (centre + d * r[:, None]).astype(np.float32)
    v = (vel * rng.normal(size=(n, 3))).astype(np.float32)
    C = (cscale * rng.normal(size=(n, 3, 3))).astype(np.float32)
    F = (np.eye(3) + fpert * rng.normal(size=(n, 3, 3))).astype(np.float32)
    return x, v, C, F


def materials(kind):
    if kind == "elastic":
        return [Material(5e4, 0.3, ELASTIC)]
    if kind == "von_mises":
        return [Material(5e4, 0.3, VON_MISES, yield_stress=50.0)]
    if kind == "drucker_prager":
        return [Material(5e4, 0.3, DRUCKER_PRAGER, friction_deg=30.0)]
    if kind == "multi":
        return [Material(1e5, 0.3, ELASTIC), Material(5e4, 0.3, DRUCKER_PRAGER),
                Material(3e4, 0.35, VON_MISES, yield_stress=80.0)]
    raise ValueError(kind)


def oracle_params(prm):
    return om.MPMParams(n_grid=prm.n_grid, space_scale=prm.space_scale, obj_scale=prm.obj_scale,
                        rho=prm.rho, dt=prm.dt, gravity=prm.gravity, bound=prm.bound,
                        eps=prm.eps, sticky_floor=prm.sticky_floor,
                        svd_exact=getattr(prm, "svd_exact", False))


def oracle_mats(mats, mat_id, dtype):
    oms = [om.Material(m.E, m.nu, m.kind, m.yield_stress) for m in mats]
    for o, m in zip(oms, mats):
        o.dp_alpha = m.dp_alpha
    return om.material_arrays(oms, mat_id, dtype)


def oracle_mats_leaf(mats, mat_id, dtype):
    """oracle_mats with the per-material (mu, lam) as leaf tensors, so reverse
    mode yields d/d mu and d/d lam of each material row."""
    mu_t = torch.tensor([m.mu for m in mats], dtype=dtype, requires_grad=True)
    lam_t = torch.tensor([m.lam for m in mats], dtype=dtype, requires_grad=True)
    _, _, kind, yld, alpha = oracle_mats(mats, mat_id, dtype)
    idx = torch.as_tensor(mat_id, dtype=torch.int64)
    return (mu_t[idx], lam_t[idx], kind, yld, alpha), mu_t, lam_t


def rel_l2(a, b):
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    den = torch.linalg.norm(b).item()
    return torch.linalg.norm(a - b).item() / max(den, 1e-30)


def default_prm(n_grid=32, dt=2e-5, rho=1000.0, sticky=False):
    return MPMParams(n_grid=n_grid, rho=rho, dt=dt, sticky_floor=sticky)
