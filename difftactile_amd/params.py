"""Host-side parameters of the MPM object and its material table.

Follows the constructor arithmetic of ``difftactile/object_model/mpm_elastic.py:19-66``:
``dx = space_scale / n_grid``, ``p_vol = (dx * obj_scale)**3``,
``p_mass = p_vol * rho * obj_scale``; Lame parameters from (E, nu) as at
``mpm_elastic.py:70-71``.
"""
import math

import numpy as np
import torch

from . import _lib

ELASTIC = 0
VON_MISES = 1
DRUCKER_PRAGER = 2


def lame(E, nu):
    return E / 2 / (1 + nu), E * nu / (1 + nu) / (1 - 2 * nu)


class MPMParams:
    """Scalar parameters of one MPM object (grid, time step, body force, walls)."""

    def __init__(self, n_grid=64, space_scale=1.0, obj_scale=1.0, rho=1.0, dt=5e-5,
                 gravity=(0.0, -9.8, 0.0), bound=3, eps=1e-5, sticky_floor=False,
                 clamp_grad=False, grad_clamp=1000.0, svd_exact=True, n_env=1):
        self.n_grid = int(n_grid)
        self.space_scale = float(space_scale)
        self.obj_scale = float(obj_scale)
        self.rho = float(rho)
        self.dx = self.space_scale / self.n_grid
        self.inv_dx = 1.0 / self.dx
        self.p_vol = (self.dx * self.obj_scale) ** 3
        self.p_mass = self.p_vol * self.rho * self.obj_scale
        self.dt = float(dt)
        self.gravity = tuple(float(g) for g in gravity)
        self.bound = int(bound)
        self.eps = float(eps)
        self.sticky_floor = bool(sticky_floor)
        self.clamp_grad = bool(clamp_grad)
        self.grad_clamp = float(grad_clamp)
        # exact SVD adjoint in the fused path (see dt_mpm_params.svd_exact)
        self.svd_exact = bool(svd_exact)
        # environments stacked along x (see dt_mpm_params.n_env): environment e
        # occupies x in [e, e + 1) * space_scale
        self.n_env = max(1, int(n_env))
        # ragged environments (set by set_ragged): device tables of each
        # environment's grid size and (dx, inv_dx, p_mass, p_vol)
        self.env_ngrid = None
        self.env_scale = None

    def set_ragged(self, env_prms, device="cuda"):
        """Stack environments of different grids: environment e keeps the
        grid size, dx, p_mass and p_vol of env_prms[e] inside an n_grid^3
        slot (n_grid must be at least every environment's); dt, gravity,
        walls and eps are this object's."""
        import torch
        if len(env_prms) != self.n_env:
            raise ValueError("one MPMParams per environment")
        if self.n_grid % 4:
            raise ValueError("the slot grid size of a ragged batch must be a multiple of 4")
        for q in env_prms:
            if q.n_grid > self.n_grid or q.n_grid % 4:
                raise ValueError("environment grids must be multiples of 4 within the slot")
            if (q.dt, q.gravity, q.bound, q.eps, q.sticky_floor) != (
                    self.dt, self.gravity, self.bound, self.eps, self.sticky_floor):
                raise ValueError("ragged environments share dt, gravity, walls and eps")
        self.env_ngrid = torch.tensor([q.n_grid for q in env_prms], dtype=torch.int32,
                                      device=device)
        self.env_scale = torch.tensor([[q.dx, q.inv_dx, q.p_mass, q.p_vol] for q in env_prms],
                                      dtype=torch.float32, device=device).contiguous()

    @property
    def n_nodes(self):
        return self.n_env * self.n_grid ** 3

    def c_struct(self, soa=False):
        p = _lib.DtMpmParams()
        p.n_grid = self.n_grid
        p.bound = self.bound
        p.sticky_floor = int(self.sticky_floor)
        p.clamp_grad = int(self.clamp_grad)
        p.dx = self.dx
        p.inv_dx = self.inv_dx
        p.dt = self.dt
        p.p_vol = self.p_vol
        p.p_mass = self.p_mass
        p.eps = self.eps
        for i in range(3):
            p.gravity[i] = self.gravity[i]
        p.grad_clamp = self.grad_clamp
        p.soa = int(soa)
        p.svd_exact = int(self.svd_exact)
        p.n_env = self.n_env
        if self.env_ngrid is not None:
            p.env_ngrid = self.env_ngrid.data_ptr()
            p.env_scale = self.env_scale.data_ptr()
        return p

    def key_bits(self):
        nb = (self.n_grid + 3) // 4
        nbx = (self.n_env * self.n_grid + 3) // 4
        return max(1, int(math.ceil(math.log2(nbx * nb * nb * 64))))


class Material:
    """One row of the material table (``multi_obj.py:65-74`` generalised to kinds)."""

    def __init__(self, E, nu, kind=ELASTIC, yield_stress=1000.0, friction_deg=30.0):
        self.E, self.nu = float(E), float(nu)
        self.mu, self.lam = lame(self.E, self.nu)
        self.kind = int(kind)
        self.yield_stress = float(yield_stress)
        s = math.sin(math.radians(friction_deg))
        self.dp_alpha = math.sqrt(2.0 / 3.0) * 2.0 * s / (3.0 - s)


def material_table(materials, device):
    """Device copy of the material table in the dt_material layout (32 B/row)."""
    rows = np.zeros((len(materials), 8), dtype=np.float32)
    ints = rows.view(np.int32)
    for i, m in enumerate(materials):
        rows[i, 0] = m.mu
        rows[i, 1] = m.lam
        rows[i, 2] = m.yield_stress
        rows[i, 3] = m.dp_alpha
        ints[i, 4] = m.kind
    return torch.from_numpy(rows).to(device)
