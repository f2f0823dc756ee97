"""Shared implementation of the MPM object facades.

Mirrors the field layout and per-kernel methods of
``difftactile/object_model/mpm_elastic.py`` (class ``MPMObj``): frame-windowed
fields ``[sub_steps, ...]`` with ``.grad`` twins, kernels ``compute_new_F``,
``svd``, ``svd_grad``, ``p2g``, ``check_grid_occupy``, ``grid_op``, ``g2p``
(each with ``.grad`` where the reference differentiates them) and the
window/cache helpers.  Every kernel runs in libdifftactile_hip.so.
"""
import numpy as np
import torch

from .. import _lib
from ..field import Field, Kernel
from ..params import MPMParams, lame


def _rotvec_deg_to_matrix(ori):
    """Rotation matrix of a rotation vector given in degrees (scipy's from_rotvec)."""
    from scipy.spatial.transform import Rotation as R
    return R.from_rotvec(np.deg2rad([ori[0], ori[1], ori[2]])).as_matrix()


class MPMBase:
    KIND = 0  # material kind of the single-material variants

    def __init__(self, dt=5e-5, sub_steps=80, obj_name=None, space_scale=1.0, obj_scale=1.0,
                 density=2, rho=1, *, particles=None, n_grid=64, bound=3, gravity=(0.0, -9.8, 0.0),
                 sticky_floor=False, device="cuda"):
        if obj_name is not None and particles is None:
            raise ValueError("mesh loading is not part of this package; pass particles=(N,3) array")
        self.sub_steps = int(sub_steps)
        self.dt = float(dt)
        self.dim = 3
        self.bound = int(bound)
        self.n_grid = int(n_grid)
        self.space_scale = space_scale
        self.obj_scale = obj_scale
        self.rho = rho * obj_scale
        self.particle_density = self.n_grid * density * obj_scale / space_scale
        self.gravity = tuple(float(g) for g in gravity)
        self.device = torch.device(device)
        self.prm = MPMParams(n_grid=n_grid, space_scale=space_scale, obj_scale=obj_scale, rho=rho,
                             dt=dt, gravity=gravity, bound=bound, sticky_floor=sticky_floor)
        self.dx_0 = self.prm.dx
        self.inv_dx_0 = self.prm.inv_dx
        self.p_vol = self.prm.p_vol
        self.p_rho = self.rho * 1.0
        self.p_mass = self.prm.p_mass
        self.eps = self.prm.eps

        pts = np.asarray(particles, dtype=np.float32) * np.float32(obj_scale)
        self.n_particles = len(pts)
        self.particles = Field(self.n_particles, (3,), device=self.device)
        self.particles.from_numpy(pts)

        S, N, n = self.sub_steps, self.n_particles, self.n_grid
        d = self.device
        self.ball_pos = Field((), (3,), device=d)
        self.ball_ori = Field((), (3,), device=d)
        self.ball_vel = Field((), (3,), device=d)
        self.rot_h = Field((), (3, 3), device=d)
        self.trans_h = Field((), (4, 4), device=d)

        self.x_0 = Field((S, N), (3,), device=d, needs_grad=True)
        self.v_0 = Field((S, N), (3,), device=d, needs_grad=True)
        self.C_0 = Field((S, N), (3, 3), device=d, needs_grad=True)
        self.F_new = Field((S, N), (3, 3), device=d, needs_grad=True)
        self.F_0 = Field((S, N), (3, 3), device=d, needs_grad=True)
        self.U_svd = Field((S, N), (3, 3), device=d, needs_grad=True)
        self.V_svd = Field((S, N), (3, 3), device=d, needs_grad=True)
        self.S_svd = Field((S, N), (3, 3), device=d, needs_grad=True)
        self.grid_v_in = Field((S, n, n, n), (3,), device=d, needs_grad=True)
        self.grid_v_out = Field((S, n, n, n), (3,), device=d, needs_grad=True)
        self.grid_m = Field((S, n, n, n), device=d, needs_grad=True)
        self.grid_f = Field((S, n, n, n), (3,), device=d, needs_grad=True)
        self.grid_occupy = Field((S, n, n, n), dtype=torch.int32, device=d)
        self.surf_f = Field(S, (3,), device=d, needs_grad=True)
        self.cache = dict()
        self._init_materials()

        self.compute_new_F = Kernel(self._compute_new_F, self._compute_new_F_grad)
        self.svd = Kernel(self._svd, self._svd_grad_unsupported)
        self.p2g = Kernel(self._p2g, self._p2g_grad)
        self.check_grid_occupy = Kernel(self._check_grid_occupy, lambda f: None)
        self.grid_op = Kernel(self._grid_op, self._grid_op_grad)
        self.g2p = Kernel(self._g2p, self._g2p_grad)

    # -- materials -------------------------------------------------------
    def _init_materials(self):
        d = self.device
        self.E_0 = Field((), device=d, needs_grad=True)
        self.nu_0 = Field((), device=d, needs_grad=True)
        self.E_0[None], self.nu_0[None] = 4e3 * self.space_scale, 0.4
        self.mu_0 = Field((), device=d, needs_grad=True)
        self.lambda_0 = Field((), device=d, needs_grad=True)
        mu, lam = lame(4e3 * self.space_scale, 0.4)
        self.mu_0[None] = mu
        self.lambda_0[None] = lam
        self.mat_id = None

    def set_material(self, E, nu):
        mu, lam = lame(E, nu)
        self.E_0[None], self.nu_0[None] = E, nu
        self.mu_0[None], self.lambda_0[None] = mu, lam

    def _mat_table(self):
        """dt_material rows built on the device from the parameter fields."""
        t = torch.zeros((1, 8), dtype=torch.float32, device=self.device)
        t[0, 0] = self.mu_0.data
        t[0, 1] = self.lambda_0.data
        t.view(torch.int32)[0, 4] = self.KIND
        return t

    # -- set-up ----------------------------------------------------------
    def init(self, position, orientation, velocity):
        self.set_object_params(position, orientation, velocity)
        self.init_object()

    def set_object_params(self, position, orientation, velocity):
        self.ball_pos[None] = position
        self.ball_ori[None] = orientation
        self.ball_vel[None] = velocity
        rot = _rotvec_deg_to_matrix(orientation)
        trans = np.eye(4)
        trans[0:3, 0:3] = rot
        trans[0:3, 3] = position
        self.rot_h[None] = rot
        self.trans_h[None] = trans

    def init_object(self):
        """``init_object`` (``mpm_elastic.py:107-114``): x0 = T p, v0 = ball_vel, F0 = I."""
        T = np.ascontiguousarray(self.trans_h.to_numpy(), np.float32).reshape(-1)
        bv = np.ascontiguousarray(self.ball_vel.to_numpy(), np.float32)
        fp = lambda a: a.ctypes.data_as(_lib.ctypes.POINTER(_lib.ctypes.c_float))
        _lib.call("dt_mpm_init_object", self.n_particles, self.particles.data, fp(T), fp(bv),
                  self.x_0.data[0], self.v_0.data[0], self.F_0.data[0], _lib.stream())

    def set_state(self, x, v, C=None, F=None, f=0):
        """Write frame f from arrays (particles-from-array entry the configs need)."""
        self.x_0.data[f] = torch.as_tensor(np.asarray(x), dtype=torch.float32, device=self.device)
        self.v_0.data[f] = torch.as_tensor(np.asarray(v), dtype=torch.float32, device=self.device)
        self.C_0.data[f] = 0.0 if C is None else torch.as_tensor(np.asarray(C), dtype=torch.float32,
                                                                  device=self.device)
        self.F_0.data[f] = torch.eye(3, device=self.device) if F is None else torch.as_tensor(
            np.asarray(F), dtype=torch.float32, device=self.device)

    def reset(self):
        """``reset`` (``mpm_elastic.py:116-123``)."""
        for fld in (self.grid_v_in, self.grid_v_out, self.grid_m, self.grid_f, self.grid_occupy,
                    self.surf_f):
            fld.fill(0)

    # -- kernels ---------------------------------------------------------
    def _n(self):
        return self.n_particles

    def _p(self, fld, f, dtype=torch.float32):
        return _lib.ptr(fld.data[f], dtype)

    def _pg(self, fld, f):
        return _lib.ptr(fld.grad.data[f])

    def _mid(self):
        return None if self.mat_id is None else _lib.ptr(self.mat_id.data, torch.int32)

    def _compute_new_F(self, f):
        _lib.call("dt_mpm_compute_new_F", self._n(), self._p(self.C_0, f), self._p(self.F_0, f),
                  self._p(self.F_new, f), self.dt, _lib.stream())

    def _compute_new_F_grad(self, f):
        _lib.call("dt_mpm_compute_new_F_adj", self._n(), self._p(self.C_0, f),
                  self._p(self.F_0, f), self._pg(self.F_new, f), self._pg(self.C_0, f),
                  self._pg(self.F_0, f), self.dt, _lib.stream())

    def _svd(self, f):
        _lib.call("dt_mpm_svd", self._n(), self._p(self.F_new, f), self._p(self.U_svd, f),
                  self._p(self.S_svd, f), self._p(self.V_svd, f), _lib.stream())

    def _svd_grad_unsupported(self, f):
        raise RuntimeError("svd.grad is not used by the reference; call svd_grad(f) "
                           "(mpm_elastic.py:144-147, grasp_plastic.py:177)")

    def svd_grad(self, f):
        """``svd_grad`` (``mpm_elastic.py:144-173``): F_new.grad += svd adjoint."""
        _lib.call("dt_mpm_svd_adj", self._n(), self._p(self.U_svd, f), self._p(self.S_svd, f),
                  self._p(self.V_svd, f), self._pg(self.U_svd, f), self._pg(self.S_svd, f),
                  self._pg(self.V_svd, f), self._pg(self.F_new, f), _lib.stream())

    def _p2g(self, f):
        mats = self._mat_table()
        _lib.call("dt_mpm_p2g", self.prm.c_struct(), _lib.ptr(mats), mats.shape[0], self._n(),
                  self._mid(), self._p(self.x_0, f), self._p(self.v_0, f), self._p(self.C_0, f),
                  self._p(self.F_new, f), self._p(self.U_svd, f), self._p(self.S_svd, f),
                  self._p(self.V_svd, f), self._p(self.grid_v_in, f), self._p(self.grid_m, f),
                  self._p(self.F_0, f + 1), _lib.stream())

    def _p2g_grad(self, f):
        mats = self._mat_table()
        # material adjoints of this launch, added into mu_0.grad / lambda_0.grad
        gm = torch.zeros((mats.shape[0], 2), dtype=torch.float32, device=self.device)
        _lib.call("dt_mpm_p2g_adj", self.prm.c_struct(), _lib.ptr(mats), mats.shape[0],
                  self._n(), self._mid(), self._p(self.x_0, f), self._p(self.v_0, f),
                  self._p(self.C_0, f), self._p(self.F_new, f), self._p(self.U_svd, f),
                  self._p(self.S_svd, f), self._p(self.V_svd, f), self._pg(self.grid_v_in, f),
                  self._pg(self.grid_m, f), self._pg(self.F_0, f + 1), self._pg(self.x_0, f),
                  self._pg(self.v_0, f), self._pg(self.C_0, f), self._pg(self.F_new, f),
                  self._pg(self.U_svd, f), self._pg(self.S_svd, f), self._pg(self.V_svd, f),
                  _lib.ptr(gm), _lib.stream())
        self.mu_0.grad.data += gm[:, 0].reshape(self.mu_0.grad.data.shape)
        self.lambda_0.grad.data += gm[:, 1].reshape(self.lambda_0.grad.data.shape)

    def _check_grid_occupy(self, f):
        _lib.call("dt_mpm_grid_occupy", self.prm.c_struct(), self._p(self.grid_m, f),
                  self._p(self.grid_occupy, f, torch.int32), _lib.stream())

    def _grid_op(self, f):
        _lib.call("dt_mpm_grid_op", self.prm.c_struct(), self._p(self.grid_v_in, f),
                  self._p(self.grid_m, f), self._p(self.grid_f, f),
                  self._p(self.grid_occupy, f, torch.int32), self._p(self.grid_v_out, f),
                  _lib.stream())

    def _grid_op_grad(self, f, clamp=False):
        prm = self.prm.c_struct()
        prm.clamp_grad = int(clamp)
        _lib.call("dt_mpm_grid_op_adj", prm, self._p(self.grid_v_in, f), self._p(self.grid_m, f),
                  self._p(self.grid_f, f), self._p(self.grid_occupy, f, torch.int32),
                  self._pg(self.grid_v_out, f), self._pg(self.grid_v_in, f),
                  self._pg(self.grid_m, f), self._pg(self.grid_f, f), _lib.stream())

    def _g2p(self, f):
        _lib.call("dt_mpm_g2p", self.prm.c_struct(), self._n(), self._p(self.x_0, f),
                  self._p(self.grid_v_out, f), self._p(self.x_0, f + 1), self._p(self.v_0, f + 1),
                  self._p(self.C_0, f + 1), _lib.stream())

    def _g2p_grad(self, f):
        _lib.call("dt_mpm_g2p_adj", self.prm.c_struct(), self._n(), self._p(self.x_0, f),
                  self._p(self.grid_v_out, f), self._pg(self.x_0, f + 1),
                  self._pg(self.v_0, f + 1), self._pg(self.C_0, f + 1), self._pg(self.x_0, f),
                  self._pg(self.grid_v_out, f), _lib.stream())

    # -- composition -----------------------------------------------------
    def substep(self, f):
        """Object-only substep in the reference order (``grasp_plastic.py:151-164``)."""
        self.compute_new_F(f)
        self.svd(f)
        self.p2g(f)
        self.check_grid_occupy(f)
        self.grid_op(f)
        self.g2p(f)

    def substep_grad(self, f):
        """Adjoint of ``substep`` in the reference order (``grasp_plastic.py:167-178``)."""
        self.g2p.grad(f)
        self.grid_op.grad(f)
        self.p2g.grad(f)
        self.svd_grad(f)
        self.compute_new_F.grad(f)

    def get_external_force(self, f):
        """``surf_f[f] += sum(grid_f[f]) / dt`` (``mpm_elastic.py:125-128``)."""
        self.surf_f.data[f] += self.grid_f.data[f].reshape(-1, 3).sum(0) / self.dt

    # -- window / cache helpers (mpm_elastic.py:262-344) ------------------
    def copy_frame(self, source, target):
        for fld in (self.x_0, self.v_0, self.C_0, self.F_0):
            fld.data[target] = fld.data[source]

    def copy_grad(self, source, target):
        for fld in (self.x_0, self.v_0, self.C_0, self.F_0):
            fld.grad.data[target] = fld.grad.data[source]

    def memory_to_cache(self, t):
        name = f"{t:06d}"
        self.cache[name] = {k: getattr(self, k).data[0].detach().to("cpu", copy=True)
                            for k in ("x_0", "v_0", "C_0", "F_0")}
        self.copy_frame(self.sub_steps - 1, 0)

    def memory_from_cache(self, t):
        name = f"{t:06d}"
        self.copy_frame(0, self.sub_steps - 1)
        self.copy_grad(0, self.sub_steps - 1)
        self.clear_step_grad(self.sub_steps - 1)
        for k, v in self.cache[name].items():
            getattr(self, k).data[0].copy_(v)

    def clear_loss_grad(self):
        for fld in (self.E_0, self.nu_0, self.mu_0, self.lambda_0):
            fld.grad.fill(0.0)

    def clear_step_grad(self, f):
        for fld in (self.grid_v_in, self.grid_v_out, self.grid_m, self.grid_f, self.F_new,
                    self.U_svd, self.V_svd, self.S_svd):
            fld.grad.fill(0.0)
        for fld in (self.x_0, self.v_0, self.C_0, self.F_0):
            fld.grad.data[:f] = 0.0
