"""Position-based rope on the GPU (``difftactile/object_model/PBD_rope.py``)
and its particle-wise contact with FEM pads
(``difftactile/tasks/cable_straightening.py:287-319``).

``PBDRope`` keeps the reference constructor and ``init`` / ``set_target``;
``step`` advances one substep (``update``, 111-180) in the HIP library and
is differentiable with respect to positions, velocities and the contact
impulse (``update.grad``).  ``rope_contact`` gives the contact impulse on the
rope and the pad's external force for one pad (closest triangle, signed
distance offset by the particle radius, penalty force), differentiable with
respect to the rope state and the pad's positions and velocities
(``collision.grad``).
"""
import math

import numpy as np
import torch

from .. import _lib


class PBDRope:
    def __init__(self, dt=5e-5, sub_steps=80, p_rad=0.05, p_rho=1.0, n_particles=10,
                 table_height=0.0, rest_length=None, device="cuda"):
        self.sub_steps, self.dt = sub_steps, float(dt)
        self.dim = 3
        self.p_rad, self.p_rho = float(p_rad), float(p_rho)
        self.p_vol = math.pi * 4 / 3 * self.p_rad ** 3
        self.p_mass = self.p_vol * self.p_rho
        self.invM = 1.0 / self.p_mass
        self.gravity = (0.0, -9.8, 0.0)
        self.eps = 1e-10
        self.table_height = float(table_height)
        self.n_particles = int(n_particles)
        self.n_segments = self.n_particles - 1
        if rest_length is None:
            self.rest_length = 2 * self.p_rad
            self.bending_length = 4 * self.p_rad
            self.length = self.rest_length * self.n_particles
        else:
            self.rest_length = float(rest_length)
            self.bending_length = 2 * self.rest_length
            self.length = self.rest_length * (self.n_particles - 1) + self.p_rad * 2
        self.stretch_compliance = 1e-5 / self.p_rho
        self.bending_compliance = 1e-5 / self.p_rho
        self.stretch_relaxation = 0.3
        self.bending_relaxation = 0.1
        self.floor_friction = 0.995
        self.device = torch.device(device)
        self.init_pos = None

    def params(self):
        p = _lib.DtRopeParams()
        p.dt, p.p_mass, p.table_height = self.dt, self.p_mass, self.table_height
        p.floor_friction = self.floor_friction
        for i in range(3):
            p.gravity[i] = self.gravity[i]
        p.rest_length, p.bending_length = self.rest_length, self.bending_length
        p.stretch_compliance, p.bending_compliance = self.stretch_compliance, self.bending_compliance
        p.stretch_relaxation, p.bending_relaxation = self.stretch_relaxation, self.bending_relaxation
        return p

    def init(self, pos, ori):
        """Straight rope from pos along ori at rest length (init, 78-88).
        Returns (positions, zero velocities) [n][3]."""
        pos = np.asarray(pos, np.float32)
        ori = np.asarray(ori, np.float32)
        x = pos[None] + ori[None] * np.float32(self.rest_length) * np.arange(
            self.n_particles, dtype=np.float32)[:, None]
        self.init_pos = torch.tensor(x, device=self.device)
        return self.init_pos.clone(), torch.zeros_like(self.init_pos)

    def set_target(self, pos):
        """Target configuration (set_target, 90-99): the rope hanging from pos
        at the angle whose sine is pos_y / length."""
        pos = np.asarray(pos, np.float64)
        whole = self.rest_length * (self.n_particles - 1)
        th = math.asin(pos[1] / whole)
        ori = np.array([0.0, -math.sin(th), math.cos(th)])
        self.target_pos = torch.tensor(
            (pos[None] + ori[None] * self.rest_length * np.arange(self.n_particles)[:, None]
             ).astype(np.float32), device=self.device)
        return self.target_pos

    def step(self, pos, vel, ext_impulse=None, fixed=None):
        """One substep; fixed: particle 0's position (default init_pos[0])."""
        if ext_impulse is None:
            ext_impulse = torch.zeros_like(pos)
        if fixed is None:
            if self.init_pos is None:
                raise RuntimeError("call init() first or pass fixed")
            fixed = self.init_pos[0]
        fx = np.ascontiguousarray(torch.as_tensor(fixed).detach().cpu().numpy(), np.float32)
        return _RopeStep.apply(self, fx, pos, vel, ext_impulse)


def _fp3(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(_lib.ctypes.POINTER(_lib.ctypes.c_float))


class _RopeStep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rope, fixed, pos, vel, ext):
        p, v, e = (t.detach().float().contiguous() for t in (pos, vel, ext))
        n = p.shape[0]
        work = torch.empty(12 * n, device=p.device)
        p1, v1 = torch.empty_like(p), torch.empty_like(v)
        keep, fp = _fp3(fixed)
        _lib.call("dt_rope_step", rope.params(), n, fp, p, v, e, work, p1, v1, _lib.stream())
        ctx.rope, ctx.fixed = rope, fixed
        ctx.save_for_backward(p, v, e)
        return p1, v1

    @staticmethod
    def backward(ctx, gp1, gv1):
        p, v, e = ctx.saved_tensors
        n = p.shape[0]
        gp1 = torch.zeros_like(p) if gp1 is None else gp1.float().contiguous()
        gv1 = torch.zeros_like(p) if gv1 is None else gv1.float().contiguous()
        gp, gv, ge = torch.zeros_like(p), torch.zeros_like(p), torch.zeros_like(p)
        work = torch.empty(12 * n, device=p.device)
        keep, fp = _fp3(ctx.fixed)
        _lib.call("dt_rope_step_adj", ctx.rope.params(), n, fp, p, v, e, gp1, gv1, work, gp, gv, ge,
                  _lib.stream())
        return None, None, gp, gv, ge


def _contact_params(kn, kd, kt, mu_f, offset):
    """The cable task's contact force (cable_straightening.py:211-223): the
    normal velocity is not clipped at zero and the shear norm has no eps."""
    cp = _lib.DtContactParams()
    cp.kn, cp.kd, cp.kt, cp.mu_f = kn, kd, kt, mu_f
    cp.norm_eps, cp.sdf_offset = 0.0, float(offset)
    cp.unclamped_vn = 1
    return cp


def rope_contact(rope, pos, vel, pad_pos, pad_vel, tris, out_dir, kn, kd, kt, mu_f):
    """Contact of the rope particles (pos, vel [n][3]) with one pad (vertex
    pad_pos, pad_vel [V][3], surface triangles tris [T][3] int32 on the GPU,
    outward direction out_dir).  Returns (impulse [n][3] = f dt on the rope,
    pad external force [V][3], contact flags [n] int32); the first two are
    differentiable w.r.t. pos, vel, pad_pos and pad_vel."""
    return _RopeContact.apply(rope, tris, np.asarray(out_dir, np.float32),
                              (kn, kd, kt, mu_f), pos, vel, pad_pos, pad_vel)


class _RopeContact(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rope, tris, od, kc, pos, vel, pad_pos, pad_vel):
        x, v, pp, pv = (t.detach().float().contiguous() for t in (pos, vel, pad_pos, pad_vel))
        n, nt = x.shape[0], tris.shape[0]
        s = _lib.stream()
        cen = torch.empty(nt, 4, device=x.device)
        _lib.call("dt_contact_centroids", nt, tris, pp, cen, s)
        idx = torch.empty(n, dtype=torch.int32, device=x.device)
        _lib.call("dt_rope_closest", n, x, nt, cen, idx, 1, 0, s)
        cp = _contact_params(*kc, rope.p_rad)
        keep, odp = _fp3(od)
        imp = torch.zeros_like(x)
        ext = torch.zeros_like(pp)
        flag = torch.zeros(n, dtype=torch.int32, device=x.device)
        _lib.call("dt_rope_collision", cp, rope.dt, n, x, v, idx, 1, 0, nt, tris, pp, pv, odp,
                  imp, ext, flag, s)
        ctx.rope, ctx.tris, ctx.od, ctx.kc = rope, tris, od, kc
        ctx.save_for_backward(x, v, pp, pv, idx)
        ctx.mark_non_differentiable(flag)
        return imp, ext, flag

    @staticmethod
    def backward(ctx, g_imp, g_ext, g_flag):
        x, v, pp, pv, idx = ctx.saved_tensors
        n, nt = x.shape[0], ctx.tris.shape[0]
        g_imp = torch.zeros_like(x) if g_imp is None else g_imp.float().contiguous()
        g_ext = torch.zeros_like(pp) if g_ext is None else g_ext.float().contiguous()
        gx, gv, gpp, gpv = (torch.zeros_like(t) for t in (x, v, pp, pv))
        cp = _contact_params(*ctx.kc, ctx.rope.p_rad)
        keep, odp = _fp3(ctx.od)
        _lib.call("dt_rope_collision_adj", cp, ctx.rope.dt, n, x, v, idx, 1, 0, nt, ctx.tris, pp,
                  pv, odp, g_imp, g_ext, gx, gv, gpp, gpv, _lib.stream())
        return None, None, None, None, gx, gv, gpp, gpv
