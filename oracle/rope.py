"""Oracle: position-based rope and its particle-wise pad contact.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Restates

  PBDRope.__init__ / init / update   difftactile/object_model/PBD_rope.py:17-75,78-88,111-180
  check_collision / collision        difftactile/tasks/cable_straightening.py:287-319

in torch (autograd gives the adjoint Taichi's ``update.grad`` computes).

One substep (``update``): v' = v + ext_f / m + g dt; a particle below the
table moving down loses its normal velocity and keeps 0.995 of the
tangential one; p* = p + v' dt; one Jacobi pass of stretch constraints
(segments i, i+1, rest length L) then one of bending constraints (i, i+2,
length 2L), each dp = -C / (w1 + w2 + a) * n / |n| * relaxation with
|n| = sqrt(n.n + 1e-10), a = compliance / dt^2; particle 0 is held at its
initial position; v_next = (p_next - p) / dt.

Pinning (tests/test_oracle_rope.py):
  * with zero compliance and relaxation 1 a single stretch constraint
    restores the rest length exactly;
  * a straight rope at rest length (no constraint active) falls freely:
    p_next = p + (v + g dt) dt, particle 0 fixed;
  * the floor removes the normal velocity of a particle crossing it and keeps
    0.995 of the tangential one;
  * FP64 central differences of a random linear functional of one substep
    w.r.t. positions, velocities and external forces;
  * contact: zero force outside the pad (sdf >= 0), Newton's third law between
    the rope's impulse and the pad's external force; a particle leaving the pad
    (n . v_rel < 0) feels the damped spring -(kn + kd n.v_rel) sdf (the cable
    form keeps the negative normal velocity).
"""
import math

import numpy as np
import torch

from . import contact as oc


class RopeParams:
    """Constants of PBDRope (PBD_rope.py:17-51)."""

    def __init__(self, n_particles, dt=5e-5, p_rad=0.05, p_rho=1.0, table_height=0.0,
                 rest_length=None, stretch_compliance=None, bending_compliance=None,
                 stretch_relaxation=0.3, bending_relaxation=0.1, gravity=(0.0, -9.8, 0.0)):
        self.n = int(n_particles)
        self.dt = float(dt)
        self.p_rad = float(p_rad)
        self.p_mass = math.pi * 4 / 3 * p_rad ** 3 * p_rho
        self.inv_m = 1.0 / self.p_mass
        self.table_height = float(table_height)
        if rest_length is None:
            self.rest_length, self.bending_length = 2 * p_rad, 4 * p_rad
        else:
            self.rest_length, self.bending_length = float(rest_length), 2 * float(rest_length)
        self.stretch_compliance = 1e-5 / p_rho if stretch_compliance is None else stretch_compliance
        self.bending_compliance = 1e-5 / p_rho if bending_compliance is None else bending_compliance
        self.stretch_relaxation = float(stretch_relaxation)
        self.bending_relaxation = float(bending_relaxation)
        self.gravity = tuple(float(g) for g in gravity)
        self.friction = 0.995
        self.eps = 1e-10


def _project(p, k0, k1, length, compliance, relaxation, rp):
    """One Jacobi pass of distance constraints between particles k0 and k1."""
    n = p[k0] - p[k1]
    nn = torch.sqrt((n * n).sum(-1, keepdim=True) + rp.eps)
    C = nn - length
    alpha = compliance / rp.dt ** 2
    dp = -C / (2 * rp.inv_m + alpha) * n / nn * relaxation
    d = torch.zeros_like(p)
    d = d.index_add(0, torch.as_tensor(k0), dp * rp.inv_m)
    d = d.index_add(0, torch.as_tensor(k1), -dp * rp.inv_m)
    return p + d


def step(rp, pos, vel, ext_f, fixed):
    """One update(f): returns (pos_next, vel_next)."""
    g = torch.tensor(rp.gravity, dtype=pos.dtype)
    v = vel + ext_f / rp.p_mass + g * rp.dt
    p_pred = pos + v * rp.dt
    floor = (p_pred[:, 1] < rp.table_height) & (v[:, 1] < 0.0)
    nrm = torch.tensor([0.0, 1.0, 0.0], dtype=pos.dtype)
    v_t = v - (v @ nrm)[:, None] * nrm
    v = torch.where(floor[:, None], v_t * rp.friction, v)
    p = pos + v * rp.dt
    idx = np.arange(rp.n)
    p = _project(p, idx[:-1], idx[1:], rp.rest_length, rp.stretch_compliance,
                 rp.stretch_relaxation, rp)
    if rp.n > 2:
        p = _project(p, idx[:-2], idx[2:], rp.bending_length, rp.bending_compliance,
                     rp.bending_relaxation, rp)
    p = torch.cat([torch.as_tensor(fixed, dtype=pos.dtype).reshape(1, 3), p[1:]], 0)
    return p, (p - pos) / rp.dt


def contact(rp, pos, vel, pad_pos, pad_vel, tris, out_dir, idx, kn, kd, kt, mu_f):
    """Particle-wise penalty contact with one pad (cable_straightening.py:
    301-319): the sdf is offset by the particle radius and the force is the
    cable task's own form (cable_straightening.py:211-223: no max(., 0) on the
    normal velocity, shear norm without eps).  Returns (rope impulse f dt per
    particle, pad external force with -f/3 per triangle vertex, flag)."""
    tris_t = torch.as_tensor(np.asarray(tris), dtype=torch.int64)
    idx = torch.as_tensor(idx, dtype=torch.int64)
    ok = idx >= 0
    idx_c = torch.where(ok, idx, torch.zeros_like(idx))
    sdf, nv, rv, flag = oc.find_sdf(pos, vel, idx_c, pad_pos, pad_vel, tris_t, out_dir,
                                    offset=rp.p_rad)
    flag = flag & ok
    f = oc.contact_force(sdf, -nv, -rv, kn, kd, kt, mu_f, norm_eps=0.0, clamp_vn=False)
    f = torch.where(flag[:, None], f, torch.zeros_like(f))
    e = torch.zeros_like(pad_pos)
    t = tris_t[idx_c]
    for c in range(3):
        e = e.index_add(0, t[:, c], -f / 3)
    return f * rp.dt, e, flag
