"""Differentiable losses on the measured path, on the HIP library.

* ``com_loss``: centre of mass (``compute_obj_pos``, ``tasks/grasp_elastic.py:240-247``).
* ``MarkerModel`` / ``marker_displacement``: tactile markers of a pad
  (``extract_markers``, ``sensor_model/fem_sensor.py:161-187`` with the fisheye
  projection of ``fisheye_model.py:13-30``) and the summed squared marker
  displacement between the deformed and the rigidly moved ("virtual") pad.
* ``marker_offset`` / ``marker_motion_loss``: the reference's marker loss,
  beta times the squared mean marker offset summed over the pads
  (``compute_marker_motion`` / ``compute_marker_loss``,
  ``tasks/grasp_plastic.py:240-246,254-256``).
* ``contact_force`` / ``force_loss``: a pad's summed contact force (1/6 of
  each contact triangle's vertex forces, ``compute_contact_force``,
  ``grasp_plastic.py:218-230``) and the sliding-ratio loss over the pads
  (``compute_force_loss``, ``grasp_plastic.py:232-238``).
* ``kabsch_angle``: rotation angle of the masked particles over one substep
  (``compute_COM`` / ``compute_H`` / ``compute_R``, ``object_model/multi_obj.py:309-340``,
  one term of ``compute_angle``, ``tasks/box_open.py:400-406``).

Every reduction is float32 in a fixed order, so losses and gradients repeat
bit for bit.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .sensor_model.markers import (FISHEYE, camera_points, interp_weights, project_points_np,
                                   synthetic_markers)


def _fptr(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _scratch(terms, device):
    return torch.empty(terms * _lib.REDUCE_BLOCKS, device=device)


class _COM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        x = x.detach().float().contiguous()
        out = torch.empty(3, device=x.device)
        _lib.call("dt_loss_com", x.shape[0], 0, x, _scratch(3, x.device), out, _lib.stream())
        ctx.n = x.shape[0]
        ctx.shape = x.shape
        return out

    @staticmethod
    def backward(ctx, g):
        gx = torch.zeros(ctx.shape, dtype=torch.float32, device=g.device)
        _lib.call("dt_loss_com_adj", ctx.n, 0, g.detach().float().contiguous(), gx, _lib.stream())
        return gx


def com_loss(x):
    """Mean particle position of x [n][3] (differentiable)."""
    return _COM.apply(x)


class MarkerModel:
    """Markers of one pad: surface vertex ids, camera transform and interpolation."""

    def __init__(self, rest_world, surface_ids, inv_T, n_markers=136, k=5, device="cuda"):
        rest_world = np.asarray(rest_world, dtype=np.float64)
        inv_T = np.asarray(inv_T, dtype=np.float64)
        local = rest_world[surface_ids] @ inv_T[:3, :3].T + inv_T[:3, 3]
        px = project_points_np(camera_points(local))
        markers = synthetic_markers(px, n_markers)
        idx, w = interp_weights(markers, px, k)
        self.k = k
        self.n_markers = len(markers)
        self.n_surface = len(surface_ids)
        self.sid = torch.as_tensor(np.asarray(surface_ids), dtype=torch.int32, device=device)
        self.idx = torch.as_tensor(idx, dtype=torch.int32, device=device).contiguous()
        self.w = torch.as_tensor(w, dtype=torch.float32, device=device).contiguous()
        # the interpolation's adjoint gathers per surface point: entries
        # m k + j of every interp_idx[m][j] == s, grouped by s in entry order
        flat = np.asarray(idx, dtype=np.int64).reshape(-1)
        order = np.argsort(flat, kind="stable")
        start = np.zeros(self.n_surface + 1, np.int64)
        np.add.at(start, flat + 1, 1)
        self.s_start = torch.as_tensor(np.cumsum(start), dtype=torch.int32, device=device)
        self.s_entry = torch.as_tensor(order, dtype=torch.int32, device=device)
        self._T, self.Tp = _fptr(inv_T.reshape(-1))
        self._K, self.Kp = _fptr(np.array(FISHEYE))
        self.device = torch.device(device)

    def predict(self, pos, vpos):
        ns, nm = self.n_surface, self.n_markers
        loc = torch.empty(ns, 2, device=self.device)
        vloc = torch.empty(ns, 2, device=self.device)
        pred = torch.empty(nm, 2, device=self.device)
        virt = torch.empty(nm, 2, device=self.device)
        pos, vpos = pos.contiguous(), vpos.contiguous()
        _lib.call("dt_markers_fwd", ns, self.sid, pos, vpos, self.Tp, self.Kp, nm, self.k,
                  self.idx, self.w, loc, vloc, pred, virt, _lib.stream())
        return pred, virt

    def adjoint(self, pos, g_pred):
        gloc = torch.empty(self.n_surface, 2, device=self.device)
        gpos = torch.zeros(pos.shape, device=self.device)
        pos, g_pred = pos.contiguous(), g_pred.float().contiguous()
        _lib.call("dt_markers_adj", self.n_surface, self.sid, pos, self.Tp, self.Kp,
                  self.n_markers, self.k, self.s_start, self.s_entry, self.w, g_pred, gloc, gpos,
                  _lib.stream())
        return gpos


class _MarkerDisp(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model, pos, vpos):
        pred, virt = model.predict(pos, vpos)
        d = pred - virt
        ctx.model = model
        ctx.save_for_backward(pos, vpos, d)
        return (d * d).sum()

    @staticmethod
    def backward(ctx, g):
        pos, vpos, d = ctx.saved_tensors
        gd = 2.0 * g * d
        return None, ctx.model.adjoint(pos, gd), ctx.model.adjoint(vpos, -gd)


def marker_displacement(model, pos, vpos):
    """Sum over markers of |pred - virtual|^2 (differentiable in pos and vpos)."""
    return _MarkerDisp.apply(model, pos, vpos)


class _MarkerOffset(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model, pos, vpos):
        pred, virt = model.predict(pos, vpos)
        out = torch.empty(2, device=model.device)
        _lib.call("dt_marker_offset", model.n_markers, pred, virt, _scratch(2, model.device), out,
                  _lib.stream())
        ctx.model = model
        ctx.save_for_backward(pos, vpos)
        return out

    @staticmethod
    def backward(ctx, g):
        pos, vpos = ctx.saved_tensors
        m = ctx.model
        gp = torch.zeros(m.n_markers, 2, device=m.device)
        gv = torch.zeros(m.n_markers, 2, device=m.device)
        _lib.call("dt_marker_offset_adj", m.n_markers, g.detach().float().contiguous(), gp, gv,
                  _lib.stream())
        return None, m.adjoint(pos, gp), m.adjoint(vpos, gv)


def marker_offset(model, pos, vpos):
    """Mean marker offset pred - virtual of one pad, (2,) (``marker_offset``
    of ``compute_marker_motion``, grasp_plastic.py:240-246), differentiable in
    pos and vpos."""
    return _MarkerOffset.apply(model, pos, vpos)


def marker_motion_loss(models, poses, beta=1.0):
    """The reference's marker loss (``compute_marker_loss``,
    grasp_plastic.py:254-256): beta sum over the pads of the squared mean
    marker offset; poses: one (pos, vpos) per model."""
    loss = None
    for m, (pos, vpos) in zip(models, poses):
        off = marker_offset(m, pos, vpos)
        t = (off * off).sum()
        loss = t if loss is None else loss + t
    return beta * loss


class _ContactForce(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tris, ext):
        ext = ext.detach().float().contiguous()
        out = torch.empty(3, device=ext.device)
        _lib.call("dt_contact_force_sum", tris.shape[0], tris, ext, _scratch(3, ext.device), out,
                  _lib.stream())
        ctx.tris, ctx.shape = tris, ext.shape
        return out

    @staticmethod
    def backward(ctx, g):
        ge = torch.zeros(ctx.shape, device=g.device)
        _lib.call("dt_contact_force_sum_adj", ctx.tris.shape[0], ctx.tris,
                  g.detach().float().contiguous(), ge, _lib.stream())
        return None, ge


def contact_force(tris, ext):
    """Summed contact force of one pad, (3,) (``compute_contact_force``,
    grasp_plastic.py:218-230): 1/6 of the external force at the three vertices
    of each contact triangle; tris: the pad's contact triangles (int32 [T][3]
    on the GPU), ext: its external forces [V][3] (e.g. the last substep's,
    ``CoupledSim.rollout(..., return_ext=True)``)."""
    return _ContactForce.apply(tris.to(torch.int32).contiguous(), ext)


class _ForceLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, inv_rot, mu, beta, *cfs):
        n = len(cfs)
        cf = torch.stack([c.detach().float() for c in cfs]).contiguous()
        out = torch.empty(1 + 3 * n, device=cf.device)
        keep, rp = _fptr(np.asarray(inv_rot, np.float32).reshape(n, 9))
        _lib.call("dt_force_loss", n, cf, rp, float(mu), float(beta), out, _lib.stream())
        ctx.save_for_backward(out)
        ctx.n = n
        return out[0].clone()

    @staticmethod
    def backward(ctx, g):
        (out,) = ctx.saved_tensors
        gc = out[1:].view(ctx.n, 3) * g
        return (None, None, None, *gc.unbind(0))


def force_loss(cfs, inv_rots, mu, beta=1.0):
    """Sliding-ratio loss of the pads' contact forces (``compute_force_loss``,
    grasp_plastic.py:232-238): p_k = inv_rot_k cf_k; beta sum_k |p_k.x| /
    |mu p_k.y| when every |p_k.x| > 1e-3, else 0.  inv_rots: one 3x3 per pad
    (the sensor's inverse rotation), mu: the friction coefficient."""
    return _ForceLoss.apply(inv_rots, mu, beta, *cfs)


class _Kabsch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, v_next, mask, dt):
        x, v = x.detach().float().contiguous(), v_next.detach().float().contiguous()
        m = mask.to(device=x.device, dtype=torch.int32).contiguous()
        state = torch.zeros(_lib.KABSCH_STATE, device=x.device)
        scratch = torch.empty(9 * _lib.REDUCE_BLOCKS, device=x.device)
        _lib.call("dt_kabsch_angle", x.shape[0], x, v, float(dt), m, scratch, state, _lib.stream())
        ctx.dt = float(dt)
        ctx.save_for_backward(x, v, m, state)
        return state[47].clone()

    @staticmethod
    def backward(ctx, g):
        x, v, m, state = ctx.saved_tensors
        st = state.clone()
        gx, gv = torch.zeros_like(x), torch.zeros_like(v)
        gg = g.detach().float().reshape(1).contiguous()
        _lib.call("dt_kabsch_angle_adj", x.shape[0], x, v, ctx.dt, m, st, gg, gx, gv, _lib.stream())
        return gx, gv, None, None


def kabsch_angle(x, v_next, mask, dt):
    """Rotation angle (radians, about z) of the masked particles between x and
    x + dt v_next; x, v_next [n][3], mask [n] (nonzero = counted).  The
    centres divide by all n particles, as the reference does.  Differentiable
    w.r.t. x and v_next."""
    return _Kabsch.apply(x, v_next, mask, dt)
