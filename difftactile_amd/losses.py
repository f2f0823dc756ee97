"""Differentiable losses on the measured path, on the HIP library.

* ``com_loss``: centre of mass (``compute_obj_pos``, ``tasks/grasp_elastic.py:240-247``).
* ``MarkerModel`` / ``marker_displacement``: tactile markers of a pad
  (``extract_markers``, ``sensor_model/fem_sensor.py:161-187`` with the fisheye
  projection of ``fisheye_model.py:13-30``) and the summed squared marker
  displacement between the deformed and the rigidly moved ("virtual") pad.
* ``kabsch_angle``: rotation angle of the masked particles over one substep
  (``compute_COM`` / ``compute_H`` / ``compute_R``, ``object_model/multi_obj.py:309-340``,
  one term of ``compute_angle``, ``tasks/box_open.py:400-406``).
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


class _COM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        x = x.contiguous()
        out = torch.zeros(3, dtype=torch.float64, device=x.device)
        _lib.call("dt_loss_sum_positions", x.shape[0], 0, x, out.data_ptr(), _lib.stream())
        ctx.n = x.shape[0]
        ctx.shape = x.shape
        return (out / max(ctx.n, 1)).float()

    @staticmethod
    def backward(ctx, g):
        gx = torch.zeros(ctx.shape, dtype=torch.float32, device=g.device)
        keep, gp = _fptr(g.detach().float().cpu().numpy())
        _lib.call("dt_loss_com_adj", ctx.n, 0, gp, gx, _lib.stream())
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
        gloc = torch.zeros(self.n_surface, 2, device=self.device)
        gpos = torch.zeros(pos.shape, device=self.device)
        pos, g_pred = pos.contiguous(), g_pred.contiguous()
        _lib.call("dt_markers_adj", self.n_surface, self.sid, pos, self.Tp, self.Kp,
                  self.n_markers, self.k, self.idx, self.w, g_pred, gloc, gpos, _lib.stream())
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
