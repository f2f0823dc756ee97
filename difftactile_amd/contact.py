"""Coupled substep of an MPM object and FEM pads with grid penalty contact.

Mirrors the ``Contact`` class of ``difftactile/tasks/grasp_plastic.py``:
``check_collision`` (269-276), ``collision`` (279-299, with ``.grad``),
``clamp_grid`` (191-202), ``update`` (151-164) and ``update_grad`` (167-178),
with the penalty constants ``kn, kd, kt, friction_coeff`` (57-64) as fields.
Works with any number of pads (the grasp task has two); ``contact_idx`` has one
column per pad.  All kernels run in libdifftactile_hip.so.
"""
import numpy as np
import torch

from . import _lib
from .field import Field, Kernel


class GridContact:
    def __init__(self, mpm_object, pads, kn=34.53, kd=269.44, kt=154.78, friction_coeff=53.85,
                 norm_eps=1e-11, clamp_grad=True):
        self.mpm_object = mpm_object
        self.pads = list(pads)
        d = mpm_object.device
        self.device = d
        self.substeps = mpm_object.sub_steps
        n = mpm_object.n_grid
        self.kn = Field((), device=d, needs_grad=True)
        self.kd = Field((), device=d, needs_grad=True)
        self.kt = Field((), device=d, needs_grad=True)
        self.friction_coeff = Field((), device=d, needs_grad=True)
        self.kn[None], self.kd[None], self.kt[None] = kn, kd, kt
        self.friction_coeff[None] = friction_coeff
        self.norm_eps = float(norm_eps)
        self.clamp_grad = bool(clamp_grad)
        npad = max(1, len(self.pads))
        self.contact_idx = Field((self.substeps, n, n, n), (npad,), dtype=torch.int32, device=d)
        self.contact_idx.fill(-1)
        self.collision = Kernel(self._collision, self._collision_grad)

    def _cparams(self):
        c = _lib.DtContactParams()
        c.kn = float(self.kn[None])
        c.kd = float(self.kd[None])
        c.kt = float(self.kt[None])
        c.mu_f = float(self.friction_coeff[None])
        c.norm_eps = self.norm_eps
        c.sdf_offset = 0.0
        return c

    def _grid(self, f):
        o = self.mpm_object
        return (_lib.ptr(o.grid_v_in.data[f]), 3, _lib.ptr(o.grid_m.data[f]), 1)

    def check_collision(self, f):
        """Closest pad triangle per occupied node (``grasp_plastic.py:269-276``)."""
        o = self.mpm_object
        npad = len(self.pads)
        for s, pad in enumerate(self.pads):
            cen = pad.centroids(f)
            _lib.call("dt_contact_closest", o.prm.c_struct(), _lib.ptr(o.grid_m.data[f]), 1,
                      pad.num_triangles, _lib.ptr(cen),
                      _lib.ptr(self.contact_idx.data[f], torch.int32), npad, s, _lib.stream())

    def _collision(self, f):
        o = self.mpm_object
        cp = self._cparams()
        mv, smv, m, sm = self._grid(f)
        for s, pad in enumerate(self.pads):
            od = pad.out_dir_host()
            _lib.call("dt_contact_collision", o.prm.c_struct(), cp, mv, smv, m, sm,
                      _lib.ptr(self.contact_idx.data[f], torch.int32), len(self.pads), s,
                      pad.num_triangles, _lib.ptr(pad.contact_seg.data, torch.int32),
                      _lib.ptr(pad.pos.data[f]), _lib.ptr(pad.vel.data[f]),
                      od.ctypes.data_as(_lib.ctypes.POINTER(_lib.ctypes.c_float)),
                      _lib.ptr(o.grid_f.data[f]), _lib.ptr(pad.external_force_field.data[f]),
                      _lib.stream())

    def _collision_grad(self, f):
        o = self.mpm_object
        cp = self._cparams()
        mv, smv, m, sm = self._grid(f)
        for s, pad in enumerate(self.pads):
            od = pad.out_dir_host()
            _lib.call("dt_contact_collision_adj", o.prm.c_struct(), cp, mv, smv, m, sm,
                      _lib.ptr(self.contact_idx.data[f], torch.int32), len(self.pads), s,
                      pad.num_triangles, _lib.ptr(pad.contact_seg.data, torch.int32),
                      _lib.ptr(pad.pos.data[f]), _lib.ptr(pad.vel.data[f]),
                      od.ctypes.data_as(_lib.ctypes.POINTER(_lib.ctypes.c_float)),
                      _lib.ptr(o.grid_f.grad.data[f]),
                      _lib.ptr(pad.external_force_field.grad.data[f]),
                      _lib.ptr(o.grid_v_in.grad.data[f]), 3, _lib.ptr(o.grid_m.grad.data[f]), 1,
                      _lib.ptr(pad.pos.grad.data[f]), _lib.ptr(pad.vel.grad.data[f]),
                      _lib.stream())

    def clamp_grid(self, f):
        """``clamp_grid`` (``grasp_plastic.py:191-202``)."""
        o = self.mpm_object
        g = o.grid_m.grad.data[f]
        _lib.call("dt_clamp", g.numel(), _lib.ptr(g), 1000.0, _lib.stream())
        for pad in self.pads:
            for t in (pad.pos.grad.data[f], pad.vel.grad.data[f]):
                _lib.call("dt_clamp", t.numel(), _lib.ptr(t), 1000.0, _lib.stream())

    def reset(self):
        """``reset`` (``grasp_plastic.py:386-402``, contact part)."""
        self.mpm_object.reset()
        for pad in self.pads:
            pad.reset_contact()
        self.contact_idx.fill(-1)

    def update(self, f):
        """One coupled substep in the reference order (``grasp_plastic.py:151-164``)."""
        o = self.mpm_object
        o.compute_new_F(f)
        o.svd(f)
        o.p2g(f)
        for pad in self.pads:
            pad.update(f)
        o.check_grid_occupy(f)
        self.check_collision(f)
        self.collision(f)
        o.grid_op(f)
        o.g2p(f)
        for pad in self.pads:
            pad.update2(f)

    def update_grad(self, f):
        """Adjoint of ``update`` in the reference order (``grasp_plastic.py:167-178``)."""
        o = self.mpm_object
        for pad in self.pads:
            pad.update2.grad(f)
        o.g2p.grad(f)
        o.grid_op.grad(f)
        if self.clamp_grad:
            self.clamp_grid(f)
        self.collision.grad(f)
        for pad in self.pads:
            pad.update.grad(f)
        o.p2g.grad(f)
        o.svd_grad(f)
        o.compute_new_F.grad(f)
