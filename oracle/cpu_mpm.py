"""ctypes binding of the oracle's C++/OpenMP restatement of one MPM substep
and its adjoint (oracle/cpu_mpm.cpp; TEST AND BASELINE INFRASTRUCTURE ONLY:
tests/ check it against oracle/mpm.py, bench.py times it as the CPU baseline).

``CpuMPM(prm, materials)`` takes the package's MPMParams and Material rows;
``substep`` / ``substep_adj`` work on float32 numpy arrays [n][3] / [n][3][3].
"""
import ctypes
import os

import numpy as np

from . import build_cpu

_lib = None


class _Prm(ctypes.Structure):
    _fields_ = [("n_grid", ctypes.c_int32), ("dx", ctypes.c_float), ("inv_dx", ctypes.c_float),
                ("dt", ctypes.c_float), ("p_vol", ctypes.c_float), ("p_mass", ctypes.c_float),
                ("eps", ctypes.c_float), ("gravity", ctypes.c_float * 3),
                ("bound", ctypes.c_int32), ("sticky_floor", ctypes.c_int32)]


class _Mat(ctypes.Structure):
    _fields_ = [("mu", ctypes.c_float), ("lam", ctypes.c_float), ("yield_stress", ctypes.c_float),
                ("dp_alpha", ctypes.c_float), ("kind", ctypes.c_int32)]


def lib():
    global _lib
    if _lib is None:
        path = build_cpu.LIB
        if not os.path.exists(path):
            build_cpu.build()
        _lib = ctypes.CDLL(path)
        fp = ctypes.c_void_p
        _lib.cpu_mpm_ws_floats.restype = ctypes.c_longlong
        _lib.cpu_mpm_ws_floats.argtypes = [ctypes.POINTER(_Prm), ctypes.c_int]
        _lib.cpu_mpm_gws_floats.restype = ctypes.c_longlong
        _lib.cpu_mpm_gws_floats.argtypes = [ctypes.POINTER(_Prm), ctypes.c_int]
        _lib.cpu_mpm_threads.restype = ctypes.c_int
        _lib.cpu_mpm_substep.argtypes = [ctypes.POINTER(_Prm), ctypes.POINTER(_Mat), ctypes.c_int,
                                         ctypes.c_int, fp] + [fp] * 9
        _lib.cpu_mpm_substep_adj.argtypes = [ctypes.POINTER(_Prm), ctypes.POINTER(_Mat),
                                             ctypes.c_int, ctypes.c_int, fp] + [fp] * 15
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class CpuMPM:
    def __init__(self, prm, materials, mat_id=None):
        self.prm = _Prm()
        self.prm.n_grid, self.prm.dx, self.prm.inv_dx = prm.n_grid, prm.dx, prm.inv_dx
        self.prm.dt, self.prm.p_vol, self.prm.p_mass, self.prm.eps = (prm.dt, prm.p_vol,
                                                                      prm.p_mass, prm.eps)
        for i in range(3):
            self.prm.gravity[i] = prm.gravity[i]
        self.prm.bound, self.prm.sticky_floor = prm.bound, int(prm.sticky_floor)
        self.mats = (_Mat * len(materials))()
        for row, m in zip(self.mats, materials):
            row.mu, row.lam, row.yield_stress = m.mu, m.lam, m.yield_stress
            row.dp_alpha, row.kind = m.dp_alpha, m.kind
        self.n_mats = len(materials)
        self.mat_id = None if mat_id is None else np.ascontiguousarray(mat_id, np.int32)

    def threads(self):
        return lib().cpu_mpm_threads()

    def workspace(self, n):
        return np.empty(lib().cpu_mpm_ws_floats(ctypes.byref(self.prm), n), np.float32)

    def substep(self, x, v, C, F, ws=None):
        """Next state (x, v, C, F) and the workspace the adjoint reads."""
        n = len(x)
        ins = [np.ascontiguousarray(a, np.float32) for a in (x, v, C, F)]
        outs = [np.empty_like(a) for a in ins]
        ws = self.workspace(n) if ws is None else ws
        lib().cpu_mpm_substep(ctypes.byref(self.prm), self.mats, self.n_mats, n, _p(self.mat_id),
                              *[_p(a) for a in ins], *[_p(a) for a in outs], _p(ws))
        return outs, ws

    def substep_adj(self, x, v, C, F, ws, g_next, gws=None):
        """Gradients of (x, v, C, F) and of each material's (mu, lam) from the
        gradients of the next state."""
        n = len(x)
        ins = [np.ascontiguousarray(a, np.float32) for a in (x, v, C, F)]
        gn = [np.ascontiguousarray(a, np.float32) for a in g_next]
        gout = [np.empty_like(a) for a in ins]
        gm = np.zeros((self.n_mats, 2), np.float32)
        if gws is None:
            gws = np.empty(lib().cpu_mpm_gws_floats(ctypes.byref(self.prm), n), np.float32)
        lib().cpu_mpm_substep_adj(ctypes.byref(self.prm), self.mats, self.n_mats, n,
                                  _p(self.mat_id), *[_p(a) for a in ins], _p(ws),
                                  *[_p(a) for a in gn], *[_p(a) for a in gout], _p(gm), _p(gws))
        return gout, gm
