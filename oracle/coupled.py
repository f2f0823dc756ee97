"""Oracle: one coupled substep of an MPM object and FEM pads (TEST INFRASTRUCTURE ONLY).

Composes the oracle kernels in the order of ``Contact.update``
(``difftactile/tasks/grasp_plastic.py:151-164``):
compute_new_F -> svd -> p2g -> pad update -> check_grid_occupy ->
check_collision -> collision -> grid_op -> g2p -> pad update2.
Reverse mode over this composition reproduces ``Contact.update_grad``
(``grasp_plastic.py:167-178``); with ``clamp=True`` the adjoint clipping of
``clamp_grid`` (``grasp_plastic.py:191-202``) is applied where the reference
applies it: to d/d grid_m as produced by grid_op's adjoint, and to the pads'
d/d pos, d/d vel as produced by update2's adjoint.
"""
import torch

from . import contact as oc
from . import fem as of
from . import mpm as om


class _ClampGrad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return x.view_as(x)

    @staticmethod
    def backward(ctx, g):
        return torch.clamp(g, -1000.0, 1000.0)


def coupled_substep(prm, state, mats, pads, cp, clamp=False, info=None):
    """state = (x, v, C, F); pads = list of dicts (pos, vel, vpos, cv, cv_next, f2v, B,
    layer, tris, out_dir, mu, lam, mass, dt); cp = (kn, kd, kt, mu_f).
    Returns (next state, list of (pos', vel', vpos'), contact_idx, nodes).
    info: a dict that receives the contact impulse grid ("grid_f") and the pads'
    external forces ("ext")."""
    x, v, C, F = state
    Fn = om.compute_new_F(C, F, prm.dt)
    U, s, V = om.svd3(Fn, clamp=1e-30 if getattr(prm, "svd_exact", False) else 1e-8)
    mv, m, F1 = om.p2g(prm, x, v, C, Fn, U, s, V, mats)
    vel_u = [of.update(p["pos"], p["vel"], p["f2v"], p["B"], p["mu"], p["lam"], p["dt"], p["mass"])
             for p in pads]
    nodes, idx = oc.check_collision(prm, m, [p["pos"] for p in pads], [p["tris"] for p in pads])
    sensors = [(p["pos"], vu, p["tris"], p["out_dir"]) for p, vu in zip(pads, vel_u)]
    kn, kd, kt, mu_f = cp
    grid_f, exts = oc.collision(prm, mv, m, nodes, idx, sensors, kn, kd, kt, mu_f)
    if info is not None:
        info["grid_f"], info["ext"] = grid_f, exts
    m_g = _ClampGrad.apply(m) if clamp else m
    gv = om.grid_op(prm, mv, m_g, grid_f)
    nx, nv, nC = om.g2p(prm, x, gv)
    out_pads = []
    for p, vu, e in zip(pads, vel_u, exts):
        pos = _ClampGrad.apply(p["pos"]) if clamp else p["pos"]
        vu2 = _ClampGrad.apply(vu) if clamp else vu
        if p.get("cv_next") is None:
            out_pads.append(of.update2_dome(pos, vu2, p["vpos"], e, p["layer"], p["cv"], p["dt"],
                                            p["mass"]))
        else:
            out_pads.append(of.update2_gripper(pos, vu2, p["vpos"], e, p["layer"], p["cv"],
                                               p["cv_next"], p["dt"], p["mass"]))
    return (nx, nv, nC, F1), out_pads, idx, nodes
