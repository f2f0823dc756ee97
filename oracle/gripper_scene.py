"""Oracle: one substep of the full gripper scene (TEST INFRASTRUCTURE ONLY).

An MPM object, FEM pads and a position-based cable, coupled by penalty
contact.  The reference has the two halves in separate tasks:

* object + two pads: ``Contact.update`` of ``difftactile/tasks/grasp_plastic.py:151-164``
  (compute_new_F -> svd -> p2g -> pad update -> check_grid_occupy ->
  check_collision -> collision -> grid_op -> g2p -> pad update2);
* cable + two pads: ``update`` of ``difftactile/tasks/cable_straightening.py:166-173``
  (pad update -> check_collision -> collision -> rope update -> pad update2),
  with the cable task's contact force (cable_straightening.py:211-223).

Here they share the pad updates: pad update runs once after p2g, both
contacts read the updated pad velocities, the cable steps after g2p, and
each pad's update2 receives the sum of the grid and cable contact forces.
Reverse mode over this composition is the scene's adjoint; with ``clamp``
the adjoint clipping of both tasks' ``clamp_grid`` (grasp_plastic.py:191-202,
cable_straightening.py:196-207) applies to d/d grid_m from grid_op, the
pads' d/d pos and d/d vel from update2 and the cable's d/d pos and d/d vel
from its step.

Pinning: each part is pinned where it is defined (oracle/mpm.py, fem.py,
contact.py, rope.py); the composition is checked by tests/test_oracle_gripper.py
(a scene without the cable equals oracle.coupled.coupled_substep, a cable far
from the pads steps as oracle.rope.step, the forces on the pads balance the
impulses on the object and the cable).
"""
import torch

from . import contact as oc
from . import coupled as ocp
from . import fem as of
from . import mpm as om
from . import rope as orp


def gripper_substep(prm, state, mats, pads, rope, rp, cp_grid, cp_rope, clamp=False,
                    info=None):
    """state = (x, v, C, F); pads: dicts as for oracle.coupled.coupled_substep;
    rope = (pos, vel, fixed) of the cable, rp its RopeParams; cp_grid,
    cp_rope = (kn, kd, kt, mu_f).  Returns (next state, [(pos', vel', vpos')
    per pad], (rope pos', rope vel')).  info receives grid_f, the pads' total
    external forces and the cable's impulse."""
    x, v, C, F = state
    Fn = om.compute_new_F(C, F, prm.dt)
    U, s, V = om.svd3(Fn, clamp=1e-30 if getattr(prm, "svd_exact", False) else 1e-8)
    mv, m, F1 = om.p2g(prm, x, v, C, Fn, U, s, V, mats)
    vel_u = [of.update(p["pos"], p["vel"], p["f2v"], p["B"], p["mu"], p["lam"], p["dt"], p["mass"])
             for p in pads]
    nodes, idx = oc.check_collision(prm, m, [p["pos"] for p in pads], [p["tris"] for p in pads])
    sensors = [(p["pos"], vu, p["tris"], p["out_dir"]) for p, vu in zip(pads, vel_u)]
    grid_f, exts = oc.collision(prm, mv, m, nodes, idx, sensors, *cp_grid)
    rpos, rvel, fixed = rope
    imp = torch.zeros_like(rpos)
    exts = list(exts)
    for k, (p, vu) in enumerate(zip(pads, vel_u)):
        ridx = oc.find_closest_np(rpos.detach().cpu().numpy(), p["pos"].detach().cpu().numpy(),
                                  p["tris"])
        imp_k, ext_k, _ = orp.contact(rp, rpos, rvel, p["pos"], vu, p["tris"], p["out_dir"], ridx,
                                      *cp_rope)
        imp = imp + imp_k
        exts[k] = exts[k] + ext_k
    cl = ocp._ClampGrad.apply if clamp else (lambda t: t)
    gv = om.grid_op(prm, mv, cl(m), grid_f)
    nx, nv, nC = om.g2p(prm, x, gv)
    rpos1, rvel1 = orp.step(rp, cl(rpos), cl(rvel), imp, fixed)
    out_pads = []
    for p, vu, e in zip(pads, vel_u, exts):
        out_pads.append(of.update2_gripper(cl(p["pos"]), cl(vu), p["vpos"], e, p["layer"], p["cv"],
                                           p["cv_next"], p["dt"], p["mass"]))
    if info is not None:
        info.update(grid_f=grid_f, ext=exts, impulse=imp)
    return (nx, nv, nC, F1), out_pads, (rpos1, rvel1)
