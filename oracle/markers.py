"""Oracle: tactile markers and losses (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

Restates ``project_3d_2d`` (``difftactile/sensor_model/fisheye_model.py:13-30``),
``extract_markers`` (``fem_sensor.py:161-187``), the marker-motion loss
(``tasks/grasp_plastic.py:240-246,254-256``), the pad contact force and the
force loss (``compute_contact_force`` / ``compute_force_loss``,
``grasp_plastic.py:218-238``) and the centre-of-mass loss
(``tasks/grasp_elastic.py:240-247``) in torch.

Pinning (tests/test_oracle_markers.py): project_3d_2d against the numpy
``project_points_to_pix`` restatement (independent formula: arccos of the
normalised dot product with e_z); the on-axis point maps to (cx, cy); radius
proportional to the polar angle; IDW weights sum to one and reproduce the
marker when it coincides with a surface point; FD of the marker loss;
contact force of a uniform field (n_tris / 2 times it) and of a single
loaded vertex (its triangle count / 6); force loss closed forms (identity and
quarter-turn inverse rotations, the 1e-3 gate on either pad, the derivative
beta sign(x) / |mu y| and -beta |x| sign(y) / (|mu| y^2)).
"""
import math

import torch


def project_3d_2d(a, f=8.627e-04, m=173913.04, cx=3.320e02, cy=2.400e02, z_offset=0.02):
    a = torch.stack([a[..., 0], a[..., 1], a[..., 2] + z_offset], -1)
    an = torch.sqrt((a * a).sum(-1) + 1e-10)
    c = torch.clamp(a[..., 2] / an, -1.0, 1.0)
    theta = torch.acos(c)
    omega = torch.atan2(a[..., 1], a[..., 0] + 1e-8) + math.pi
    r = m * f * theta
    return torch.stack([r * torch.cos(omega) + cx, r * torch.sin(omega) + cy], -1)


def to_cam(inv_T, p):
    q = p @ inv_T[:3, :3].T + inv_T[:3, 3]
    return torch.stack([q[..., 0], q[..., 2], q[..., 1]], -1)


def extract_markers(pos, vpos, surface_id, inv_T, interp_idx, interp_w, fisheye=None):
    kw = {} if fisheye is None else dict(zip(("f", "m", "cx", "cy", "z_offset"), fisheye))
    loc = project_3d_2d(to_cam(inv_T, pos[surface_id]), **kw)
    vloc = project_3d_2d(to_cam(inv_T, vpos[surface_id]), **kw)
    pred = (interp_w[..., None] * loc[interp_idx]).sum(1)
    virt = (interp_w[..., None] * vloc[interp_idx]).sum(1)
    return pred, virt


def marker_motion_loss(pred, virt, beta=1.0):
    off = (pred - virt).mean(0)
    return beta * (off * off).sum()


def com(x):
    return x.mean(0)


def contact_force(ext, tris):
    """1/6 of the external force at each contact triangle's three vertices,
    summed (compute_contact_force, grasp_plastic.py:219-230)."""
    t = torch.as_tensor(tris, dtype=torch.int64)
    return (ext[t[:, 0]] + ext[t[:, 1]] + ext[t[:, 2]]).sum(0) / 6.0


def force_loss(cfs, inv_rots, mu, beta=1.0):
    """compute_force_loss (grasp_plastic.py:232-238) over the pads: p = inv_rot
    cf; beta sum |p.x| / |mu p.y| when every |p.x| > 1e-3, else 0."""
    ps = [torch.as_tensor(R, dtype=cf.dtype) @ cf for cf, R in zip(cfs, inv_rots)]
    if all(float(abs(p[0].detach())) > 1e-3 for p in ps):
        return beta * sum(abs(p[0]) / abs(mu * p[1]) for p in ps)
    return sum(0.0 * p.sum() for p in ps)
