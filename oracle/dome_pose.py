"""Oracle: pose control of the dome pad (TEST INFRASTRUCTURE ONLY, see
oracle/__init__.py).

Restates ``difftactile/sensor_model/fem_sensor.py``:

  eul2mat             196-210  Rz Ry Rx of x-y-z angles in degrees + translation
  set_trans_h         217-227  world-frame step: trans_h = T trans_h, rot_h = R rot_h
  set_pose_control    229-245  local-frame step: dtrans_h = W T W^-1,
                               trans_local = T trans_local, trans_h = W trans_local,
                               rot_local = R rot_local, rot_h = rot_world rot_local
  set_pose_control_bp 248-256  dtrans_h and the inverses only
  set_control_vel     258-266  ((dtrans_h [v; 1]) - v)_{0:3} / (dt (sub_steps - 1))

with T, R = eul2mat(d_ori dt (sub_steps - 1), d_pos dt (sub_steps - 1)) and
every step ending with inv_trans_h = trans_h^-1, inv_rot = rot_h^-1; in torch
(float64 on the CPU by default) so autograd gives the adjoints the reference
takes from Taichi's AD (set_pose_control.grad, set_control_vel.grad).

Pinning (tests/test_oracle_kinematics.py): eul2mat is orthonormal with det 1,
single-axis quarter turns follow the right-hand rule and two axes compose as
Rz Ry Rx (not Rx Ry Rz); zero controls keep the pose and give zero velocity;
a pure d_pos gives every vertex R_world d_pos; a pure d_ori about z turns the
vertices about the pad frame's z axis through its origin (closed form); two
local steps compose as T2 T1; the world step left-multiplies.
"""
import math

import torch

POSE_FIELDS = ("trans_world", "trans_local", "trans_h", "inv_trans_h", "dtrans_h",
               "rot_world", "rot_local", "rot_h", "inv_rot")


def eul2mat(rot_deg, trans):
    """(T 4x4, R 3x3) with R = Rz Ry Rx of the angles (degrees)."""
    r = rot_deg * (math.pi / 180.0)
    cx, sx = torch.cos(r[0]), torch.sin(r[0])
    cy, sy = torch.cos(r[1]), torch.sin(r[1])
    cz, sz = torch.cos(r[2]), torch.sin(r[2])
    o, z = torch.ones_like(cx), torch.zeros_like(cx)
    mx = torch.stack([torch.stack([o, z, z]), torch.stack([z, cx, -sx]), torch.stack([z, sx, cx])])
    my = torch.stack([torch.stack([cy, z, sy]), torch.stack([z, o, z]), torch.stack([-sy, z, cy])])
    mz = torch.stack([torch.stack([cz, -sz, z]), torch.stack([sz, cz, z]), torch.stack([z, z, o])])
    R = mz @ my @ mx
    T = torch.cat([torch.cat([R, trans.reshape(3, 1)], 1),
                   torch.cat([torch.zeros(1, 3, dtype=R.dtype), torch.ones(1, 1, dtype=R.dtype)],
                             1)], 0)
    return T, R


def init_pose(W, dtype=torch.float64):
    """Pose after ``init`` (fem_sensor.py:112-128) from the pad -> world 4x4 W."""
    W = torch.as_tensor(W, dtype=dtype)
    I4, I3 = torch.eye(4, dtype=dtype), torch.eye(3, dtype=dtype)
    Rw = W[:3, :3]
    return dict(trans_world=W, trans_local=I4, trans_h=W @ I4, inv_trans_h=torch.linalg.inv(W),
                dtrans_h=I4, rot_world=Rw, rot_local=I3, rot_h=Rw @ I3,
                inv_rot=torch.linalg.inv(Rw))


def pose_step(pose, d_pos, d_ori, step_time, mode=0):
    """mode 0 set_pose_control, 1 set_trans_h, 2 set_pose_control_bp; returns
    the new pose dict."""
    T, R = eul2mat(d_ori * step_time, d_pos * step_time)
    p = dict(pose)
    W = p["trans_world"]
    if mode != 1:
        p["dtrans_h"] = W @ T @ torch.linalg.inv(W)
    if mode == 0:
        p["trans_local"] = T @ p["trans_local"]
        p["trans_h"] = W @ p["trans_local"]
        p["rot_local"] = R @ p["rot_local"]
        p["rot_h"] = p["rot_world"] @ p["rot_local"]
    elif mode == 1:
        p["trans_h"] = T @ p["trans_h"]
        p["rot_h"] = R @ p["rot_h"]
    p["inv_trans_h"] = torch.linalg.inv(p["trans_h"])
    p["inv_rot"] = torch.linalg.inv(p["rot_h"])
    return p


def control_vel(vpos, dtrans_h, step_time):
    xh = torch.cat([vpos, torch.ones(len(vpos), 1, dtype=vpos.dtype)], 1)
    return ((xh @ dtrans_h.T)[:, :3] - vpos) / step_time
