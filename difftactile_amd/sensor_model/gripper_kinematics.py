"""Parallel-jaw kinematics and the pads' control velocities on the GPU.

Follows ``difftactile/sensor_model/gripper_kinematics.py``: the gripper
geometry (70-73) and frame 0 (``init_params``, 122-161) are set up on the
host; the chain over a trajectory (``pos_transformation`` 186-192 and
``grip_kinematic`` 169-183 per substep) and every finger pad's per-substep
control velocity (``fem_sensor{1,2}_set_control_vel`` 194-216) run in the HIP
library (``dt_jaw_*``), with hand-written adjoints (``kinematic_grad``
219-229).  ``control_velocities`` is differentiable with respect to the
controls, so a loss on a coupled rollout (``coupled.CoupledSim``) reaches the
jaw trajectory.

Controls: one row of seven values per substep, (d_pos xyz, d_ori xyz in
degrees per unit time, d_gripper) -- the reference's ``d_pos``, ``d_ori`` and
``d_gripper`` fields, which it holds per outer step.
"""
import math

import numpy as np
import torch

from .. import _lib

BASE_HEIGHT = 2.0
BASE_WIDTH = 14.0
FINGER_HEIGHT = 7.5
FINGER_WIDTH = 1.4
F32 = np.float32


def eul2mat(rot_deg, trans):
    """[Rx Ry Rz, t; 0 1] of x-y-z angles in degrees (float32, host), as
    ``eul2mat`` (289-303) builds the constant transforms of ``init_params``."""
    r = np.asarray(rot_deg, F32) * F32(math.pi / 180.0)
    cx, sx, cy, sy, cz, sz = (np.cos(r[0]), np.sin(r[0]), np.cos(r[1]), np.sin(r[1]),
                              np.cos(r[2]), np.sin(r[2]))
    mx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]], F32)
    my = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]], F32)
    mz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]], F32)
    m = np.eye(4, dtype=F32)
    m[:3, :3] = (mx @ my) @ mz
    m[:3, 3] = np.asarray(trans, F32)
    return m


class Gripper:
    """Kinematic state of the parallel-jaw gripper (``Gripper`` of the
    reference without its meshes and pads)."""

    def __init__(self, dt=5e-5, device="cuda"):
        self.dt = float(dt)
        self.device = torch.device(device)
        self.base_height, self.base_width = BASE_HEIGHT, BASE_WIDTH
        self.finger_height, self.finger_width = FINGER_HEIGHT, FINGER_WIDTH

    def init_params(self, rot, trans, grip_width):
        """Frame 0 from the end-effector pose (Euler degrees, translation) and
        the grip width (``init_params``, 122-161)."""
        self.ee_world = eul2mat(rot, trans)
        self.finger1_t = eul2mat([-90.0, 0.0, -90.0],
                                 [-self.base_width / 2 + grip_width,
                                  -self.base_height - self.finger_height, 0.0])
        self.finger2_t = eul2mat([-90.0, 0.0, 90.0],
                                 [self.base_width / 2 - grip_width,
                                  -self.base_height - self.finger_height, 0.0])
        eye = np.eye(4, dtype=F32)
        ee = self.ee_world @ eye
        self.frame0 = np.stack([eye, ee, eye, eye, (ee @ self.finger1_t) @ eye,
                                (ee @ self.finger2_t) @ eye]).astype(F32)
        self.prm = _lib.DtJawParams()
        for i, v in enumerate(self.ee_world.reshape(-1)):
            self.prm.ee_world[i] = float(v)
        for k, t in enumerate((self.finger1_t, self.finger2_t)):
            for i, v in enumerate(t.reshape(-1)):
                self.prm.finger_t[k][i] = float(v)
        self.prm.dt = self.dt

    def finger_pose(self, finger):
        """Frame-0 pose of finger 1 or 2 (``finger{1,2}_ee[0]``): the transform
        that places its pad (``init_fem_sensor``)."""
        return self.frame0[3 + finger].copy()

    def frames(self, controls):
        """Frames 0..T of the chain for controls (T, 7): (T + 1, 6, 4, 4) with
        ee_local, ee, finger1_local, finger2_local, finger1_ee, finger2_ee."""
        c = controls.detach().to(self.device, torch.float32).contiguous()
        fr = self._chain(c)
        return fr.view(-1, 6, 4, 4)

    def _chain(self, c):
        T = c.shape[0]
        fr = torch.empty(T + 1, _lib.JAW_FRAME_FLOATS, device=self.device)
        fr[0] = torch.as_tensor(self.frame0.reshape(-1), device=self.device)
        _lib.call("dt_jaw_chain", self.prm, T, c, fr, _lib.stream())
        return fr

    def control_velocities(self, controls, init_x1, init_x2, dt=None):
        """Per-substep control velocities of the two finger pads, (T, V_k, 3)
        each, for controls (T, 7); init_x_k are the pads' rest vertices in
        their finger frame (``init_x``).  dt defaults to the gripper's (the
        pads' time step in the reference).  Differentiable w.r.t. controls."""
        dt = self.dt if dt is None else float(dt)
        return _JawControlVel.apply(self, dt, controls, init_x1, init_x2)


class _JawControlVel(torch.autograd.Function):
    @staticmethod
    def forward(ctx, grip, dt, controls, x1, x2):
        c = controls.detach().to(grip.device, torch.float32).contiguous()
        xs = [x.detach().to(grip.device, torch.float32).contiguous() for x in (x1, x2)]
        T = c.shape[0]
        fr = grip._chain(c)
        outs = []
        for k, x in ((1, xs[0]), (2, xs[1])):
            cv = torch.empty(T, x.shape[0], 3, device=grip.device)
            _lib.call("dt_jaw_control_vel", T, x.shape[0], k, fr, x, dt, cv, _lib.stream())
            outs.append(cv)
        ctx.grip, ctx.dt = grip, dt
        ctx.save_for_backward(c, fr, xs[0], xs[1])
        return tuple(outs)

    @staticmethod
    def backward(ctx, g1, g2):
        c, fr, x1, x2 = ctx.saved_tensors
        grip = ctx.grip
        T = c.shape[0]
        gfr = torch.zeros_like(fr)
        for k, x, g in ((1, x1, g1), (2, x2, g2)):
            if g is not None:
                _lib.call("dt_jaw_control_vel_adj", T, x.shape[0], k, x, ctx.dt,
                          g.float().contiguous(), gfr, _lib.stream())
        gc = torch.zeros_like(c)
        _lib.call("dt_jaw_chain_adj", grip.prm, T, c, fr, gfr, gc, _lib.stream())
        return None, None, gc, None, None
