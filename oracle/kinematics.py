"""Oracle: parallel-jaw kinematic chain and pad control velocities.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Restates
``difftactile/sensor_model/gripper_kinematics.py``:

  eul2mat            289-303  x-y-z Euler angles in degrees -> 4x4 transform
  init_params        122-161  frame 0 (end effector and both finger frames)
  pos_transformation 186-192  end-effector step from (d_pos, d_ori) * dt
  grip_kinematic     169-183  finger opening step from d_gripper * dt
  set_control_vel    194-216  per-vertex velocity of a finger's rigid motion

in torch (float64 on the CPU by default) so autograd gives the adjoint that
``kinematic_grad`` (219-229) computes with Taichi's AD.

Controls: row f of ``controls`` is (d_pos xyz, d_ori xyz in degrees per unit
time, d_gripper) for the step from frame f to frame f+1; the reference holds
one such set per outer step (``d_pos``, ``d_ori``, ``d_gripper`` fields) and
calls ``kinematic(f)`` once per substep.

Pinning (tests/test_oracle_kinematics.py):
  eul2mat         -- orthonormal with det 1; single-axis quarter turns map the
                     axes as the right-hand rule says; zero angles give the
                     pure translation.
  chain + control_vel
                  -- zero controls keep every frame and give zero velocity;
                     a pure d_pos gives every vertex the velocity R_w d_pos
                     (R_w the end-effector world rotation); a pure d_gripper
                     moves each finger along its own local y at d_gripper;
                     a pure d_ori about z turns every vertex on the circle
                     about the end-effector axis (closed form with cos/sin).
  gradients       -- FP64 central differences of a random linear functional.
"""
import math

import torch

# gripper geometry (cm), gripper_kinematics.py:70-73
BASE_HEIGHT = 2.0
BASE_WIDTH = 14.0
FINGER_HEIGHT = 7.5


def eul2mat(rot_deg, trans):
    """4x4 transform [R t; 0 1] with R = Rx Ry Rz of the angles (degrees)."""
    r = rot_deg * (math.pi / 180.0)
    cx, sx = torch.cos(r[0]), torch.sin(r[0])
    cy, sy = torch.cos(r[1]), torch.sin(r[1])
    cz, sz = torch.cos(r[2]), torch.sin(r[2])
    o, z = torch.ones_like(cx), torch.zeros_like(cx)
    mx = torch.stack([torch.stack([o, z, z]), torch.stack([z, cx, -sx]), torch.stack([z, sx, cx])])
    my = torch.stack([torch.stack([cy, z, sy]), torch.stack([z, o, z]), torch.stack([-sy, z, cy])])
    mz = torch.stack([torch.stack([cz, -sz, z]), torch.stack([sz, cz, z]), torch.stack([z, z, o])])
    R = mx @ my @ mz
    top = torch.cat([R, trans.reshape(3, 1)], dim=1)
    bottom = torch.cat([torch.zeros(1, 3, dtype=R.dtype), torch.ones(1, 1, dtype=R.dtype)], dim=1)
    return torch.cat([top, bottom], dim=0)


def init_frames(rot, trans, grip_width, dtype=torch.float64):
    """Frame 0 of init_params: (ee_world, finger1_t, finger2_t) and the six
    per-frame matrices (ee_local, ee, finger1_local, finger2_local,
    finger1_ee, finger2_ee)."""
    t = lambda v: torch.as_tensor(v, dtype=dtype)
    W = eul2mat(t(rot), t(trans))
    f1t = eul2mat(t([-90.0, 0.0, -90.0]),
                  t([-BASE_WIDTH / 2 + grip_width, -BASE_HEIGHT - FINGER_HEIGHT, 0.0]))
    f2t = eul2mat(t([-90.0, 0.0, 90.0]),
                  t([BASE_WIDTH / 2 - grip_width, -BASE_HEIGHT - FINGER_HEIGHT, 0.0]))
    eye = torch.eye(4, dtype=dtype)
    frame0 = [eye, W @ eye, eye, eye, W @ f1t @ eye, W @ f2t @ eye]
    return W, f1t, f2t, frame0


def chain(W, f1t, f2t, frame0, controls, dt):
    """Frames 0..T for T = len(controls): each a list of six 4x4 matrices."""
    frames = [list(frame0)]
    for c in controls:
        ee_local, _, l1, l2, _, _ = frames[-1]
        D = eul2mat(c[3:6] * dt, c[0:3] * dt)                 # pos_transformation
        ee_local1 = D @ ee_local
        ee1 = W @ ee_local1
        gd = 2.0 * c[6] * dt                                  # grip_delta
        G = eul2mat(torch.zeros(3, dtype=c.dtype), torch.stack(
            [torch.zeros((), dtype=c.dtype), gd / 2, torch.zeros((), dtype=c.dtype)]))
        l1n, l2n = G @ l1, G @ l2                             # grip_kinematic
        frames.append([ee_local1, ee1, l1n, l2n, ee1 @ f1t @ l1n, ee1 @ f2t @ l2n])
    return frames


def control_vel(frames, finger, init_x, dt):
    """Per-vertex control velocity of one finger (k = 1, 2) for steps
    0..T-1: ((E[f+1] - E[f]) [x; 1])_{0:3} / dt with E the finger pose."""
    xh = torch.cat([init_x, torch.ones(len(init_x), 1, dtype=init_x.dtype)], dim=1)  # (V,4)
    out = []
    for f in range(len(frames) - 1):
        E0, E1 = frames[f][3 + finger], frames[f + 1][3 + finger]
        out.append(((xh @ E1.T) - (xh @ E0.T))[:, :3] / dt)
    return torch.stack(out)
