"""Pins the jaw-kinematics oracle (CPU): closed forms of the chain and the
control velocities, and FP64 central differences of its autograd adjoint."""
import math

import numpy as np
import torch

from oracle import kinematics as ok

D64 = torch.float64


def _t(v):
    return torch.as_tensor(v, dtype=D64)


def test_eul2mat_is_a_proper_rotation_and_follows_the_right_hand_rule():
    rng = np.random.default_rng(0)
    for _ in range(20):
        M = ok.eul2mat(_t(rng.uniform(-180, 180, 3)), _t(rng.normal(size=3)))
        R = M[:3, :3]
        assert torch.allclose(R @ R.T, torch.eye(3, dtype=D64), atol=1e-12)
        assert abs(torch.linalg.det(R).item() - 1.0) < 1e-12
        assert torch.equal(M[3], _t([0, 0, 0, 1]))
    ex, ey, ez = _t([1, 0, 0]), _t([0, 1, 0]), _t([0, 0, 1])
    z3 = _t([0, 0, 0])
    assert torch.allclose(ok.eul2mat(_t([90, 0, 0]), z3)[:3, :3] @ ey, ez, atol=1e-12)
    assert torch.allclose(ok.eul2mat(_t([0, 90, 0]), z3)[:3, :3] @ ez, ex, atol=1e-12)
    assert torch.allclose(ok.eul2mat(_t([0, 0, 90]), z3)[:3, :3] @ ex, ey, atol=1e-12)
    # order: R = Rx Ry Rz, so a vector is turned about z first
    M = ok.eul2mat(_t([90, 0, 90]), z3)[:3, :3]
    assert torch.allclose(M @ ex, ez, atol=1e-12)   # z: ex -> ey, then x: ey -> ez
    T = ok.eul2mat(_t([0, 0, 0]), _t([1.5, -2, 3]))
    assert torch.allclose(T[:3, :3], torch.eye(3, dtype=D64))
    assert torch.equal(T[:3, 3], _t([1.5, -2, 3]))


def _setup(T=6):
    W, f1t, f2t, frame0 = ok.init_frames([10.0, -20.0, 30.0], [1.0, 2.0, 3.0], 2.5)
    rng = np.random.default_rng(1)
    x1 = _t(rng.normal(size=(40, 3)))
    x2 = _t(rng.normal(size=(30, 3)))
    return W, f1t, f2t, frame0, x1, x2


def test_zero_controls_keep_the_frames_and_give_zero_velocity():
    W, f1t, f2t, frame0, x1, x2 = _setup()
    frames = ok.chain(W, f1t, f2t, frame0, torch.zeros(5, 7, dtype=D64), 1e-3)
    for fr in frames:
        for a, b in zip(fr, frames[0]):
            assert torch.allclose(a, b, atol=1e-14)
    assert ok.control_vel(frames, 1, x1, 1e-3).abs().max() < 1e-12


def test_pure_translation_moves_every_vertex_at_the_world_rotated_velocity():
    W, f1t, f2t, frame0, x1, x2 = _setup()
    dp = _t([0.3, -0.7, 1.1])
    c = torch.zeros(4, 7, dtype=D64)
    c[:, 0:3] = dp
    frames = ok.chain(W, f1t, f2t, frame0, c, 1e-2)
    want = W[:3, :3] @ dp
    for k, x in ((1, x1), (2, x2)):
        cv = ok.control_vel(frames, k, x, 1e-2)
        assert torch.allclose(cv, want.expand_as(cv), atol=1e-10)


def test_pure_grip_moves_each_finger_along_its_own_y():
    W, f1t, f2t, frame0, x1, x2 = _setup()
    c = torch.zeros(3, 7, dtype=D64)
    c[:, 6] = 0.8
    frames = ok.chain(W, f1t, f2t, frame0, c, 1e-2)
    for k, x, Ft in ((1, x1, f1t), (2, x2, f2t)):
        cv = ok.control_vel(frames, k, x, 1e-2)
        want = (W @ Ft)[:3, :3] @ _t([0.0, 0.8, 0.0])
        assert torch.allclose(cv, want.expand_as(cv), atol=1e-10)


def test_pure_rotation_about_z_turns_vertices_on_circles():
    W, f1t, f2t, frame0, x1, x2 = _setup()
    w, dt = 45.0, 1e-2
    c = torch.zeros(1, 7, dtype=D64)
    c[0, 5] = w
    frames = ok.chain(W, f1t, f2t, frame0, c, dt)
    th = math.radians(w * dt)
    for k, x, Ft in ((1, x1, f1t), (2, x2, f2t)):
        cv = ok.control_vel(frames, k, x, dt)[0]
        # vertex in the end-effector frame, turned about its z axis
        p = (torch.cat([x, torch.ones(len(x), 1, dtype=D64)], 1) @ (Ft).T)[:, :3]
        d = torch.stack([p[:, 0] * (math.cos(th) - 1) - p[:, 1] * math.sin(th),
                         p[:, 0] * math.sin(th) + p[:, 1] * (math.cos(th) - 1),
                         torch.zeros(len(p), dtype=D64)], 1)
        want = d @ W[:3, :3].T / dt
        assert torch.allclose(cv, want, atol=1e-9)


def test_adjoint_matches_central_differences():
    W, f1t, f2t, frame0, x1, x2 = _setup()
    rng = np.random.default_rng(3)
    c0 = _t(rng.normal(size=(5, 7)) * [0.5, 0.5, 0.5, 20, 20, 20, 0.3])
    w1 = _t(rng.normal(size=(5, len(x1), 3)))
    w2 = _t(rng.normal(size=(5, len(x2), 3)))
    dt = 1e-2

    def loss(c):
        fr = ok.chain(W, f1t, f2t, frame0, c, dt)
        return (ok.control_vel(fr, 1, x1, dt) * w1).sum() + (ok.control_vel(fr, 2, x2, dt) * w2).sum()

    c = c0.clone().requires_grad_(True)
    g = torch.autograd.grad(loss(c), c)[0]
    # the velocities difference nearly equal poses and divide by dt, so the
    # step is moderate and the bar is 1e-5 relative
    h = 1e-4
    for (f, j) in [(0, 0), (1, 2), (2, 3), (3, 4), (4, 5), (0, 6), (4, 6), (2, 1)]:
        e = torch.zeros_like(c0)
        e[f, j] = h
        fd = (loss(c0 + e) - loss(c0 - e)) / (2 * h)
        assert abs(fd.item() - g[f, j].item()) <= 1e-5 * max(1.0, abs(fd.item())), (f, j)


# ---- dome pad pose control -------------------------------------------------
from oracle import dome_pose as odp  # noqa: E402


def _rot(axis, deg):
    t = torch.zeros(3, dtype=torch.float64)
    t[axis] = deg
    return odp.eul2mat(t, torch.zeros(3, dtype=torch.float64))[1]


def test_dome_eul2mat_is_rz_ry_rx():
    e = torch.eye(3, dtype=torch.float64)
    # right-hand quarter turns: x -> y about z, z -> x about y, y -> z about x
    assert torch.allclose(_rot(2, 90.0) @ e[0], e[1], atol=1e-15)
    assert torch.allclose(_rot(1, 90.0) @ e[2], e[0], atol=1e-15)
    assert torch.allclose(_rot(0, 90.0) @ e[1], e[2], atol=1e-15)
    ang = torch.tensor([30.0, -50.0, 70.0], dtype=torch.float64)
    T, R = odp.eul2mat(ang, torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64))
    want = _rot(2, 70.0) @ _rot(1, -50.0) @ _rot(0, 30.0)
    assert torch.allclose(R, want, atol=1e-15)
    assert not torch.allclose(R, _rot(0, 30.0) @ _rot(1, -50.0) @ _rot(2, 70.0), atol=1e-3)
    assert torch.allclose(R @ R.T, e, atol=1e-14) and abs(float(torch.det(R)) - 1) < 1e-14
    assert torch.allclose(T[:3, 3], torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64))


def _pad_pose(seed=0):
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(seed)
    W = np.eye(4)
    W[:3, :3] = Rotation.from_rotvec(np.deg2rad(rng.uniform(-40, 40, 3))).as_matrix()
    W[:3, 3] = rng.uniform(-1, 1, 3)
    return odp.init_pose(W), rng


def test_dome_pose_control_closed_forms():
    pose, rng = _pad_pose()
    st = 5e-5 * 49
    v = torch.tensor(rng.normal(size=(50, 3)))
    z3 = torch.zeros(3, dtype=torch.float64)
    # zero controls: the pose stays, the velocity is zero
    p0 = odp.pose_step(pose, z3, z3, st)
    for k in odp.POSE_FIELDS:
        assert torch.allclose(p0[k], pose[k], atol=1e-14), k
    assert torch.allclose(odp.control_vel(v, p0["dtrans_h"], st), torch.zeros(50, 3, dtype=D64),
                          atol=1e-12)
    # a pure translation (local frame): every vertex moves at R_world d_pos
    d = torch.tensor([0.3, -0.2, 0.5], dtype=torch.float64)
    p1 = odp.pose_step(pose, d, z3, st)
    cv = odp.control_vel(v, p1["dtrans_h"], st)
    assert torch.allclose(cv, (pose["rot_world"] @ d).expand(50, 3), atol=1e-10)
    # a pure turn about the pad's local z by theta: rotation about the pad
    # frame's z axis through its origin
    w = 40.0
    p2 = odp.pose_step(pose, z3, torch.tensor([0.0, 0.0, w], dtype=torch.float64), st)
    th = math.radians(w * st)
    W = pose["trans_world"]
    Rz = torch.tensor([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0],
                       [0, 0, 1]], dtype=torch.float64)
    local = (v - W[:3, 3]) @ W[:3, :3]          # W^-1 v for a rigid W
    moved = local @ Rz.T @ W[:3, :3].T + W[:3, 3]
    assert torch.allclose(odp.control_vel(v, p2["dtrans_h"], st), (moved - v) / st, atol=1e-9)
    # two local steps compose as T2 T1; the world step left-multiplies trans_h
    a = torch.tensor([0.1, 0.2, -0.3], dtype=torch.float64)
    b = torch.tensor([10.0, -20.0, 5.0], dtype=torch.float64)
    q = odp.pose_step(odp.pose_step(pose, a, b, st), b * 0.01, a * 50, st)
    T1 = odp.eul2mat(b * st, a * st)[0]
    T2 = odp.eul2mat(a * 50 * st, b * 0.01 * st)[0]
    assert torch.allclose(q["trans_local"], T2 @ T1, atol=1e-13)
    assert torch.allclose(q["trans_h"], W @ T2 @ T1, atol=1e-13)
    assert torch.allclose(q["inv_trans_h"] @ q["trans_h"], torch.eye(4, dtype=torch.float64),
                          atol=1e-13)
    qw = odp.pose_step(pose, a, b, st, mode=1)
    assert torch.allclose(qw["trans_h"], T1 @ W, atol=1e-13)
    assert torch.allclose(qw["rot_h"], T1[:3, :3] @ pose["rot_h"], atol=1e-13)
