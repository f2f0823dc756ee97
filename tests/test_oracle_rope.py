"""Pins the rope oracle (CPU)."""
import numpy as np
import torch

from oracle import rope as orp

D64 = torch.float64


def _t(a):
    return torch.as_tensor(np.asarray(a, np.float64))


def test_a_stiff_stretch_constraint_restores_the_rest_length():
    rp = orp.RopeParams(2, dt=1e-3, stretch_compliance=0.0, stretch_relaxation=1.0)
    p = _t([[0.0, 1.0, 0.0], [0.17, 1.05, -0.02]])
    q = orp._project(p, np.array([0]), np.array([1]), rp.rest_length, 0.0, 1.0, rp)
    assert abs(torch.linalg.norm(q[0] - q[1]).item() - rp.rest_length) < 1e-8
    # the correction is along the segment and splits equally (equal masses)
    assert torch.allclose(q.mean(0), p.mean(0), atol=1e-14)


def test_a_straight_rope_at_rest_length_falls_freely_with_its_end_held():
    rp = orp.RopeParams(12, dt=1e-3, table_height=-10.0)
    p = _t([[0.0, 1.0, 0.1 * i] for i in range(12)])
    v = torch.zeros_like(p)
    p1, v1 = orp.step(rp, p, v, torch.zeros_like(p), p[0])
    g = torch.tensor(rp.gravity, dtype=D64)
    want = p + g * rp.dt * rp.dt
    assert torch.allclose(p1[1:], want[1:], atol=1e-12)
    assert torch.equal(p1[0], p[0])
    # the 1e-10 in |n| leaves corrections of ~1e-11, 1e-8 once divided by dt
    assert torch.allclose(v1[1:], (want[1:] - p[1:]) / rp.dt, atol=1e-7)


def test_the_floor_removes_normal_velocity_and_applies_friction():
    # particle 0 is the held one: a 2-particle rope without constraint action
    rp2 = orp.RopeParams(2, dt=1e-3, table_height=0.0, gravity=(0.0, 0.0, 0.0),
                         stretch_relaxation=0.0)
    p2 = _t([[0.0, 1.0, 0.0], [0.0, 0.0005, 0.0]])
    v2 = _t([[0.0, 0.0, 0.0], [0.3, -2.0, -0.1]])
    q, w = orp.step(rp2, p2, v2, torch.zeros_like(p2), p2[0])
    assert torch.allclose(w[1], _t([0.3 * 0.995, 0.0, -0.1 * 0.995]), atol=1e-12)


def test_step_gradient_matches_central_differences():
    rp = orp.RopeParams(9, dt=1e-3, table_height=0.0)
    rng = np.random.default_rng(0)
    p0 = _t([[0.0, 0.3, 0.1 * i] for i in range(9)]) + 0.01 * _t(rng.normal(size=(9, 3)))
    p0[3, 1] = 0.001                          # one particle near the floor
    v0 = 0.5 * _t(rng.normal(size=(9, 3)))
    v0[3, 1] = -3.0
    e0 = 1e-4 * _t(rng.normal(size=(9, 3)))
    w1 = _t(rng.normal(size=(9, 3)))
    w2 = _t(rng.normal(size=(9, 3)))

    def loss(p, v, e):
        a, b = orp.step(rp, p, v, e, p0[0].detach())
        return (a * w1).sum() + 1e-3 * (b * w2).sum()

    args = [t.clone().requires_grad_(True) for t in (p0, v0, e0)]
    gs = torch.autograd.grad(loss(*args), args)
    h = 1e-7
    for k, base in enumerate((p0, v0, e0)):
        for (i, a) in [(1, 0), (3, 1), (5, 2), (8, 1)]:
            e = torch.zeros_like(base)
            e[i, a] = h
            pl = [p0, v0, e0]
            pl[k] = base + e
            mi = [p0, v0, e0]
            mi[k] = base - e
            fd = (loss(*pl) - loss(*mi)) / (2 * h)
            assert abs(fd.item() - gs[k][i, a].item()) <= 1e-6 * max(1.0, abs(fd.item())), (k, i, a)


def test_contact_is_zero_outside_and_obeys_action_reaction():
    from difftactile_amd.sensor_model.meshes import slab_mesh
    nodes, f2v, tris, layer, od = slab_mesh(8, 8, 3, h=0.05, origin=(0.0, 0.2, 0.0))
    rp = orp.RopeParams(20, dt=1e-3, p_rad=0.02)
    rng = np.random.default_rng(1)
    x = _t(rng.uniform([0.05, 0.15, 0.05], [0.3, 0.23, 0.3], size=(20, 3)))
    v = _t(rng.normal(size=(20, 3)))
    pos = _t(nodes)
    vel = _t(0.1 * rng.normal(size=nodes.shape))
    from oracle import contact as oc
    idx = oc.find_closest_np(x.numpy().astype(np.float32), nodes, tris)
    imp, ext, flag = orp.contact(rp, x, v, pos, vel, tris, _t(od), idx, 1e3, 10.0, 50.0, 0.5)
    assert flag.any() and (~flag).any()
    assert (imp[~flag] == 0).all()
    assert torch.allclose(imp.sum(0) / rp.dt + ext.sum(0), torch.zeros(3, dtype=D64), atol=1e-9)
    far = x + _t([0.0, -1.0, 0.0])
    idx_f = oc.find_closest_np(far.numpy().astype(np.float32), nodes, tris)
    imp_f, ext_f, flag_f = orp.contact(rp, far, v, pos, vel, tris, _t(od), idx_f,
                                       1e3, 10.0, 50.0, 0.5)
    assert not flag_f.any() and (imp_f == 0).all() and (ext_f == 0).all()


def test_contact_keeps_the_cable_forms_negative_normal_velocity():
    """A particle under a flat pad moving away from it (n . v_rel < 0 in the
    force's frame) feels -(kn + kd n.v_rel) sdf along n: the cable task's
    force (cable_straightening.py:211-223) does not clip the normal velocity
    at zero; a shear speed just under 1e-4 adds no tangential force (no eps)."""
    from oracle import contact as oc
    rp = orp.RopeParams(2, dt=1e-3, p_rad=0.01)
    pos = _t([[-1.0, 0.2, -1.0], [3.0, 0.2, -1.0], [-1.0, 0.2, 3.0]])
    tris = np.array([[0, 2, 1]])
    od = _t([0.0, -1.0, 0.0])                         # contact face looks down
    x = _t([[0.3, 0.195, 0.3], [0.4, 0.195, 0.4]])   # 5 mm below: sdf = -0.005 - r
    kn, kd, kt, mu = 1e3, 30.0, 50.0, 0.5
    pad_v = torch.zeros(3, 3, dtype=D64)
    # particle 0 falls (moves away from the face), particle 1 slides slowly
    v = _t([[0.0, -2.0, 0.0], [0.99975e-4, 0.0, 0.0]])
    idx = np.zeros(2, np.int64)
    imp, ext, flag = orp.contact(rp, x, v, pos, pad_v, tris, od, idx, kn, kd, kt, mu)
    assert flag.all()
    sdf = 0.2 - 0.195 - rp.p_rad                      # n = -y: sdf = (x - p1) . n - r
    # the force f = -(kn + kd v_n) sdf n acts along the face normal n = -y with
    # r = -(v - v_pad) = -v, so v_n = n . r = -2 for the falling particle
    vn0 = -2.0
    np.testing.assert_allclose(imp[0].numpy() / rp.dt, [0.0, (kn + kd * vn0) * sdf, 0.0],
                               rtol=1e-12)
    np.testing.assert_allclose(imp[1].numpy() / rp.dt, [0.0, kn * sdf, 0.0], rtol=1e-12)
