"""GPU parity of the FEM pad and grid contact kernels against the oracle.

Forward within 1e-5 relative L2 of the FP64 oracle on the same float32
inputs; adjoints within 1e-4; closest-triangle indices bit-exact against the
float32 brute force of the oracle.
"""
import ctypes

import numpy as np
import pytest
import torch

from difftactile_amd import _lib
from difftactile_amd.contact import GridContact
from difftactile_amd.object_model.mpm_elastic import MPMObj
from difftactile_amd.params import MPMParams
from difftactile_amd.sensor_model.fem_sensor import FEMDomeSensor
from difftactile_amd.sensor_model.meshes import dome_mesh, slab_mesh
from oracle import contact as oc
from oracle import coupled as ocp
from oracle import fem as of
from oracle import mpm as om
from tests.mpm_cases import oracle_params, rel_l2

pytestmark = pytest.mark.gpu
F64 = torch.float64
P = _lib.ptr


def dev(a, dtype=torch.float32):
    return torch.as_tensor(np.asarray(a)).to(dtype).cuda().contiguous()


def h64(t):
    return t.detach().cpu().to(F64)


def fem_params(mu, lam, dt, mass):
    p = _lib.DtFemParams()
    p.mu, p.lam, p.dt, p.mass = mu, lam, dt, mass
    return p


def _dome(seed=0, pert=0.01):
    nodes, f2v, surf, layer = dome_mesh()
    rng = np.random.default_rng(seed)
    pos = (nodes + pert * rng.normal(size=nodes.shape)).astype(np.float32)
    return nodes.astype(np.float32), pos, f2v, surf, layer


def test_fem_update_on_the_slab_is_within_1e5_of_fp64(gpu):
    """The C2 slab (40 x 40 x 6 vertices, Kuhn tetrahedra, cond(Dm) < 10) with
    C2's material: B, the updated velocity and its increment dt f / m are
    within 1e-5 relative L2 of the FP64 oracle, the adjoint within 1e-5."""
    from difftactile_amd.scenes import config2_pad_press
    c = config2_pad_press(n_particles=1000, n_grid=32)
    nodes, f2v, tris, layer, od = c["mesh"]
    nodes = nodes.astype(np.float32)
    nt, nv = len(f2v), len(nodes)
    rng = np.random.default_rng(4)
    h = float(np.abs(nodes[1] - nodes[0]).max())
    pos = (nodes + 0.05 * h * rng.normal(size=nodes.shape)).astype(np.float32)
    B = torch.empty(nt, 3, 3, device="cuda")
    _lib.call("dt_fem_init_B", nt, dev(f2v, torch.int32), dev(nodes), P(B), _lib.stream())
    B_ref = of.init_B(torch.tensor(nodes, dtype=F64), torch.tensor(f2v).long())
    assert float(torch.linalg.cond(torch.linalg.inv(B_ref)).max()) < 10
    assert rel_l2(h64(B), B_ref) < 1e-6
    mu, lam, mass, dt = c["pad_mu"], c["pad_lam"], c["pad_mass"], c["prm"].dt
    prm = fem_params(mu, lam, dt, mass)
    vel0 = (1e-4 * rng.normal(size=(nv, 3))).astype(np.float32)
    vel = dev(vel0)
    _lib.call("dt_fem_update", prm, nt, dev(f2v, torch.int32), B, dev(pos), P(vel), _lib.stream())
    posd = torch.tensor(pos, dtype=F64, requires_grad=True)
    vref = of.update(posd, torch.tensor(vel0, dtype=F64), torch.tensor(f2v).long(), h64(B), mu,
                     lam, dt, mass)
    assert rel_l2(h64(vel), vref.detach()) < 1e-5
    dv = vref.detach() - torch.tensor(vel0, dtype=F64)
    assert float(dv.norm()) > float(torch.tensor(vel0).norm())   # the forces dominate
    assert rel_l2(h64(vel) - torch.tensor(vel0, dtype=F64), dv) < 1e-5
    gvel = rng.normal(size=(nv, 3)).astype(np.float32)
    gpos = torch.zeros(nv, 3, device="cuda")
    _lib.call("dt_fem_update_adj", prm, nt, dev(f2v, torch.int32), B, dev(pos), dev(gvel),
              P(gpos), _lib.stream())
    g_ref = torch.autograd.grad((vref * torch.tensor(gvel, dtype=F64)).sum(), posd)[0]
    assert rel_l2(h64(gpos), g_ref) < 1e-5


def test_fem_init_B_update_and_adjoint(gpu):
    nodes, pos, f2v, surf, layer = _dome()
    nt, nv = len(f2v), len(nodes)
    B = torch.empty(nt, 3, 3, device="cuda")
    _lib.call("dt_fem_init_B", nt, dev(f2v, torch.int32), dev(nodes), P(B),
              _lib.stream())
    B_ref = of.init_B(torch.tensor(nodes, dtype=F64), torch.tensor(f2v).long())
    # float32 inverse: per-tet error bounded by cond(Dm) * eps32
    cond = torch.linalg.cond(torch.linalg.inv(B_ref))
    err = (h64(B) - B_ref).flatten(1).norm(dim=1) / B_ref.flatten(1).norm(dim=1)
    assert (err <= 20 * cond * 1.2e-7).all()
    assert rel_l2(h64(B), B_ref) < 1e-4
    mu, lam, dt, mass = 1294.01, 9201.11, 5e-5, 0.0147
    prm = fem_params(mu, lam, dt, mass)
    rng = np.random.default_rng(1)
    vel0 = (0.01 * rng.normal(size=(nv, 3))).astype(np.float32)
    vel = dev(vel0)
    Bf = dev(h64(B).numpy())
    _lib.call("dt_fem_update", prm, nt, dev(f2v, torch.int32), P(Bf), dev(pos),
              P(vel), _lib.stream())
    posd = torch.tensor(pos, dtype=F64, requires_grad=True)
    vref = of.update(posd, torch.tensor(vel0, dtype=F64), torch.tensor(f2v).long(), h64(Bf), mu,
                     lam, dt, mass)
    # the dome's Delaunay mesh has slivers (cond up to ~3e4): the tolerance is
    # three times the float32 oracle's own distance from float64
    v32 = of.update(torch.tensor(pos), torch.tensor(vel0), torch.tensor(f2v).long(), Bf.cpu(), mu,
                    lam, dt, mass)
    tol_v = max(1e-5, 3 * rel_l2(v32, vref.detach()))
    assert rel_l2(h64(vel), vref.detach()) < tol_v
    # adjoint
    gvel = rng.normal(size=(nv, 3)).astype(np.float32)
    gpos = torch.zeros(nv, 3, device="cuda")
    _lib.call("dt_fem_update_adj", prm, nt, dev(f2v, torch.int32), P(Bf),
              dev(pos), dev(gvel), P(gpos), _lib.stream())
    g_ref = torch.autograd.grad((vref * torch.tensor(gvel, dtype=F64)).sum(), posd)[0]
    p32 = torch.tensor(pos, requires_grad=True)
    v32 = of.update(p32, torch.tensor(vel0), torch.tensor(f2v).long(), Bf.cpu(), mu, lam, dt, mass)
    g32 = torch.autograd.grad((v32 * torch.tensor(gvel)).sum(), p32)[0]
    tol_g = max(1e-4, 3 * rel_l2(g32, g_ref))
    assert rel_l2(h64(gpos), g_ref) < tol_g
    # atomic-free CSR form: same values up to float summation order
    from difftactile_amd.coupled import vertex_tet_csr
    vs, vt = vertex_tet_csr(f2v, nv)
    vs, vt = dev(vs, torch.int32), dev(vt, torch.int32)
    tbuf = torch.empty(nt, 4, 4, device="cuda")
    vel2 = dev(vel0)
    _lib.call("dt_fem_update_csr", prm, nt, nv, dev(f2v, torch.int32), Bf, dev(pos), vel2, vs, vt,
              tbuf, _lib.stream())
    assert rel_l2(h64(vel2), h64(vel)) < 1e-6
    assert rel_l2(h64(vel2), vref.detach()) < tol_v
    gpos2 = torch.zeros(nv, 3, device="cuda")
    _lib.call("dt_fem_update_adj_csr", prm, nt, nv, dev(f2v, torch.int32), Bf, dev(pos),
              dev(gvel), gpos2, vs, vt, tbuf, _lib.stream())
    assert rel_l2(h64(gpos2), h64(gpos)) < 1e-5
    assert rel_l2(h64(gpos2), g_ref) < tol_g


@pytest.mark.parametrize("gripper", [False, True])
def test_fem_update2_and_adjoint(gpu, gripper):
    rng = np.random.default_rng(2)
    nv = 1910
    arrs = [rng.normal(size=(nv, 3)).astype(np.float32) for _ in range(6)]
    pos, vel, vpos, ext, cv, cv1 = arrs
    layer = rng.integers(0, 4, nv).astype(np.int32)
    prm = fem_params(1.0, 2.0, 5e-5, 0.0147)
    outs = [torch.empty(nv, 3, device="cuda") for _ in range(3)]
    _lib.call("dt_fem_update2", prm, nv, dev(layer, torch.int32), dev(pos),
              dev(vel), dev(vpos), dev(ext), dev(cv),
              dev(cv1) if gripper else None, *[P(o) for o in outs], _lib.stream())
    t = [torch.tensor(a, dtype=F64, requires_grad=True) for a in arrs]
    L = torch.tensor(layer)
    if gripper:
        ref = of.update2_gripper(t[0], t[1], t[2], t[3], L, t[4], t[5], 5e-5, 0.0147)
    else:
        ref = of.update2_dome(t[0], t[1], t[2], t[3], L, t[4], 5e-5, 0.0147)
    for o, r in zip(outs, ref):
        assert rel_l2(h64(o), r.detach()) < 1e-6
    seeds = [rng.normal(size=(nv, 3)).astype(np.float32) for _ in range(3)]
    g = [torch.zeros(nv, 3, device="cuda") for _ in range(6)]
    _lib.call("dt_fem_update2_adj", prm, nv, dev(layer, torch.int32),
              *[dev(s) for s in seeds], P(g[0]), P(g[1]), P(g[2]), P(g[3]), P(g[4]),
              P(g[5]) if gripper else None, _lib.stream())
    loss = sum((r * torch.tensor(s, dtype=F64)).sum() for r, s in zip(ref, seeds))
    gr = torch.autograd.grad(loss, t, allow_unused=True)
    for k, (got, want) in enumerate(zip(g, gr)):
        if want is None or (k == 5 and not gripper):
            continue
        assert rel_l2(h64(got), want) < 1e-6, k


def _contact_scene(n_grid=64, n_particles=20000, seed=0):
    prm = MPMParams(n_grid=n_grid, rho=1000.0, dt=2e-5)
    h = 1.0 / 128
    nodes, f2v, surf, layer, od = slab_mesh(24, 24, 4, h=h, origin=(0.4, 0.52, 0.4))
    rng = np.random.default_rng(seed)
    pos = (nodes + 2e-4 * rng.normal(size=nodes.shape)).astype(np.float32)
    x = rng.uniform([0.4, 0.44, 0.4], [0.58, 0.56, 0.58], size=(n_particles, 3)).astype(np.float32)
    v = rng.normal(0, 0.2, size=(n_particles, 3)).astype(np.float32)
    return prm, (nodes, f2v, surf, layer, od), pos, x, v


def test_closest_triangle_bit_exact(gpu):
    prm, (nodes, f2v, surf, layer, od), pos, x, v = _contact_scene()
    op = oracle_params(prm)
    mats = om.material_arrays([om.Material(5e4, 0.3)], np.zeros(len(x), int), F64)
    Fi = torch.eye(3, dtype=F64).repeat(len(x), 1, 1)
    U, s, V = om.svd3(Fi)
    mv, m, _ = om.p2g(op, torch.tensor(x, dtype=F64), torch.tensor(v, dtype=F64),
                      torch.zeros_like(Fi), Fi, U, s, V, mats)
    m32 = m.float()
    nn = prm.n_nodes
    cen = torch.empty(len(surf), 4, device="cuda")
    _lib.call("dt_contact_centroids", len(surf), dev(surf, torch.int32),
              dev(pos), P(cen), _lib.stream())
    nodes_o, ref = oc.check_collision(op, m32, [torch.tensor(pos)], [surf])
    assert len(nodes_o) > 1000
    # a previous result to start from: the answer itself on occupied nodes
    # (rolled by one so most starts are not the answer), garbage elsewhere
    ref_full = np.random.default_rng(4).integers(-3, len(surf) + 3, nn).astype(np.int32)
    ref_full[nodes_o] = np.roll(ref[:, 0], 1)
    occ_mask = np.zeros(nn, bool)
    occ_mask[nodes_o] = True
    bnd = torch.empty(_lib.tri_bounds_count(len(surf)), 4, device="cuda")
    _lib.call("dt_contact_tri_batches", len(surf), P(cen), P(bnd), _lib.stream())
    for batched in (False, True):
        idx = torch.full((nn, 2), -7, dtype=torch.int32, device="cuda")
        if batched:
            # the output as its own hint, holding the previous result
            idx[:, 1] = torch.as_tensor(ref_full, device="cuda")
            _lib.call("dt_contact_closest_batched", prm.c_struct(), dev(m32), 1, None, len(surf),
                      P(cen), P(bnd), P(idx, torch.int32), P(idx, torch.int32), 2, 1,
                      _lib.stream())
        else:
            _lib.call("dt_contact_closest", prm.c_struct(), dev(m32), 1, len(surf), P(cen),
                      P(idx, torch.int32), 2, 1, _lib.stream())
        got = idx.cpu().numpy()
        np.testing.assert_array_equal(got[nodes_o, 1], ref[:, 0])
        untouched = ~occ_mask
        before = ref_full[untouched] if batched else -7
        assert (got[untouched, 1] == before).all() and (got[:, 0] == -7).all()


@pytest.mark.parametrize("node_set,batched", [(False, False), (True, False), (False, True),
                                              (True, True)])
def test_closest_triangle_ties_envs_and_node_set(gpu, node_set, batched):
    """Closest triangle for two stacked environments, each searching only its
    own pad, on lattice-aligned slabs (the two triangles of a cell are
    equidistant from many node centres: the first minimum must win), over
    every node or over the active bricks of a binning, by the plain search or
    the one that skips runs of triangles by their bounding boxes; bit-exact
    against the oracle's float32 brute force per environment."""
    import ctypes
    from difftactile_amd.batch import concat_pad_meshes
    from tests.test_gpu_mpm import _workspace
    n_env, n_grid = 2, 32
    prm = MPMParams(n_grid=n_grid, rho=1000.0, dt=1e-4)
    prm.n_env = n_env
    dx = np.float32(prm.dx)
    meshes = []
    for e in range(n_env):
        nodes, f2v, tris, layer, _ = slab_mesh(12 + 4 * e, 10, 3, h=float(dx),
                                               origin=(float(dx) * (9 + e), 0.5, float(dx) * 8))
        meshes.append((nodes.astype(np.float32), f2v, tris, layer))
    nodes, _, tris, _, _, tri_start = concat_pad_meshes(meshes)
    rng = np.random.default_rng(5)
    npart = 6000
    x = rng.uniform(0.2, 0.62, size=(npart, 3)).astype(np.float32)
    env = rng.integers(0, n_env, npart).astype(np.int32)
    # occupancy: the nodes of the particles' 3x3x3 stencils
    base = np.clip(om.cell_base_np(x, prm.inv_dx).astype(np.int64), 0, n_grid - 3)
    base[:, 0] += env.astype(np.int64) * n_grid
    off = np.stack(np.meshgrid(*[np.arange(3)] * 3, indexing="ij"), -1).reshape(-1, 3)
    st = (base[:, None, :] + off[None]).reshape(-1, 3)
    occ = np.unique((st[:, 0] * n_grid + st[:, 1]) * n_grid + st[:, 2])
    nn = n_env * n_grid ** 3
    m = np.zeros(nn, np.float32)
    m[occ] = 1.0
    c_prm = prm.c_struct()
    env_t = dev(env, torch.int32)
    c_prm.env_id = env_t.data_ptr()
    if node_set:
        ws, nbytes = _workspace(c_prm, npart)
        _lib.call("dt_mpm_bin", c_prm, npart, dev(x), ws, nbytes, _lib.stream())
        a, c = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(_lib.lib().dt_mpm_active_bricks(c_prm, npart, ws.data_ptr(), nbytes,
                                                   ctypes.byref(a), ctypes.byref(c)))
        c_prm.active_bricks, c_prm.n_active_bricks = a.value, c.value
    cen = torch.empty(len(tris), 4, device="cuda")
    _lib.call("dt_contact_centroids", len(tris), dev(tris, torch.int32), dev(nodes), P(cen),
              _lib.stream())
    idx = torch.full((nn,), -7, dtype=torch.int32, device="cuda")
    ts = dev(tri_start, torch.int32)
    if batched:
        bnd = torch.empty(_lib.tri_bounds_count(len(tris)), 4, device="cuda")
        _lib.call("dt_contact_tri_batches", len(tris), P(cen), P(bnd), _lib.stream())
        # a hint array of arbitrary entries (in range or not) changes nothing
        hint = dev(np.random.default_rng(9).integers(-5, len(tris) + 5, nn), torch.int32)
        _lib.call("dt_contact_closest_batched", c_prm, dev(m), 1, ts, len(tris), P(cen), P(bnd),
                  P(hint, torch.int32), P(idx, torch.int32), 1, 0, _lib.stream())
    else:
        _lib.call("dt_contact_closest_env", c_prm, dev(m), 1, ts, len(tris), P(cen),
                  P(idx, torch.int32), 1, 0, _lib.stream())
    got = idx.cpu().numpy()
    i, j, k = occ // (n_grid * n_grid), (occ // n_grid) % n_grid, occ % n_grid
    q = (np.stack([i % n_grid, j, k], 1).astype(np.float32) + np.float32(0.5)) * dx
    ties = 0
    for e in range(n_env):
        sel = (i // n_grid) == e
        t0, t1 = tri_start[e], tri_start[e + 1]
        ref = oc.find_closest_np(q[sel], nodes, tris[t0:t1])
        np.testing.assert_array_equal(got[occ[sel]], ref + t0)
        # count nodes whose best distance is attained more than once
        cc = (nodes[tris[t0:t1, 0]] + nodes[tris[t0:t1, 1]]) + nodes[tris[t0:t1, 2]]
        cc = np.float32(1.0 / 3.0) * cc
        dd = cc[None] - q[sel][:, None]
        dist = np.sqrt(((dd[..., 0] * dd[..., 0] + dd[..., 1] * dd[..., 1]) +
                        dd[..., 2] * dd[..., 2]) + np.float32(1e-10))
        ties += int(((dist == dist.min(1, keepdims=True)).sum(1) > 1).sum())
    assert ties > 100
    untouched = np.ones(nn, bool)
    untouched[occ] = False
    assert (got[untouched] == -7).all()


def _build_facade(prm, mesh, pos, x, v, S=3):
    nodes, f2v, surf, layer, od = mesh
    obj = MPMObj(dt=prm.dt, sub_steps=S, particles=x, n_grid=prm.n_grid, rho=prm.rho,
                 gravity=prm.gravity)
    obj.set_material(5e4, 0.3)
    obj.set_state(x, v)
    pad = FEMDomeSensor(dt=prm.dt, sub_steps=S, mesh=(nodes, f2v, surf, layer, od), E=1e5, nu=0.45,
                        mass=2e-4)
    pad.init()
    pad.pos.data[0] = torch.tensor(pos).cuda()
    pad.virtual_pos.data[0] = torch.tensor(pos).cuda()
    cv = np.zeros((len(nodes), 3), np.float32)
    cv[:, 1] = -0.5
    pad.control_vel.from_numpy(cv)
    pad.set_vel(0)
    return obj, pad


@pytest.mark.parametrize("clamp", [False, True])
def test_coupled_substep_forward_and_adjoint_match_oracle(gpu, clamp):
    """One coupled substep through the per-kernel facade (GridContact.update /
    update_grad) vs the oracle; with clamp, clamp_grid (grasp_plastic.py:191-202)
    clips d/d grid_m after grid_op.grad and the pads' d/d pos, d/d vel after
    update2.grad, and the seeds are large enough that it changes the result."""
    prm, mesh, pos, x, v = _contact_scene(n_particles=12000, seed=3)
    nodes, f2v, surf, layer, od = mesh
    obj, pad = _build_facade(prm, mesh, pos, x, v)
    sc = GridContact(obj, [pad], kn=1e3, kd=10.0, kt=50.0, friction_coeff=0.5, clamp_grad=clamp)
    sc.reset()
    # update() adds the tet forces into vel[0] in place (fem_sensor.py:439-440):
    # keep the pre-update velocity as the oracle's input
    vel_before = h64(pad.vel.data[0])
    sc.update(0)
    torch.cuda.synchronize()
    # oracle, FP64, same inputs (B and the contact indices from float32 data); the
    # per-kernel facade chain uses the reference's svd_grad
    op = oracle_params(prm)
    op.svd_exact = False
    n = len(x)
    mats = om.material_arrays([om.Material(5e4, 0.3)], np.zeros(n, int), F64)
    # (mu, lam) as leaves: p2g.grad accumulates mu_0.grad / lambda_0.grad
    mu_t = torch.tensor([mats[0][0].item()], dtype=F64, requires_grad=True)
    lam_t = torch.tensor([mats[1][0].item()], dtype=F64, requires_grad=True)
    zi = torch.zeros(n, dtype=torch.int64)
    mats = (mu_t[zi], lam_t[zi]) + tuple(mats[2:])
    st = [torch.tensor(x, dtype=F64, requires_grad=True), torch.tensor(v, dtype=F64, requires_grad=True),
          torch.zeros(n, 3, 3, dtype=F64, requires_grad=True),
          torch.eye(3, dtype=F64).repeat(n, 1, 1).requires_grad_(True)]
    padd = dict(pos=torch.tensor(pos, dtype=F64, requires_grad=True),
                vel=vel_before.clone().requires_grad_(True),
                vpos=torch.tensor(pos, dtype=F64), cv=h64(pad.control_vel.data), cv_next=None,
                f2v=torch.tensor(f2v).long(), B=h64(pad.B.data), layer=torch.tensor(layer),
                tris=surf, out_dir=torch.tensor(od, dtype=F64), mu=float(pad.mu[None]),
                lam=float(pad.lam[None]), mass=pad.mass, dt=prm.dt)
    out, pads_out, idx_ref, nodes_ref = ocp.coupled_substep(op, tuple(st), mats, [padd],
                                                           (1e3, 10.0, 50.0, 0.5), clamp=clamp)
    got_idx = sc.contact_idx.data[0].reshape(-1, 1).cpu().numpy()
    np.testing.assert_array_equal(got_idx[nodes_ref, 0], idx_ref[:, 0])
    assert (obj.grid_f.data[0].abs().sum(-1) > 0).sum() > 50, "scene must make contact"
    for fld, want in zip((obj.x_0, obj.v_0, obj.C_0, obj.F_0), out):
        assert rel_l2(h64(fld.data[1]), want.detach()) < 1e-5
    for fld, want in zip((pad.pos, pad.vel, pad.virtual_pos), pads_out[0]):
        assert rel_l2(h64(fld.data[1]), want.detach()) < 1e-5
    # adjoint: seed frame-1 gradients, run update_grad, compare frame-0 gradients
    rng = np.random.default_rng(9)
    seeds = {}
    scale = 30.0 if clamp else 1.0
    for name, fld in (("x", obj.x_0), ("v", obj.v_0), ("C", obj.C_0), ("F", obj.F_0),
                      ("pos", pad.pos), ("vel", pad.vel), ("vpos", pad.virtual_pos)):
        s = (scale * rng.normal(size=fld.data[1].shape)).astype(np.float32)
        fld.grad.data[1] = torch.tensor(s).cuda()
        seeds[name] = torch.tensor(s, dtype=F64)
    sc.update_grad(0)
    torch.cuda.synchronize()
    loss = sum((o * seeds[k]).sum() for o, k in zip(out, ("x", "v", "C", "F")))
    loss = loss + sum((o * seeds[k]).sum() for o, k in zip(pads_out[0], ("pos", "vel", "vpos")))
    leaves = st + [padd["pos"], padd["vel"], mu_t, lam_t]
    gr = torch.autograd.grad(loss, leaves, retain_graph=clamp)
    if clamp:
        # the clamp must bite: the unclamped oracle gradient differs
        out_u, pads_u, _, _ = ocp.coupled_substep(op, tuple(st), mats, [padd],
                                                  (1e3, 10.0, 50.0, 0.5), clamp=False)
        loss_u = sum((o * seeds[k]).sum() for o, k in zip(out_u, ("x", "v", "C", "F")))
        loss_u = loss_u + sum((o * seeds[k]).sum()
                              for o, k in zip(pads_u[0], ("pos", "vel", "vpos")))
        gu = torch.autograd.grad(loss_u, leaves)
        assert max(rel_l2(a, b) for a, b in zip(gr, gu)) > 1e-2
    gots = [obj.x_0.grad.data[0], obj.v_0.grad.data[0], obj.C_0.grad.data[0],
            obj.F_0.grad.data[0], pad.pos.grad.data[0], pad.vel.grad.data[0],
            obj.mu_0.grad.data.reshape(1), obj.lambda_0.grad.data.reshape(1)]
    for k, (got, want) in enumerate(zip(gots, gr)):
        err = rel_l2(h64(got), want)
        assert err < 1e-4, (k, err)


@pytest.mark.parametrize("sort,sparse,resident,clamp",
                         [(False, False, "none", False), (False, True, "none", False),
                          (True, True, "none", False), (True, True, "all", False),
                          (True, True, "first", False), (False, False, "none", True),
                          (True, True, "first", True)])
def test_coupled_rollout_gradient_matches_oracle(gpu, sort, sparse, resident, clamp):
    """Fused coupled rollout (object + pad + contact) vs the oracle's reverse
    mode over the same composition, including d/d(pad velocity sequence);
    dense object stages, binned brick passes on the caller's order, and binned
    passes on cell-sorted particles.  resident: which of the two segments the
    forward pass keeps for the backward pass (none: both recomputed from
    checkpoints; first: the first kept, the second recomputed).  clamp: the
    +-1000 adjoint clipping of clamp_grid (grasp_plastic.py:191-202), which
    this scene's gradients exceed (checked against the unclamped oracle)."""
    from difftactile_amd.coupled import CoupledSim, Pad
    from difftactile_amd.params import Material
    from difftactile_amd.sim import MPMSim
    prm, mesh, pos, x, v = _contact_scene(n_particles=6000, seed=5)
    nodes, f2v, surf, layer, od = mesh
    n = len(x)
    steps = 6
    mu, lam = 1e5 / 2 / 1.45, 1e5 * 0.45 / 1.45 / 0.1
    mass = 2e-4
    sim = MPMSim(prm, [Material(5e4, 0.3)], n, checkpoint_every=4, sort=sort, sparse_grid=sparse)
    pad = Pad(nodes.astype(np.float32), f2v, surf, layer, od, mu, lam, mass, prm.dt)
    cs = CoupledSim(sim, [pad], kn=1e3, kd=10.0, kt=50.0, mu_f=0.5, clamp_grad=clamp)
    # segments of 4 and 2 substeps: 5 and 3 kept substeps' worth of bytes (a
    # compact store's size follows the active bricks of its segment's start)
    cap = cs.store_capacity(dev(x).t().contiguous()) if sparse else None
    sim.resident_bytes = {"none": 0, "all": 8 * cs.substep_bytes(),
                          "first": 5 * cs.substep_bytes(cap)}[resident]
    rng = np.random.default_rng(11)
    cv = np.zeros((steps + 1, len(nodes), 3), np.float32)
    cv[:, :, 1] = -0.5 + 0.05 * rng.normal(size=(steps + 1, 1))
    st = [dev(x).requires_grad_(True), dev(v).requires_grad_(True),
          torch.zeros(n, 3, 3, device="cuda", requires_grad=True),
          torch.eye(3, device="cuda").repeat(n, 1, 1).requires_grad_(True)]
    pvel0 = cv[0].copy()
    pst = (dev(pos).requires_grad_(True), dev(pvel0).requires_grad_(True), dev(pos))
    cvt = dev(cv).requires_grad_(True)
    out = cs.rollout(*st, [pst], [cvt], steps)
    w = [rng.normal(size=t.shape).astype(np.float32) for t in out[:4]]
    wp = [rng.normal(size=(len(nodes), 3)).astype(np.float32) for _ in range(3)]
    loss = sum((o * dev(wi)).sum() for o, wi in zip(out[:4], w))
    loss = loss + sum((o * dev(wi)).sum() for o, wi in zip(out[4][0], wp))
    leaves = st + [pst[0], pst[1], cvt]
    grads = torch.autograd.grad(loss, leaves)
    assert cs.last_resident[0] == {"none": 0, "all": 2, "first": 1}[resident]
    # oracle
    op = oracle_params(prm)
    mats = om.material_arrays([om.Material(5e4, 0.3)], np.zeros(n, int), F64)
    so = [torch.tensor(x, dtype=F64, requires_grad=True), torch.tensor(v, dtype=F64, requires_grad=True),
          torch.zeros(n, 3, 3, dtype=F64, requires_grad=True),
          torch.eye(3, dtype=F64).repeat(n, 1, 1).requires_grad_(True)]
    ppos = torch.tensor(pos, dtype=F64, requires_grad=True)
    pvel = torch.tensor(pvel0, dtype=F64, requires_grad=True)
    cvo = torch.tensor(cv, dtype=F64, requires_grad=True)
    Bo = h64(pad.B)

    def oracle_rollout(clamp_o):
        state, padstate = tuple(so), (ppos, pvel, torch.tensor(pos, dtype=F64))
        for f in range(steps):
            pd = dict(pos=padstate[0], vel=padstate[1], vpos=padstate[2], cv=cvo[f],
                      cv_next=cvo[f + 1], f2v=torch.tensor(f2v).long(), B=Bo,
                      layer=torch.tensor(layer), tris=surf, out_dir=torch.tensor(od, dtype=F64),
                      mu=mu, lam=lam, mass=mass, dt=prm.dt)
            state, pads_o, _, _ = ocp.coupled_substep(op, state, mats, [pd],
                                                      (1e3, 10.0, 50.0, 0.5), clamp=clamp_o)
            padstate = pads_o[0]
        return state, padstate

    state, padstate = oracle_rollout(clamp)
    # trajectory error after 6 stiff contact substeps (the per-substep 1e-5 bound
    # is checked by the single-substep tests)
    for got, want in zip(out[:4], state):
        assert rel_l2(h64(got), want.detach()) < 1e-4
    for got, want in zip(out[4][0], padstate):
        assert rel_l2(h64(got), want.detach()) < 1e-4
    loss_o = sum((o * torch.tensor(wi, dtype=F64)).sum() for o, wi in zip(state, w))
    loss_o = loss_o + sum((o * torch.tensor(wi, dtype=F64)).sum() for o, wi in zip(padstate, wp))
    go = torch.autograd.grad(loss_o, so + [ppos, pvel, cvo])
    for k, (a, b) in enumerate(zip(grads, go)):
        err = rel_l2(h64(a), b)
        assert err < 1e-3, (k, err)
    if clamp:
        su, pu = oracle_rollout(False)
        loss_u = sum((o * torch.tensor(wi, dtype=F64)).sum() for o, wi in zip(su, w))
        loss_u = loss_u + sum((o * torch.tensor(wi, dtype=F64)).sum() for o, wi in zip(pu, wp))
        gu = torch.autograd.grad(loss_u, so + [ppos, pvel, cvo])
        assert max(rel_l2(a, b) for a, b in zip(go, gu)) > 1e-2, "the clamp must bite"
