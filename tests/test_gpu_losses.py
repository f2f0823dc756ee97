"""GPU parity of the loss kernels (centre of mass, tactile markers) with the oracle."""
import ctypes

import numpy as np
import pytest
import torch

from difftactile_amd import _lib
from difftactile_amd.sensor_model.markers import (FISHEYE, camera_points, interp_weights,
                                                   project_points_np, synthetic_markers)
from difftactile_amd.sensor_model.meshes import dome_mesh
from oracle import markers as ok
from tests.mpm_cases import rel_l2

pytestmark = pytest.mark.gpu
F64 = torch.float64


def fptr(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


@pytest.mark.parametrize("soa", [0, 1])
def test_com_and_adjoint(gpu, soa):
    """Float32 fixed-order mean: within float32 summation error of the FP64
    mean, the same bits on a second call, the adjoint g / n."""
    n = 300_001
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1, size=(n, 3)).astype(np.float32)
    xt = torch.tensor(x.T.copy() if soa else x).cuda()
    scratch = torch.empty(3 * _lib.REDUCE_BLOCKS, device="cuda")
    out = torch.empty(3, device="cuda")
    _lib.call("dt_loss_com", n, soa, xt, scratch, out, _lib.stream())
    np.testing.assert_allclose(out.cpu().numpy(), x.astype(np.float64).mean(0), rtol=2e-6)
    again = torch.empty(3, device="cuda")
    _lib.call("dt_loss_com", n, soa, xt, scratch, again, _lib.stream())
    assert torch.equal(out, again)
    g = torch.tensor([0.5, -2.0, 3.0], device="cuda")
    gx = torch.zeros_like(xt)
    _lib.call("dt_loss_com_adj", n, soa, g, gx, _lib.stream())
    want = np.tile(g.cpu().numpy() / n, (n, 1))
    got = gx.cpu().numpy()
    np.testing.assert_allclose(got.T if soa else got, want, rtol=1e-6)


def test_markers_forward_and_adjoint(gpu):
    nodes, f2v, surf, layer = dome_mesh()
    sid = np.where(layer == layer.max())[0].astype(np.int32)
    cam = project_points_np(camera_points(nodes[sid]))
    mk = synthetic_markers(cam, 136)
    idx, w = interp_weights(mk, cam)
    rng = np.random.default_rng(2)
    T = np.eye(4)
    T[:3, 3] = [0.01, -0.02, 0.005]
    pos = (nodes + 0.01 * rng.normal(size=nodes.shape)).astype(np.float32)
    vpos = nodes.astype(np.float32)
    Tk, Tp = fptr(T.reshape(-1))
    Kk, Kp = fptr(np.array(FISHEYE))
    ns, nm = len(sid), len(mk)
    d = lambda a, t=torch.float32: torch.tensor(a).to(t).cuda().contiguous()
    sid_d, idx_d, w_d, pos_d, vpos_d = d(sid, torch.int32), d(idx, torch.int32), d(w), d(pos), d(vpos)
    loc = torch.empty(ns, 2, device="cuda")
    vloc = torch.empty(ns, 2, device="cuda")
    pred = torch.empty(nm, 2, device="cuda")
    virt = torch.empty(nm, 2, device="cuda")
    _lib.call("dt_markers_fwd", ns, sid_d, pos_d, vpos_d, Tp, Kp, nm, 5, idx_d, w_d, loc, vloc,
              pred, virt, _lib.stream())
    pt = torch.tensor(pos, dtype=F64, requires_grad=True)
    args = (torch.as_tensor(sid).long(), torch.tensor(T, dtype=F64), torch.as_tensor(idx).long(),
            torch.tensor(w, dtype=F64))
    pr, vr = ok.extract_markers(pt, torch.tensor(vpos, dtype=F64), *args)
    assert rel_l2(pred.cpu().double(), pr.detach()) < 1e-6
    assert rel_l2(virt.cpu().double(), vr) < 1e-6
    gpred = rng.normal(size=(nm, 2)).astype(np.float32)
    gloc = torch.zeros(ns, 2, device="cuda")
    gpos = torch.zeros(len(nodes), 3, device="cuda")
    gpred_d = d(gpred)
    flat = idx.reshape(-1).astype(np.int64)
    start = np.concatenate([[0], np.cumsum(np.bincount(flat, minlength=ns))])
    entry = np.argsort(flat, kind="stable")
    _lib.call("dt_markers_adj", ns, sid_d, pos_d, Tp, Kp, nm, 5, d(start, torch.int32),
              d(entry, torch.int32), w_d, gpred_d, gloc, gpos, _lib.stream())
    gref = torch.autograd.grad((pr * torch.tensor(gpred, dtype=F64)).sum(), pt)[0]
    assert rel_l2(gpos.cpu().double(), gref) < 1e-4


def _dome_markers(seed=2):
    from difftactile_amd.losses import MarkerModel
    nodes, f2v, surf, layer = dome_mesh()
    sid = np.where(layer == layer.max())[0].astype(np.int32)
    T = np.eye(4)
    T[:3, 3] = [0.01, -0.02, 0.005]
    mm = MarkerModel(nodes, sid, T)
    rng = np.random.default_rng(seed)
    pos = (nodes + 0.01 * rng.normal(size=nodes.shape)).astype(np.float32)
    vpos = (nodes + 0.002 * rng.normal(size=nodes.shape)).astype(np.float32)
    return mm, nodes, sid, T, pos, vpos


def test_marker_motion_loss_and_gradient_match_the_oracle(gpu):
    """The reference's marker loss (beta times the squared mean marker offset,
    summed over two pads) and its gradient w.r.t. both pads' positions and
    virtual positions, against the FP64 oracle; the gradient repeats bit for
    bit."""
    from difftactile_amd.losses import marker_motion_loss
    mm, nodes, sid, T, pos, vpos = _dome_markers()
    mm2, _, _, _, pos2, vpos2 = _dome_markers(seed=5)
    beta = 3.0
    dev = torch.device("cuda")
    leaves = [torch.tensor(a, device=dev, requires_grad=True) for a in (pos, vpos, pos2, vpos2)]

    def run():
        loss = marker_motion_loss([mm, mm2], [(leaves[0], leaves[1]), (leaves[2], leaves[3])],
                                  beta)
        return loss, torch.autograd.grad(loss, leaves)

    loss, grads = run()
    loss_b, grads_b = run()
    assert torch.equal(loss, loss_b) and all(torch.equal(a, b) for a, b in zip(grads, grads_b))
    lo = [torch.tensor(a, dtype=F64, requires_grad=True) for a in (pos, vpos, pos2, vpos2)]
    args = (torch.as_tensor(sid).long(), torch.tensor(T, dtype=F64), mm.idx.long().cpu(),
            mm.w.cpu().double())
    ref = 0
    for p, vp in ((lo[0], lo[1]), (lo[2], lo[3])):
        pr, vr = ok.extract_markers(p, vp, *args)
        ref = ref + ok.marker_motion_loss(pr, vr, beta)
    assert abs(float(loss) - float(ref)) <= 1e-4 * abs(float(ref))
    gref = torch.autograd.grad(ref, lo)
    for got, want in zip(grads, gref):
        assert rel_l2(got.cpu().double(), want) < 1e-3


def test_contact_force_and_force_loss_match_the_oracle(gpu):
    """Two pads' contact forces (1/6 per contact-triangle vertex) and the
    sliding-ratio force loss with their gradients w.r.t. the external force
    fields, against the FP64 oracle; the pad-local frames are the gripper
    pads' quarter turns about z."""
    from difftactile_amd.losses import contact_force, force_loss
    from difftactile_amd.sensor_model.meshes import slab_mesh
    nodes, f2v, tris, layer, od = slab_mesh(20, 20, 4, h=0.01)
    rng = np.random.default_rng(8)
    dev = torch.device("cuda")
    q = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    inv_rots = [q.T, q]
    mu, beta = 4.9, 2.0
    exts = [(rng.normal(size=(len(nodes), 3)) + np.array(b)).astype(np.float32)
            for b in ((0.5, 2.0, 0.1), (-0.3, -1.5, 0.2))]
    tri_d = torch.tensor(np.asarray(tris), dtype=torch.int32, device=dev)
    el = [torch.tensor(e, device=dev, requires_grad=True) for e in exts]
    cfs = [contact_force(tri_d, e) for e in el]
    loss = force_loss(cfs, inv_rots, mu, beta)
    grads = torch.autograd.grad(loss, el)
    eo = [torch.tensor(e, dtype=F64, requires_grad=True) for e in exts]
    cfo = [ok.contact_force(e, tris) for e in eo]
    for got, want in zip(cfs, cfo):
        assert rel_l2(got.detach().cpu().double(), want.detach()) < 1e-5
    ref = ok.force_loss(cfo, inv_rots, mu, beta)
    assert float(ref) > 0
    assert abs(float(loss) - float(ref)) <= 1e-5 * float(ref)
    gref = torch.autograd.grad(ref, eo)
    for got, want in zip(grads, gref):
        assert float(want.abs().max()) > 0
        assert rel_l2(got.cpu().double(), want) < 1e-5
    # below the shear gate: zero loss and zero gradient
    small = [torch.zeros(len(nodes), 3, device=dev, requires_grad=True) for _ in range(2)]
    l0 = force_loss([contact_force(tri_d, e) for e in small], inv_rots, mu, beta)
    g0 = torch.autograd.grad(l0, small)
    assert float(l0) == 0.0 and all(float(g.abs().max()) == 0.0 for g in g0)
