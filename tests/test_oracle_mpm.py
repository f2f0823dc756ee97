"""Pins the MPM oracle (CPU only): each check is one a plausible mistake would fail.

Closed forms, conservation identities, np.linalg.svd, brute force and FP64
central differences -- not a retyping of the oracle's own formulas.
"""
import numpy as np
import pytest
import torch

from oracle import mpm as om

F64 = torch.float64


def _blob(n, seed=0, n_grid=32):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.3, 0.7, size=(n, 3))
    v = rng.normal(0, 0.3, size=(n, 3))
    C = rng.normal(0, 2.0, size=(n, 3, 3))
    F = np.eye(3) + 0.1 * rng.normal(size=(n, 3, 3))
    return [torch.tensor(a, dtype=F64) for a in (x, v, C, F)]


# ---------------------------------------------------------------------------
# cell keys / sort
# ---------------------------------------------------------------------------
def test_cell_base_rounding_rule_brute_force():
    rng = np.random.default_rng(0)
    inv_dx = 6.4                         # space_scale 10 / 64: inv_dx not exact in f32
    x = rng.uniform(0, 10, size=(20000, 3)).astype(np.float32)
    b = om.cell_base_np(x, inv_dx)
    # brute force: the product and the subtraction rounded to float32 one at a time
    for i in range(0, 20000, 97):
        for a in range(3):
            t = np.float32(np.float64(x[i, a]) * np.float64(np.float32(inv_dx)))
            u = np.float32(np.float64(t) - 0.5)
            assert b[i, a] == int(u)
    # truncation toward zero, not floor: -0.3 -> 0
    assert om.cell_base_np(np.array([[0.01, 0.0, 0.07]], np.float32), 10.0).tolist() == [[0, 0, 0]]


def test_cell_keys_brick_order_and_stable_sort():
    n = 64
    x = np.array([[0.5, 0.5, 0.5], [0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [0.99, 0.0, 0.0]], np.float32)
    keys, perm = om.sort_by_cell_np(x, 64.0, n)
    b = np.clip(om.cell_base_np(x, 64.0), 0, n - 1)
    brick = ((b[:, 0] // 4) * 16 + b[:, 1] // 4) * 16 + b[:, 2] // 4
    local = ((b[:, 0] % 4) * 4 + b[:, 1] % 4) * 4 + b[:, 2] % 4
    assert keys.tolist() == (brick * 64 + local).tolist()
    # equal keys keep input order
    assert perm.tolist() == sorted(range(4), key=lambda i: (keys[i], i))
    # keys are a bijection of cells
    ii = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    k = om.cell_keys_np((ii + 0.75).astype(np.float32) / n, float(n), n)
    assert len(np.unique(k)) == n ** 3 and k.max() == n ** 3 - 1


# ---------------------------------------------------------------------------
# weights
# ---------------------------------------------------------------------------
def test_bspline_partition_of_unity_and_zero_first_moment():
    x = torch.rand(5000, 3, dtype=F64) * 0.5 + 0.25
    base, fx, w = om.stencil(x, 32.0)
    assert torch.allclose(w.sum(1), torch.ones_like(fx), atol=1e-14)
    off = torch.arange(3, dtype=F64)[None, :, None]
    assert torch.allclose((w * (off - fx[:, None, :])).sum(1), torch.zeros_like(fx), atol=1e-13)
    assert (fx >= 0.5).all() and (fx < 1.5 + 1e-12).all()


# ---------------------------------------------------------------------------
# SVD
# ---------------------------------------------------------------------------
def test_svd_convention_against_numpy():
    rng = np.random.default_rng(1)
    F = rng.normal(size=(500, 3, 3))
    F[:50] = np.diag([1.0, 2.0, -0.5])                # negative determinant
    U, s, V = om.svd3_np(F)
    np.testing.assert_allclose(U @ (s[:, :, None] * np.swapaxes(V, 1, 2)), F, atol=1e-12)
    np.testing.assert_allclose(np.linalg.det(U), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.linalg.det(V), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.abs(s), np.linalg.svd(F, compute_uv=False), atol=1e-12)
    assert (np.sign(s[:, 2]) == np.sign(np.linalg.det(F))).all()
    assert (s[:, 0] >= s[:, 1]).all() and (s[:, 1] >= np.abs(s[:, 2])).all()


def test_svd_grad_reference_formula_matches_fd_and_torch():
    """svd_grad of mpm_elastic.py:157-173 is the exact SVD adjoint for distinct sigma."""
    rng = np.random.default_rng(2)
    F = torch.tensor(np.diag([2.0, 1.3, 0.6]) + 0.1 * rng.normal(size=(3, 3)), dtype=F64)
    F = F[None].requires_grad_(True)
    wU, ws, wV = (torch.tensor(rng.normal(size=s), dtype=F64) for s in ((3, 3), (3,), (3, 3)))

    def loss_sign_invariant(F):
        U, s, V = om.svd3(F)
        # a loss invariant to the sign/gauge of singular vectors
        R = U @ V.transpose(-1, -2)
        return (R[0] * wU).sum() + (s[0] * ws).sum() + ((U[0] * s[0]) @ V[0].T * wV).sum()

    g = torch.autograd.grad(loss_sign_invariant(F), F)[0][0]
    # central differences
    fd = torch.zeros(3, 3, dtype=F64)
    h = 1e-6
    for i in range(3):
        for j in range(3):
            E = torch.zeros(1, 3, 3, dtype=F64)
            E[0, i, j] = h
            fd[i, j] = (loss_sign_invariant(F.detach() + E) - loss_sign_invariant(F.detach() - E)) / (2 * h)
    assert torch.allclose(g, fd, rtol=1e-6, atol=1e-7)
    # and against torch's own SVD backward
    Ft = F.detach().clone().requires_grad_(True)
    Ut, st, Vht = torch.linalg.svd(Ft[0])
    d = torch.sign(torch.det(Ut)) * torch.sign(torch.det(Vht))
    Rt = Ut @ Vht
    lt = (Rt * wU).sum() + (st * torch.tensor([1.0, 1.0, d.item()], dtype=F64) * ws).sum() + (Ft[0] * wV).sum()
    gt = torch.autograd.grad(lt, Ft)[0][0]
    assert torch.allclose(g, gt, rtol=1e-8, atol=1e-10)


# ---------------------------------------------------------------------------
# constitutive models
# ---------------------------------------------------------------------------
def _mats(kind, n, **kw):
    m = om.Material(5e4, 0.3, kind, **kw)
    return om.material_arrays([m], np.zeros(n, int), F64), m


def test_fixed_corotated_rotation_is_stress_free_and_is_energy_gradient():
    n = 64
    mats, m = _mats(om.KIND_ELASTIC, n)
    rng = np.random.default_rng(3)
    q = torch.tensor(np.linalg.qr(rng.normal(size=(n, 3, 3)))[0], dtype=F64)
    q = q * torch.sign(torch.linalg.det(q))[:, None, None]
    U, s, V = om.svd3(q)
    tau, _ = om.particle_stress(q, U, s, V, *mats)
    assert tau.abs().max() < 1e-9
    # tau = P F^T with P = dPsi/dF, Psi = mu |F - R|^2 + lam/2 (J - 1)^2
    F = torch.eye(3, dtype=F64) + 0.2 * torch.tensor(rng.normal(size=(n, 3, 3)), dtype=F64)

    def psi(F):
        U, s, V = om.svd3(F)
        R = U @ V.transpose(-1, -2)
        J = torch.linalg.det(F)
        return (m.mu * ((F - R) ** 2).sum((-1, -2)) + m.lam / 2 * (J - 1) ** 2).sum()

    P = torch.zeros_like(F)
    h = 1e-6
    for i in range(3):
        for j in range(3):
            E = torch.zeros_like(F)
            E[:, i, j] = h
            # per-particle FD (psi is a sum of independent terms)
            P[:, i, j] = (om_psi_each(F + E, m) - om_psi_each(F - E, m)) / (2 * h)
    U, s, V = om.svd3(F)
    tau, _ = om.particle_stress(F, U, s, V, *_mats(om.KIND_ELASTIC, n)[0])
    assert torch.allclose(tau, P @ F.transpose(-1, -2), rtol=1e-6, atol=1e-3)


def om_psi_each(F, m):
    U, s, V = om.svd3(F)
    R = U @ V.transpose(-1, -2)
    J = torch.linalg.det(F)
    return m.mu * ((F - R) ** 2).sum((-1, -2)) + m.lam / 2 * (J - 1) ** 2


def test_von_mises_return_map_closed_forms():
    n = 400
    mats, m = _mats(om.KIND_VON_MISES, n, yield_stress=50.0)
    rng = np.random.default_rng(4)
    scale = np.where(np.arange(n) < n // 2, 0.3, 2e-4)[:, None, None]
    F = torch.eye(3, dtype=F64) + torch.tensor(scale * rng.normal(size=(n, 3, 3)), dtype=F64)
    F = F * torch.sign(torch.linalg.det(F))[:, None, None]
    U, s, V = om.svd3(F)
    Fp, yielded = om.von_mises(F, U, s, V, mats[0], mats[3])
    assert yielded.any() and (~yielded).any()
    _, sp, _ = om.svd3(Fp)
    e = torch.log(torch.clamp(s, min=0.05))
    ep = torch.log(sp)
    # sort-independent checks: trace preserved, deviatoric norm on the yield surface
    tr, trp = e.sum(1), ep.sum(1)
    assert torch.allclose(tr[yielded], trp[yielded], atol=1e-10)
    dev = ep - trp[:, None] / 3
    # |dev eps'| = tau_y/(2 mu) * |eh| / |eh|_eps  (the norm carries eps = 1e-8)
    eh = e - tr[:, None] / 3
    ehn = torch.sqrt((eh ** 2).sum(1) + 1e-8)
    target = m.yield_stress / (2 * m.mu) * torch.sqrt((eh ** 2).sum(1)) / ehn
    nrm = torch.sqrt((dev ** 2).sum(1))
    assert torch.allclose(nrm[yielded], target[yielded], rtol=1e-8)
    assert torch.equal(Fp[~yielded], F[~yielded])


def test_drucker_prager_return_map_closed_forms():
    n = 600
    mats, m = _mats(om.KIND_DRUCKER_PRAGER, n, friction_deg=30.0)
    rng = np.random.default_rng(5)
    F = torch.eye(3, dtype=F64) + 0.25 * torch.tensor(rng.normal(size=(n, 3, 3)), dtype=F64)
    F = F * torch.sign(torch.linalg.det(F))[:, None, None]
    U, s, V = om.svd3(F)
    Fp, changed = om.drucker_prager(F, U, s, V, mats[0], mats[1], mats[4])
    e = torch.log(torch.clamp(s, min=0.05))
    tr = e.sum(1)
    tension = tr >= 0
    # tension: projected to the cone tip, F_p = U V^T (a rotation)
    R = U @ V.transpose(-1, -2)
    assert torch.allclose(Fp[tension], R[tension], atol=1e-12)
    # shear-yielded: the returned strain sits on the cone |dev| = -alpha*(3 lam+2 mu)/(2 mu) tr
    sh = changed & ~tension
    assert sh.any()
    _, sp, _ = om.svd3(Fp[sh])
    ep = torch.log(sp)
    trp = ep.sum(1)
    dev = ep - trp[:, None] / 3
    nrm = torch.sqrt((dev ** 2).sum(1))
    k = (3 * m.lam + 2 * m.mu) / (2 * m.mu) * m.dp_alpha
    eh = e[sh] - tr[sh][:, None] / 3
    scale = torch.sqrt((eh ** 2).sum(1)) / torch.sqrt((eh ** 2).sum(1) + 1e-8)
    assert torch.allclose(nrm, -k * trp * scale, rtol=1e-8, atol=1e-12)
    assert torch.allclose(trp, tr[sh], atol=1e-10)   # volume-preserving projection
    # inside the cone: unchanged
    assert torch.equal(Fp[~changed], F[~changed])


# ---------------------------------------------------------------------------
# transfers
# ---------------------------------------------------------------------------
def test_p2g_conserves_mass_and_momentum():
    prm = om.MPMParams(n_grid=32, rho=1000.0, dt=1e-4)
    x, v, C, F = _blob(3000)
    mats = om.material_arrays([om.Material(5e4, 0.3)], np.zeros(3000, int), F64)
    Fn = om.compute_new_F(C, F, prm.dt)
    U, s, V = om.svd3(Fn)
    mv, m, _ = om.p2g(prm, x, v, C, Fn, U, s, V, mats)
    assert abs(m.sum().item() - 3000 * prm.p_mass) < 1e-12
    # affine and stress terms cancel in the total by the zero first moment
    assert torch.allclose(mv.sum(0), prm.p_mass * v.sum(0), rtol=1e-10, atol=1e-14)


def test_p2g_second_moment_fixes_the_stress_and_affine_terms():
    """For quadratic B-splines sum_i w_i d_i d_i^T = dx^2/4 I and sum_i w_i d_i = 0
    (d_i = x_i - x_p), so one particle's scatter has the second moment
    sum_i mv_i (x_i - x_p)^T = -dt p_vol tau + p_mass C dx^2/4 and
    sum_i m_i d_i d_i^T = p_mass dx^2/4 I (mpm_elastic.py:183-200).  tau comes
    from scipy's polar decomposition, not from the oracle."""
    import scipy.linalg
    prm = om.MPMParams(n_grid=16, rho=1000.0, dt=1e-4)
    n = prm.n_grid
    rng = np.random.default_rng(7)
    table = [om.Material(5e4, 0.3), om.Material(2e5, 0.25), om.Material(1e4, 0.45)]
    idx = torch.arange(n ** 3)
    node_x = torch.stack([idx // (n * n), (idx // n) % n, idx % n], 1).to(F64) * prm.dx
    for trial in range(24):
        mid = trial % 3
        x = torch.tensor(rng.uniform(0.3, 0.7, size=(1, 3)), dtype=F64)
        v = torch.tensor(rng.normal(0, 0.5, size=(1, 3)), dtype=F64)
        C = torch.tensor(rng.normal(0, 3.0, size=(1, 3, 3)), dtype=F64)
        F = np.eye(3) + 0.15 * rng.normal(size=(3, 3))
        assert np.linalg.det(F) > 0
        Ft = torch.tensor(F[None], dtype=F64)
        U, s, V = om.svd3(Ft)
        mats = om.material_arrays(table, np.array([mid]), F64)
        mv, m, _ = om.p2g(prm, x, v, C, Ft, U, s, V, mats)
        R, _ = scipy.linalg.polar(F)
        J = np.linalg.det(F)
        mu, lam = table[mid].mu, table[mid].lam
        tau = 2 * mu * (F - R) @ F.T + lam * J * (J - 1) * np.eye(3)
        d = node_x - x                                   # (n^3, 3)
        mom2 = (mv[:, :, None] * d[:, None, :]).sum(0).numpy()
        want = -prm.dt * prm.p_vol * tau + prm.p_mass * C[0].numpy() * prm.dx ** 2 / 4
        np.testing.assert_allclose(mom2, want, rtol=1e-9, atol=1e-12 * np.abs(want).max())
        m2 = (m[:, None, None] * d[:, :, None] * d[:, None, :]).sum(0).numpy()
        np.testing.assert_allclose(m2, prm.p_mass * prm.dx ** 2 / 4 * np.eye(3), atol=1e-18)


def test_g2p_reproduces_affine_fields_exactly():
    prm = om.MPMParams(n_grid=32, dt=1e-4)
    n = prm.n_grid
    idx = torch.arange(n ** 3)
    node_x = torch.stack([idx // (n * n), (idx // n) % n, idx % n], 1).to(F64) * prm.dx
    A = torch.tensor([[0.3, -1.0, 0.2], [0.5, 0.1, -0.7], [0.0, 2.0, 0.4]], dtype=F64)
    b = torch.tensor([0.1, -0.2, 0.3], dtype=F64)
    field = node_x @ A.T + b
    x = torch.rand(2000, 3, dtype=F64) * 0.5 + 0.25
    nx, nv, nC = om.g2p(prm, x, field)
    assert torch.allclose(nv, x @ A.T + b, atol=1e-12)
    assert torch.allclose(nC, A.expand(2000, 3, 3), atol=1e-10)
    assert torch.allclose(nx, x + prm.dt * nv, atol=1e-15)


def test_grid_op_closed_form_and_walls():
    prm = om.MPMParams(n_grid=8, dt=1e-3, gravity=(0.0, -9.8, 0.0), bound=3)
    n = 8
    m = torch.full((n ** 3,), 2e-3, dtype=F64)
    m[5] = 0.0                                   # unoccupied
    mv = torch.full((n ** 3, 3), 1e-3, dtype=F64)
    mv[:, 1] = -1e-3
    gf = torch.full((n ** 3, 3), 1e-4, dtype=F64)
    v = om.grid_op(prm, mv, m, gf)
    inv = 1 / (2e-3 + 1e-5)
    expect = torch.tensor([1e-3 * inv + 1e-4 * inv, -1e-3 * inv + 1e-4 * inv - 9.8e-3,
                           1e-3 * inv + 1e-4 * inv], dtype=F64)
    i, j, k = 4, 4, 4
    assert torch.allclose(v[(i * n + j) * n + k], expect)
    assert (v[5] == 0).all()
    # x: i<3 zeroes negative x-velocity only; i > n-3 zeroes positive
    assert v[(0 * n + 4) * n + 4, 0] > 0           # moving inward: kept
    assert v[(7 * n + 4) * n + 4, 0] == 0          # i=7 > 5 moving outward: zeroed
    assert v[(6 * n + 4) * n + 4, 0] == 0          # i=6 > 5
    assert v[(5 * n + 4) * n + 4, 0] != 0          # i=5 not > 5: kept
    assert v[(4 * n + 2) * n + 4, 1] == 0          # j=2 < 3 with v_y < 0: zeroed
    assert v[(4 * n + 3) * n + 4, 1] != 0


# ---------------------------------------------------------------------------
# the composed substep: reverse mode vs FP64 central differences
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("kind", [om.KIND_ELASTIC, om.KIND_VON_MISES, om.KIND_DRUCKER_PRAGER])
def test_substep_reverse_mode_matches_central_differences(kind, exact):
    prm = om.MPMParams(n_grid=16, rho=1000.0, dt=2e-5, svd_exact=exact)
    n = 60
    rng = np.random.default_rng(6)
    x = torch.tensor(0.5 + 0.08 * rng.normal(size=(n, 3)), dtype=F64)
    v = torch.tensor(0.3 * rng.normal(size=(n, 3)), dtype=F64)
    C = torch.tensor(rng.normal(size=(n, 3, 3)), dtype=F64)
    F = torch.tensor(np.eye(3) + 0.2 * rng.normal(size=(n, 3, 3)), dtype=F64)
    F = F * torch.sign(torch.linalg.det(F))[:, None, None]
    mats = om.material_arrays([om.Material(5e4, 0.3, kind, yield_stress=30.0)], np.zeros(n, int), F64)
    w = [torch.tensor(rng.normal(size=t.shape), dtype=F64) for t in (x, v, C, F)]

    def L(state):
        out = om.rollout(prm, state, mats, 2)
        return sum((o * wi).sum() for o, wi in zip(out, w))

    st = [t.clone().requires_grad_(True) for t in (x, v, C, F)]
    grads = torch.autograd.grad(L(tuple(st)), st)
    d = [torch.tensor(rng.normal(size=t.shape), dtype=F64) for t in (x, v, C, F)]
    ad = sum((g * di).sum() for g, di in zip(grads, d)).item()
    h = 1e-6
    with torch.no_grad():
        lp = L(tuple(t + h * di for t, di in zip((x, v, C, F), d))).item()
        lm = L(tuple(t - h * di for t, di in zip((x, v, C, F), d))).item()
    fd = (lp - lm) / (2 * h)
    assert abs(ad - fd) <= 1e-6 * max(1.0, abs(fd)), (ad, fd)


@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("kind", [om.KIND_ELASTIC, om.KIND_VON_MISES, om.KIND_DRUCKER_PRAGER])
def test_material_reverse_mode_matches_central_differences(kind, exact):
    """d loss / d (mu, lam) of a 2-substep rollout, reverse mode against FP64
    central differences (covers the stress terms and the return maps' yield
    threshold and cone slope)."""
    prm = om.MPMParams(n_grid=16, rho=1000.0, dt=2e-5, svd_exact=exact)
    n = 60
    rng = np.random.default_rng(8)
    x = torch.tensor(0.5 + 0.08 * rng.normal(size=(n, 3)), dtype=F64)
    v = torch.tensor(0.3 * rng.normal(size=(n, 3)), dtype=F64)
    C = torch.tensor(rng.normal(size=(n, 3, 3)), dtype=F64)
    F = torch.tensor(np.eye(3) + 0.2 * rng.normal(size=(n, 3, 3)), dtype=F64)
    F = F * torch.sign(torch.linalg.det(F))[:, None, None]
    mat = om.Material(5e4, 0.3, kind, yield_stress=30.0)
    _, _, kinds, yld, alpha = om.material_arrays([mat], np.zeros(n, int), F64)
    w = [torch.tensor(rng.normal(size=t.shape), dtype=F64) for t in (x, v, C, F)]

    def L(mu, lam):
        mats = (mu.expand(n), lam.expand(n), kinds, yld, alpha)
        out = om.rollout(prm, (x, v, C, F), mats, 2)
        return sum((o * wi).sum() for o, wi in zip(out, w))

    mu = torch.tensor(mat.mu, dtype=F64, requires_grad=True)
    lam = torch.tensor(mat.lam, dtype=F64, requires_grad=True)
    g_mu, g_lam = torch.autograd.grad(L(mu, lam), (mu, lam))
    with torch.no_grad():
        for g, which in ((g_mu, 0), (g_lam, 1)):
            h = 1e-6 * (mat.mu if which == 0 else mat.lam)
            dp = torch.tensor(h if which == 0 else 0.0, dtype=F64)
            dl = torch.tensor(h if which == 1 else 0.0, dtype=F64)
            fd = (L(mu + dp, lam + dl) - L(mu - dp, lam - dl)).item() / (2 * h)
            assert abs(g.item() - fd) <= 1e-5 * max(abs(fd), 1e-12), (which, g.item(), fd)
            assert abs(fd) > 0.0


def test_polar_rotation_adjoint_exact_at_near_repeated_singular_values():
    """_Polar's adjoint (M_ij - M_ji)/(s_i + s_j) against FP64 central differences,
    including F = I + O(1e-9) where the reference's clamp changes svd_grad."""
    rng = np.random.default_rng(8)
    for scale in (1e-9, 1e-3, 0.2):
        F = torch.eye(3, dtype=F64) + scale * torch.tensor(rng.normal(size=(1, 3, 3)), dtype=F64)
        W = torch.tensor(rng.normal(size=(1, 3, 3)), dtype=F64)

        def loss(F):
            U, s, V = om.svd3(F.detach())
            return (om._Polar.apply(F, U, s, V) * W).sum()

        def rot(F):
            U, s, Vh = np.linalg.svd(F[0].numpy())
            R = U @ Vh
            return torch.tensor(R, dtype=F64)

        Fv = F.clone().requires_grad_(True)
        g = torch.autograd.grad(loss(Fv), Fv)[0][0]
        fd = torch.zeros(3, 3, dtype=F64)
        h = 1e-6  # R(F) is smooth at F = I; rounding / h stays ~1e-10
        for i in range(3):
            for j in range(3):
                E = torch.zeros(1, 3, 3, dtype=F64)
                E[0, i, j] = h
                fd[i, j] = ((rot(F + E) - rot(F - E)) * W[0]).sum() / (2 * h)
        assert torch.allclose(g, fd, rtol=1e-5, atol=1e-6), scale


@pytest.mark.parametrize("kind", [om.KIND_ELASTIC, om.KIND_VON_MISES, om.KIND_DRUCKER_PRAGER])
def test_exact_adjoint_from_the_undeformed_state(kind):
    """From F = I (all singular values equal; Drucker-Prager particles start in the
    tension branch) the exact adjoint stays finite and matches FP64 central
    differences over 3 substeps."""
    prm = om.MPMParams(n_grid=16, rho=1000.0, dt=1e-5, svd_exact=True)
    n = 80
    rng = np.random.default_rng(9)
    x = torch.tensor(0.5 + 0.06 * rng.normal(size=(n, 3)), dtype=F64)
    v = torch.tensor(0.3 * rng.normal(size=(n, 3)), dtype=F64)
    C = torch.zeros(n, 3, 3, dtype=F64)
    F = torch.eye(3, dtype=F64).repeat(n, 1, 1)
    mats = om.material_arrays([om.Material(5e4, 0.3, kind, yield_stress=30.0)], np.zeros(n, int), F64)
    w = torch.tensor(rng.normal(size=(n, 3)), dtype=F64)

    def L(vv):
        out = om.rollout(prm, (x, vv, C, F), mats, 3)
        return (out[0] * w).sum() + (out[3] * out[3]).sum()

    vv = v.clone().requires_grad_(True)
    g = torch.autograd.grad(L(vv), vv)[0]
    assert torch.isfinite(g).all()
    d = torch.tensor(rng.normal(size=(n, 3)), dtype=F64)
    h = 1e-6
    with torch.no_grad():
        fd = (L(v + h * d) - L(v - h * d)) / (2 * h)
    ad = (g * d).sum()
    assert abs(ad - fd) <= 1e-5 * max(1.0, abs(fd.item())), (ad.item(), fd.item())


def test_bin_keys_name_the_cell_of_the_stencil_base():
    """The sparse path's binning key decodes to the stencil base the scatter
    uses (oracle stencil, no clamp needed inside the grid) and to the clamped
    base [0, n-3] outside it."""
    n = 32
    inv_dx = float(n)
    rng = np.random.default_rng(7)
    x = rng.uniform(2.0 / n, 1.0 - 3.0 / n, size=(5000, 3)).astype(np.float32)
    keys = om.bin_keys_np(x, inv_dx, n)
    nb = (n + 3) // 4
    brick, local = keys >> 6, keys & 63
    bx, by, bz = brick // (nb * nb), (brick // nb) % nb, brick % nb
    base = np.stack([bx * 4 + (local >> 4), by * 4 + ((local >> 2) & 3), bz * 4 + (local & 3)], 1)
    want, _, _ = om.stencil(torch.tensor(x), inv_dx)
    assert np.array_equal(base, want.numpy())
    assert np.array_equal(keys, om.cell_keys_np(x, inv_dx, n))
    out = rng.uniform(-0.2, 1.2, size=(5000, 3)).astype(np.float32)
    k2 = om.bin_keys_np(out, inv_dx, n)
    b2, l2 = k2 >> 6, k2 & 63
    dec = np.stack([(b2 // (nb * nb)) * 4 + (l2 >> 4), ((b2 // nb) % nb) * 4 + ((l2 >> 2) & 3),
                    (b2 % nb) * 4 + (l2 & 3)], 1)
    assert dec.min() >= 0 and dec.max() <= n - 3


def test_init_object_closed_forms():
    rng = np.random.default_rng(3)
    p = torch.tensor(rng.normal(size=(40, 3)))
    bv = torch.tensor([0.5, -1.0, 2.0], dtype=torch.float64)
    x, v, F = om.init_object(p, torch.eye(4, dtype=torch.float64), bv)
    assert torch.equal(x, p) and torch.equal(v, bv.expand(40, 3))
    assert torch.equal(F, torch.eye(3, dtype=torch.float64).expand(40, 3, 3))
    T = torch.eye(4, dtype=torch.float64)
    T[:3, 3] = torch.tensor([1.0, 2.0, -3.0])
    assert torch.allclose(om.init_object(p, T, bv)[0], p + T[:3, 3], atol=1e-15)
    T = torch.eye(4, dtype=torch.float64)
    T[:2, :2] = torch.tensor([[0.0, -1.0], [1.0, 0.0]])
    want = torch.stack([-p[:, 1], p[:, 0], p[:, 2]], 1)
    assert torch.allclose(om.init_object(p, T, bv)[0], want, atol=1e-15)
