"""Oracle: MLS-MPM object path (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

Restates, in torch on the CPU (float32 or float64), the per-substep kernels of
``difftactile/object_model/mpm_elastic.py`` (elastic, fixed-corotated),
``mpm_plastic.py`` (von Mises return map) and ``multi_obj.py`` (per-particle
material id), plus a Drucker-Prager return map that has no reference
implementation (Klar et al. 2016, "Drucker-Prager Elastoplasticity for Sand
Animation"); that one is pinned only by its closed forms and FP64 central
differences.

Layout follows the reference's ``to_numpy()`` shapes: particles ``(N, 3)`` and
``(N, 3, 3)``; grid flattened row-major ``(n**3, 3)`` / ``(n**3,)`` with node
``(i, j, k)`` at ``(i * n + j) * n + k``.

Pinning (tests/test_oracle_mpm.py):
  init_object    -- identity pose keeps the rest particles; a pure translation
                    shifts them; a quarter turn about z maps (x, y, z) to
                    (-y, x, z); v is ball_vel and F the identity everywhere.
  cell_base      -- closed form trunc(f32(x) * f32(inv_dx) - 0.5f) checked by a
                    brute-force float64 re-derivation on lattices; sort order
                    against np.argsort(kind="stable").
  weights        -- partition of unity and zero first moment (exact identities).
  svd3           -- U S V^T = F, U, V in SO(3), ordering, vs np.linalg.svd.
  svd_grad_ref   -- FP64 central differences of scalar losses of (U, S, V) and
                    the dot-product test; agreement with torch's own SVD
                    backward on well-separated singular values.
  p2g            -- mass and momentum conservation (affine term cancels by the
                    zero first moment), FP64 FD of the adjoint.
  grid_op        -- per-node closed form and the boundary truth table.
  g2p            -- uniform field -> (v, 0); affine field v = A x + b -> C = A.
  fixed-corotated stress -- tau(R) = 0 for rotations; tau = P F^T with
                    P = dPsi/dF by FP64 FD of Psi = mu |F-R|^2 + lam/2 (J-1)^2.
  von Mises      -- after yield |dev eps'| = tau_y / (2 mu), tr eps' = tr eps;
                    identity inside the yield surface.
  Drucker-Prager -- tension -> tip (sigma' = 1); inside cone -> identity;
                    projected point lies on the cone; FP64 FD.
"""
import math

import numpy as np
import torch

# ---------------------------------------------------------------------------
# parameters
# ---------------------------------------------------------------------------

KIND_ELASTIC = 0      # fixed corotated, J = det(Sigma)      (mpm_elastic.py:183-185)
KIND_VON_MISES = 1    # von Mises in Hencky strain, J = det(F_p) (mpm_plastic.py:150-166,221-256)
KIND_DRUCKER_PRAGER = 2  # sand return map, J = det(F_p)       (no reference implementation)


class MPMParams:
    """Scalar parameters of one MPM object (``mpm_elastic.py:19-66``).

    ``dx = space_scale / n_grid``; ``p_vol = (dx * obj_scale) ** 3``;
    ``p_mass = p_vol * rho * obj_scale``.  Python floats, as the reference
    keeps them; kernels receive them rounded to the working dtype.
    """

    def __init__(self, n_grid=64, space_scale=1.0, obj_scale=1.0, rho=1.0, dt=1e-4,
                 gravity=(0.0, -9.8, 0.0), bound=3, eps=1e-5, sticky_floor=False,
                 svd_exact=False):
        self.n_grid = int(n_grid)
        self.dx = float(space_scale / n_grid)
        self.inv_dx = 1.0 / self.dx
        self.p_vol = (self.dx * obj_scale) ** 3
        self.p_mass = self.p_vol * rho * obj_scale
        self.dt = float(dt)
        self.gravity = tuple(float(g) for g in gravity)
        self.bound = int(bound)
        self.eps = float(eps)
        self.sticky_floor = bool(sticky_floor)
        # False: the reference's svd_grad (clamp |s_j^2 - s_i^2| >= 1e-8).
        # True: exact adjoint -- rotation path in the closed form
        # (M_ij - M_ji) / (s_i + s_j), other SVD terms without the 1e-8 clamp.
        self.svd_exact = bool(svd_exact)


def lame(E, nu):
    """Lame parameters, ``mpm_elastic.py:70-71``."""
    mu = E / 2 / (1 + nu)
    lam = E * nu / (1 + nu) / (1 - 2 * nu)
    return mu, lam


class Material:
    """One row of the per-particle material table (``multi_obj.py:65-74``)."""

    def __init__(self, E, nu, kind=KIND_ELASTIC, yield_stress=1000.0, friction_deg=30.0):
        self.E, self.nu = float(E), float(nu)
        self.mu, self.lam = lame(self.E, self.nu)
        self.kind = int(kind)
        self.yield_stress = float(yield_stress)
        s = math.sin(math.radians(friction_deg))
        # Klar et al. 2016 eq. (27): alpha = sqrt(2/3) * 2 sin(phi) / (3 - sin(phi))
        self.dp_alpha = math.sqrt(2.0 / 3.0) * 2.0 * s / (3.0 - s)


# ---------------------------------------------------------------------------
# cell keys (bit-exact rule shared with the HIP kernels)
# ---------------------------------------------------------------------------

def cell_base_np(x, inv_dx):
    """``base = int(x * inv_dx - 0.5)`` (``mpm_elastic.py:178``) in float32.

    Declared rounding rule: ``t = fl32(fl32(x) * fl32(inv_dx))``,
    ``u = fl32(t - 0.5f)`` (no fused multiply-add), ``base = trunc(u)``
    (C cast toward zero, as Taichi's ``cast(int)``).
    """
    x = np.asarray(x, dtype=np.float32)
    t = x * np.float32(inv_dx)
    u = t - np.float32(0.5)
    return np.trunc(u).astype(np.int32)


def cell_keys_np(x, inv_dx, n_grid):
    """Cell key of each particle's stencil base (the sort key of the particle order).

    Bases are clamped to [0, n-1]; cells are grouped in 4x4x4 bricks:
    ``brick = ((bx>>2)*nb + (by>>2))*nb + (bz>>2)`` with ``nb = ceil(n/4)``,
    ``local = ((bx&3)*4 + (by&3))*4 + (bz&3)``, ``key = brick*64 + local``.
    The reference processes particles in index order and has no key; this
    rule is the build's own and is stated here as the parity target.
    """
    return _brick_keys(cell_base_np(x, inv_dx), int(n_grid), int(n_grid) - 1)


def bin_keys_np(x, inv_dx, n_grid):
    """Binning key of the sparse scatter: as ``cell_keys_np`` but with the base
    clamped to [0, n-3], the range the 3x3x3 stencil is clamped to, so the key
    names the cell whose stencil the particle scatters through."""
    return _brick_keys(cell_base_np(x, inv_dx), int(n_grid), int(n_grid) - 3)


def _brick_keys(base, n, hi):
    b = np.clip(base.astype(np.int64), 0, hi)
    nb = (n + 3) // 4
    brick = ((b[:, 0] >> 2) * nb + (b[:, 1] >> 2)) * nb + (b[:, 2] >> 2)
    local = ((b[:, 0] & 3) * 4 + (b[:, 1] & 3)) * 4 + (b[:, 2] & 3)
    return (brick * 64 + local).astype(np.int64)


def sort_by_cell_np(x, inv_dx, n_grid):
    """Stable sort order of particles by cell key (``np.argsort(kind='stable')``)."""
    keys = cell_keys_np(x, inv_dx, n_grid)
    return keys, np.argsort(keys, kind="stable")


# ---------------------------------------------------------------------------
# B-spline weights
# ---------------------------------------------------------------------------

_OFFS = torch.tensor([[i, j, k] for i in range(3) for j in range(3) for k in range(3)],
                     dtype=torch.int64)


def stencil(x, inv_dx):
    """base, fx and quadratic weights (``mpm_elastic.py:178-181``).

    Returns ``base (N,3) int64``, ``fx (N,3)`` (differentiable in x) and
    ``w (N,3,3)`` with ``w[p, a, axis]`` the weight of stencil offset ``a``.
    """
    inv = torch.tensor(inv_dx, dtype=x.dtype)
    with torch.no_grad():
        base = (x * inv - 0.5).to(torch.int64)    # truncation toward zero, like cast(int)
    fx = x * inv - base.to(x.dtype)
    w = torch.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1) ** 2, 0.5 * (fx - 0.5) ** 2], dim=1)
    return base, fx, w


def stencil_nodes(base, w, n_grid):
    """Flat node index ``(N,27)`` and weight ``(N,27)`` = w[i][0]*w[j][1]*w[k][2]."""
    n = int(n_grid)
    node = base[:, None, :] + _OFFS[None, :, :]
    flat = (node[..., 0] * n + node[..., 1]) * n + node[..., 2]
    wi = w[:, _OFFS[:, 0], 0]
    wj = w[:, _OFFS[:, 1], 1]
    wk = w[:, _OFFS[:, 2], 2]
    return flat, wi * wj * wk


# ---------------------------------------------------------------------------
# F update and SVD
# ---------------------------------------------------------------------------

def init_object(particles, trans_h, ball_vel):
    """``init_object`` (``mpm_elastic.py:107-114``): x = (trans_h [p; 1])_{0:3},
    v = ball_vel, F = I."""
    ph = torch.cat([particles, torch.ones(len(particles), 1, dtype=particles.dtype)], 1)
    x = (ph @ trans_h.T)[:, :3]
    v = ball_vel.reshape(1, 3).expand(len(particles), 3).clone()
    F = torch.eye(3, dtype=particles.dtype).repeat(len(particles), 1, 1)
    return x, v, F


def compute_new_F(C, F, dt):
    """``F_new = (I + dt C) F`` (``mpm_elastic.py:134-137``)."""
    eye = torch.eye(3, dtype=C.dtype)
    return (eye + dt * C) @ F


def svd3_np(F):
    """3x3 SVD with the rotation convention: U, V in SO(3), sigma_3 carries the sign.

    |sigma_1| >= |sigma_2| >= |sigma_3|, sigma_1, sigma_2 >= 0.  This is the
    convention of the SVD the reference calls (``ti.svd``, ``mpm_elastic.py:142``),
    for which ``R = U V^T`` is the polar rotation.
    """
    F = np.asarray(F)
    U, s, Vh = np.linalg.svd(F)
    V = np.swapaxes(Vh, -1, -2).copy()
    U = U.copy()
    s = s.copy()
    dU = np.linalg.det(U)
    dV = np.linalg.det(V)
    fu = dU < 0
    fv = dV < 0
    U[fu, :, 2] *= -1
    s[fu, 2] *= -1
    V[fv, :, 2] *= -1
    s[fv, 2] *= -1
    return U, s, V


def _clamp_denominator(a, lim=1e-8):
    """``clamp`` of ``mpm_elastic.py:149-155``: |a| >= 1e-8 keeping the sign (0 -> +1e-8)."""
    return torch.where(a >= 0, torch.clamp(a, min=lim), torch.clamp(a, max=-lim))


def svd_grad_ref(U, sig, V, gU, gsig, gV, clamp=1e-8):
    """Hand-written SVD adjoint of ``mpm_elastic.py:157-173`` (``single_svd_grad``).

    F_bar = U S_bar V^T + U [(ff o (U^T U_bar - U_bar^T U)) S] V^T
                        + U [S (ff o (V^T V_bar - V_bar^T V))] V^T,
    ff_ij = 1 / clamp(s_j^2 - s_i^2), ff_ii = 0.
    Returned in the dtype of the inputs (the HIP kernels evaluate the ff terms
    in float64).
    """
    S = torch.diag_embed(sig)
    Vt = V.transpose(-1, -2)
    Ut = U.transpose(-1, -2)
    s_term = U @ torch.diag_embed(gsig) @ Vt
    s2 = sig * sig
    diff = s2[..., None, :] - s2[..., :, None]          # [i, j] = s_j - s_i
    ff = 1.0 / _clamp_denominator(diff, clamp)
    eye = torch.eye(3, dtype=torch.bool)
    ff = torch.where(eye, torch.zeros_like(ff), ff)
    u_term = U @ ((ff * (Ut @ gU - gU.transpose(-1, -2) @ U)) @ S) @ Vt
    v_term = U @ (S @ ((ff * (Vt @ gV - gV.transpose(-1, -2) @ V)) @ Vt))
    return u_term + v_term + s_term


class _SVD3(torch.autograd.Function):
    """``ti.svd`` forward with ``svd_grad`` as its adjoint (``grasp_plastic.py:177``)."""

    @staticmethod
    def forward(ctx, F, clamp):
        U, s, V = svd3_np(F.detach().cpu().numpy())
        U = torch.from_numpy(U).to(F.dtype)
        s = torch.from_numpy(s).to(F.dtype)
        V = torch.from_numpy(V).to(F.dtype)
        ctx.save_for_backward(U, s, V)
        ctx.clamp = clamp
        return U, s, V

    @staticmethod
    def backward(ctx, gU, gs, gV):
        U, s, V = ctx.saved_tensors
        gU = torch.zeros_like(U) if gU is None else gU
        gs = torch.zeros_like(s) if gs is None else gs
        gV = torch.zeros_like(V) if gV is None else gV
        return svd_grad_ref(U, s, V, gU, gs, gV, ctx.clamp), None


class _UDV(torch.autograd.Function):
    """F_p = U diag(d) V^T of the return map with an exact adjoint.

    d_k = G s_k^a (s = max(sigma, 0.05)) for every yielded branch.  The
    off-diagonal (U, V) part of the adjoint is N_ij A_ij + N_ji B_ij with
    N = U^T F_p_bar V, A/B = +-(d_i + d_j) / (2 (s_i + s_j)) + Dd_ij / 2,
    Dd_ij = (d_j - d_i) / (s_j - s_i) (its limit at s_i = s_j); equal to the
    SVD adjoint formula for distinct singular values without its small
    denominators.  The adjoint of d is diag(N) and flows back through d(s).
    """

    @staticmethod
    def forward(ctx, F, U, V, d, s, a, G):
        ctx.save_for_backward(U, V, d, s, a, G)
        return U @ torch.diag_embed(d) @ V.transpose(-1, -2)

    @staticmethod
    def backward(ctx, g):
        U, V, d, s, a, G = ctx.saved_tensors
        N = U.transpose(-1, -2) @ g @ V
        si, sj = s[..., :, None], s[..., None, :]
        di, dj = d[..., :, None], d[..., None, :]
        a_, G_ = a[..., None, None], G[..., None, None]
        den = si + sj
        tiny = torch.where(den < 0, torch.full_like(den, -1e-12), torch.full_like(den, 1e-12))
        den = torch.where(den.abs() < 1e-12, tiny, den)
        delta = sj - si
        both = (si > 0.05) & (sj > 0.05)
        safe_si = torch.where(both, si, torch.ones_like(si))
        safe_del = torch.where(delta != 0, delta, torch.ones_like(delta))
        dd_smooth = torch.where(delta != 0,
                                G_ * safe_si ** a_ * torch.expm1(a_ * torch.log1p(delta / safe_si)) / safe_del,
                                G_ * a_ * safe_si ** (a_ - 1))
        sci, scj = torch.clamp(si, min=0.05), torch.clamp(sj, min=0.05)
        dd_clamped = torch.where(delta != 0, G_ * (scj ** a_ - sci ** a_) / safe_del,
                                 torch.zeros_like(delta))
        Dd = torch.where(both, dd_smooth, dd_clamped)
        hs = 0.5 * (di + dj) / den
        Mo = N * (hs + 0.5 * Dd) + N.transpose(-1, -2) * (-hs + 0.5 * Dd)
        Mo = Mo * (1 - torch.eye(3, dtype=Mo.dtype))
        gF = U @ Mo @ V.transpose(-1, -2)
        gd = torch.diagonal(N, dim1=-2, dim2=-1)
        return gF, None, None, gd, None, None, None


def svd3(F, clamp=1e-8):
    """(U, sigma, V) of each 3x3 in ``F (N,3,3)``; differentiable with ``svd_grad_ref``."""
    return _SVD3.apply(F, clamp)


class _Polar(torch.autograd.Function):
    """R = U V^T (rotation of the polar decomposition) with its exact adjoint:
    F_bar = U K V^T, K_ij = (M_ij - M_ji) / (s_i + s_j), M = U^T R_bar V.
    Equal to svd_grad's rotation terms wherever |s_j^2 - s_i^2| >= 1e-8, and
    the correct limit where the reference's clamp is active."""

    @staticmethod
    def forward(ctx, F, U, s, V):
        ctx.save_for_backward(U.detach(), s.detach(), V.detach())
        return U.detach() @ V.detach().transpose(-1, -2)

    @staticmethod
    def backward(ctx, gR):
        U, s, V = ctx.saved_tensors
        M = U.transpose(-1, -2) @ gR @ V
        den = s[..., :, None] + s[..., None, :]
        tiny = torch.where(den < 0, torch.full_like(den, -1e-12), torch.full_like(den, 1e-12))
        den = torch.where(den.abs() < 1e-12, tiny, den)
        K = (M - M.transpose(-1, -2)) / den
        K = K * (1 - torch.eye(3, dtype=K.dtype))
        return U @ K @ V.transpose(-1, -2), None, None, None


# ---------------------------------------------------------------------------
# constitutive models
# ---------------------------------------------------------------------------

def _norm_eps(x, eps=1e-8):
    """``norm(x, eps) = sqrt(x.x + eps)`` (``mpm_plastic.py:139-141``)."""
    return torch.sqrt((x * x).sum(-1) + eps)


def von_mises(F, U, sig, V, mu, yield_stress, exact=False):
    """Return map of ``compute_von_mises`` (``mpm_plastic.py:150-166``).

    sigma <- max(sigma, 0.05); eps = log sigma; eps_hat = eps - tr(eps)/3;
    dg = |eps_hat|_eps - tau_y / (2 mu); if dg > 0:
    eps -= dg / |eps_hat| * eps_hat and F = U diag(exp eps) V^T.
    """
    s = torch.clamp(sig, min=0.05)
    eps = torch.log(s)
    eh = eps - eps.sum(-1, keepdim=True) / 3
    ehn = _norm_eps(eh)
    dg = ehn - yield_stress / (2 * mu)
    yielded = dg > 0
    eps_y = eps - (dg / ehn)[..., None] * eh
    if exact:
        a = (1 - dg / ehn).detach()
        G = torch.exp((1 - a) * eps.sum(-1) / 3).detach()
        Fy = _UDV.apply(F, U.detach(), V.detach(), torch.exp(eps_y), sig.detach(), a, G)
    else:
        Fy = U @ torch.diag_embed(torch.exp(eps_y)) @ V.transpose(-1, -2)
    return torch.where(yielded[..., None, None], Fy, F), yielded


def drucker_prager(F, U, sig, V, mu, lam, alpha, exact=False):
    """Drucker-Prager return map in Hencky strain (Klar et al. 2016, sec. 7.3).

    No reference implementation exists (the reference has no sand model).
    eps = log max(sigma, 0.05); tr = sum eps; eps_hat = eps - tr/3;
    tension (tr >= 0): project to the cone tip, F = U V^T;
    else dg = |eps_hat|_eps + (3 lam + 2 mu) / (2 mu) * tr * alpha;
    dg <= 0: unchanged; else eps -= dg / |eps_hat|_eps * eps_hat.
    """
    s = torch.clamp(sig, min=0.05)
    eps = torch.log(s)
    tr = eps.sum(-1)
    eh = eps - tr[..., None] / 3
    ehn = _norm_eps(eh)
    dg = ehn + (3 * lam + 2 * mu) / (2 * mu) * tr * alpha
    tension = tr >= 0
    shear = (~tension) & (dg > 0)
    eps_y = eps - (dg / ehn)[..., None] * eh
    eps_new = torch.where(tension[..., None], torch.zeros_like(eps), eps_y)
    if exact:
        a = torch.where(tension, torch.zeros_like(dg), 1 - dg / ehn).detach()
        G = torch.where(tension, torch.ones_like(dg),
                        torch.exp((1 - a) * tr / 3)).detach()
        Fy = _UDV.apply(F, U.detach(), V.detach(), torch.exp(eps_new), sig.detach(), a, G)
    else:
        Fy = U @ torch.diag_embed(torch.exp(eps_new)) @ V.transpose(-1, -2)
    changed = tension | shear
    return torch.where(changed[..., None, None], Fy, F), changed


def particle_stress(F_new, U, sig, V, mat_mu, mat_lam, mat_kind, mat_yield, mat_alpha,
                    R=None, exact=False):
    """Fixed-corotated Kirchhoff stress and the post-return-map F.

    ``mpm_elastic.py:183-185`` (elastic: J = det Sigma, F_p = F_new);
    ``mpm_plastic.py:90-95`` (plastic: F_p = return map, J = det F_p,
    R = U V^T of the pre-return-map SVD).
    tau = 2 mu (F_p - R) F_p^T + lam J (J - 1) I.
    """
    mu = mat_mu[..., None, None]
    lam = mat_lam[..., None, None]
    Vt = V.transpose(-1, -2)
    if R is None:
        R = U @ Vt
    F_vm, _ = von_mises(F_new, U, sig, V, mat_mu, mat_yield, exact)
    F_dp, _ = drucker_prager(F_new, U, sig, V, mat_mu, mat_lam, mat_alpha, exact)
    k = mat_kind[..., None, None]
    F_p = torch.where(k == KIND_VON_MISES, F_vm, torch.where(k == KIND_DRUCKER_PRAGER, F_dp, F_new))
    J_el = sig[..., 0] * sig[..., 1] * sig[..., 2]
    J_pl = torch.linalg.det(F_p)
    J = torch.where(mat_kind == KIND_ELASTIC, J_el, J_pl)[..., None, None]
    eye = torch.eye(3, dtype=F_new.dtype)
    tau = 2 * mu * (F_p - R) @ F_p.transpose(-1, -2) + eye * lam * J * (J - 1)
    return tau, F_p


def material_arrays(materials, mat_id, dtype):
    """Gather per-particle (mu, lam, kind, yield, alpha) from a material table."""
    mat_id = torch.as_tensor(mat_id, dtype=torch.int64)
    mu = torch.tensor([m.mu for m in materials], dtype=dtype)[mat_id]
    lam = torch.tensor([m.lam for m in materials], dtype=dtype)[mat_id]
    kind = torch.tensor([m.kind for m in materials], dtype=torch.int64)[mat_id]
    yld = torch.tensor([m.yield_stress for m in materials], dtype=dtype)[mat_id]
    alpha = torch.tensor([m.dp_alpha for m in materials], dtype=dtype)[mat_id]
    return mu, lam, kind, yld, alpha


# ---------------------------------------------------------------------------
# P2G, grid_op, G2P
# ---------------------------------------------------------------------------

def p2g(prm, x, v, C, F_new, U, sig, V, mats):
    """Particle-to-grid scatter (``mpm_elastic.py:175-208`` / ``mpm_plastic.py:221-256``).

    Returns ``grid_mv (n^3, 3)``, ``grid_m (n^3,)`` and ``F_next`` (the value
    stored in ``F_0[f+1]``).  ``mats`` is the tuple from ``material_arrays``.
    """
    n = prm.n_grid
    base, fx, w = stencil(x, prm.inv_dx)
    R = _Polar.apply(F_new, U, sig, V) if getattr(prm, "svd_exact", False) else None
    tau, F_p = particle_stress(F_new, U, sig, V, *mats, R=R,
                               exact=getattr(prm, "svd_exact", False))
    stress = (-prm.dt * prm.p_vol * 4 * prm.inv_dx * prm.inv_dx) * tau
    affine = stress + prm.p_mass * C
    flat, weight = stencil_nodes(base, w, n)
    dpos = (_OFFS.to(x.dtype)[None, :, :] - fx[:, None, :]) * prm.dx          # (N,27,3)
    mv = weight[..., None] * (prm.p_mass * v[:, None, :] + (affine[:, None, :, :] @ dpos[..., None])[..., 0])
    m = weight * prm.p_mass
    grid_mv = torch.zeros((n ** 3, 3), dtype=x.dtype).index_add(0, flat.reshape(-1), mv.reshape(-1, 3))
    grid_m = torch.zeros((n ** 3,), dtype=x.dtype).index_add(0, flat.reshape(-1), m.reshape(-1))
    return grid_mv, grid_m, F_p


def check_grid_occupy(prm, grid_m):
    """``grid_m > eps`` (``mpm_elastic.py:210-214``)."""
    return grid_m > prm.eps


def grid_op(prm, grid_mv, grid_m, grid_f=None):
    """Grid update (``mpm_elastic.py:216-240``).

    For occupied nodes: v = mv/(m+eps) + f/(m+eps) + dt g, then the one-sided
    wall conditions: component a is zeroed when (idx_a < bound and v_a < 0) or
    (idx_a > n - bound and v_a > 0).  Unoccupied nodes stay 0 (``reset``).
    ``sticky_floor`` (build-defined, no reference) zeroes all components at
    nodes with j < bound.
    """
    n = prm.n_grid
    occ = check_grid_occupy(prm, grid_m)
    inv_m = 1.0 / (grid_m + prm.eps)
    v = inv_m[:, None] * grid_mv
    if grid_f is not None:
        v = v + inv_m[:, None] * grid_f
    g = torch.tensor(prm.gravity, dtype=grid_mv.dtype)
    v = v + prm.dt * g
    idx = torch.arange(n ** 3)
    ijk = torch.stack([idx // (n * n), (idx // n) % n, idx % n], dim=1)
    lo = (ijk < prm.bound) & (v < 0)
    hi = (ijk > n - prm.bound) & (v > 0)
    v = torch.where(lo | hi, torch.zeros_like(v), v)
    if prm.sticky_floor:
        v = torch.where((ijk[:, 1] < prm.bound)[:, None], torch.zeros_like(v), v)
    return torch.where(occ[:, None], v, torch.zeros_like(v))


def g2p(prm, x, grid_v):
    """Grid-to-particle gather (``mpm_elastic.py:243-260``).

    v' = sum w g_v;  C' = sum 4 inv_dx w g_v (offset - fx)^T;  x' = x + dt v'.
    """
    n = prm.n_grid
    base, fx, w = stencil(x, prm.inv_dx)
    flat, weight = stencil_nodes(base, w, n)
    dpos = _OFFS.to(x.dtype)[None, :, :] - fx[:, None, :]
    gv = grid_v[flat]                                          # (N,27,3)
    new_v = (weight[..., None] * gv).sum(1)
    new_C = (4 * prm.inv_dx * weight[..., None, None] * (gv[..., :, None] * dpos[..., None, :])).sum(1)
    new_x = x + prm.dt * new_v
    return new_x, new_v, new_C


# ---------------------------------------------------------------------------
# substep composition (object only; contact lives in oracle/contact.py)
# ---------------------------------------------------------------------------

def substep(prm, state, mats, grid_f=None, return_grid=False):
    """One MPM substep without sensors (``grasp_plastic.py:151-164`` minus FEM/contact).

    ``state = (x, v, C, F)``; returns the next state.  The order is
    compute_new_F -> svd -> p2g -> check_grid_occupy -> grid_op -> g2p.
    """
    x, v, C, F = state
    F_new = compute_new_F(C, F, prm.dt)
    U, sig, V = svd3(F_new, clamp=1e-30 if getattr(prm, "svd_exact", False) else 1e-8)
    grid_mv, grid_m, F_next = p2g(prm, x, v, C, F_new, U, sig, V, mats)
    grid_v = grid_op(prm, grid_mv, grid_m, grid_f)
    nx, nv, nC = g2p(prm, x, grid_v)
    out = (nx, nv, nC, F_next)
    if return_grid:
        return out, (grid_mv, grid_m, grid_v)
    return out


def rollout(prm, state, mats, n_substeps):
    """Run ``n_substeps`` substeps; returns the final state."""
    for _ in range(n_substeps):
        state = substep(prm, state, mats)
    return state
