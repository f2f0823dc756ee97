"""Oracle: grid-mass deformation losses.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Restates

  compute_grid_m_kernel   difftactile/object_model/mpm_plastic.py:168-178
  update_target_sdf       difftactile/tasks/deform_loss.py:61-86
  density / sdf losses    difftactile/tasks/deform_loss.py:99-111
  iou                     difftactile/tasks/deform_loss.py:139-153

Pinning (tests/test_oracle_deform.py):
  grid_mass   -- total mass = n p_mass (partition of unity); first moment
                 sum m_node x_node = p_mass sum x_p (the quadratic B-spline
                 reproduces linear functions); FP64 central differences of a
                 random linear functional w.r.t. x.
  target_sdf  -- one occupied node: every node's distance is the Euclidean
                 distance to it (sqrt(d^2 + 1e-8)); a plane of occupied nodes:
                 the distance to the plane; a pure-Python scalar sweep on a
                 tiny grid (bit-exact, including the first-minimum tie rule).
  losses      -- zero density loss at m = target; zero sdf loss for mass on
                 the target; FP64 central differences through grid_mass.
  iou         -- 1 for identical grids, I / (U - I) of two disjoint-free
                 boxes by counting.
  kabsch_angle (multi_obj.py:309-340, box_open.py:400-406) -- see its
                 docstring.
"""
import numpy as np
import torch

from .mpm import stencil, stencil_nodes

SDF_INF = np.float32(1000.0)


def grid_mass(prm, x):
    """Mass-only p2g: sum_p w(node) p_mass into an n^3 grid (differentiable in x)."""
    n = prm.n_grid
    base, fx, w = stencil(x, prm.inv_dx)
    flat, weight = stencil_nodes(base, w, n)
    return torch.zeros((n ** 3,), dtype=x.dtype).index_add(
        0, flat.reshape(-1), (weight * prm.p_mass).reshape(-1))


_OFFS = [(ox, oy, oz) for ox in range(-3, 3) for oy in range(-3, 3) for oz in range(-3, 3)
         if (ox, oy, oz) != (0, 0, 0)]


def target_sdf_np(density, n, dx, n_sweeps):
    """update_target: n_sweeps Jacobi sweeps in float32 (numpy), offsets in
    x-y-z lexicographic order, strict < (the first minimum keeps its point),
    distance sqrt(((d0^2 + d1^2) + d2^2) + 1e-8) with d = I dx - point."""
    f = np.float32
    dens = np.asarray(density, f).reshape(n, n, n)
    dxf = f(dx)
    ii = np.arange(n, dtype=f)
    gp = np.stack(np.meshgrid(ii * dxf, ii * dxf, ii * dxf, indexing="ij"), -1)   # (n,n,n,3)
    sdf = np.full((n, n, n), SDF_INF, f)
    pt = np.zeros((n, n, n, 3), f)
    occ = dens > f(1e-4)
    pad = 3
    for _ in range(n_sweeps):
        best = np.full((n, n, n), SDF_INF, f)
        bp = pt.copy()
        sp = np.full((n + 2 * pad,) * 3, SDF_INF, f)
        sp[pad:pad + n, pad:pad + n, pad:pad + n] = sdf
        pp = np.zeros((n + 2 * pad,) * 3 + (3,), f)
        pp[pad:pad + n, pad:pad + n, pad:pad + n] = pt
        for ox, oy, oz in _OFFS:
            s_v = sp[pad + ox:pad + ox + n, pad + oy:pad + oy + n, pad + oz:pad + oz + n]
            p_v = pp[pad + ox:pad + ox + n, pad + oy:pad + oy + n, pad + oz:pad + oz + n]
            d = gp - p_v
            dist = np.sqrt(((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
                           + f(1e-8))
            take = (s_v < SDF_INF) & (dist < best)
            best = np.where(take, dist, best)
            bp = np.where(take[..., None], p_v, bp)
        best = np.where(occ, f(0.0), best)
        bp = np.where(occ[..., None], gp, bp)
        sdf, pt = best.astype(f), bp.astype(f)
    return sdf.reshape(-1), pt.reshape(-1, 3)


def target_sdf_scalar(density, n, dx, n_sweeps):
    """Pure-Python restatement of the sweep loop (tiny grids only)."""
    f = np.float32
    dens = np.asarray(density, f).reshape(n, n, n)
    sdf = np.full((n, n, n), SDF_INF, f)
    pt = np.zeros((n, n, n, 3), f)
    for _ in range(n_sweeps):
        s2, p2 = sdf.copy(), pt.copy()
        for i in range(n):
            for j in range(n):
                for k in range(n):
                    g = [f(i) * f(dx), f(j) * f(dx), f(k) * f(dx)]
                    s2[i, j, k] = SDF_INF
                    if dens[i, j, k] > f(1e-4):
                        s2[i, j, k] = f(0.0)
                        p2[i, j, k] = g
                        continue
                    for ox, oy, oz in _OFFS:
                        a, b, c = i + ox, j + oy, k + oz
                        if min(a, b, c) < 0 or max(a, b, c) >= n:
                            continue
                        if sdf[a, b, c] < SDF_INF:
                            q = pt[a, b, c]
                            d = [f(g[t] - q[t]) for t in range(3)]
                            dd = f(f(f(d[0] * d[0]) + f(d[1] * d[1])) + f(d[2] * d[2]))
                            dist = f(np.sqrt(f(dd + f(1e-8))))
                            if dist < s2[i, j, k]:
                                s2[i, j, k] = dist
                                p2[i, j, k] = q
        sdf, pt = s2, p2
    return sdf.reshape(-1), pt.reshape(-1, 3)


def deform_losses(m, target, sdf):
    """(sum |m - t|, sum sdf m) (compute_density_loss_kernel, compute_sdf_loss_kernel)."""
    return (m - target).abs().sum(), (sdf * m).sum()


def iou(m, target):
    """deform_loss.py:139-153."""
    m = np.asarray(m, np.float64)
    t = np.asarray(target, np.float64)
    inter = (m * t).sum() / m.max() / t.max()
    union = m.sum() / m.max() + t.sum() / t.max()
    return inter / (union - inter)


def kabsch_angle(x, v_next, mask, dt):
    """Rotation angle of the masked particles over one substep
    (``multi_obj.py:309-340``, ``box_open.py:400-406``): c0, c1 are the masked
    sums of x and y = x + dt v_next divided by all n particles (the
    reference's quirk), H = sum_masked (x - c0)(y - c1)^T, R = V U^T of its
    SVD (the transposed polar rotation, exact adjoint), angle = atan2(R_10, R_11).

    Pinned (tests/test_oracle_deform.py): a rigid turn by theta about z gives
    theta, for every particle masked with a translation, and for a partial
    mask without one (then H = A R^T with A symmetric positive definite);
    FP64 central differences of the angle w.r.t. x and v_next."""
    from .mpm import _Polar, svd3_np
    n = x.shape[0]
    msk = torch.as_tensor(mask, dtype=torch.bool)
    y = x + dt * v_next
    c0 = x[msk].sum(0) / n
    c1 = y[msk].sum(0) / n
    H = (x[msk] - c0).T @ (y[msk] - c1)
    U, s, V = (torch.as_tensor(a, dtype=x.dtype) for a in svd3_np(H.detach().numpy()))
    Q = _Polar.apply(H, U, s, V)
    R = Q.T
    return torch.atan2(R[1, 0], R[1, 1])
