"""Host-side logic of the package (CPU only)."""
import numpy as np


def test_vertex_tet_csr_lists_every_corner_once():
    from difftactile_amd.coupled import vertex_tet_csr
    rng = np.random.default_rng(3)
    f2v = rng.integers(0, 50, size=(200, 4))
    start, ent = vertex_tet_csr(f2v, 50)
    assert start[0] == 0 and start[-1] == 800 and np.all(np.diff(start) >= 0)
    assert np.array_equal(np.sort(ent), np.arange(800))
    for v in range(50):
        e = ent[start[v]:start[v + 1]]
        assert np.all(f2v.reshape(-1)[e] == v)


def _fma32(a, b, c):
    """float32 fused multiply-add: the product of two float32 is exact in
    float64, so one float64 sum then one float32 rounding."""
    return (a.astype(np.float64) * b + c).astype(np.float32)


def test_closest_search_lower_bound_never_exceeds_the_compared_distance():
    """The closest-triangle search skips a triangle when
    lb = (1 - 4e-6)(|c|^2 + |q|^2) - 2 c.q (float32, three FMAs, as the
    kernel evaluates it) reaches the node's best squared distance.  That is
    exact only if lb stays below the float32 squared distance the search
    compares, ((d0 d0 + d1 d1) + d2 d2) + 1e-10 with d = c - q, for every
    centroid and node: checked here over coordinates from 1e-3 to 1e3, both
    near-coincident and far-apart pairs."""
    f = np.float32
    shrink = f(1.0) - f(4e-6)
    rng = np.random.default_rng(0)
    for scale in (1e-3, 1e-1, 1.0, 10.0, 1e3):
        for spread in (1e-7, 1e-4, 1e-2, 1.0):
            q = (rng.uniform(0, scale, size=(200_000, 3))).astype(f)
            c = (q + rng.normal(0, spread * scale, size=q.shape)).astype(f)
            cw = shrink * ((c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1]) + c[:, 2] * c[:, 2])
            qb = shrink * ((q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1]) + q[:, 2] * q[:, 2])
            m = f(-2.0) * q
            lb = _fma32(c[:, 0], m[:, 0],
                        _fma32(c[:, 1], m[:, 1], _fma32(c[:, 2], m[:, 2], cw + qb)))
            d = c - q
            s = ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) + f(1e-10)
            assert lb.dtype == np.float32 and s.dtype == np.float32
            assert np.all(lb < s), (scale, spread, np.max(lb - s))
