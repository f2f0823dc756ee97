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
