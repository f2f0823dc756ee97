"""dt_mpm_advance_static (static scratch, no allocation) against dt_mpm_advance
(caller-provided scratch and workspace) through the C-ABI: two calls in a row
on the same device, the second with fewer particles, and the size check."""
import ctypes

import numpy as np
import pytest
import torch

from difftactile_amd import _lib
from difftactile_amd.params import material_table
from difftactile_amd.sim import workspace_bytes
from tests.mpm_cases import blob, default_prm, materials, rel_l2

pytestmark = pytest.mark.gpu
P = _lib.ptr


def _soa_state(n, seed, n_grid):
    x, v, C, F = blob(n, n_grid=n_grid, seed=seed, fpert=0.05)
    C = 0.2 * C
    return [torch.tensor(a.reshape(n, -1).T.copy()).cuda().contiguous() for a in (x, v, C, F)]


def _advance(prm, mats, nm, n, mid, state, k):
    """dt_mpm_advance with its own scratch state, grids and workspace."""
    nn = prm.n_grid ** 3
    tmp = [torch.empty_like(t) for t in state]
    g4 = torch.empty(nn, 4, device="cuda")
    gv4 = torch.empty(nn, 4, device="cuda")
    c = prm.c_struct(soa=True)
    nb = workspace_bytes(c, n)
    ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
    _lib.call("dt_mpm_advance", c, P(mats), nm, n, P(mid, torch.int32), *[P(t) for t in state],
              *[P(t) for t in tmp], k, P(g4), P(gv4), P(ws, torch.uint8), nb, _lib.stream())


def test_static_scratch_advance_matches_the_allocating_one(gpu):
    mk = materials("multi")
    mats = material_table(mk, "cuda")
    prm = default_prm(64)
    c = prm.c_struct(soa=True)
    for n, seed, k in ((30000, 3, 7), (6000, 4, 5)):
        mid = torch.tensor(np.random.default_rng(seed).integers(0, len(mk), n).astype(np.int32)).cuda()
        ref = _soa_state(n, seed, 64)
        got = [t.clone() for t in ref]
        _advance(prm, mats, len(mk), n, mid, ref, k)
        _lib.call("dt_mpm_advance_static", c, P(mats), len(mk), n, P(mid, torch.int32),
                  *[P(t) for t in got], k, _lib.stream())
        torch.cuda.synchronize()
        moved = float((ref[0] - _soa_state(n, seed, 64)[0]).abs().max())
        assert moved > 0
        for a, b in zip(got, ref):
            assert torch.isfinite(a).all()
            assert rel_l2(a.cpu(), b.cpu()) <= 1e-6, (n, rel_l2(a.cpu(), b.cpu()))


def test_static_scratch_rejects_what_does_not_fit(gpu):
    prm = default_prm(128)
    c = prm.c_struct(soa=True)
    mats = material_table(materials("elastic"), "cuda")
    x = torch.zeros(3, 10, device="cuda")
    rc = _lib.lib().dt_mpm_advance_static(ctypes.byref(c), P(mats), 1, 10, None, P(x), P(x),
                                          P(x), P(x), 1, _lib.stream())
    assert rc != 0
    msg = _lib.lib().dt_last_error().decode()
    assert "static scratch holds" in msg
