"""Pins the composed gripper substep of the oracle (CPU): without cable
contact it is the object-pad substep and a free cable step; with contact the
impulses on the object and the cable and the forces on the pads balance."""
import numpy as np
import torch

from oracle import coupled as ocp
from oracle import gripper_scene as ogs
from oracle import mpm as om
from oracle import rope as orp
from tests.gripper_cases import (constant_controls, oracle_B, oracle_pads, oracle_rope_params,
                                 rope_start, small_scene)
from tests.mpm_cases import oracle_mats, oracle_params

F64 = torch.float64


def _setup(c):
    prm, n = c["prm"], len(c["x"])
    op = oracle_params(prm)
    mats = oracle_mats(c["mats"], c["mat_id"], F64)
    st = (torch.tensor(c["x"], dtype=F64), torch.tensor(c["v"], dtype=F64),
          torch.zeros(n, 3, 3, dtype=F64), torch.eye(3, dtype=F64).repeat(n, 1, 1))
    nv = len(c["pads"][0][0])
    cvs = [torch.tensor(cv, dtype=F64) for cv in constant_controls(c, nv)]
    pos = [torch.tensor(p[0], dtype=F64) for p in c["pads"]]
    pads = oracle_pads(c, pos, [cv[0].clone() for cv in cvs], [p.clone() for p in pos], cvs, 0,
                       oracle_B(c))
    return op, mats, st, pads


def test_without_cable_contact_it_is_the_object_pad_substep():
    c = small_scene()
    op, mats, st, pads = _setup(c)
    rp = oracle_rope_params(c)
    rx = torch.tensor(rope_start(c), dtype=F64) + torch.tensor([0.0, 0.0, 100.0], dtype=F64)
    rv = torch.zeros_like(rx)
    cp = tuple(c["contact"].values())
    s1, p1, r1 = ogs.gripper_substep(op, st, mats, pads, (rx, rv, rx[0]), rp, cp,
                                     tuple(c["rope_contact"].values()), clamp=True)
    s2, p2, _, _ = ocp.coupled_substep(op, st, mats, pads, cp, clamp=True)
    for a, b in zip(s1, s2):
        assert torch.equal(a, b)
    for pa, pb in zip(p1, p2):
        for a, b in zip(pa, pb):
            assert torch.equal(a, b)
    r2 = orp.step(rp, rx, rv, torch.zeros_like(rx), rx[0])
    for a, b in zip(r1, r2):
        assert torch.equal(a, b)


def test_impulses_on_object_and_cable_balance_the_pad_forces():
    c = small_scene()
    op, mats, st, pads = _setup(c)
    rp = oracle_rope_params(c)
    rx = torch.tensor(rope_start(c), dtype=F64)
    rv = torch.zeros_like(rx)
    info = {}
    ogs.gripper_substep(op, st, mats, pads, (rx, rv, rx[0]), rp, tuple(c["contact"].values()),
                        tuple(c["rope_contact"].values()), info=info)
    gf, imp, ext = info["grid_f"], info["impulse"], info["ext"]
    N = c["prm"].n_grid
    i_node = torch.nonzero(gf.abs().sum(1) > 0).squeeze(1) // (N * N)
    assert (i_node < N // 2).sum() > 20 and (i_node >= N // 2).sum() > 20, "both pads touch"
    assert (imp.abs().sum(1) > 0).sum() > 5, "cable - pad contact"
    total = gf.sum(0) / op.dt + imp.sum(0) / op.dt + sum(e.sum(0) for e in ext)
    scale = gf.abs().sum() / op.dt + imp.abs().sum() / op.dt
    assert float(total.abs().max()) <= 1e-10 * float(scale)
