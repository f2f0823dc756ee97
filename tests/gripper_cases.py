"""The scaled gripper scene (two pads, a cable, an object) for the oracle and
GPU parity tests."""
import numpy as np
import torch

from difftactile_amd import scenes
from oracle import fem as of
from oracle import rope as orp

F64 = torch.float64


def small_scene(steps=6, n_particles=6000):
    return scenes.config5_gripper(n_particles=n_particles, n_grid=32, n_substeps=steps,
                                  pad_verts=(12, 12, 3), n_rope=101, rope_length=6.0)


def constant_controls(c, n_verts):
    """Per-pad control velocities (steps + 1, V, 3): each pad's base layer
    closes along its face normal at the scene's closing speed."""
    T = c["n_substeps"]
    out = []
    for (_, _, _, _, od) in c["pads"]:
        cv = np.zeros((T + 1, n_verts, 3), np.float32)
        cv[:, :, :] = (10.0 * np.asarray(od, np.float32))[None, None, :]
        out.append(cv)
    return out


def rope_start(c):
    r = c["rope"]
    x = (np.asarray(r["start"], np.float32)[None] + np.asarray(r["ori"], np.float32)[None] *
         np.float32(r["rest_length"]) * np.arange(r["n"], dtype=np.float32)[:, None])
    return x.astype(np.float32)


def oracle_rope_params(c):
    r = c["rope"]
    return orp.RopeParams(r["n"], dt=c["prm"].dt, p_rad=r["p_rad"], p_rho=r["p_rho"],
                          rest_length=r["rest_length"],
                          stretch_compliance=r["compliance"], bending_compliance=r["compliance"])


def oracle_pads(c, pos, vel, vpos, cvs, f, B, dtype=F64):
    out = []
    for k, (nodes, f2v, tris, layer, od) in enumerate(c["pads"]):
        out.append(dict(pos=pos[k], vel=vel[k], vpos=vpos[k], cv=cvs[k][f], cv_next=cvs[k][f + 1],
                        f2v=torch.tensor(f2v).long(), B=B[k], layer=torch.tensor(layer),
                        tris=tris, out_dir=torch.tensor(od, dtype=dtype), mu=c["pad_mu"],
                        lam=c["pad_lam"], mass=c["pad_mass"], dt=c["prm"].dt))
    return out


def oracle_B(c, dtype=F64):
    return [of.init_B(torch.tensor(nodes.astype(np.float32)).to(dtype), torch.tensor(f2v).long())
            for nodes, f2v, *_ in c["pads"]]
