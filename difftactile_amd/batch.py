"""Batched populations: independent MPM environments advanced together.

Environments that share a grid size and a time step are stacked along the
grid's x axis (``dt_mpm_params.n_env``).  One binning, one brick-owned P2G,
one grid pass and one G2P per substep then serve all of them, and the sparse
path visits only the bricks their particles occupy.  So an environment costs
its own particles and nodes, not the launch sequence of a separate
simulation.  Positions stay local to each environment: a per-particle
environment id shifts the stencils into the environment's slab of the stacked
grid, the walls act per environment, and environments never exchange
momentum.

The reference runs one scene per process (``tasks/*.py`` build one
``Contact`` each); a population is a loop over such scenes.
"""
import copy

import numpy as np
import torch

from .sim import MPMSim


class EnvBatch:
    """E independent MPM environments in one ``MPMSim``.

    prm:  ``MPMParams`` of one environment (grid, time step, gravity, walls).
    envs: one ``(materials, n_particles, mat_id or None)`` per environment,
          optionally with a fourth entry, the environment's own ``MPMParams``:
          environments of different grids then share one launch sequence
          ("ragged": each keeps its grid size, dx and particle mass inside a
          slot of the largest grid; they share dt, gravity, walls and eps).
    Further keyword arguments go to ``MPMSim`` (``checkpoint_every``,
    ``sort``, ``resident_bytes``, ``material_grad``); ragged batches run the
    sparse path.
    """

    def __init__(self, prm, envs, device="cuda", **sim_kwargs):
        if len(envs) < 1:
            raise ValueError("a batch needs at least one environment")
        env_prms = [e[3] if len(e) > 3 and e[3] is not None else prm for e in envs]
        ragged = any((q.n_grid, q.dx, q.p_mass) != (prm.n_grid, prm.dx, prm.p_mass)
                     for q in env_prms)
        self.prm = copy.copy(prm)
        self.prm.n_env = len(envs)
        if ragged:
            self.prm.n_grid = max(q.n_grid for q in env_prms)
            if sim_kwargs.get("sparse_grid") is False:
                raise ValueError("ragged batches run the sparse path")
            self.prm.set_ragged(env_prms, device=device)
        self.env_prms = env_prms
        self.ragged = ragged
        envs = [e[:3] for e in envs]
        mats, mids, ids, counts, n_mats = [], [], [], [], []
        for e, (materials, n, mat_id) in enumerate(envs):
            n = int(n)
            base = len(mats)
            mats.extend(materials)
            mid = np.zeros(n, np.int32) if mat_id is None else np.asarray(mat_id, np.int32)
            if mid.shape != (n,):
                raise ValueError("mat_id of environment %d must have %d entries" % (e, n))
            if n and (mid.min() < 0 or mid.max() >= len(materials)):
                raise ValueError("mat_id of environment %d out of its material table" % e)
            mids.append(mid + base)
            ids.append(np.full(n, e, np.int32))
            counts.append(n)
            n_mats.append(len(materials))
        self.counts = counts
        self.offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.mat_offsets = np.concatenate([[0], np.cumsum(n_mats)]).astype(np.int64)
        self.sim = MPMSim(self.prm, mats, int(self.offsets[-1]), mat_id=np.concatenate(mids),
                          device=device, env_id=np.concatenate(ids), **sim_kwargs)

    @property
    def n_env(self):
        return len(self.counts)

    def pack(self, states):
        """Per-environment ``(x, v, C, F)`` tuples -> the stacked state."""
        if len(states) != self.n_env:
            raise ValueError("expected %d environment states" % self.n_env)
        for e, st in enumerate(states):
            if any(t.shape[0] != self.counts[e] for t in st):
                raise ValueError("environment %d: state rows != its particle count" % e)
        return tuple(torch.cat([st[i] for st in states]) for i in range(4))

    def unpack(self, x, v, C, F):
        """Stacked state -> per-environment ``(x, v, C, F)`` views."""
        return [tuple(t[a:b] for t in (x, v, C, F))
                for a, b in zip(self.offsets[:-1], self.offsets[1:])]

    def rollout(self, states, n_substeps):
        """Differentiable rollout of every environment by n_substeps."""
        return self.unpack(*self.sim.rollout(*self.pack(states), n_substeps))

    def step_no_grad(self, states, n_substeps):
        return self.unpack(*self.sim.step_no_grad(*self.pack(states), n_substeps))

    def material_grad(self):
        """Per-environment ``(n_materials, 2)`` views of the accumulated d/d mu,
        d/d lam (``MPMSim(material_grad=True)``)."""
        g = self.sim.material_grad
        if g is None:
            return None
        return [g[a:b] for a, b in zip(self.mat_offsets[:-1], self.mat_offsets[1:])]


def concat_pad_meshes(meshes):
    """One mesh of the per-environment pads: ``meshes`` holds one
    ``(nodes, f2v, tris, layer)`` per environment (positions local to it).
    Returns ``(nodes, f2v, tris, layer, vert_start, tri_start)`` with the
    vertex indices offset into the concatenation."""
    nodes, f2v, tris, layer = [], [], [], []
    vs, ts = [0], [0]
    for nd, fv, tr, ly in meshes:
        off = vs[-1]
        nodes.append(np.asarray(nd, np.float32))
        f2v.append(np.asarray(fv, np.int32) + off)
        tris.append(np.asarray(tr, np.int32) + off)
        layer.append(np.asarray(ly, np.int32))
        vs.append(off + len(nd))
        ts.append(ts[-1] + len(tr))
    return (np.concatenate(nodes), np.concatenate(f2v), np.concatenate(tris),
            np.concatenate(layer), np.asarray(vs, np.int64), np.asarray(ts, np.int32))


class CoupledEnvBatch:
    """E environments, each an MPM object and one FEM pad in penalty contact,
    advanced together: the objects as an ``EnvBatch``, the pads concatenated
    into one ``Pad`` whose triangles each environment's grid nodes search
    within their own pad (``dt_contact_closest_env``).  The pads share one
    material and one outward direction.

    envs:       one ``(materials, n_particles, mat_id or None)`` per environment
    pad_meshes: one ``(nodes_rest, f2v, tris, layer)`` per environment
    pad:        dict(out_dir, mu, lam, mass) of the pads
    contact:    dict(kn, kd, kt, mu_f) of the penalty contact
    graphs:     capture each checkpoint segment into a HIP graph (CoupledSim)
    """

    def __init__(self, prm, envs, pad_meshes, pad, contact, device="cuda", graphs=False,
                 **sim_kwargs):
        from .coupled import CoupledSim, Pad
        if len(pad_meshes) != len(envs):
            raise ValueError("one pad mesh per environment")
        self.batch = EnvBatch(prm, envs, device=device, **sim_kwargs)
        nodes, f2v, tris, layer, vs, ts = concat_pad_meshes(pad_meshes)
        self.vert_offsets = vs
        self.pad = Pad(nodes, f2v, tris, layer, pad["out_dir"], pad["mu"], pad["lam"],
                       pad["mass"], prm.dt, device=device, tri_start=ts)
        self.cs = CoupledSim(self.batch.sim, [self.pad], **contact, graphs=graphs)

    def rollout(self, states, pad_states, control_vels, n_substeps):
        """states: per environment (x, v, C, F); pad_states: per environment
        (pos, vel, vpos); control_vels: per environment (n_substeps + 1, V_e, 3).
        Returns per environment ((x, v, C, F), (pos, vel, vpos))."""
        pst = tuple(torch.cat([p[i] for p in pad_states]) for i in range(3))
        cv = torch.cat(list(control_vels), dim=1)
        out = self.cs.rollout(*self.batch.pack(states), [pst], [cv], n_substeps)
        objs = self.batch.unpack(*out[:4])
        pads = out[4][0]
        vo = self.vert_offsets
        return [(o, tuple(t[vo[e]:vo[e + 1]] for t in pads)) for e, o in enumerate(objs)]
