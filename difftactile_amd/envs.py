"""Batched tactile environments behind the reference's task protocol.

The reference's RL and CMA-ES drivers (``baseline/tactile_env.py:26-33``,
``baseline/cma_baseline.py:116-125``) drive one ``Contact`` scene through
``prepare_env``, ``init`` / ``reset``, ``apply_action(action, ts)``,
``calculate_force(ts)``, ``compute_loss(ts)`` and ``loss[None]``, one scene
per process.  ``PadPressEnvs`` keeps that protocol for a population of
environments advanced together (``batch.CoupledEnvBatch``: one MPM object and
one FEM pad each, environments of one grid size stacked along x): actions and
losses carry a leading environment axis, and one ``calculate_force`` advances
every environment by ``sub_steps`` substeps.

An action is a velocity added to each pad's base control velocity for the
next ``sub_steps`` substeps (the reference's ``d_pos`` per outer step); the
loss of an environment is its pad's summed squared vertex displacement from
the rigidly driven ("virtual") pad, the quantity its markers measure.
Gradients through a whole trajectory are available from
``CoupledEnvBatch.rollout``; this shim is the forward, step-by-step driver.
"""
import numpy as np
import torch

from .batch import CoupledEnvBatch


class _Loss:
    """``loss[None]`` as in the Taichi scalar fields the drivers read."""

    def __init__(self, n):
        self.value = np.zeros(n)

    def __getitem__(self, key):
        return self.value


class PadPressEnvs:
    """envs: per environment a dict as ``scenes.config4_population(pads=True)``
    gives (prm, mats, x, v, mat_id, pad_mesh, pad_out, pad, pad_cv); all must
    share one grid size.  contact: dict(kn, kd, kt, mu_f)."""

    def __init__(self, envs, contact, sub_steps=10, device="cuda"):
        grids = {d["prm"].n_grid for d in envs}
        if len(grids) != 1:
            raise ValueError("one grid size per PadPressEnvs (group environments by grid)")
        self.envs = envs
        self.n_env = len(envs)
        self.sub_steps = int(sub_steps)
        self.device = torch.device(device)
        d0 = envs[0]
        spec = [(d["mats"], len(d["x"]), d["mat_id"]) for d in envs]
        self.batch = CoupledEnvBatch(d0["prm"], spec, [d["pad_mesh"] for d in envs],
                                     dict(out_dir=d0["pad_out"], **d0["pad"]), contact,
                                     checkpoint_every=self.sub_steps, sort=True, resident_bytes=0)
        self.loss = _Loss(self.n_env)
        self.action_dim = 3
        self.prepare_env()

    # -- protocol -----------------------------------------------------------
    def prepare_env(self):
        dev = self.device
        self._x0 = [torch.tensor(d["x"], device=dev) for d in self.envs]
        self._v0 = [torch.tensor(d["v"], device=dev) for d in self.envs]
        self._nodes = [torch.tensor(d["pad_mesh"][0], device=dev) for d in self.envs]
        self._cv = [torch.tensor(d["pad_cv"], device=dev) for d in self.envs]
        self.init()

    def init(self):
        """Initial object and pad states (the reference's init)."""
        self.states = []
        for x, v in zip(self._x0, self._v0):
            n = x.shape[0]
            self.states.append([x.clone(), v.clone(), torch.zeros(n, 3, 3, device=self.device),
                                torch.eye(3, device=self.device).repeat(n, 1, 1)])
        self.pads = [(nd.clone(), cv.clone(), nd.clone()) for nd, cv in zip(self._nodes, self._cv)]
        self.actions = torch.zeros(self.n_env, 3, device=self.device)
        self.time_step = 0

    def reset(self):
        self.init()
        self.loss.value[:] = 0.0

    def clear_all_grad(self):
        """No gradient state is kept between steps (kept for the protocol)."""

    def apply_action(self, actions, ts=None):
        """actions [n_env, 3] (or [3] for every environment): velocity added
        to each pad's base control velocity for the next step."""
        a = torch.as_tensor(np.asarray(actions, np.float32), device=self.device)
        self.actions = a.expand(self.n_env, 3).clone() if a.dim() == 1 else a.reshape(self.n_env, 3)

    def calculate_force(self, ts=None):
        """Advance every environment by sub_steps substeps under the current
        actions."""
        T = self.sub_steps
        cvs = [(cv + a).expand(T + 1, -1, -1).contiguous() for cv, a in zip(self._cv, self.actions)]
        with torch.no_grad():
            out = self.batch.rollout(self.states, self.pads, cvs, T)
        self.states = [list(o) for o, _ in out]
        self.pads = [tuple(p) for _, p in out]
        self.time_step += 1

    def compute_loss(self, ts=None):
        """Per environment: sum over pad vertices of |pos - virtual pos|^2."""
        vals = [((p[0] - p[2]) ** 2).sum() for p in self.pads]
        self.loss.value = torch.stack(vals).double().cpu().numpy()
        return self.loss.value

    def observations(self):
        """[n_env, 9]: object centre of mass, pad mean displacement from its
        virtual pose, pad mean vertex velocity."""
        rows = []
        for st, p in zip(self.states, self.pads):
            rows.append(torch.cat([st[0].mean(0), (p[0] - p[2]).mean(0), p[1].mean(0)]))
        return torch.stack(rows).cpu().numpy()

    def step(self, actions):
        """apply_action + calculate_force + compute_loss; returns (observations,
        rewards = -loss, dones, truncated) with a leading environment axis."""
        self.apply_action(actions, self.time_step + 1)
        self.calculate_force(self.time_step + 1)
        loss = self.compute_loss(self.time_step)
        obs = self.observations()
        bad = ~np.isfinite(obs).all(axis=1)
        return obs, -loss, np.zeros(self.n_env, bool), bad
