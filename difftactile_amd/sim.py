"""Checkpointed, differentiable MPM trajectories on the fused HIP path.

The reference keeps a window of ``sub_steps`` frames on the device, caches
frame 0 of every window on the host and recomputes each window in the
backward pass (``mpm_elastic.py:300-320``, ``grasp_plastic.py:594-648``).
Here the same structure lives entirely in HBM:

* the forward pass keeps only one state every ``checkpoint_every`` substeps;
* the backward pass recomputes one segment at a time, keeping that segment's
  frames and grids, and sweeps the fused adjoint kernels over it in reverse;
* segments whose frames and grids fit ``resident_bytes`` (HBM set aside for the
  trajectory) are kept from the forward pass instead, so the backward pass
  does not recompute them.  Both give the same gradients.

Particles are ordered by cell key (a stable integer radix sort) at the start
of each call; the order is undone on output so callers see their own order.
"""
import ctypes

import torch

from . import _lib
from .params import material_table


def workspace_bytes(c_prm, n, history_substeps=0):
    """Device workspace bytes of the sparse path for n particles (plus the
    binning history of a history_substeps trajectory)."""
    out = ctypes.c_size_t(0)
    _lib.check(_lib.lib().dt_mpm_workspace_bytes(c_prm, int(n), ctypes.byref(out)))
    total = int(out.value)
    if history_substeps:
        _lib.check(_lib.lib().dt_mpm_workspace_history_bytes(c_prm, int(n), int(history_substeps),
                                                             ctypes.byref(out)))
        total += int(out.value)
    return total


class MPMSim:
    """One MPM object on the GPU: parameters, material table, per-particle ids."""

    def __init__(self, prm, materials, n_particles, mat_id=None, device="cuda",
                 checkpoint_every=50, sort=True, sparse_grid=True, resident_bytes=None,
                 material_grad=False, env_id=None):
        self.prm = prm
        if resident_bytes is None:
            # a quarter of the device's memory for resident segments
            dev = torch.device(device)
            total = torch.cuda.get_device_properties(dev).total_memory if dev.type == "cuda" else 0
            resident_bytes = total // 4
        self.resident_bytes = int(resident_bytes)
        self.sparse_grid = bool(sparse_grid)
        # fused kernels run on the [k][n] (SoA) particle layout; the caller's
        # [n][k] arrays are transposed (and cell-sorted) at the boundary
        self.c_prm = prm.c_struct(soa=True)
        self.c_prm_aos = prm.c_struct(soa=False)
        self.n = int(n_particles)
        self.device = torch.device(device)
        self.materials = list(materials)
        self.mats = material_table(self.materials, self.device)
        if mat_id is None:
            self.mat_id = None
        else:
            self.mat_id = torch.as_tensor(mat_id, dtype=torch.int32, device=self.device).contiguous()
            if self.mat_id.numel() != self.n:
                raise ValueError("mat_id must have one entry per particle")
            mx = int(self.mat_id.max().item()) if self.n else 0
            mn = int(self.mat_id.min().item()) if self.n else 0
            if mn < 0 or mx >= len(self.materials):
                raise ValueError("mat_id out of range of the material table")
        # environments stacked along x (prm.n_env > 1): env_id[p] names particle
        # p's environment; positions stay local to it
        if prm.n_env > 1:
            if env_id is None:
                raise ValueError("prm.n_env > 1 needs env_id (one id per particle)")
            self.env_id = torch.as_tensor(env_id, dtype=torch.int32,
                                          device=self.device).contiguous()
            if self.env_id.numel() != self.n:
                raise ValueError("env_id must have one entry per particle")
            if self.n and (int(self.env_id.min()) < 0 or int(self.env_id.max()) >= prm.n_env):
                raise ValueError("env_id out of range [0, n_env)")
        else:
            self.env_id = None
        self.checkpoint_every = int(checkpoint_every)
        self.sort = bool(sort)
        self._ws = {}
        # material_grad: backward passes accumulate d loss / d (mu, lam) of each
        # material row here (cleared by zero_material_grad), as the reference's
        # mu_0.grad / lambda_0.grad accumulate
        self.material_grad = (torch.zeros((len(self.materials), 2), device=self.device)
                              if material_grad else None)

    def _set_env(self, env):
        """Points the parameter structs at the environment ids of the particle
        order the next launches see (and keeps that tensor alive as long as the
        structs name it)."""
        self._env_ref = env
        p = None if env is None else _lib.ptr(env, torch.int32)
        self.c_prm.env_id = p
        self.c_prm_aos.env_id = p

    def zero_material_grad(self):
        if self.material_grad is not None:
            self.material_grad.zero_()

    def _mat_grad_ptr(self):
        return None if self.material_grad is None else _lib.ptr(self.material_grad)

    # -- workspace -------------------------------------------------------
    def _buf(self, name, shape, dtype=torch.float32):
        t = self._ws.get(name)
        if t is None or tuple(t.shape) != tuple(shape) or t.dtype != dtype:
            t = torch.empty(shape, dtype=dtype, device=self.device)
            self._ws[name] = t
        return t

    def _grid_scratch(self):
        nn = self.prm.n_nodes
        return self._buf("g4", (nn, 4)), self._buf("gv4", (nn, 4))

    def _workspace(self, history_substeps=0):
        """(pointer, bytes) of the sparse-path workspace; (None, 0) selects the
        dense grid passes.  With history_substeps the workspace also holds the
        binning history of a segment that long (reused by its adjoint sweep)."""
        if not self.sparse_grid:
            return None, 0
        need = workspace_bytes(self.c_prm, self.n, history_substeps)
        t = self._ws.get("bin_ws")
        if t is None or t.numel() < need:
            t = torch.empty(need, dtype=torch.uint8, device=self.device)
            self._ws["bin_ws"] = t
        return _lib.ptr(t, torch.uint8), t.numel()

    # -- ordering --------------------------------------------------------
    def sort_order(self, x):
        """(sorted keys, permutation) of the particles at positions x by cell key."""
        n = self.n
        keys = self._buf("keys", (n,), torch.int32)
        keys_out = torch.empty(n, dtype=torch.int32, device=self.device)
        perm = torch.empty(n, dtype=torch.int32, device=self.device)
        self._set_env(self.env_id)
        _lib.call("dt_mpm_cell_keys", self.c_prm_aos, n, _lib.ptr(x), _lib.ptr(keys, torch.int32),
                  _lib.stream())
        tb = int(_lib.lib().dt_sort_tmp_bytes(n))
        tmp = self._buf("sort_tmp", (max(tb, 1),), torch.uint8)
        _lib.call("dt_sort_by_key", n, self.prm.key_bits(), _lib.ptr(keys, torch.int32),
                  _lib.ptr(keys_out, torch.int32), _lib.ptr(perm, torch.int32),
                  tmp.data_ptr(), tb, _lib.stream())
        return keys_out, perm

    def _gather(self, perm, t, width):
        """[n][width] in caller order -> [width][n] in sorted order."""
        out = torch.empty((width, self.n), dtype=torch.float32, device=self.device)
        src = t.contiguous()
        _lib.call("dt_gather_rows_soa", self.n, width, perm, src, out, _lib.stream())
        return out

    def _scatter(self, perm, t, width, shape):
        """[width][n] in sorted order -> caller-order tensor of `shape`."""
        out = torch.empty(shape, dtype=torch.float32, device=self.device)
        _lib.call("dt_scatter_rows_soa", self.n, width,
                  None if perm is None else _lib.ptr(perm, torch.int32), _lib.ptr(t),
                  _lib.ptr(out), _lib.stream())
        return out

    def _gather_i(self, perm, t):
        # int32 ids gathered through the float path (bit pattern preserved)
        out = torch.empty_like(t)
        _lib.call("dt_gather_rows", self.n, 1, _lib.ptr(perm, torch.int32),
                  _lib.ptr(t.view(torch.float32)), _lib.ptr(out.view(torch.float32)),
                  _lib.stream())
        return out

    # -- raw segment kernels ----------------------------------------------
    def _mid(self, mat_id):
        return None if mat_id is None else _lib.ptr(mat_id, torch.int32)

    def segment_bytes(self, k):
        """HBM bytes of one k-substep segment's frames and grids."""
        return 4 * (self.n * 24 * (k + 1) + 2 * 4 * k * self.prm.n_nodes)

    def run_frames(self, X, V, C, F, k, mat_id, keep_grids, grids=None):
        """Advance frames 0..k of frame-major buffers X[k+1,n,3] ... in place.
        keep_grids: keep every substep's grids (in `grids` when given)."""
        nn = self.prm.n_nodes
        if keep_grids:
            if grids is not None:
                G4, GV4 = grids
            else:
                G4 = self._buf("G4", (k, nn, 4))
                GV4 = self._buf("GV4", (k, nn, 4))
            s4, sv4 = None, None
            g4p, gv4p = _lib.ptr(G4), _lib.ptr(GV4)
        else:
            G4 = GV4 = None
            s4, sv4 = self._grid_scratch()
            g4p, gv4p = None, None
        _lib.call("dt_mpm_run_fwd", self.c_prm, _lib.ptr(self.mats), len(self.materials), self.n,
                  self._mid(mat_id), _lib.ptr(X), _lib.ptr(V), _lib.ptr(C), _lib.ptr(F), k,
                  g4p, gv4p, _lib.ptr(s4), _lib.ptr(sv4),
                  *self._workspace(k if keep_grids else 0), _lib.stream())
        return G4, GV4

    def advance(self, state, k, mat_id):
        """Advance state = [x, v, C, F] by k substeps in place (no history kept)."""
        s4, sv4 = self._grid_scratch()
        tmp = [self._buf("adv%d" % i, t.shape) for i, t in enumerate(state)]
        _lib.call("dt_mpm_advance", self.c_prm, _lib.ptr(self.mats), len(self.materials), self.n,
                  self._mid(mat_id), *[_lib.ptr(t) for t in state], *[_lib.ptr(t) for t in tmp],
                  int(k), _lib.ptr(s4), _lib.ptr(sv4), *self._workspace(), _lib.stream())

    def adj_frames(self, X, V, C, F, k, mat_id, G4, GV4, g):
        """Reverse sweep over k substeps; g = [gx, gv, gC, gF] seeded at frame k, in place."""
        nn = self.prm.n_nodes
        gg4 = self._buf("gg4", (nn, 4))
        ggv4 = self._buf("ggv4", (nn, 4))
        scratch = [self._buf("ga%d" % i, t.shape) for i, t in enumerate(g)]
        _lib.call("dt_mpm_run_adj", self.c_prm, _lib.ptr(self.mats), len(self.materials), self.n,
                  self._mid(mat_id), _lib.ptr(X), _lib.ptr(V), _lib.ptr(C), _lib.ptr(F), k,
                  _lib.ptr(G4), _lib.ptr(GV4), _lib.ptr(gg4), _lib.ptr(ggv4),
                  *[_lib.ptr(t) for t in g], *[_lib.ptr(t) for t in scratch],
                  *self._workspace(k), self._mat_grad_ptr(), _lib.stream())

    def _frames(self, k, prefix="F"):
        n = self.n
        return (self._buf(prefix + "x", (k + 1, 3, n)), self._buf(prefix + "v", (k + 1, 3, n)),
                self._buf(prefix + "C", (k + 1, 9, n)), self._buf(prefix + "F", (k + 1, 9, n)))

    # -- public API ------------------------------------------------------
    def rollout(self, x, v, C, F, n_substeps):
        """Differentiable rollout: returns (x, v, C, F) after n_substeps."""
        return _Rollout.apply(self, int(n_substeps), x, v, C, F)

    def step_no_grad(self, x, v, C, F, n_substeps):
        with torch.no_grad():
            return _Rollout.forward(None, self, int(n_substeps), x, v, C, F)


def _segments(n_substeps, k):
    s = 0
    out = []
    while s < n_substeps:
        out.append((s, min(k, n_substeps - s)))
        s += k
    return out


class _Rollout(torch.autograd.Function):
    WIDTHS = (3, 3, 9, 9)

    @staticmethod
    def forward(ctx, sim, n_substeps, x, v, C, F):
        for t in (x, v, C, F):
            if t.device.type != "cuda" or t.dtype != torch.float32:
                raise ValueError("state tensors must be float32 on the GPU")
        x, v, C, F = (t.contiguous() for t in (x, v, C, F))
        shapes = [t.shape for t in (x, v, C, F)]
        perm = sim.sort_order(x)[1] if (sim.sort and sim.n > 0) else None
        state = [sim._gather(perm, t, w) for t, w in zip((x, v, C, F), _Rollout.WIDTHS)]
        if sim.mat_id is not None and perm is not None:
            mat_id = sim._gather_i(perm, sim.mat_id)
        else:
            mat_id = sim.mat_id
        env = sim.env_id
        if env is not None and perm is not None:
            env = sim._gather_i(perm, env)
        sim._set_env(env)
        ckpts = []
        used = 0
        for (s, k) in _segments(n_substeps, sim.checkpoint_every):
            nbytes = sim.segment_bytes(k)
            if ctx is not None and used + nbytes <= sim.resident_bytes:
                # resident segment: keep frames and grids for the backward pass
                used += nbytes
                n, nn = sim.n, sim.prm.n_nodes
                fr = [torch.empty((k + 1, w, n), device=sim.device) for w in _Rollout.WIDTHS]
                grids = (torch.empty((k, nn, 4), device=sim.device),
                         torch.empty((k, nn, 4), device=sim.device))
                for f, t in zip(fr, state):
                    f[0].copy_(t)
                sim.run_frames(*fr, k, mat_id, keep_grids=True, grids=grids)
                for f, t in zip(fr, state):
                    t.copy_(f[k])
                ckpts.append((fr, grids))
            else:
                if ctx is not None:
                    ckpts.append([t.clone() for t in state])
                sim.advance(state, k, mat_id)
        out = [sim._scatter(perm, t, w, sh) for t, w, sh in zip(state, _Rollout.WIDTHS, shapes)]
        if ctx is not None:
            ctx.sim = sim
            ctx.n_substeps = n_substeps
            ctx.perm = perm
            ctx.mat_id = mat_id
            ctx.env = env
            ctx.ckpts = ckpts
            ctx.shapes = shapes
        return tuple(out)

    @staticmethod
    def backward(ctx, gx, gv, gC, gF):
        sim = ctx.sim
        perm = ctx.perm
        g = [t if t is not None else torch.zeros(sh, device=sim.device)
             for t, sh in zip((gx, gv, gC, gF), ctx.shapes)]
        g = [sim._gather(perm, t.float(), w) for t, w in zip(g, _Rollout.WIDTHS)]
        segs = _segments(ctx.n_substeps, sim.checkpoint_every)
        sim._set_env(ctx.env)
        for (s, k), ck in zip(reversed(segs), reversed(ctx.ckpts)):
            if isinstance(ck, tuple):           # resident segment
                (X, V, Cb, Fb), (G4, GV4) = ck
            else:                               # recompute from the checkpoint
                X, V, Cb, Fb = sim._frames(k, "R%d" % k)
                X[0].copy_(ck[0]); V[0].copy_(ck[1]); Cb[0].copy_(ck[2]); Fb[0].copy_(ck[3])
                G4, GV4 = sim.run_frames(X, V, Cb, Fb, k, ctx.mat_id, keep_grids=True)
            sim.adj_frames(X, V, Cb, Fb, k, ctx.mat_id, G4, GV4, g)
        ctx.ckpts = None
        g = [sim._scatter(perm, t, w, sh) for t, w, sh in zip(g, _Rollout.WIDTHS, ctx.shapes)]
        return (None, None, *g)
