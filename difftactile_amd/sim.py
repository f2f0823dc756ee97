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
of each call and again at every checkpoint (``resort``); the order is undone
on output, and the backward pass undoes each re-sort, so callers see their own
order.
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
                 material_grad=False, env_id=None, resort=True, launch=None):
        self.prm = prm
        # launch shapes of the persistent passes (process-wide): "cache" sets
        # the autotuner's cached shapes of this workload, a dict sets those
        if launch == "cache":
            from . import autotune
            autotune.apply_cached(prm, int(n_particles))
        elif isinstance(launch, dict):
            from . import autotune
            autotune.set_config(launch)
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
        # resort: with sort, the particles are sorted by cell again at every
        # checkpoint of a rollout (the order decays as they move: equal brick
        # keys stop being neighbours and grid reads stop being coalesced)
        self.resort = bool(resort)
        self._ws = {}
        self._free = {}
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

    # -- buffer pool ------------------------------------------------------
    # A rollout takes its state, id, frame and checkpoint buffers from here
    # and returns them after its backward pass, so the next rollout runs on
    # the same device addresses and the library's graph cache replays its
    # captured segments instead of capturing them again.
    def _take(self, taken, key, make):
        free = self._free.setdefault(key, [])
        bufs = free.pop() if free else make()
        taken.append((key, bufs))
        return bufs

    def _give_back(self, taken):
        # last taken, first back: a key taken several times in one rollout
        # hands out its buffers in the same order in the next one (graph
        # keys name the addresses)
        for key, bufs in reversed(taken):
            self._free.setdefault(key, []).append(bufs)
        taken.clear()

    def release_buffers(self):
        """Frees the pooled rollout buffers (resident frames included)."""
        self._free.clear()

    def _state_bufs(self, widths):
        return [torch.empty((w, self.n), device=self.device) for w in widths]

    def _id_buf(self):
        return torch.empty(self.n, dtype=torch.int32, device=self.device)

    def _permute(self, q, t, width, scatter=False, out=None):
        """[width][n] reordered by q: gather out[:, i] = t[:, q[i]], or the
        inverse, scatter out[:, q[i]] = t[:, i] (int32 arrays as bit patterns)."""
        out = torch.empty_like(t) if out is None else out
        src, dst = (t, out) if t.dtype == torch.float32 else (t.view(torch.float32),
                                                              out.view(torch.float32))
        _lib.call("dt_permute_soa", self.n, width, _lib.ptr(q, torch.int32), int(scatter),
                  _lib.ptr(src), _lib.ptr(dst), _lib.stream())
        return out

    def resort_state(self, state, mat_id, env, order, out=None):
        """Sorts the SoA state (x first) by cell key again.  Returns the
        permuted (state, mat_id, env, order) and q, the permutation applied
        (new position i holds the particle at old position q[i]).  out: the
        buffers to permute into (state list, mat_id, env, order, q)."""
        n = self.n
        keys = self._buf("keys", (n,), torch.int32)
        keys_out = self._buf("keys_out", (n,), torch.int32)
        if out is None:
            out = ([torch.empty_like(t) for t in state],
                   None if mat_id is None else torch.empty_like(mat_id),
                   None if env is None else torch.empty_like(env),
                   None if order is None else torch.empty_like(order), self._id_buf())
        st_o, mid_o, env_o, ord_o, q = out
        self._set_env(env)
        _lib.call("dt_mpm_cell_keys", self.c_prm, n, _lib.ptr(state[0]),
                  _lib.ptr(keys, torch.int32), _lib.stream())
        tb = int(_lib.lib().dt_sort_tmp_bytes(n))
        tmp = self._buf("sort_tmp", (max(tb, 1),), torch.uint8)
        _lib.call("dt_sort_by_key", n, self.prm.key_bits(), _lib.ptr(keys, torch.int32),
                  _lib.ptr(keys_out, torch.int32), _lib.ptr(q, torch.int32), tmp.data_ptr(), tb,
                  _lib.stream())
        state = [self._permute(q, t, t.shape[0], out=o) for t, o in zip(state, st_o)]
        mat_id = None if mat_id is None else self._permute(q, mat_id, 1, out=mid_o)
        env = None if env is None else self._permute(q, env, 1, out=env_o)
        order = None if order is None else self._permute(q, order, 1, out=ord_o)
        self._set_env(env)
        return state, mat_id, env, order, q

    def _gather(self, perm, t, width, out=None):
        """[n][width] in caller order -> [width][n] in sorted order."""
        if out is None:
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

    def _gather_i(self, perm, t, out=None):
        # int32 ids gathered through the float path (bit pattern preserved)
        out = torch.empty_like(t) if out is None else out
        _lib.call("dt_gather_rows", self.n, 1, _lib.ptr(perm, torch.int32),
                  _lib.ptr(t.view(torch.float32)), _lib.ptr(out.view(torch.float32)),
                  _lib.stream())
        return out

    # -- raw segment kernels ----------------------------------------------
    def _mid(self, mat_id):
        return None if mat_id is None else _lib.ptr(mat_id, torch.int32)

    def segment_bytes(self, k):
        """HBM bytes of one resident k-substep segment: frames, grids and (on
        the sparse path) its own workspace with the binning history."""
        # the sparse path keeps the (mv, m) grid of each substep only: its
        # adjoint remakes the velocities from it
        grids = 1 if self.sparse_grid else 2
        b = 4 * (self.n * 24 * (k + 1) + grids * 4 * k * self.prm.n_nodes)
        if self.sparse_grid:
            b += workspace_bytes(self.c_prm, self.n, k)
        return b

    def _kept_grids(self, k):
        """Fresh kept grids of a k-substep segment: (G4, GV4), GV4 None on the
        sparse path."""
        nn = self.prm.n_nodes
        return (torch.empty((k, nn, 4), device=self.device),
                None if self.sparse_grid else torch.empty((k, nn, 4), device=self.device))

    def segment_workspace(self, k):
        """A workspace with room for the binning history of a k-substep
        segment (a resident segment keeps its own, so its adjoint sweep reuses
        the forward binning after later segments ran)."""
        if not self.sparse_grid:
            return None
        return torch.empty(workspace_bytes(self.c_prm, self.n, k), dtype=torch.uint8,
                           device=self.device)

    def _ws_args(self, ws, history_substeps):
        if ws is None:
            return self._workspace(history_substeps)
        return _lib.ptr(ws, torch.uint8), ws.numel()

    def run_frames(self, X, V, C, F, k, mat_id, keep_grids, grids=None, ws=None):
        """Advance frames 0..k of frame-major buffers X[k+1,n,3] ... in place.
        keep_grids: keep every substep's grids (in `grids` when given); ws: the
        workspace to bin in (default the simulation's)."""
        nn = self.prm.n_nodes
        if keep_grids:
            if grids is not None:
                G4, GV4 = grids
            else:
                G4 = self._buf("G4", (k, nn, 4))
                GV4 = None if self.sparse_grid else self._buf("GV4", (k, nn, 4))
            # sparse path: the velocity grid of each substep is scratch
            s4, sv4 = None, (self._grid_scratch()[1] if GV4 is None else None)
            g4p, gv4p = _lib.ptr(G4), _lib.ptr(GV4)
        else:
            G4 = GV4 = None
            s4, sv4 = self._grid_scratch()
            g4p, gv4p = None, None
        _lib.call("dt_mpm_run_fwd", self.c_prm, _lib.ptr(self.mats), len(self.materials), self.n,
                  self._mid(mat_id), _lib.ptr(X), _lib.ptr(V), _lib.ptr(C), _lib.ptr(F), k,
                  g4p, gv4p, _lib.ptr(s4), _lib.ptr(sv4),
                  *self._ws_args(ws, k if keep_grids else 0), _lib.stream())
        return G4, GV4

    def advance(self, state, k, mat_id):
        """Advance state = [x, v, C, F] by k substeps in place (no history kept)."""
        s4, sv4 = self._grid_scratch()
        tmp = [self._buf("adv%d" % i, t.shape) for i, t in enumerate(state)]
        _lib.call("dt_mpm_advance", self.c_prm, _lib.ptr(self.mats), len(self.materials), self.n,
                  self._mid(mat_id), *[_lib.ptr(t) for t in state], *[_lib.ptr(t) for t in tmp],
                  int(k), _lib.ptr(s4), _lib.ptr(sv4), *self._workspace(), _lib.stream())

    def adj_frames(self, X, V, C, F, k, mat_id, G4, GV4, g, ws=None):
        """Reverse sweep over k substeps; g = [gx, gv, gC, gF] seeded at frame k, in place;
        ws: the workspace the forward binned in (its history is reused)."""
        nn = self.prm.n_nodes
        gg4 = self._buf("gg4", (nn, 4))
        ggv4 = self._buf("ggv4", (nn, 4))
        scratch = [self._buf("ga%d" % i, t.shape) for i, t in enumerate(g)]
        _lib.call("dt_mpm_run_adj", self.c_prm, _lib.ptr(self.mats), len(self.materials), self.n,
                  self._mid(mat_id), _lib.ptr(X), _lib.ptr(V), _lib.ptr(C), _lib.ptr(F), k,
                  _lib.ptr(G4), _lib.ptr(GV4), _lib.ptr(gg4), _lib.ptr(ggv4),
                  *[_lib.ptr(t) for t in g], *[_lib.ptr(t) for t in scratch],
                  *self._ws_args(ws, k), self._mat_grad_ptr(), _lib.stream())

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
        W = _Rollout.WIDTHS
        taken = []
        perm = sim.sort_order(x)[1] if (sim.sort and sim.n > 0) else None
        sbuf = [sim._take(taken, ("state", 0), lambda: sim._state_bufs(W)),
                sim._take(taken, ("state", 1), lambda: sim._state_bufs(W))]
        state = [sim._gather(perm, t, w, out=o) for t, w, o in zip((x, v, C, F), W, sbuf[0])]
        ids0 = sim._take(taken, ("ids", 0), lambda: [sim._id_buf() for _ in range(4)])
        if sim.mat_id is not None and perm is not None:
            mat_id = sim._gather_i(perm, sim.mat_id, out=ids0[0])
        else:
            mat_id = sim.mat_id
        env = sim.env_id
        if env is not None and perm is not None:
            env = sim._gather_i(perm, env, out=ids0[1])
        sim._set_env(env)
        # order[i]: the caller's index of the particle at position i
        order = perm
        ckpts, meta = [], []
        used = 0
        n, nn = sim.n, sim.prm.n_nodes
        for i, (s, k) in enumerate(_segments(n_substeps, sim.checkpoint_every)):
            q = None
            if s > 0 and order is not None and sim.resort:
                ids = sim._take(taken, ("ids", i), lambda: [sim._id_buf() for _ in range(4)])
                out = (sbuf[i % 2], None if mat_id is None else ids[0],
                       None if env is None else ids[1], ids[2], ids[3])
                state, mat_id, env, order, q = sim.resort_state(state, mat_id, env, order,
                                                                out=out)
            meta.append((mat_id, env, q))
            nbytes = sim.segment_bytes(k)
            if ctx is not None and used + nbytes <= sim.resident_bytes:
                # resident segment: keep frames and grids for the backward pass
                used += nbytes
                fr = sim._take(taken, ("frames", i, k), lambda: (
                    [torch.empty((k + 1, w, n), device=sim.device) for w in W],
                    sim._kept_grids(k), sim.segment_workspace(k)))
                frames, grids, seg_ws = fr
                for f, t in zip(frames, state):
                    f[0].copy_(t)
                sim.run_frames(*frames, k, mat_id, keep_grids=True, grids=grids, ws=seg_ws)
                for f, t in zip(frames, state):
                    t.copy_(f[k])
                ckpts.append((frames, grids, seg_ws))
            else:
                if ctx is not None:
                    ck = sim._take(taken, ("ckpt", i), lambda: sim._state_bufs(W))
                    for c, t in zip(ck, state):
                        c.copy_(t)
                    ckpts.append(ck)
                sim.advance(state, k, mat_id)
        out = [sim._scatter(order, t, w, sh) for t, w, sh in zip(state, W, shapes)]
        if ctx is not None:
            ctx.sim = sim
            ctx.n_substeps = n_substeps
            ctx.perm = perm
            ctx.order = order
            ctx.meta = meta
            ctx.ckpts = ckpts
            ctx.shapes = shapes
            ctx.taken = taken
        else:
            sim._give_back(taken)
        return tuple(out)

    @staticmethod
    def backward(ctx, gx, gv, gC, gF):
        sim = ctx.sim
        W = _Rollout.WIDTHS
        taken = ctx.taken
        gbuf = [sim._take(taken, ("grad", 0), lambda: sim._state_bufs(W)),
                sim._take(taken, ("grad", 1), lambda: sim._state_bufs(W))]
        g = [t if t is not None else torch.zeros(sh, device=sim.device)
             for t, sh in zip((gx, gv, gC, gF), ctx.shapes)]
        g = [sim._gather(ctx.order, t.float(), w, out=o) for t, w, o in zip(g, W, gbuf[0])]
        cur = 0
        segs = _segments(ctx.n_substeps, sim.checkpoint_every)
        for (s, k), ck, (mat_id, env, q) in zip(reversed(segs), reversed(ctx.ckpts),
                                                reversed(ctx.meta)):
            sim._set_env(env)
            seg_ws = None
            if isinstance(ck, tuple):           # resident segment
                (X, V, Cb, Fb), (G4, GV4), seg_ws = ck
            else:                               # recompute from the checkpoint
                X, V, Cb, Fb = sim._frames(k, "R%d" % k)
                X[0].copy_(ck[0]); V[0].copy_(ck[1]); Cb[0].copy_(ck[2]); Fb[0].copy_(ck[3])
                G4, GV4 = sim.run_frames(X, V, Cb, Fb, k, mat_id, keep_grids=True)
            sim.adj_frames(X, V, Cb, Fb, k, mat_id, G4, GV4, g, ws=seg_ws)
            if q is not None:
                # back to the order before this segment's re-sort
                cur = 1 - cur
                g = [sim._permute(q, t, w, scatter=True, out=o)
                     for t, w, o in zip(g, W, gbuf[cur])]
        ctx.ckpts = None
        ctx.meta = None
        g = [sim._scatter(ctx.perm, t, w, sh) for t, w, sh in zip(g, W, ctx.shapes)]
        sim._give_back(taken)
        return (None, None, *g)
