"""Differentiable coupled rollouts: MPM object + FEM pads + grid penalty contact.

Composes the fused object stages with the pad and contact kernels in the
order of ``Contact.update`` (``difftactile/tasks/grasp_plastic.py:151-164``):

  zero grid -> p2g (fused compute_new_F/svd/return map) -> pad update ->
  closest triangle -> collision -> grid_op -> g2p -> pad update2

With the object's sparse grid (the default), p2g and the g2p adjoint run as
the brick-owned binned passes and their tiles are summed into the dense grid
the contact kernels read; particles are cell-sorted at the start of a
rollout (undone on output) when the object sorts.

and its adjoint in the order of ``Contact.update_grad`` (167-178), including
``clamp_grid`` (191-202).  Pads use the gripper form of update2 (a control
velocity per frame, ``gripper_fem.py:403-418``), so the gradient with respect
to the pad velocity sequence is available.  Checkpointing as in ``sim.py``:
one object + pad state every ``checkpoint_every`` substeps in HBM; the backward
pass recomputes each segment keeping its frames, grids, contact impulses and
closest-triangle indices.  Segments whose frames and per-substep stores fit
the object's ``resident_bytes`` (HBM set aside for the trajectory; a quarter
of the device by default) are kept from the forward pass instead, so the
backward pass does not recompute them.
"""
import numpy as np
import torch

from . import _lib
from .sim import MPMSim, _segments

P = _lib.ptr
I32 = torch.int32


def vertex_tet_csr(f2v, n_verts):
    """(start[n_verts+1], entries[4 n_tets]) with entries 4*tet + corner grouped
    by vertex (stable in tet order)."""
    f2v = np.asarray(f2v, dtype=np.int64).reshape(-1, 4)
    key = f2v.reshape(-1)
    order = np.argsort(key, kind="stable")
    start = np.searchsorted(key[order], np.arange(n_verts + 1), side="left").astype(np.int32)
    return start, order.astype(np.int32)


class Pad:
    """Static data of one FEM pad on the GPU (mesh, rest shapes, material).

    tri_start: for environments stacked along x (one pad per environment, the
    pads' vertices, tets and triangles concatenated), the first triangle of
    each environment's pad, [n_env + 1]; environment e's grid nodes contact
    triangles [tri_start[e], tri_start[e+1]) only."""

    def __init__(self, nodes_rest, f2v, tris, layer, out_dir, mu, lam, mass, dt, device="cuda",
                 tri_start=None):
        self.device = torch.device(device)
        self.tri_start = (None if tri_start is None else
                          torch.as_tensor(np.asarray(tri_start), dtype=I32,
                                          device=self.device).contiguous())
        self.n_verts = len(nodes_rest)
        self.n_tets = len(f2v)
        self.n_tris = len(tris)
        self.f2v = torch.as_tensor(np.asarray(f2v), dtype=I32, device=self.device).contiguous()
        self.tris = torch.as_tensor(np.asarray(tris), dtype=I32, device=self.device).contiguous()
        self.layer = torch.as_tensor(np.asarray(layer), dtype=I32, device=self.device).contiguous()
        self.out_dir = np.ascontiguousarray(np.asarray(out_dir, dtype=np.float32))
        rest = torch.as_tensor(np.asarray(nodes_rest), dtype=torch.float32,
                               device=self.device).contiguous()
        self.B = torch.empty(self.n_tets, 3, 3, device=self.device)
        _lib.call("dt_fem_init_B", self.n_tets, self.f2v, rest, self.B, _lib.stream())
        self.fem = _lib.DtFemParams()
        self.fem.mu, self.fem.lam, self.fem.dt, self.fem.mass = mu, lam, dt, mass
        self.cen = torch.empty(self.n_tris, 4, device=self.device)
        # bounding boxes of runs and groups of centroids (the accelerated search)
        self.tri_bounds = torch.empty(max(1, _lib.tri_bounds_count(self.n_tris)), 4,
                                      device=self.device)
        # vertex -> (tet, corner) CSR for the atomic-free tet-force passes
        vs, vt = vertex_tet_csr(np.asarray(f2v), self.n_verts)
        self.vert_start = torch.as_tensor(vs, device=self.device)
        self.vert_tets = torch.as_tensor(vt, device=self.device)
        self.tet_buf = torch.empty(self.n_tets, 4, 4, device=self.device)
        # exact (fixed-point) vertex sums of the contact terms; left zero by
        # every call
        self.fx_acc = torch.zeros(6 * self.n_verts, dtype=torch.int64, device=self.device)
        self.fx_scale = torch.zeros(4, dtype=torch.int32, device=self.device)

    def od(self):
        return self.out_dir.ctypes.data_as(_lib.ctypes.POINTER(_lib.ctypes.c_float))


class CoupledSim:
    """An MPM object, FEM pads and optionally a position-based cable (``rope``,
    a ``PBDRope`` whose particle 0 stays at ``rope.init_pos[0]``), coupled by
    grid penalty contact (object - pads) and particle-wise contact (cable -
    pads, the cable task's force with ``rope_contact`` = dict(kn, kd, kt,
    mu_f), default the grid constants).  With a cable the substep composes the
    grasp and cable tasks' updates (grasp_plastic.py:151-164,
    cable_straightening.py:166-173): p2g -> pad update -> grid contact ->
    cable contact -> grid_op -> g2p -> cable step -> pad update2 with the sum
    of both contacts' pad forces; the adjoint reverses it, clipping the
    cable's d/d pos, d/d vel as the cable task's clamp_grid does."""

    def __init__(self, sim: MPMSim, pads, kn, kd, kt, mu_f, norm_eps=1e-11, clamp_grad=True,
                 rope=None, rope_contact=None, graphs=False):
        self.sim = sim
        # graphs: each checkpoint segment's forward and adjoint launch
        # sequences (a lone substep's, outside a rollout) are captured once
        # into a HIP graph and replayed while the pointers they touch repeat
        # (they do across rollouts of a steady caller); a caller whose
        # buffers keep moving turns it off after 512 captures
        self.graphs = bool(graphs)
        self.graph_hits = self.graph_misses = 0
        self._in_segment = False
        self.pads = list(pads)
        self.rope = rope
        if rope is not None:
            from .object_model.pbd_rope import _contact_params
            rc = rope_contact or dict(kn=kn, kd=kd, kt=kt, mu_f=mu_f)
            self.rope_cp = _contact_params(rc["kn"], rc["kd"], rc["kt"], rc["mu_f"], rope.p_rad)
            self.rope_prm = rope.params()
            if rope.init_pos is None:
                raise ValueError("the cable needs init() (its fixed end) before coupling")
            self.rope_fixed = np.ascontiguousarray(rope.init_pos[0].detach().cpu().numpy(),
                                                   np.float32)
            nr = rope.n_particles
            self.rope_work = torch.empty(12 * nr, device=sim.device)
            self.rope_gimp = torch.empty(nr, 3, device=sim.device)
        self.cp = _lib.DtContactParams()
        self.cp.kn, self.cp.kd, self.cp.kt, self.cp.mu_f = kn, kd, kt, mu_f
        self.cp.norm_eps, self.cp.sdf_offset = norm_eps, 0.0
        self.clamp_grad = bool(clamp_grad)
        self.c_prm_clamp = sim.prm.c_struct(soa=True)
        self.c_prm_clamp.clamp_grad = int(clamp_grad)
        self.c_prm_clamp.grad_clamp = 1000.0
        nn = sim.prm.n_nodes
        d = sim.device
        self.npad = max(1, len(self.pads))
        self.grid_f = torch.zeros(nn, 3, device=d)
        self.idx = torch.full((nn, self.npad), -1, dtype=I32, device=d)
        if rope is not None:
            self.ridx = torch.full((rope.n_particles, self.npad), -1, dtype=I32, device=d)
            self.rimp = torch.zeros(rope.n_particles, 3, device=d)
        # per-pad buffers of the substeps that keep no store (no allocation
        # inside a substep: it may be captured)
        self._vu = [torch.empty(pd.n_verts, 3, device=d) for pd in self.pads]
        self._ext = [torch.empty(pd.n_verts, 3, device=d) for pd in self.pads]
        self._ge = [torch.empty(pd.n_verts, 3, device=d) for pd in self.pads]
        self._last_ext = self._ext
        # compact stores (sparse grids): a kept substep keeps the grids of its
        # active bricks only (the adjoint's node passes read no other node);
        # the substep runs on the same working grids as one that keeps nothing.
        # None: when whole-grid stores do not fit the resident budget; True:
        # always; False: never
        self.compact_stores = None
        self.overflow = torch.zeros(1, dtype=I32, device=d)
        self.last_resident = (0, 0, 0)

    def _node_set(self, base, ws, nb):
        """Copy of the parameter struct base whose grid-node passes run over
        the active bricks of the workspace's last binning."""
        p = type(base).from_buffer_copy(base)
        a, c = _lib.ctypes.c_void_p(), _lib.ctypes.c_void_p()
        _lib.call("dt_mpm_active_bricks", base, self.sim.n, ws, nb, _lib.ctypes.byref(a),
                  _lib.ctypes.byref(c))
        p.active_bricks, p.n_active_bricks = a.value, c.value
        return p

    def substep_bytes(self, cap=None):
        """HBM bytes one kept substep holds for the backward pass: object and
        pad frames, the packed grids, contact forces, closest triangles and the
        pads' updated velocities.  cap: the store keeps the nodes of at most
        cap active bricks (None: every node)."""
        nn = self.sim.prm.n_nodes if cap is None else 64 * cap
        b = 4 * self.sim.n * 24 + nn * (16 + 16 + 12 + 4 * self.npad) + self._bin_slot_bytes()
        for pd in self.pads:
            b += 4 * 3 * pd.n_verts * 4
        if self.rope is not None:
            b += self.rope.n_particles * (4 * 9 + 4 * self.npad)
        return b

    def _fixed_ptr(self):
        return self.rope_fixed.ctypes.data_as(_lib.ctypes.POINTER(_lib.ctypes.c_float))

    def _bin_slot_bytes(self):
        """Bytes of one substep's saved binning (0 on the dense path)."""
        if not self.sim.sparse_grid:
            return 0
        nb = _lib.ctypes.c_size_t()
        _lib.call("dt_mpm_bin_slot_bytes", self.sim.c_prm, self.sim.n, _lib.ctypes.byref(nb))
        return nb.value

    def _pool(self, taken, key, shape, dtype=torch.float32):
        """A rollout buffer from the simulation's pool (taken records it for
        the return); None taken: a fresh tensor."""
        make = lambda: torch.empty(tuple(shape), dtype=dtype, device=self.sim.device)
        if taken is None:
            return make()
        return self.sim._take(taken, ("coupled", id(self)) + tuple(key) + (tuple(shape), dtype),
                              make)

    def _copy_in(self, taken, key, t):
        return self._pool(taken, key, t.shape, t.dtype).copy_(t)

    def store_capacity(self, x):
        """Active bricks a compact store of a segment starting at the sorted
        positions x has room for: the bricks active there, a quarter more and
        64, rounded up to 1/8 of their power of two (so a steady caller's
        segments reuse the same pooled buffers); None without compact stores.
        Reads one count from the device."""
        sim = self.sim
        if self.compact_stores is False or not sim.sparse_grid:
            return None
        ws, nb = sim._workspace()
        s = _lib.stream()
        _lib.call("dt_mpm_bin", sim.c_prm, sim.n, P(x), ws, nb, s)
        a, c = _lib.ctypes.c_void_p(), _lib.ctypes.c_void_p()
        _lib.call("dt_mpm_active_bricks", sim.c_prm, sim.n, ws, nb, _lib.ctypes.byref(a),
                  _lib.ctypes.byref(c))
        wst = sim._ws["bin_ws"]
        off = c.value - wst.data_ptr()
        n_act = int(wst[off:off + 4].view(I32).item())
        want = int(n_act * 1.25) + 64
        g = max(64, 1 << max(0, want.bit_length() - 4))
        total = -(-sim.prm.n_nodes // 64)
        return min(-(-want // g) * g, total + 64)

    def _store_copy(self, prm, store, direction, grids):
        """Pack (direction 0) the active bricks' nodes of the working grids
        (g4, gv4, gf, idx) into a compact store, or unpack (1) them back."""
        sv = _lib.ctypes.c_void_p * 4
        names = ("g4", "gv4", "gf", "idx")
        dense = sv(*[t.data_ptr() for t in grids])
        comp = sv(*[store[c].data_ptr() for c in names])
        widths = (_lib.ctypes.c_int32 * 4)(4, 4, 3, self.npad)
        _lib.call("dt_mpm_brick_nodes_copy", prm, 4, dense, comp, widths, store["cap"], direction,
                  P(self.overflow, I32) if direction == 0 else None, _lib.stream())

    def _overflowed(self):
        hit = bool(self.overflow.item())
        self.overflow.zero_()
        return hit

    def _new_store(self, hint, taken=None, skey=(), cap=None):
        sim = self.sim
        nn = sim.prm.n_nodes if cap is None else 64 * cap
        b = lambda name, shape, dtype=torch.float32: self._pool(taken, ("store",) + skey + (name,),
                                                                shape, dtype)
        return {"g4": b("g4", (nn, 4)), "gv4": b("gv4", (nn, 4)), "gf": b("gf", (nn, 3)),
                "idx": b("idx", (nn, self.npad), I32), "hint": hint, "cap": cap,
                "vel_u": [b(("vel_u", k), (pd.n_verts, 3)) for k, pd in enumerate(self.pads)],
                "ext": [b(("ext", k), (pd.n_verts, 3)) for k, pd in enumerate(self.pads)],
                "ridx": (None if self.rope is None else
                         b("ridx", (self.rope.n_particles, self.npad), I32)),
                "rimp": (None if self.rope is None else b("rimp", (self.rope.n_particles, 3))),
                "bin": (b("bin", (self._bin_slot_bytes(),), torch.uint8)
                        if sim.sparse_grid else None)}

    def _run_keep(self, st0, pst0, cvs, s0, k, mat_id, taken=None, seg=(), cap=None):
        """Advance substeps s0 .. s0+k-1 from (st0, pst0) keeping every frame
        and each substep's store.  Returns (frames, pad frames, stores); st0
        and pst0 are copied, not kept.  taken / seg: the frames and stores
        come from the buffer pool under keys of segment seg.  cap: compact
        stores with room for cap active bricks (None: dense stores)."""
        fb = lambda j, a, t: self._pool(taken, ("frame",) + seg + (j, a), t.shape, t.dtype)
        pb = lambda j, b_, c, t: self._pool(taken, ("pframe",) + seg + (j, b_, c), t.shape, t.dtype)
        frames = [[fb(0, a, t).copy_(t) for a, t in enumerate(st0)]]
        pframes = [[[pb(0, b_, c, t).copy_(t) for c, t in enumerate(p)] for b_, p in enumerate(pst0)]]
        stores = []
        # every buffer first (nothing is allocated while a segment is captured)
        for j in range(k):
            # closest-triangle starts: the previous substep's result
            stores.append(self._new_store(stores[-1]["idx"] if stores else self.idx, taken,
                                          seg + (j,), cap))
            frames.append([fb(j + 1, a, t) for a, t in enumerate(frames[-1])])
            pframes.append([[pb(j + 1, b_, c, t) for c, t in enumerate(p)]
                            for b_, p in enumerate(pframes[-1])])
        calls = [(frames[j], frames[j + 1], pframes[j], pframes[j + 1],
                  [c[s0 + j] for c in cvs], [c[s0 + j + 1] for c in cvs], mat_id, stores[j])
                 for j in range(k)]
        self._fwd_segment(calls)
        return frames, pframes, stores

    def _fwd_segment(self, calls):
        """The substeps _fwd(*c) for c in calls, in order, as one captured
        graph when graphs are on."""
        if not calls:
            return
        rid = self.ridx if self.rope is not None else None

        def body():
            for c in calls:
                self._fwd(*c)

        self._segment([1, rid] + calls, body)
        last = calls[-1]
        self._last_ext = last[7]["ext"] if len(last) > 7 and last[7] else self._ext

    def _adj_segment(self, calls):
        """The substep adjoints _adj(*c) for c in calls, in order, as one
        captured graph when graphs are on."""
        if not calls:
            return

        def body():
            for c in calls:
                self._adj(*c)

        self._segment([2] + calls, body)

    def _segment(self, keyobjs, body):
        if not self.graphs or self._in_segment:
            return body()
        key = self._graph_key(3, *keyobjs)
        self._in_segment = True
        try:
            self._graphed(key, body)
        finally:
            self._in_segment = False

    # -- graph capture of a substep -------------------------------------------
    def _graph_key(self, tag, *objs):
        # lazily made buffers first, so the key names the ones a capture uses
        self._warm()
        words = [tag, id(self)]

        def add(o):
            if o is None:
                words.append(0)
            elif isinstance(o, torch.Tensor):
                words.extend((o.data_ptr(), o.numel()))
            elif isinstance(o, dict):
                for k in sorted(o):
                    add(o[k])
            elif isinstance(o, (list, tuple)):
                words.append(len(o))
                for t in o:
                    add(t)
            else:
                words.append(int(o) & 0xFFFFFFFFFFFFFFFF)

        for o in objs:
            add(o)
        # the simulation's scratch buffers (a regrown one moves)
        for name in sorted(self.sim._ws, key=str):
            add(self.sim._ws[name])
        for prm in (self.sim.c_prm, self.c_prm_clamp):
            b = bytes(prm)
            b += bytes(-len(b) % 8)
            words.extend(np.frombuffer(b, np.uint64).tolist())
        return np.asarray(words, dtype=np.uint64)

    def _warm(self):
        """Allocate the lazily made buffers a substep uses (before a capture)."""
        sim = self.sim
        nn = sim.prm.n_nodes
        if sim.sparse_grid:
            sim._workspace()
        sim._grid_scratch()
        sim._buf("cgg4", (nn, 4))
        sim._buf("cggv4", (nn, 4))
        sim._buf("cggf", (nn, 3))

    def _graphed(self, key, body):
        main = _lib.stream()
        hit = _lib.ctypes.c_int32(0)
        _lib.call("dt_graph_replay", main, key.ctypes.data, len(key), _lib.ctypes.byref(hit))
        if hit.value:
            self.graph_hits += 1
            return
        self._warm()
        # captured on a stream of its own (the default stream cannot be);
        # nothing runs during the capture, the graph then runs on main
        if getattr(self, "_cap_stream", None) is None:
            self._cap_stream = torch.cuda.Stream(device=self.sim.device)
        side = self._cap_stream
        cap = _lib.ctypes.c_int32(0)
        _lib.call("dt_graph_begin", side.cuda_stream, _lib.ctypes.byref(cap))
        if not cap.value:
            body()
            return
        try:
            with torch.cuda.stream(side):
                body()
        except BaseException:
            _lib.call("dt_graph_abort", side.cuda_stream)
            raise
        _lib.call("dt_graph_end", side.cuda_stream, main, key.ctypes.data, len(key))
        self.graph_misses += 1
        # more distinct sequences than the library's graph cache holds
        # (256) never replay: stop capturing
        if self.graph_misses >= 512 and self.graph_hits < self.graph_misses:
            self.graphs = False

    # -- one substep --------------------------------------------------------
    def _fwd(self, st, nxt, pads_in, pads_out, cv, cv1, mat_id, store=None):
        """st/nxt: [x, v, C, F] SoA; pads_in/out: per pad [pos, vel, vpos];
        store: dict receiving what the adjoint needs (or None)."""
        exts = store["ext"] if store else self._ext
        self._last_ext = exts
        if not self.graphs or self._in_segment:
            return self._fwd_body(st, nxt, pads_in, pads_out, cv, cv1, mat_id, store)
        key = self._graph_key(1, st, nxt, pads_in, pads_out, cv, cv1, mat_id, store,
                              self.ridx if self.rope is not None else None)
        self._graphed(key, lambda: self._fwd_body(st, nxt, pads_in, pads_out, cv, cv1, mat_id,
                                                  store))

    def _fwd_body(self, st, nxt, pads_in, pads_out, cv, cv1, mat_id, store=None):
        sim, s = self.sim, _lib.stream()
        nn = sim.prm.n_nodes
        prm = sim.c_prm
        compact = bool(store) and store.get("cap") is not None
        dense = bool(store) and not compact
        g4 = store["g4"] if dense else sim._grid_scratch()[0]
        gv4 = store["gv4"] if dense else sim._grid_scratch()[1]
        gf = store["gf"] if dense else self.grid_f
        idx = store["idx"] if dense else self.idx
        # closest-triangle starts: the previous substep's result
        hint = store.get("hint", idx) if dense else idx
        mats, nm = P(sim.mats), len(sim.materials)
        mid = None if mat_id is None else P(mat_id, I32)
        if sim.sparse_grid:
            # every grid pass below runs over the active bricks only: the
            # tiles sum writes each of their nodes, nodes elsewhere are neither
            # written nor read
            ws, nb = sim._workspace()
            _lib.call("dt_mpm_bin", prm, sim.n, P(st[0]), ws, nb, s)
            if store and store.get("bin") is not None:
                # kept for the adjoint, which then need not bin again
                _lib.call("dt_mpm_bin_copy", prm, sim.n, ws, nb, P(store["bin"], torch.uint8), 1, s)
            prm = self._node_set(prm, ws, nb)
            _lib.call("dt_mpm_p2g_binned", prm, mats, nm, sim.n, mid, *[P(t) for t in st],
                      P(nxt[3]), ws, nb, s)
            # the tile sums, with the contact force grid cleared on the same nodes
            _lib.call("dt_mpm_tiles_sum_zero", prm, sim.n, ws, nb, P(g4), P(gf), 3, s)
        else:
            _lib.call("dt_mpm_zero4", nn, P(g4), s)
            _lib.call("dt_mpm_p2g_fused", prm, mats, nm, sim.n, mid, *[P(t) for t in st], P(g4),
                      P(nxt[3]), s)
        vel_u = []
        for k, pad in enumerate(self.pads):
            pos, vel, _ = pads_in[k]
            vu = store["vel_u"][k] if store else self._vu[k]
            vu.copy_(vel)
            _lib.call("dt_fem_update_csr", pad.fem, pad.n_tets, pad.n_verts, pad.f2v, pad.B,
                      pos, vu, pad.vert_start, pad.vert_tets, pad.tet_buf, s)
            vel_u.append(vu)
        if not sim.sparse_grid:
            idx.fill_(-1)
        for k, pad in enumerate(self.pads):
            _lib.call("dt_contact_centroids", pad.n_tris, P(pad.tris, I32), P(pads_in[k][0]),
                      P(pad.cen), s)
            _lib.call("dt_contact_tri_batches", pad.n_tris, P(pad.cen), P(pad.tri_bounds), s)
            # grid mass = 4th float of each packed node
            _lib.call("dt_contact_closest_batched", prm, g4.data_ptr() + 12, 4,
                      None if pad.tri_start is None else P(pad.tri_start, I32), pad.n_tris,
                      P(pad.cen), P(pad.tri_bounds), P(hint, I32), P(idx, I32), self.npad, k, s)
        if not sim.sparse_grid:
            gf.zero_()
        exts = store["ext"] if store else self._ext
        for k, pad in enumerate(self.pads):
            ext = exts[k]
            ext.zero_()
            _lib.call("dt_contact_collision_fx", prm, self.cp, P(g4), 4, g4.data_ptr() + 12, 4,
                      P(idx, I32), self.npad, k, pad.n_tris, P(pad.tris, I32), P(pads_in[k][0]),
                      P(vel_u[k]), pad.od(), P(gf), pad.n_verts, P(ext),
                      pad.fx_acc.data_ptr(), pad.fx_scale.data_ptr(), s)
        npd = len(self.pads)
        if self.rope is not None:
            # cable - pad contact (the cable's state follows the pads' in pads_in)
            rpos, rvel = pads_in[npd]
            nr = self.rope.n_particles
            ridx = store["ridx"] if store else self.ridx
            rimp = store["rimp"] if store else self.rimp
            rimp.zero_()
            for k, pad in enumerate(self.pads):
                # the entries' previous contents seed the search (any start
                # gives the same result)
                _lib.call("dt_rope_closest_batched", nr, P(rpos), pad.n_tris, P(pad.cen),
                          P(pad.tri_bounds), P(ridx, I32), P(ridx, I32), self.npad, k, s)
                _lib.call("dt_rope_collision_fx", self.rope_cp, sim.prm.dt, nr, P(rpos), P(rvel),
                          P(ridx, I32), self.npad, k, pad.n_tris, P(pad.tris, I32),
                          P(pads_in[k][0]), P(vel_u[k]), pad.od(), P(rimp), pad.n_verts,
                          P(exts[k]), None, pad.fx_acc.data_ptr(), pad.fx_scale.data_ptr(), s)
        _lib.call("dt_mpm_grid_op4", prm, P(g4), P(gf), P(gv4), s)
        if compact:
            # keep the active bricks' nodes of the grids the adjoint reads
            self._store_copy(prm, store, 0, (g4, gv4, gf, idx))
        _lib.call("dt_mpm_g2p4", prm, sim.n, P(st[0]), P(gv4), P(nxt[0]), P(nxt[1]), P(nxt[2]), s)
        if self.rope is not None:
            ro = pads_out[npd]
            _lib.call("dt_rope_step", self.rope_prm, nr, self._fixed_ptr(), P(rpos), P(rvel),
                      P(rimp), P(self.rope_work), P(ro[0]), P(ro[1]), s)
        for k, pad in enumerate(self.pads):
            pos, _, vpos = pads_in[k]
            o = pads_out[k]
            _lib.call("dt_fem_update2", pad.fem, pad.n_verts, P(pad.layer, I32), P(pos),
                      P(vel_u[k]), P(vpos), P(exts[k]), P(cv[k]), P(cv1[k]), P(o[0]), P(o[1]),
                      P(o[2]), s)

    def _adj(self, st, pads_in, store, g_next, g_pads_next, g, g_pads, g_cv, g_cv1, mat_id,
             g_ext_out=None):
        """Adjoint of one substep.  g_next/g_pads_next: adjoints of the outputs;
        writes g (object, overwritten), g_pads (overwritten), accumulates g_cv, g_cv1;
        g_ext_out: adjoints of this substep's pad external forces as outputs
        (or None)."""
        args = (st, pads_in, store, g_next, g_pads_next, g, g_pads, g_cv, g_cv1, mat_id,
                g_ext_out)
        if not self.graphs or self._in_segment:
            return self._adj_body(*args)
        key = self._graph_key(2, *args)
        self._graphed(key, lambda: self._adj_body(*args))

    def _adj_body(self, st, pads_in, store, g_next, g_pads_next, g, g_pads, g_cv, g_cv1, mat_id,
                  g_ext_out=None):
        sim, s = self.sim, _lib.stream()
        nn = sim.prm.n_nodes
        prm = sim.c_prm
        compact = store.get("cap") is not None
        if compact:
            # the working grids, filled from the store below
            (g4, gv4), gf, idx = sim._grid_scratch(), self.grid_f, self.idx
        else:
            g4, gv4, gf, idx = store["g4"], store["gv4"], store["gf"], store["idx"]
        gg4 = sim._buf("cgg4", (nn, 4))
        ggv4 = sim._buf("cggv4", (nn, 4))
        ggf = sim._buf("cggf", (nn, 3))
        mats, nm = P(sim.mats), len(sim.materials)
        mid = None if mat_id is None else P(mat_id, I32)
        g_ext = []
        for k, pad in enumerate(self.pads):
            gp, gvu, gvp = g_pads[k]
            ge = self._ge[k]
            # the adjoint kernels below accumulate: one launch clears all four
            torch._foreach_zero_([gp, gvu, gvp, ge])
            gpn, gvn, gvpn = g_pads_next[k]
            _lib.call("dt_fem_update2_adj", pad.fem, pad.n_verts, P(pad.layer, I32), P(gpn),
                      P(gvn), P(gvpn), P(gp), P(gvu), P(gvp), P(ge), P(g_cv[k]), P(g_cv1[k]), s)
            if g_ext_out is not None and g_ext_out[k] is not None:
                ge.add_(g_ext_out[k])
            g_ext.append(ge)
        npd = len(self.pads)
        if self.rope is not None:
            rpos, rvel = pads_in[npd]
            nr = self.rope.n_particles
            grp, grv = g_pads[npd]
            gimp = self.rope_gimp
            torch._foreach_zero_([grp, grv, gimp])
            gpn, gvn = g_pads_next[npd]
            _lib.call("dt_rope_step_adj", self.rope_prm, nr, self._fixed_ptr(), P(rpos), P(rvel),
                      P(store["rimp"]), P(gpn), P(gvn), P(self.rope_work), P(grp), P(grv),
                      P(gimp), s)
            if self.clamp_grad:
                for t in (grp, grv):
                    _lib.call("dt_clamp", t.numel(), P(t), 1000.0, s)
        prm_clamp = self.c_prm_clamp
        if sim.sparse_grid:
            # the forward substep's binning (saved, or made again from its
            # positions): the same active bricks as the stored grids
            ws, nb = sim._workspace()
            if store.get("bin") is not None:
                _lib.call("dt_mpm_bin_copy", prm, sim.n, ws, nb, P(store["bin"], torch.uint8), 0, s)
            else:
                _lib.call("dt_mpm_bin", prm, sim.n, P(st[0]), ws, nb, s)
            prm = self._node_set(prm, ws, nb)
            prm_clamp = self._node_set(prm_clamp, ws, nb)
            if compact:
                # the active bricks of this binning are the ones the forward kept
                self._store_copy(prm, store, 1, (g4, gv4, gf, idx))
            _lib.call("dt_mpm_g2p4_adj_binned", prm, sim.n, P(st[0]), P(gv4), P(g_next[0]),
                      P(g_next[1]), P(g_next[2]), P(g[0]), ws, nb, s)
            # tile sums fused into the grid_op adjoint
            _lib.call("dt_mpm_grid_op4_adj_tiles_f", prm_clamp, sim.n, ws, nb, P(g4), P(gf), None,
                      P(gg4), P(ggf), s)
        else:
            _lib.call("dt_mpm_zero4", nn, P(ggv4), s)
            _lib.call("dt_mpm_g2p4_adj", prm, sim.n, P(st[0]), P(gv4), P(g_next[0]), P(g_next[1]),
                      P(g_next[2]), P(g[0]), P(ggv4), s)
            _lib.call("dt_mpm_grid_op4_adj", prm_clamp, P(g4), P(gf), P(ggv4), P(gg4), P(ggf), s)
        if self.clamp_grad:
            for k in range(len(self.pads)):
                for t in (g_pads[k][0], g_pads[k][1]):
                    _lib.call("dt_clamp", t.numel(), P(t), 1000.0, s)
        for k, pad in enumerate(self.pads):
            _lib.call("dt_contact_collision_adj_fx", prm, self.cp, P(g4), 4, g4.data_ptr() + 12,
                      4, P(idx, I32), self.npad, k, pad.n_tris, P(pad.tris, I32),
                      P(pads_in[k][0]), P(store["vel_u"][k]), pad.od(), P(ggf), P(g_ext[k]),
                      P(gg4), 4, gg4.data_ptr() + 12, 4, pad.n_verts, P(g_pads[k][0]),
                      P(g_pads[k][1]), pad.fx_acc.data_ptr(), pad.fx_scale.data_ptr(), s)
        if self.rope is not None:
            for k, pad in enumerate(self.pads):
                _lib.call("dt_rope_collision_adj_fx", self.rope_cp, sim.prm.dt, nr, P(rpos),
                          P(rvel), P(store["ridx"], I32), self.npad, k, pad.n_tris,
                          P(pad.tris, I32), P(pads_in[k][0]), P(store["vel_u"][k]), pad.od(),
                          P(gimp), P(g_ext[k]), P(grp), P(grv), pad.n_verts, P(g_pads[k][0]),
                          P(g_pads[k][1]), pad.fx_acc.data_ptr(), pad.fx_scale.data_ptr(), s)
        for k, pad in enumerate(self.pads):
            _lib.call("dt_fem_update_adj_csr", pad.fem, pad.n_tets, pad.n_verts, pad.f2v, pad.B,
                      pads_in[k][0], g_pads[k][1], g_pads[k][0], pad.vert_start, pad.vert_tets,
                      pad.tet_buf, s)
        _lib.call("dt_mpm_p2g_fused_adj", prm, mats, nm, sim.n, mid, *[P(t) for t in st], P(gg4),
                  P(g_next[3]), P(g[0]), P(g[1]), P(g[2]), P(g[3]), sim._mat_grad_ptr(), s)

    # -- public --------------------------------------------------------------
    def rollout(self, x, v, C, F, pad_states, control_vel, n_substeps, rope_state=None,
                return_ext=False):
        """x, v, C, F: object state ([n][k] layout); pad_states: list of (pos, vel, vpos)
        per pad; control_vel: list of per-pad (n_substeps+1, n_verts, 3) velocity
        sequences; rope_state: (pos, vel) of the cable (with a cable).  Returns
        (x, v, C, F, [(pos, vel, vpos) per pad]) after n_substeps, then the cable's
        (pos, vel) with a cable, then with return_ext the pads' external forces of
        the last substep (the reference's ``external_force_field[sub_steps - 2]``,
        read by ``compute_contact_force``), all differentiable with respect to
        every tensor input."""
        if return_ext and n_substeps < 1:
            raise ValueError("return_ext needs at least one substep")
        flat = [x, v, C, F]
        for ps in pad_states:
            flat += list(ps)
        flat += list(control_vel)
        if self.rope is not None:
            if rope_state is None:
                raise ValueError("a scene with a cable needs rope_state")
            flat += list(rope_state)
        out = _CoupledRollout.apply(self, int(n_substeps), bool(return_ext), *flat)
        npd = len(self.pads)
        obj = out[:4]
        pads = [tuple(out[4 + 3 * k: 7 + 3 * k]) for k in range(npd)]
        res = [*obj, pads]
        o = 4 + 3 * npd
        if self.rope is not None:
            res.append(tuple(out[o:o + 2]))
            o += 2
        if return_ext:
            res.append(list(out[o:o + npd]))
        return tuple(res)


class _CoupledRollout(torch.autograd.Function):
    # Buffers come from the simulation's pool and return to it after the
    # backward pass (or at the end of a forward pass without one), so the
    # next rollout runs on the same addresses: graphed substeps replay.
    @staticmethod
    def forward(ctx, cs, n_substeps, return_ext, *flat):
        sim = cs.sim
        npd = len(cs.pads)
        taken = []
        P = lambda key, t: cs._copy_in(taken, key, t)
        x, v, C, F = (t.contiguous() for t in flat[:4])
        pads0 = [[P(("p0", k, c), flat[4 + 3 * k + c].contiguous()) for c in range(3)]
                 for k in range(npd)]
        cvs = [P(("cv", k), t.contiguous()) for k, t in enumerate(flat[4 + 3 * npd: 4 + 4 * npd])]
        if cs.rope is not None:
            # the cable's state rides along as one more body after the pads
            pads0.append([P(("r0", c), t.contiguous().float())
                          for c, t in enumerate(flat[4 + 4 * npd:])])
        shapes = [t.shape for t in (x, v, C, F)]
        W = (3, 3, 9, 9)
        perm = sim.sort_order(x)[1] if (sim.sort and sim.n > 0) else None
        st = [sim._gather(perm, t, w, out=cs._pool(taken, ("st", a), (w, sim.n)))
              for a, (t, w) in enumerate(zip((x, v, C, F), W))]
        if sim.mat_id is not None and perm is not None:
            mat_id = sim._gather_i(perm, sim.mat_id, out=cs._pool(taken, ("mid",), (sim.n,), I32))
        else:
            mat_id = sim.mat_id
        env = sim.env_id
        if env is not None and perm is not None:
            env = sim._gather_i(perm, env, out=cs._pool(taken, ("env",), (sim.n,), I32))
        sim._set_env(env)
        nxt = [cs._pool(taken, ("nx", a), t.shape) for a, t in enumerate(st)]
        pst = pads0
        pnx = [[cs._pool(taken, ("pn", k, c), t.shape, t.dtype) for c, t in enumerate(p)]
               for k, p in enumerate(pst)]
        ckpts, kept = [], []
        keep_ok = any(ctx.needs_input_grad)
        used = 0
        segs = _segments(n_substeps, sim.checkpoint_every)
        # compact stores only when whole-grid stores of every segment do not
        # fit the budget (packing them costs a launch per substep each way)
        compact = keep_ok and sim.sparse_grid and (
            cs.compact_stores is True or (cs.compact_stores is None and sum(
                (k + 1) * cs.substep_bytes() for _, k in segs) > sim.resident_bytes))
        ctx.compact = compact
        for i, (s0, k) in enumerate(segs):
            cap = cs.store_capacity(st[0]) if compact else None
            nbytes = (k + 1) * cs.substep_bytes(cap)
            if keep_ok and used + nbytes <= sim.resident_bytes:
                # resident segment: frames and stores kept for the backward pass
                used += nbytes
                seg = cs._run_keep(st, pst, cvs, s0, k, mat_id, taken, ("seg", i), cap)
                if cap is not None and cs._overflowed():
                    # more active bricks than the stores hold: the backward
                    # pass recomputes the segment from its first frame
                    ckpts.append((seg[0][0], seg[1][0]))
                    kept.append(None)
                else:
                    ckpts.append(None)
                    kept.append(seg)
                st = [P(("sc", a), t) for a, t in enumerate(seg[0][-1])]
                pst = [[P(("psc", b_, c), t) for c, t in enumerate(p)]
                       for b_, p in enumerate(seg[1][-1])]
                if cap is None:
                    cs.idx.copy_(seg[2][-1]["idx"])
                continue
            if keep_ok:
                ckpts.append(([P(("ck", i, a), t) for a, t in enumerate(st)],
                              [[P(("ckp", i, b_, c), t) for c, t in enumerate(p)]
                               for b_, p in enumerate(pst)]))
            else:
                ckpts.append(None)
            kept.append(None)
            calls = []
            for f in range(s0, s0 + k):
                calls.append((st, nxt, pst, pnx, [c[f] for c in cvs], [c[f + 1] for c in cvs],
                              mat_id))
                st, nxt = nxt, st
                pst, pnx = pnx, pst
            cs._fwd_segment(calls)
        # (segments kept for the backward pass, segments, bytes they hold)
        cs.last_resident = (sum(kp is not None for kp in kept), len(kept), used)
        out = [sim._scatter(perm, t, w, sh) for t, w, sh in zip(st, W, shapes)]
        for p in pst:
            out += [t.clone() for t in p]
        if return_ext:
            out += [e.clone() for e in cs._last_ext]
        ctx.return_ext = return_ext
        ctx.cs, ctx.n_substeps, ctx.ckpts, ctx.shapes = cs, n_substeps, ckpts, shapes
        ctx.kept = kept
        ctx.perm, ctx.mat_id, ctx.env = perm, mat_id, env
        ctx.cv_shapes = [c.shape for c in cvs]
        ctx.cvs = cvs
        ctx.taken = taken
        if not keep_ok:
            sim._give_back(taken)
        return tuple(out)

    @staticmethod
    def backward(ctx, *gout):
        cs = ctx.cs
        sim = cs.sim
        npd = len(cs.pads)
        cvs = ctx.cvs
        taken = ctx.taken
        W = (3, 3, 9, 9)
        dev = sim.device
        n = sim.n

        def pooled(key, t, shape, dtype=torch.float32):
            b = cs._pool(taken, key, shape, dtype)
            return b.zero_() if t is None else b.copy_(t)

        g = [sim._gather(ctx.perm, (t if t is not None else torch.zeros(sh, device=dev)).float(), w,
                         out=cs._pool(taken, ("g", a), (w, n)))
             for a, (t, w, sh) in enumerate(zip(gout[:4], W, ctx.shapes))]
        gp = []
        for k in range(npd):
            trio = gout[4 + 3 * k: 7 + 3 * k]
            gp.append([pooled(("gp", k, c), t, (cs.pads[k].n_verts, 3))
                       for c, t in enumerate(trio)])
        o = 4 + 3 * npd
        if cs.rope is not None:
            nr = cs.rope.n_particles
            gp.append([pooled(("gr", c), None if t is None else t.float(), (nr, 3))
                       for c, t in enumerate(gout[o:o + 2])])
            o += 2
        g_ext_out = None
        if ctx.return_ext:
            g_ext_out = [None if t is None else pooled(("gext", k), t.float(), t.shape)
                         for k, t in enumerate(gout[o:o + npd])]
        g_cv = [pooled(("gcv", k), None, sh) for k, sh in enumerate(ctx.cv_shapes)]
        gbuf = [cs._pool(taken, ("gb", a), t.shape) for a, t in enumerate(g)]
        gpbuf = [[cs._pool(taken, ("gpb", k, c), t.shape) for c, t in enumerate(p)]
                 for k, p in enumerate(gp)]
        segs = _segments(ctx.n_substeps, sim.checkpoint_every)
        sim._set_env(ctx.env)
        for i in reversed(range(len(segs))):
            s0, k = segs[i]
            seg_taken = None
            if ctx.kept[i] is not None:
                frames, pframes, stores = ctx.kept[i]
                ctx.kept[i] = None
            else:
                if ctx.ckpts[i] is None:
                    raise RuntimeError("a coupled rollout's backward runs once (it releases the "
                                       "segments it kept); run the rollout again for another "
                                       "gradient")
                # recompute the segment, keeping what the adjoint needs (one
                # set of recompute buffers, back to the pool after each segment)
                ck, ckp = ctx.ckpts[i]
                seg_taken = []
                cap = cs.store_capacity(ck[0]) if ctx.compact else None
                frames, pframes, stores = cs._run_keep(ck, ckp, cvs, s0, k, ctx.mat_id,
                                                       seg_taken, ("recompute",), cap)
                if cap is not None and cs._overflowed():
                    # the recomputed segment outgrew its compact stores: again
                    # with dense ones
                    del frames, pframes, stores
                    sim._give_back(seg_taken)
                    seg_taken = []
                    frames, pframes, stores = cs._run_keep(ck, ckp, cvs, s0, k, ctx.mat_id,
                                                           seg_taken, ("recompute",))
            calls = []
            for j in reversed(range(k)):
                f = s0 + j
                calls.append((frames[j], pframes[j], stores[j], g, gp, gbuf, gpbuf,
                              [gc[f] for gc in g_cv], [gc[f + 1] for gc in g_cv], ctx.mat_id,
                              g_ext_out if f == ctx.n_substeps - 1 else None))
                g, gbuf = gbuf, g
                gp, gpbuf = gpbuf, gp
            cs._adj_segment(calls)
            if seg_taken is not None:
                del frames, pframes, stores
                sim._give_back(seg_taken)
        res = [sim._scatter(ctx.perm, t, w, sh) for t, w, sh in zip(g, W, ctx.shapes)]
        for p in gp[:npd]:
            res += [t.clone() for t in p]
        res += [t.clone() for t in g_cv]
        for p in gp[npd:]:
            res += [t.clone() for t in p]
        sim._give_back(taken)
        ctx.taken = []
        return (None, None, None, *res)
