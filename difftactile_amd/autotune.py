"""Launch-shape autotuner for the MPM passes.

The brick-owned scatter / gather passes, the active-brick node passes and
the per-brick segment sort run as persistent grids whose block counts are
process-wide settings of the HIP library (``dt_set_launch_config``).  The
best counts depend on the grid size, the number of environments and the
particle count, so this module times a forward + adjoint trajectory of the
caller's own simulation under candidate shapes (coordinate descent over the
three knobs, graphs off while timing), keeps the fastest and caches it in a
JSON file keyed by (n_grid, n_env, particle-count bucket).

Cache files: the package's ``tuned_launch.json`` (read-only defaults that
ship with it) and a user cache (``$DT_TUNE_CACHE``, else
``~/.cache/difftactile_amd/tuned_launch.json``) that ``tune`` writes; user
entries win.  Results do not depend on the shapes (every brick and node is
processed by one lane group whatever the grid size), so tuning changes
timing only.
"""
import json
import math
import os
import time

import torch

from . import _lib

CANDIDATES = {
    "brick_blocks": [512, 768, 1024, 1536, 2048, 3072, 4096],
    "node_blocks": [256, 512, 1024, 2048, 4096],
    "sort_blocks": [128, 512, 1024, 2048, 4096],
}
DEFAULT = {"brick_blocks": 2048, "node_blocks": 1024, "sort_blocks": 2048}
PACKAGE_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_launch.json")


def user_cache_path():
    return os.environ.get("DT_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "difftactile_amd", "tuned_launch.json")


def key(prm, n_particles):
    """Cache key: grid size, environments, particle count to the nearest power of 2."""
    b = 0 if n_particles <= 0 else int(round(math.log2(n_particles)))
    return f"g{prm.n_grid}_e{max(1, prm.n_env)}_n2^{b}"


def _read(path):
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def lookup(prm, n_particles):
    """The cached shapes of this workload (user cache first), or None."""
    k = key(prm, n_particles)
    for path in (user_cache_path(), PACKAGE_CACHE):
        e = _read(path).get(k)
        if e and all(isinstance(e.get(f), int) for f in DEFAULT):
            return {f: e[f] for f in DEFAULT}
    return None


def store(prm, n_particles, cfg, timing=None, path=None):
    path = path or user_cache_path()
    d = _read(path)
    e = dict(cfg)
    if timing is not None:
        e["ms"] = timing
    d[key(prm, n_particles)] = e
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def set_config(cfg):
    c = _lib.DtLaunchConfig()
    c.brick_blocks, c.node_blocks, c.sort_blocks = (int(cfg.get(f, 0)) for f in
                                                    ("brick_blocks", "node_blocks", "sort_blocks"))
    _lib.call("dt_set_launch_config", _lib.ctypes.byref(c))


def get_config():
    c = _lib.DtLaunchConfig()
    _lib.call("dt_get_launch_config", _lib.ctypes.byref(c))
    return {"brick_blocks": c.brick_blocks, "node_blocks": c.node_blocks,
            "sort_blocks": c.sort_blocks}


def apply_cached(prm, n_particles):
    """Set the cached shapes of this workload if there are any; returns them."""
    cfg = lookup(prm, n_particles)
    if cfg is not None:
        set_config(cfg)
    return cfg


def coordinate_descent(measure, candidates=None, start=None, rounds=2):
    """Minimise measure(cfg) -> seconds one knob at a time; returns (cfg, seconds)."""
    candidates = candidates or CANDIDATES
    best = dict(start or DEFAULT)
    best_t = measure(best)
    for _ in range(rounds):
        improved = False
        for knob, values in candidates.items():
            for v in values:
                if v == best[knob]:
                    continue
                trial = dict(best)
                trial[knob] = v
                t = measure(trial)
                if t < best_t:
                    best, best_t, improved = trial, t, True
        if not improved:
            break
    return best, best_t


def tune(sim, state, n_substeps=4, reps=3, save=True, candidates=None):
    """Tune the shapes on sim's own workload: each candidate is timed as the
    best of reps forward + adjoint rollouts of n_substeps from state.  Sets
    and (with save) caches the winner; returns (cfg, ms per rollout)."""
    x, v, C, F = state

    def run():
        leaves = [t.detach().requires_grad_(True) for t in (x, v, C, F)]
        out = sim.rollout(*leaves, n_substeps)
        g = torch.autograd.grad(out[0].sum() + out[1].sum(), leaves[1])[0]
        return g

    def measure(cfg):
        set_config(cfg)
        run()
        torch.cuda.synchronize()
        best = math.inf
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    _lib.call("dt_graphs_enable", 0)
    try:
        cfg, t = coordinate_descent(measure, candidates)
    finally:
        _lib.call("dt_graphs_enable", 1)
    set_config(cfg)
    if save:
        store(sim.prm, sim.n, cfg, timing=round(t * 1e3, 4))
    return cfg, t * 1e3
