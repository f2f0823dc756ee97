"""Host logic of the launch-shape autotuner (no GPU): cache keys, the cache
files and the coordinate descent."""
import json

from difftactile_amd import autotune
from difftactile_amd.params import MPMParams


def test_key_buckets_particle_counts_by_power_of_two():
    prm = MPMParams(n_grid=128)
    assert autotune.key(prm, 1_000_000) == "g128_e1_n2^20"
    assert autotune.key(prm, 800_000) == autotune.key(prm, 1_300_000) == "g128_e1_n2^20"
    assert autotune.key(MPMParams(n_grid=64), 5000) == "g64_e1_n2^12"


def test_store_then_lookup_prefers_the_user_cache(tmp_path, monkeypatch):
    path = tmp_path / "tuned.json"
    monkeypatch.setenv("DT_TUNE_CACHE", str(path))
    prm = MPMParams(n_grid=64)
    assert autotune.lookup(prm, 4096) is None
    cfg = {"brick_blocks": 768, "node_blocks": 512, "sort_blocks": 1024}
    autotune.store(prm, 4096, cfg, timing=1.25)
    assert autotune.lookup(prm, 4096) == cfg
    d = json.loads(path.read_text())
    assert d[autotune.key(prm, 4096)]["ms"] == 1.25
    # a corrupt file reads as empty
    path.write_text("{not json")
    assert autotune.lookup(prm, 4096) is None


def test_coordinate_descent_finds_the_separable_minimum():
    cands = {"brick_blocks": [512, 1024, 2048, 4096], "node_blocks": [256, 512, 1024],
             "sort_blocks": [1024, 2048]}
    target = {"brick_blocks": 1024, "node_blocks": 256, "sort_blocks": 1024}
    calls = []

    def measure(cfg):
        calls.append(dict(cfg))
        return sum(abs(cfg[k] - target[k]) / target[k] for k in target) + 1.0

    best, t = autotune.coordinate_descent(measure, cands)
    assert best == target and abs(t - 1.0) < 1e-12
    assert len(calls) < 4 * 3 * 2 * 2 + 1
