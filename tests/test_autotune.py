"""The autotuner's host logic with a stand-in context (no GPU): the sweep
keeps the fastest candidate, skips candidates the library refuses, restores
the built-in heuristic afterwards; shape buckets and the JSON cache."""
import pytest

from pycpet_amd import autotune as AT
from pycpet_amd._native import CpetError


class _Ctx:
    """Records tuning calls; phase_times reports a preset time per shape."""

    def __init__(self, times):
        self.times = times
        self.cur = (0, 0)
        self.calls = []
        self.profiling = False

    def set_profiling(self, on):
        self.profiling = on

    def set_tuning(self, kind, a, b):
        self.calls.append((kind, a, b))
        self.cur = (a, b)

    def phase_times(self, n):
        return [self.times[self.cur]] * n


def test_sweep_picks_fastest_and_restores_default():
    times = {(1, 0): 5.0, (4, 256): 3.0, (16, 0): 4.0, (8, 512): 9.0}
    ctx = _Ctx(times)
    best, table = AT.sweep(ctx, 3, list(times), lambda: None, phase=0)
    assert best == (4, 256)
    assert table == {"1,0": 5.0, "4,256": 3.0, "16,0": 4.0, "8,512": 9.0}
    assert ctx.calls[-1] == (3, 0, 0) and ctx.profiling


def test_sweep_skips_refused_candidates():
    ctx = _Ctx({(1, 0): 2.0, (16, 0): 4.0})

    def run():
        if ctx.cur == (1, 0):
            raise CpetError(-1, "shape refused")

    best, table = AT.sweep(ctx, 2, [(1, 0), (16, 0)], run, phase=1)
    assert best == (16, 0) and list(table) == ["16,0"]


def test_buckets_and_cache_roundtrip(tmp_path):
    assert AT.key(3, 5000, 100_000) == "3:8192:131072"
    assert AT.key(0, 1, 0) == "0:1:1"
    assert AT.key(2, 4096, 4096) == "2:4096:4096"
    path = tmp_path / "tune.json"
    assert AT.load_cache(path) == {}
    AT.save_cache({"3:16:1024": [1, 0]}, path)
    assert AT.load_cache(path) == {"3:16:1024": [1, 0]}
    path.write_text("not json")
    assert AT.load_cache(path) == {}


def test_topo_candidates_include_one_wave(monkeypatch):
    """Both streamline tuners try one-wave workgroups (bitwise the 16-wave
    results) besides the wave/workgroup grid."""
    seen = {}

    def fake_sweep(ctx, kind, cands, run, phase, repeats=2):
        seen[kind] = list(cands)
        return cands[0], {}

    monkeypatch.setattr(AT, "sweep", fake_sweep)
    ctx = _Ctx({})
    import numpy as np

    AT.tune_topo(ctx, np.zeros((10, 3), np.float32), np.ones(10, np.int32), 0.01,
                 np.ones(3, np.float32), np.zeros((4, 3), np.float32), np.zeros(4, np.float32))
    AT.tune_batch_topo(ctx, [{"q": np.zeros(4), "starts": np.zeros((10, 3))}])
    for kind in (2, 3):
        assert seen[kind][0] == (1, 0)
        assert (16, 0) in seen[kind] and len(seen[kind]) == 13
    assert ctx.calls[-1] == (3, 1, 0)  # the chosen shape is applied
