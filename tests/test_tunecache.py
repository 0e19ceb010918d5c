"""The persisted autotuner cache: keys follow everything that determines a
launch shape, and entries survive a save/load round trip (CPU only)."""

import numpy as np

from echidna_amd import workloads as wl
from echidna_amd.tape import GpuTapeData
from echidna_amd.tunecache import TuneCache


def test_keys_follow_the_tape_and_the_evaluation(tmp_path):
    a = GpuTapeData.from_tape(wl.tanh_mlp(8, 6, 1))
    b = GpuTapeData.from_tape(wl.tanh_mlp(8, 6, 2))  # other weights, same structure
    k = TuneCache.key(a, False, 1, 0, 1024)
    assert k == TuneCache.key(GpuTapeData.from_tape(wl.tanh_mlp(8, 6, 1)), False, 1, 0, 1024)
    others = {TuneCache.key(b, False, 1, 0, 1024), TuneCache.key(a, True, 1, 0, 1024),
              TuneCache.key(a, False, 2, 0, 1024), TuneCache.key(a, False, 3, 3, 1024),
              TuneCache.key(a, False, 1, 0, 2048), TuneCache.key(a, False, 1, 0, 1024, "gfx942")}
    assert k not in others and len(others) == 6


def test_save_and_load(tmp_path):
    p = str(tmp_path / "cache.json")
    c = TuneCache(p)
    c.entries["k"] = {"wg": 128, "lds_budget": 0, "fetch": 2, "points_per_lane": 4, "us": 210}
    c.save()
    d = TuneCache(p)
    assert d.entries == c.entries
    assert TuneCache(str(tmp_path / "missing.json")).entries == {}
