"""Persisted autotuner results.

`ehip_autotune_dev` times every launch shape of one evaluation of one tape
and keeps the fastest on the tape handle; `ehip_set_tuned` restores a shape.
This cache stores those choices in a JSON file keyed by what determines
them -- the tape's arrays (a SHA-256 of opcodes, arguments, constants and
outputs), the scalar type, the evaluation (TUNE_*), the Taylor order, the
batch size and the device architecture -- so a later process restores the
shape instead of timing every candidate again.  A shape never changes a
result (every launch shape computes the same bits), so a stale or foreign
entry can cost speed, never correctness.
"""

from __future__ import annotations

import hashlib
import json
import os
from typing import Optional

import numpy as np

from .backend import HipContext, HipTapeBuffers
from .tape import GpuTapeData

VERSION = 1


def tape_digest(data: GpuTapeData) -> str:
    h = hashlib.sha256()
    for a in (data.opcodes, data.arg0, data.arg1, data.output_indices):
        h.update(np.ascontiguousarray(a, dtype=np.uint32).tobytes())
    h.update(np.ascontiguousarray(data.constants, dtype=np.float64).tobytes())
    h.update(f"{data.num_inputs}/{data.output_index}".encode())
    return h.hexdigest()[:32]


class TuneCache:
    def __init__(self, path: str):
        self.path = path
        self.entries = {}
        if os.path.exists(path):
            with open(path) as f:
                doc = json.load(f)
            if doc.get("version") == VERSION:
                self.entries = doc.get("entries", {})

    @staticmethod
    def key(data: GpuTapeData, f64: bool, what: int, order: int, batch: int, arch: str = "gfx950") -> str:
        return f"{tape_digest(data)}:{'f64' if f64 else 'f32'}:{what}:{order}:{batch}:{arch}"

    def restore(self, tape: HipTapeBuffers, key: str, what: int, order: int = 0) -> Optional[dict]:
        """Pin the cached shape on the tape; None when there is none (or the
        specialised kernel it names is not built)."""
        e = self.entries.get(key)
        if not e:
            return None
        try:
            HipContext.set_tuned(tape, what, e["wg"], e["lds_budget"], order=order, fetch=e["fetch"],
                                 points_per_lane=e.get("points_per_lane", 0))
        except Exception:
            return None
        return dict(e, cached=True)

    def autotune(self, ctx: HipContext, tape: HipTapeBuffers, key: str, what: int, batch: int, d_x, order: int = 0,
                 **bufs) -> dict:
        """The cached shape if there is one, else the autotuner's choice
        (stored for the next process)."""
        got = self.restore(tape, key, what, order)
        if got is not None:
            return got
        res = ctx.autotune_dev(tape, what, batch, d_x, order=order, **bufs)
        self.entries[key] = {k: res[k] for k in ("wg", "lds_budget", "fetch", "points_per_lane")}
        self.entries[key]["us"] = res.get("us")
        return dict(res, cached=False)

    def save(self) -> None:
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump({"version": VERSION, "entries": self.entries}, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)
