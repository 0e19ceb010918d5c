"""Serde wire formats (Python mirror).

``tape_to_json`` / ``tape_from_json`` (and the CBOR pair ``tape_to_cbor`` /
``tape_from_cbor``, the binary form the reference round-trips through
ciborium) call the library's tape readers and writers, which follow
``impl Serialize / Deserialize for BytecodeTape``
(``src/bytecode_tape/serde_support.rs:11-100``): the same field names and
order, opcodes as variant names, and the same rejections with the same
messages (Custom opcodes, a non-empty ``custom_second_args``, a
``num_variables`` stamp that disagrees with ``opcodes.len()``, then
``BytecodeTape::validate``, ``src/bytecode_tape/mod.rs:575-703``).

The derived serde forms of the sparsity patterns (``src/sparse.rs:13-20,
245-256``) and of ``NonsmoothInfo`` / ``KinkEntry``
(``src/nonsmooth.rs:45-74``) are plain field-name objects and live here.
"""

from __future__ import annotations

import ctypes
import json
import math
from dataclasses import dataclass
from typing import List

import numpy as np

from .backend import EHIP_F32, EHIP_F64, _check, _vp, load_library
from .sparse import JacobianSparsityPattern, NonsmoothBatch, SparsityPattern
from .tape import BytecodeTape, GpuError, OpCode

_c_u64 = ctypes.c_uint64
_c_u32 = ctypes.c_uint32


def _p(a: np.ndarray):
    return _vp(a.ctypes.data) if a.size else _vp(None)


def _encode(fn: str, tape: BytecodeTape, f32: bool) -> bytes:
    n = len(tape.opcodes)
    op = np.asarray(tape.opcodes, np.uint32)
    a0 = np.asarray(tape.arg0, np.uint32)
    a1 = np.asarray(tape.arg1, np.uint32)
    with np.errstate(over="ignore"):
        vals = np.asarray(tape.values, np.float64).astype(np.float32 if f32 else np.float64)
    outs = np.asarray(tape.output_indices, np.uint32)
    fn = getattr(load_library(), fn)
    args = (EHIP_F32 if f32 else EHIP_F64, _p(op), _p(a0), _p(a1), _p(vals), n, tape.num_inputs, tape.output_index,
            _p(outs), len(outs))
    ln = _c_u64(0)
    _check(fn(*args, None, 0, ctypes.byref(ln)))
    buf = ctypes.create_string_buffer(int(ln.value) + 1)
    _check(fn(*args, buf, int(ln.value) + 1, ctypes.byref(ln)))
    return buf.raw[:int(ln.value)]


def tape_to_json(tape: BytecodeTape, f32: bool = False) -> str:
    """``serde_json::to_string(&tape)`` for a ``BytecodeTape<f64>`` (``f32``:
    the values rounded to f32, as a ``BytecodeTape<f32>`` would carry them)."""
    return _encode("ehip_tape_to_json", tape, f32).decode()


def tape_to_cbor(tape: BytecodeTape, f32: bool = False) -> bytes:
    """``ciborium::into_writer(&tape, ..)``: the same model as CBOR (RFC 8949)."""
    return _encode("ehip_tape_to_cbor", tape, f32)


def tape_from_json(text, f32: bool = False) -> BytecodeTape:
    """``serde_json::from_str::<BytecodeTape<F>>``; a rejected payload raises
    ``GpuError(OTHER)`` carrying the reference's message."""
    return _decode("ehip_tape_from_json", text.encode() if isinstance(text, str) else bytes(text), f32)


def tape_from_cbor(data: bytes, f32: bool = False) -> BytecodeTape:
    """``ciborium::from_reader::<BytecodeTape<F>>``, with the same checks."""
    return _decode("ehip_tape_from_cbor", bytes(data), f32)


def _decode(fn: str, raw: bytes, f32: bool) -> BytecodeTape:
    fn = getattr(load_library(), fn)
    dt = EHIP_F32 if f32 else EHIP_F64
    n, ni, oi, m = _c_u64(0), _c_u32(0), _c_u32(0), _c_u64(0)
    counts = (ctypes.byref(n), ctypes.byref(ni), ctypes.byref(oi), ctypes.byref(m))
    _check(fn(raw, len(raw), dt, None, None, None, None, 0, None, 0, *counts))
    N, M = int(n.value), int(m.value)
    op = np.zeros(max(N, 1), np.uint32)
    a0 = np.zeros(max(N, 1), np.uint32)
    a1 = np.zeros(max(N, 1), np.uint32)
    vals = np.zeros(max(N, 1), np.float32 if f32 else np.float64)
    outs = np.zeros(max(M, 1), np.uint32)
    _check(fn(raw, len(raw), dt, _vp(op.ctypes.data), _vp(a0.ctypes.data), _vp(a1.ctypes.data),
              _vp(vals.ctypes.data), max(N, 1), _vp(outs.ctypes.data), max(M, 1), *counts))
    t = BytecodeTape()
    t.opcodes = op[:N].tolist()
    t.arg0 = a0[:N].tolist()
    t.arg1 = a1[:N].tolist()
    t.values = [float(v) for v in vals[:N]]
    t.num_inputs = int(ni.value)
    t.output_index = int(oi.value)
    t.output_indices = outs[:M].tolist()
    return t


# ---- derived serde forms ------------------------------------------------------

def _f(v: float):
    """serde_json writes non-finite floats as null."""
    v = float(v)
    return v if math.isfinite(v) else None


def _need(d: dict, key: str):
    if key not in d:
        raise GpuError(GpuError.OTHER, f"missing field `{key}`")
    return d[key]


def _u32_list(v, field: str) -> np.ndarray:
    if not isinstance(v, list) or not all(isinstance(x, int) and not isinstance(x, bool) and 0 <= x <= 0xFFFFFFFF
                                          for x in v):
        raise GpuError(GpuError.OTHER, f"invalid type for `{field}`: expected a sequence of u32")
    return np.asarray(v, np.uint32)


def _usize(v, field: str) -> int:
    if not isinstance(v, int) or isinstance(v, bool) or v < 0:
        raise GpuError(GpuError.OTHER, f"invalid type for `{field}`: expected usize")
    return v


def pattern_to_json(p) -> str:
    """Derived Serialize of ``SparsityPattern`` / ``JacobianSparsityPattern``."""
    if isinstance(p, SparsityPattern):
        d = {"dim": p.dim, "rows": [int(x) for x in p.rows], "cols": [int(x) for x in p.cols]}
    else:
        d = {"num_outputs": p.num_outputs, "num_inputs": p.num_inputs, "rows": [int(x) for x in p.rows],
             "cols": [int(x) for x in p.cols]}
    return json.dumps(d, separators=(",", ":"))


def pattern_from_json(text: str):
    d = json.loads(text)
    if not isinstance(d, dict):
        raise GpuError(GpuError.OTHER, "invalid type: expected a struct")
    rows, cols = _u32_list(_need(d, "rows"), "rows"), _u32_list(_need(d, "cols"), "cols")
    if "dim" in d:
        return SparsityPattern(_usize(d["dim"], "dim"), rows, cols)
    return JacobianSparsityPattern(_usize(_need(d, "num_outputs"), "num_outputs"),
                                   _usize(_need(d, "num_inputs"), "num_inputs"), rows, cols)


@dataclass
class KinkEntry:
    """``KinkEntry<F>`` (src/nonsmooth.rs:45-66)."""

    tape_index: int
    opcode: int
    switching_value: float
    branch: int


@dataclass
class NonsmoothInfo:
    """``NonsmoothInfo<F>`` of one point (src/nonsmooth.rs:69-74)."""

    kinks: List[KinkEntry]

    def signature(self) -> List[int]:
        """Branch signature, one entry per kink (``NonsmoothInfo::signature``)."""
        return [k.branch for k in self.kinks]

    def to_json(self) -> str:
        return json.dumps({"kinks": [{"tape_index": k.tape_index, "opcode": OpCode(k.opcode).name,
                                      "switching_value": _f(k.switching_value), "branch": k.branch}
                                     for k in self.kinks]}, separators=(",", ":"))

    @staticmethod
    def from_json(text: str) -> "NonsmoothInfo":
        d = json.loads(text)
        kinks = []
        for e in _need(d, "kinks"):
            name = _need(e, "opcode")
            if name not in OpCode.__members__:
                raise GpuError(GpuError.OTHER, f"unknown variant `{name}`")
            sv = _need(e, "switching_value")
            if sv is None or isinstance(sv, bool) or not isinstance(sv, (int, float)):
                raise GpuError(GpuError.OTHER, "invalid type for `switching_value`: expected f64")
            br = _need(e, "branch")
            if not isinstance(br, int) or isinstance(br, bool) or not -128 <= br <= 127:
                raise GpuError(GpuError.OTHER, "invalid value for `branch`: expected i8")
            kinks.append(KinkEntry(_usize(_need(e, "tape_index"), "tape_index"), int(OpCode[name]), float(sv), br))
        return NonsmoothInfo(kinks)


def point_info(batch: NonsmoothBatch, b: int) -> NonsmoothInfo:
    """The ``NonsmoothInfo`` of point ``b`` of a GPU kink batch."""
    return NonsmoothInfo([KinkEntry(int(batch.tape_index[k]), int(batch.opcode[k]), float(batch.switching[b, k]),
                                    int(batch.branch[b, k])) for k in range(len(batch.tape_index))])
