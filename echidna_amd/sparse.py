"""Sparse derivative structure and nonsmooth kink records (Python mirror).

Pattern types and colorings follow the reference's ``src/sparse.rs`` and
``src/bytecode_tape/sparse.rs``; the GPU sparse Jacobian / Hessian follow
``GpuBackend::sparse_jacobian`` / ``sparse_hessian``
(``src/gpu/cuda_backend.rs:306-441``); kink records follow
``BytecodeTape::forward_nonsmooth`` (``src/bytecode_tape/forward.rs:87-187``).
Everything runs in ``libechidna_hip.so``: structure on the host, sweeps on the
GPU.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from .backend import EHIP_F64, HipContext, HipTapeBuffers, _check, _u32, _vp, load_library
from .tape import GpuError, GpuTapeData

UNSET = 0xFFFFFFFF


def _p(a: np.ndarray):
    return _vp(a.ctypes.data) if a.size else _vp(None)


@dataclass
class SparsityPattern:
    """Lower-triangle COO Hessian pattern (``SparsityPattern``, src/sparse.rs:9-46)."""

    dim: int
    rows: np.ndarray
    cols: np.ndarray

    def nnz(self) -> int:
        return len(self.rows)

    def is_empty(self) -> bool:
        return len(self.rows) == 0

    def contains(self, i: int, j: int) -> bool:
        r, c = (i, j) if i >= j else (j, i)
        return bool(np.any((self.rows == r) & (self.cols == c)))


@dataclass
class JacobianSparsityPattern:
    """(output, input) COO pattern (``JacobianSparsityPattern``, src/sparse.rs:233-266)."""

    num_outputs: int
    num_inputs: int
    rows: np.ndarray
    cols: np.ndarray

    def nnz(self) -> int:
        return len(self.rows)

    def contains(self, o: int, i: int) -> bool:
        return bool(np.any((self.rows == o) & (self.cols == i)))


def _two_call(fn, *args):
    nnz = ctypes.c_uint64(0)
    fn(*args, None, None, 0, ctypes.byref(nnz))
    n = int(nnz.value)
    rows = np.zeros(max(n, 1), np.uint32)
    cols = np.zeros(max(n, 1), np.uint32)
    _check(fn(*args, _vp(rows.ctypes.data), _vp(cols.ctypes.data), n, ctypes.byref(nnz)))
    return rows[:n].copy(), cols[:n].copy()


def _tape_arrays(data: GpuTapeData):
    op, a0, a1 = _u32(data.opcodes), _u32(data.arg0), _u32(data.arg1)
    return op, a0, a1


def detect_sparsity(data: GpuTapeData) -> SparsityPattern:
    """``BytecodeTape::detect_sparsity`` on the flattened tape."""
    op, a0, a1 = _tape_arrays(data)
    lib = load_library()
    nnz = ctypes.c_uint64(0)
    args = (_p(op), _p(a0), _p(a1), data.num_ops, data.num_inputs)
    rc = lib.ehip_hessian_sparsity(*args, None, None, 0, ctypes.byref(nnz))
    if rc != 0 and nnz.value == 0:
        _check(rc)
    rows, cols = _two_call(lib.ehip_hessian_sparsity, *args)
    return SparsityPattern(data.num_inputs, rows, cols)


def detect_jacobian_sparsity(data: GpuTapeData) -> JacobianSparsityPattern:
    """``BytecodeTape::detect_jacobian_sparsity`` on the flattened tape."""
    op, a0, a1 = _tape_arrays(data)
    outs = _u32(data.output_indices)
    lib = load_library()
    args = (_p(op), _p(a0), _p(a1), data.num_ops, data.num_inputs, _p(outs), len(outs))
    nnz = ctypes.c_uint64(0)
    rc = lib.ehip_jacobian_sparsity(*args, None, None, 0, ctypes.byref(nnz))
    if rc != 0 and nnz.value == 0:
        _check(rc)
    rows, cols = _two_call(lib.ehip_jacobian_sparsity, *args)
    return JacobianSparsityPattern(len(outs), data.num_inputs, rows, cols)


def greedy_coloring(pattern: SparsityPattern) -> Tuple[np.ndarray, int]:
    rows, cols = _u32(pattern.rows), _u32(pattern.cols)
    colors = np.zeros(max(pattern.dim, 1), np.uint32)
    nc = ctypes.c_uint32(0)
    _check(load_library().ehip_greedy_coloring(pattern.dim, _p(rows), _p(cols), len(rows), _p(colors),
                                                ctypes.byref(nc)))
    return colors[:pattern.dim].copy(), int(nc.value)


def column_coloring(pattern: JacobianSparsityPattern) -> Tuple[np.ndarray, int]:
    rows, cols = _u32(pattern.rows), _u32(pattern.cols)
    colors = np.zeros(max(pattern.num_inputs, 1), np.uint32)
    nc = ctypes.c_uint32(0)
    _check(load_library().ehip_column_coloring(pattern.num_outputs, pattern.num_inputs, _p(rows), _p(cols),
                                                len(rows), _p(colors), ctypes.byref(nc)))
    return colors[:pattern.num_inputs].copy(), int(nc.value)


def row_coloring(pattern: JacobianSparsityPattern) -> Tuple[np.ndarray, int]:
    rows, cols = _u32(pattern.rows), _u32(pattern.cols)
    colors = np.zeros(max(pattern.num_outputs, 1), np.uint32)
    nc = ctypes.c_uint32(0)
    _check(load_library().ehip_row_coloring(pattern.num_outputs, pattern.num_inputs, _p(rows), _p(cols),
                                             len(rows), _p(colors), ctypes.byref(nc)))
    return colors[:pattern.num_outputs].copy(), int(nc.value)


def _dtype(bufs: HipTapeBuffers):
    return np.float64 if bufs.f64 else np.float32


def sparse_hessian(ctx: HipContext, bufs: HipTapeBuffers, x):
    """(value, gradient, pattern, values) at one point, GPU HVP per color."""
    F = _dtype(bufs)
    x = np.ascontiguousarray(x, dtype=F)
    n = bufs.num_inputs
    value = np.zeros(1, F)
    grad = np.zeros(max(n, 1), F)
    lib = load_library()
    nnz = ctypes.c_uint64(0)
    lib.ehip_sparse_hessian(ctx.handle, bufs.handle, _p(x), _p(value), _p(grad), None, None, None, 0,
                            ctypes.byref(nnz))
    m = int(nnz.value)
    rows = np.zeros(max(m, 1), np.uint32)
    cols = np.zeros(max(m, 1), np.uint32)
    vals = np.zeros(max(m, 1), F)
    _check(lib.ehip_sparse_hessian(ctx.handle, bufs.handle, _p(x), _p(value), _p(grad), _p(rows), _p(cols),
                                   _p(vals), m, ctypes.byref(nnz)))
    return F(value[0]), grad[:n].copy(), SparsityPattern(n, rows[:m].copy(), cols[:m].copy()), vals[:m].copy()


def sparse_jacobian(ctx: HipContext, bufs: HipTapeBuffers, x):
    """(outputs, pattern, values) at one point, GPU tangent sweep per color."""
    F = _dtype(bufs)
    x = np.ascontiguousarray(x, dtype=F)
    out = np.zeros(max(bufs.num_outputs, 1), F)
    lib = load_library()
    nnz = ctypes.c_uint64(0)
    lib.ehip_sparse_jacobian(ctx.handle, bufs.handle, _p(x), _p(out), None, None, None, 0, ctypes.byref(nnz))
    m = int(nnz.value)
    rows = np.zeros(max(m, 1), np.uint32)
    cols = np.zeros(max(m, 1), np.uint32)
    vals = np.zeros(max(m, 1), F)
    _check(lib.ehip_sparse_jacobian(ctx.handle, bufs.handle, _p(x), _p(out), _p(rows), _p(cols), _p(vals), m,
                                    ctypes.byref(nnz)))
    pat = JacobianSparsityPattern(bufs.num_outputs, bufs.num_inputs, rows[:m].copy(), cols[:m].copy())
    return out[:bufs.num_outputs].copy(), pat, vals[:m].copy()


@dataclass
class NonsmoothBatch:
    """Kink records of a batch (``NonsmoothInfo`` per point, src/nonsmooth.rs)."""

    outputs: np.ndarray      # [B, m]
    tape_index: np.ndarray   # [K]
    opcode: np.ndarray       # [K]
    switching: np.ndarray    # [B, K]
    branch: np.ndarray       # [B, K] int8, +1 / -1


def kink_info(bufs: HipTapeBuffers):
    lib = load_library()
    K = int(lib.ehip_kink_count(bufs.handle))
    ti = np.zeros(max(K, 1), np.uint32)
    op = np.zeros(max(K, 1), np.uint32)
    _check(lib.ehip_kink_info(bufs.handle, _p(ti), _p(op), K))
    return ti[:K].copy(), op[:K].copy()


def forward_nonsmooth(ctx: HipContext, bufs: HipTapeBuffers, inputs, batch: int) -> NonsmoothBatch:
    F = _dtype(bufs)
    x = np.ascontiguousarray(inputs, dtype=F)
    if x.size != batch * bufs.num_inputs:
        raise GpuError(GpuError.OTHER, "GPU error: inputs length must be batch * num_inputs")
    ti, op = kink_info(bufs)
    K = len(ti)
    out = np.zeros(max(batch * bufs.num_outputs, 1), F)
    sw = np.zeros(max(batch * K, 1), F)
    br = np.zeros(max(batch * K, 1), np.int8)
    _check(load_library().ehip_forward_nonsmooth(ctx.handle, bufs.handle, _p(x), batch, _p(out), _p(sw), _p(br)))
    return NonsmoothBatch(out[:batch * bufs.num_outputs].reshape(batch, bufs.num_outputs), ti, op,
                          sw[:batch * K].reshape(batch, K), br[:batch * K].reshape(batch, K))
