"""ctypes binding of the CPU oracle (test infrastructure only).

Loads ``liborc_echidna.so`` built by ``oracle/Makefile`` (``build()`` in the
repository entry point runs it).  Only the tests, the smoke check and
bench.py's cpu_baseline leg import this module.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc_echidna.so")
_lib = None

u32p = ctypes.POINTER(ctypes.c_uint32)
f64p = ctypes.POINTER(ctypes.c_double)
f32p = ctypes.POINTER(ctypes.c_float)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


class OracleTape:
    """A flattened tape held as contiguous numpy arrays for the oracle."""

    def __init__(self, tape):
        self.opcodes = np.ascontiguousarray(tape.opcodes, dtype=np.uint32)
        self.arg0 = np.ascontiguousarray(tape.arg0, dtype=np.uint32)
        self.arg1 = np.ascontiguousarray(tape.arg1, dtype=np.uint32)
        self.values = np.ascontiguousarray(tape.values, dtype=np.float64)
        self.num_ops = len(self.opcodes)
        self.num_inputs = tape.num_inputs
        self.outs = np.ascontiguousarray(tape.all_output_indices(), dtype=np.uint32)
        self.num_outputs = len(self.outs)

    def args(self):
        return (_ptr(self.opcodes, u32p), _ptr(self.arg0, u32p), _ptr(self.arg1, u32p),
                _ptr(self.values, f64p), ctypes.c_uint32(self.num_ops),
                ctypes.c_uint32(self.num_inputs), _ptr(self.outs, u32p),
                ctypes.c_uint32(self.num_outputs))


def _dt(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "f64", f64p
    if dtype == np.float32:
        return "f32", f32p
    raise TypeError(dtype)


def forward(tape, x, dtype=np.float64, nthreads=1):
    t = OracleTape(tape)
    sfx, p = _dt(dtype)
    x = np.ascontiguousarray(x, dtype=dtype).reshape(-1, t.num_inputs)
    B = x.shape[0]
    y = np.zeros((B, t.num_outputs), dtype=dtype)
    getattr(lib(), "orc_forward_" + sfx)(*t.args(), _ptr(x, p), ctypes.c_uint64(B), _ptr(y, p),
                                          ctypes.c_int(nthreads))
    return y


def gradient(tape, x, dtype=np.float64, nthreads=1):
    t = OracleTape(tape)
    sfx, p = _dt(dtype)
    x = np.ascontiguousarray(x, dtype=dtype).reshape(-1, t.num_inputs)
    B = x.shape[0]
    y = np.zeros((B, t.num_outputs), dtype=dtype)
    g = np.zeros((B, t.num_inputs), dtype=dtype)
    getattr(lib(), "orc_gradient_" + sfx)(*t.args(), _ptr(x, p), ctypes.c_uint64(B), _ptr(y, p),
                                           _ptr(g, p), ctypes.c_int(nthreads))
    return y, g


def hvp(tape, x, v, dtype=np.float64, nthreads=1):
    """Per-point x when ``x`` has B rows; one shared x when it is a single row."""
    t = OracleTape(tape)
    sfx, p = _dt(dtype)
    n = t.num_inputs
    v = np.ascontiguousarray(v, dtype=dtype).reshape(-1, n)
    B = v.shape[0]
    x = np.ascontiguousarray(x, dtype=dtype).reshape(-1, n)
    xstride = n if x.shape[0] == B and B > 1 else 0
    if x.shape[0] not in (1, B):
        raise ValueError("x must have 1 or B rows")
    g = np.zeros((B, n), dtype=dtype)
    hv = np.zeros((B, n), dtype=dtype)
    getattr(lib(), "orc_hvp_" + sfx)(*t.args(), _ptr(x, p), ctypes.c_uint64(xstride), _ptr(v, p),
                                      ctypes.c_uint64(B), _ptr(g, p), _ptr(hv, p), ctypes.c_int(nthreads))
    return g, hv


def taylor(tape, x, v, K=3, dtype=np.float64, nthreads=1):
    """Jets [B, m, K] with inputs seeded (x_i, v_i, 0, ...)."""
    t = OracleTape(tape)
    sfx, p = _dt(dtype)
    n = t.num_inputs
    v = np.ascontiguousarray(v, dtype=dtype).reshape(-1, n)
    B = v.shape[0]
    x = np.ascontiguousarray(x, dtype=dtype).reshape(-1, n)
    xstride = n if x.shape[0] == B and B > 1 else 0
    out = np.zeros((B, t.num_outputs, K), dtype=dtype)
    rc = getattr(lib(), "orc_taylor_" + sfx)(*t.args(), _ptr(x, p), ctypes.c_uint64(xstride), _ptr(v, p),
                                              ctypes.c_uint64(B), ctypes.c_int(K), _ptr(out, p),
                                              ctypes.c_int(nthreads))
    if rc != 0:
        raise ValueError(f"unsupported Taylor order {K}")
    return out


def welford(samples, dtype=np.float64):
    sfx, p = _dt(dtype)
    s = np.ascontiguousarray(samples, dtype=dtype)
    out = np.zeros(4, dtype=dtype)
    getattr(lib(), "orc_welford_" + sfx)(_ptr(s, p), ctypes.c_uint64(len(s)), _ptr(out, p))
    return float(out[0]), float(out[1]), float(out[2]), int(out[3])


def laplacian_with_stats(tape, x, dirs):
    t = OracleTape(tape)
    x = np.ascontiguousarray(x, dtype=np.float64)
    d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, t.num_inputs)
    res = np.zeros(5)
    lib().orc_laplacian_with_stats_f64(*t.args(), _ptr(x, f64p), _ptr(d, f64p), ctypes.c_uint64(d.shape[0]),
                                       _ptr(res, f64p))
    return {"value": res[0], "estimate": res[1], "sample_variance": res[2],
            "standard_error": res[3], "num_samples": int(res[4])}


def laplacian_with_control(tape, x, dirs, diag):
    t = OracleTape(tape)
    x = np.ascontiguousarray(x, dtype=np.float64)
    d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, t.num_inputs)
    dg = np.ascontiguousarray(diag, dtype=np.float64)
    res = np.zeros(5)
    lib().orc_laplacian_with_control_f64(*t.args(), _ptr(x, f64p), _ptr(d, f64p), ctypes.c_uint64(d.shape[0]),
                                         _ptr(dg, f64p), _ptr(res, f64p))
    return {"value": res[0], "estimate": res[1], "sample_variance": res[2],
            "standard_error": res[3], "num_samples": int(res[4])}


def hessian_diagonal(tape, x):
    t = OracleTape(tape)
    x = np.ascontiguousarray(x, dtype=np.float64)
    val = np.zeros(1)
    diag = np.zeros(t.num_inputs)
    lib().orc_hessian_diagonal_f64(*t.args(), _ptr(x, f64p), _ptr(val, f64p), _ptr(diag, f64p))
    return float(val[0]), diag


def hessian(tape, x):
    t = OracleTape(tape)
    x = np.ascontiguousarray(x, dtype=np.float64)
    H = np.zeros((t.num_inputs, t.num_inputs))
    lib().orc_hessian_f64(*t.args(), _ptr(x, f64p), _ptr(H, f64p))
    return H


# ---------------------------------------------------------------------------
# Structure and kinks (src/bytecode_tape/sparse.rs, src/sparse.rs,
# src/bytecode_tape/forward.rs:87-187)
# ---------------------------------------------------------------------------
def _pattern(fn, tape):
    t = OracleTape(tape)
    nnz = ctypes.c_uint64(0)
    fn(*t.args(), None, None, ctypes.c_uint64(0), ctypes.byref(nnz))
    n = int(nnz.value)
    rows = np.zeros(max(n, 1), np.uint32)
    cols = np.zeros(max(n, 1), np.uint32)
    fn(*t.args(), _ptr(rows, u32p), _ptr(cols, u32p), ctypes.c_uint64(n), ctypes.byref(nnz))
    return rows[:n].copy(), cols[:n].copy()


def hessian_pattern(tape):
    """detect_sparsity: (rows, cols) lower triangle, sorted, deduplicated."""
    return _pattern(lib().orc_hessian_pattern, tape)


def jacobian_pattern(tape):
    """detect_jacobian_sparsity: (output rows, input cols)."""
    return _pattern(lib().orc_jacobian_pattern, tape)


def greedy_coloring(n, rows, cols):
    rows = np.ascontiguousarray(rows, np.uint32)
    cols = np.ascontiguousarray(cols, np.uint32)
    colors = np.zeros(max(n, 1), np.uint32)
    f = lib().orc_greedy_coloring
    f.restype = ctypes.c_uint32
    k = f(ctypes.c_uint32(n), _ptr(rows, u32p), _ptr(cols, u32p), ctypes.c_uint64(len(rows)), _ptr(colors, u32p))
    return colors[:n].copy(), int(k)


def _intersection(gk, ck, gdim, cdim):
    gk = np.ascontiguousarray(gk, np.uint32)
    ck = np.ascontiguousarray(ck, np.uint32)
    colors = np.zeros(max(cdim, 1), np.uint32)
    f = lib().orc_intersection_coloring
    f.restype = ctypes.c_uint32
    k = f(_ptr(gk, u32p), _ptr(ck, u32p), ctypes.c_uint64(len(gk)), ctypes.c_uint32(gdim), ctypes.c_uint32(cdim),
          _ptr(colors, u32p))
    return colors[:cdim].copy(), int(k)


def column_coloring(num_outputs, num_inputs, rows, cols):
    return _intersection(rows, cols, num_outputs, num_inputs)


def row_coloring(num_outputs, num_inputs, rows, cols):
    return _intersection(cols, rows, num_inputs, num_outputs)


def nonsmooth(tape, X, dtype=np.float64):
    """forward_nonsmooth per point: (outputs [B,m], switching [B,K], branch [B,K])."""
    t = OracleTape(tape)
    sfx, p = _dt(dtype)
    X = np.ascontiguousarray(X, dtype=dtype).reshape(-1, t.num_inputs)
    B = X.shape[0]
    nonsmooth_ops = {10, 11, 36, 37, 38, 39, 40, 41, 42}
    K = int(sum(1 for o in t.opcodes[t.num_inputs:] if int(o) in nonsmooth_ops))
    y = np.zeros((B, t.num_outputs), dtype)
    sw = np.zeros((B, max(K, 1)), dtype)
    br = np.zeros((B, max(K, 1)), np.int8)
    rc = getattr(lib(), "orc_nonsmooth_" + sfx)(*t.args(), _ptr(X, p), ctypes.c_uint64(B), ctypes.c_uint32(K),
                                                 _ptr(y, p), _ptr(sw, p), _ptr(br, ctypes.POINTER(ctypes.c_int8)))
    if rc != 0:
        raise ValueError("kink count mismatch")
    return y, sw[:, :K], br[:, :K]


def magnitudes(tape, x, v=None, nthreads=1):
    """Sums of the absolute terms behind each result (the cancellation floor
    of a parity bar): dict with value [B,m], grad [B,n], hv [B,n], c1 [B],
    c2 [B] (hv, c1 and c2 along v; v = 0 when not given)."""
    t = OracleTape(tape)
    n = t.num_inputs
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, n)
    B = x.shape[0]
    v = np.zeros((B, n)) if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(-1, n)
    if v.shape[0] != B:
        x = np.ascontiguousarray(np.broadcast_to(x, (v.shape[0], n)))
        B = v.shape[0]
    val = np.zeros((B, t.num_outputs))
    g = np.zeros((B, n))
    hv = np.zeros((B, n))
    c12 = np.zeros((B, 2))
    lib().orc_magnitudes_f64(*t.args(), _ptr(x, f64p), ctypes.c_uint64(n), _ptr(v, f64p), ctypes.c_uint64(B),
                             _ptr(val, f64p), _ptr(g, f64p), _ptr(hv, f64p), _ptr(c12, f64p),
                             ctypes.c_int(nthreads))
    return {"value": val, "grad": g, "hv": hv, "c1": c12[:, 0], "c2": c12[:, 1]}
