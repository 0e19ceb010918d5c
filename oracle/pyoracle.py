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
