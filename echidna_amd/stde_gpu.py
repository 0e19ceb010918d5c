"""GPU STDE estimators — mirror of the reference's ``stde_gpu`` module
(``src/gpu/stde_gpu.rs:51-200``) over the HIP backend, plus batched
many-point variants.

The Taylor samples are reduced on the device in binary64 (one wavefront per
point, Welford within lanes, Chan merges across lanes) instead of the
reference's host-side f32 Welford; results are ``EstimatorResult`` values.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from .backend import HipContext, HipTapeBuffers, _check, _ptr, load_library

DIR_SUPPLIED, DIR_RADEMACHER, DIR_GAUSSIAN = 0, 1, 2


@dataclass
class EstimatorResult:
    """``EstimatorResult`` (src/stde/types.rs:9-22)."""

    value: float
    estimate: float
    sample_variance: float
    standard_error: float
    num_samples: int

    @staticmethod
    def from_row(r) -> "EstimatorResult":
        return EstimatorResult(float(r[0]), float(r[1]), float(r[2]), float(r[3]), int(r[4]))


def _dirs(tape: HipTapeBuffers, directions) -> np.ndarray:
    d = np.ascontiguousarray(np.asarray(directions), dtype=tape.dtype)
    if d.ndim != 2 or d.shape[1] != tape.num_inputs:
        raise ValueError("direction length must match x")
    return d


def laplacian_gpu(ctx: HipContext, tape: HipTapeBuffers, x, directions) -> EstimatorResult:
    """Hutchinson Laplacian at one point (``laplacian_gpu``)."""
    xv = np.ascontiguousarray(x, dtype=tape.dtype)
    d = _dirs(tape, directions) if len(directions) else np.zeros((0, tape.num_inputs), tape.dtype)
    res = np.zeros(5)
    _check(load_library().ehip_laplacian(ctx.handle, tape.handle, _ptr(xv), _ptr(d), d.shape[0],
                                         res.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return EstimatorResult.from_row(res)


def laplacian_with_control_gpu(ctx: HipContext, tape: HipTapeBuffers, x, directions,
                               control_diagonal) -> EstimatorResult:
    """Diagonal control variate Laplacian (``laplacian_with_control_gpu``)."""
    xv = np.ascontiguousarray(x, dtype=tape.dtype)
    d = _dirs(tape, directions)
    cd = np.ascontiguousarray(control_diagonal, dtype=tape.dtype)
    if cd.size != tape.num_inputs:
        raise ValueError("control diagonal length must match x")
    res = np.zeros(5)
    _check(load_library().ehip_laplacian_with_control(ctx.handle, tape.handle, _ptr(xv), _ptr(d), d.shape[0],
                                                      _ptr(cd), res.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return EstimatorResult.from_row(res)


def hessian_diagonal_gpu(ctx: HipContext, tape: HipTapeBuffers, x):
    """Exact Hessian diagonal via n basis jets (``hessian_diagonal_gpu``)."""
    xv = np.ascontiguousarray(x, dtype=tape.dtype)
    val = np.zeros(1, tape.dtype)
    diag = np.zeros(tape.num_inputs, tape.dtype)
    _check(load_library().ehip_hessian_diagonal(ctx.handle, tape.handle, _ptr(xv), _ptr(val), _ptr(diag)))
    return float(val[0]), diag


def laplacian_batch(ctx: HipContext, tape: HipTapeBuffers, xs, S: int, kind: int = DIR_RADEMACHER,
                    seed: int = 0, directions=None) -> np.ndarray:
    """Per-point Laplacian estimates for many points; returns [npoints, 5]."""
    X = np.ascontiguousarray(xs, dtype=tape.dtype).reshape(-1, tape.num_inputs)
    npts = X.shape[0]
    d = None
    if kind == DIR_SUPPLIED:
        d = np.ascontiguousarray(directions, dtype=tape.dtype).reshape(npts * S, tape.num_inputs)
    res = np.zeros((npts, 5))
    _check(load_library().ehip_laplacian_batch(ctx.handle, tape.handle, _ptr(X), npts, S, kind, seed,
                                               _ptr(d) if d is not None else None,
                                               res.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return res


def laplacian_batch_dev(ctx: HipContext, tape: HipTapeBuffers, d_xs, npoints: int, S: int, kind: int, seed: int,
                        d_results, d_dirs=None, d_diag=None, diag_stride: int = 0, stream: int = 0) -> None:
    vp = ctypes.c_void_p
    _check(load_library().ehip_laplacian_batch_dev(
        ctx.handle, tape.handle, vp(d_xs.data_ptr()), npoints, S, kind, seed,
        vp(d_dirs.data_ptr()) if d_dirs is not None else None,
        vp(d_diag.data_ptr()) if d_diag is not None else None, diag_stride,
        vp(d_results.data_ptr()), vp(stream or None)))
