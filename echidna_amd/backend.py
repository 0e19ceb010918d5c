"""Python mirror of the reference's ``GpuBackend`` over the HIP C ABI.

``HipContext`` plays the role of the reference's ``CudaContext``
(``src/gpu/cuda_backend.rs:477-953``): ``new()`` acquires the device,
``upload_tape`` validates and uploads a ``GpuTapeData`` (raising like the
trait's panic on invalid data), and the batch methods take and return numpy
arrays in the reference's row-major layouts.  f32 and f64 tapes share the
methods; the ``*_f64`` names exist as aliases for readers of the reference.

Every call goes through ``libechidna_hip.so``; there is no CPU fallback.  If
the library is missing, importing this module raises.
"""

from __future__ import annotations

import ctypes
import os
import weakref
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from .tape import GpuError, GpuTapeData

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libechidna_hip.so")

EHIP_F32 = 0
EHIP_F64 = 1

_c_u32p = ctypes.POINTER(ctypes.c_uint32)
_c_u64p = ctypes.POINTER(ctypes.c_uint64)
_c_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

# Every symbol the C header declares, with its ctypes signature.
_SIGS = {
    "ehip_last_error": (ctypes.c_char_p, []),
    "ehip_version": (ctypes.c_char_p, []),
    "ehip_ctx_create": (ctypes.c_int32, [ctypes.c_int32, ctypes.POINTER(_vp)]),
    "ehip_ctx_destroy": (None, [_vp]),
    "ehip_tape_upload": (ctypes.c_int32, [_vp, ctypes.c_int32, _c_u32p, _c_u32p, _c_u32p, _vp, ctypes.c_uint32,
                                          ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _c_u32p,
                                          ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "ehip_tape_free": (None, [_vp]),
    "ehip_num_outputs": (ctypes.c_uint32, [_vp]),
    "ehip_num_inputs": (ctypes.c_uint32, [_vp]),
    "ehip_tape_dtype": (ctypes.c_int32, [_vp]),
    "ehip_forward_batch": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint32, _vp]),
    "ehip_gradient_batch": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "ehip_hvp_batch": (ctypes.c_int32, [_vp, _vp, _vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "ehip_hvp_batch_points": (ctypes.c_int32, [_vp, _vp, _vp, _vp, ctypes.c_uint32, _vp, _vp]),
    "ehip_taylor_kth_batch": (ctypes.c_int32, [_vp, _vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32, _vp]),
    "ehip_forward_batch_dev": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint64, _vp, _vp]),
    "ehip_gradient_batch_dev": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint64, _vp, _vp, _vp]),
    "ehip_hvp_batch_dev": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64, _vp, _vp, _vp]),
    "ehip_taylor_kth_batch_dev": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64,
                                                   ctypes.c_uint32, _vp, _vp]),
    "ehip_program_info": (ctypes.c_int32, [_vp, ctypes.c_int32, _c_u64p]),
    "ehip_compile_program": (ctypes.c_int32, [ctypes.c_int32, _c_u32p, _c_u32p, _c_u32p, _c_f64p, ctypes.c_uint32,
                                              ctypes.c_uint32, ctypes.c_uint32, _c_u32p, ctypes.c_uint32,
                                              ctypes.c_int32, ctypes.c_int32, _c_u32p, ctypes.c_uint64, _c_u32p,
                                              ctypes.c_uint64, _c_u32p, ctypes.c_uint64, _c_u64p]),
    "ehip_set_dense": (ctypes.c_int32, [_vp, ctypes.c_int32]),
    "ehip_set_launch": (ctypes.c_int32, [_vp, ctypes.c_uint32, ctypes.c_uint32]),
    "ehip_laplacian": (ctypes.c_int32, [_vp, _vp, _vp, _vp, ctypes.c_uint32, _c_f64p]),
    "ehip_laplacian_with_control": (ctypes.c_int32, [_vp, _vp, _vp, _vp, ctypes.c_uint32, _vp, _c_f64p]),
    "ehip_hessian_diagonal": (ctypes.c_int32, [_vp, _vp, _vp, _vp, _vp]),
    "ehip_laplacian_batch": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                              ctypes.c_uint64, _vp, _c_f64p]),
    "ehip_laplacian_batch_dev": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                                  ctypes.c_uint64, _vp, _vp, ctypes.c_uint64, _vp, _vp]),
    "ehip_set_fusion": (ctypes.c_int32, [_vp, ctypes.c_int32]),
    "ehip_cv_create": (ctypes.c_int32, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                        ctypes.c_float, ctypes.POINTER(_vp)]),
    "ehip_cv_destroy": (None, [_vp]),
    "ehip_cv_info": (ctypes.c_int32, [_vp, _c_u64p]),
    "ehip_cv_get": (ctypes.c_int32, [_vp, ctypes.c_int32, _vp]),
    "ehip_cv_set_params": (ctypes.c_int32, [_vp, _vp]),
    "ehip_cv_set_gain": (ctypes.c_int32, [_vp, ctypes.c_float]),
    "ehip_cv_gain": (ctypes.c_float, [_vp]),
    "ehip_cv_set_signs": (ctypes.c_int32, [_vp, ctypes.c_uint32]),
    "ehip_cv_forward_dev": (ctypes.c_int32, [_vp, _vp, ctypes.c_uint64, _vp, _vp]),
    "ehip_cv_loss_grad_dev": (ctypes.c_int32, [_vp, _vp, _vp, _vp, ctypes.c_uint64, _vp, _vp]),
    "ehip_cv_adam_step": (ctypes.c_int32, [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                           _vp]),
    "ehip_cv_train_dev": (ctypes.c_int32, [_vp, _vp, _vp, _vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_float, _vp, _vp]),
    "ehip_cv_make_tuples_dev": (ctypes.c_int32, [_vp, _vp, ctypes.c_uint64, ctypes.c_uint64, _vp, _vp, _vp, _vp]),
    "ehip_laplacian_cv_batch_dev": (ctypes.c_int32, [_vp, _vp, _vp, _vp, ctypes.c_uint64, ctypes.c_uint32,
                                                     ctypes.c_uint32, ctypes.c_uint64, _vp, _vp, _vp]),
    "ehip_hessian_trace_batch_dev": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint64, _vp, _vp]),
    "ehip_hessian_sparsity": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp,
                                               ctypes.c_uint64, _c_u64p]),
    "ehip_jacobian_sparsity": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32, _vp,
                                                ctypes.c_uint32, _vp, _vp, ctypes.c_uint64, _c_u64p]),
    "ehip_greedy_coloring": (ctypes.c_int32, [ctypes.c_uint32, _vp, _vp, ctypes.c_uint64, _vp, _c_u32p]),
    "ehip_column_coloring": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, ctypes.c_uint64, _vp,
                                              _c_u32p]),
    "ehip_row_coloring": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, ctypes.c_uint64, _vp,
                                           _c_u32p]),
    "ehip_sparse_hessian": (ctypes.c_int32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_uint64, _c_u64p]),
    "ehip_sparse_jacobian": (ctypes.c_int32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_uint64, _c_u64p]),
    "ehip_kink_count": (ctypes.c_uint32, [_vp]),
    "ehip_kink_info": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint32]),
    "ehip_forward_nonsmooth_dev": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint64, _vp, _vp, _vp, _vp]),
    "ehip_forward_nonsmooth": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint32, _vp, _vp, _vp]),
    "ehip_multi_batch_dev": (ctypes.c_int32, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _vp]),
    "ehip_graph_capture": (ctypes.c_int32, [_vp, _vp, ctypes.c_int32, ctypes.c_uint32, _vp, ctypes.c_uint64, _vp,
                                            ctypes.c_uint64, _vp, _vp, _vp, ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "ehip_graph_launch": (ctypes.c_int32, [_vp, _vp]),
    "ehip_graph_destroy": (None, [_vp]),
    "ehip_autotune_dev": (ctypes.c_int32, [_vp, _vp, ctypes.c_int32, ctypes.c_uint32, _vp, ctypes.c_uint64, _vp,
                                           ctypes.c_uint64, _vp, _vp, _vp, _c_u32p]),
    "ehip_tape_specialize": (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.c_uint32, _c_u32p]),
    "ehip_set_specialize": (ctypes.c_int32, [_vp, ctypes.c_uint32, ctypes.c_uint32]),
    "ehip_set_jet_mma": (ctypes.c_int32, [_vp, ctypes.c_int32]),
    "ehip_gradient_batch_dev_noalloc": (ctypes.c_int32, [_vp, _vp, _vp, ctypes.c_uint64, _vp, _vp, _vp]),
    "ehip_noalloc_scratch_bytes": (ctypes.c_uint64, []),
    "ehip_spec_error": (ctypes.c_char_p, [_vp, ctypes.c_int32]),
    "ehip_spec_source": (ctypes.c_int32, [ctypes.c_int32, _c_u32p, _c_u32p, _c_u32p, _c_f64p, ctypes.c_uint32,
                                          ctypes.c_uint32, ctypes.c_uint32, _c_u32p, ctypes.c_uint32, ctypes.c_int32,
                                          ctypes.c_uint32, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, _c_u64p]),
    "ehip_set_tuned": (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_uint32]),
    "ehip_set_fetch": (ctypes.c_int32, [_vp, ctypes.c_int32]),
    "ehip_get_tuned": (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.c_uint32, _c_u32p]),
    "ehip_tape_to_json": (ctypes.c_int32, [ctypes.c_int32, _vp, _vp, _vp, _vp, ctypes.c_uint64, ctypes.c_uint32,
                                           ctypes.c_uint32, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64, _c_u64p]),
    "ehip_tape_from_json": (ctypes.c_int32, [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int32, _vp, _vp, _vp, _vp,
                                             ctypes.c_uint64, _vp, ctypes.c_uint64, _c_u64p, _c_u32p, _c_u32p,
                                             _c_u64p]),
    "ehip_tape_upload_json": (ctypes.c_int32, [_vp, ctypes.c_int32, ctypes.c_char_p, ctypes.c_uint64,
                                               ctypes.POINTER(_vp)]),
    "ehip_tape_to_cbor": (ctypes.c_int32, [ctypes.c_int32, _vp, _vp, _vp, _vp, ctypes.c_uint64, ctypes.c_uint32,
                                           ctypes.c_uint32, _vp, ctypes.c_uint64, _vp, ctypes.c_uint64, _c_u64p]),
    "ehip_tape_from_cbor": (ctypes.c_int32, [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int32, _vp, _vp, _vp, _vp,
                                             ctypes.c_uint64, _vp, ctypes.c_uint64, _c_u64p, _c_u32p, _c_u32p,
                                             _c_u64p]),
}

class EhipJob(ctypes.Structure):
    """``ehip_job``: one tape's share of a multi-tape queue."""

    _fields_ = [("tape", _vp), ("what", ctypes.c_int32), ("d_x", _vp), ("batch", ctypes.c_uint64),
                ("d_y", _vp), ("d_g", _vp)]


# Evaluation kinds of the autotuner.
TUNE_FORWARD, TUNE_GRADIENT, TUNE_HVP, TUNE_TAYLOR = 0, 1, 2, 3
# specialize(what=SPEC_GRAD_ACC64): f32 gradient kernel with binary64 adjoints
SPEC_GRAD_ACC64 = 8

_lib = None


def load_library(path: str = LIB_PATH):
    """Load the HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: run echidna_amd/build.py (or __graft_entry__.build()) first")
    # torch ships its own libamdhip64.so.7 with the same SONAME as ROCm's; the
    # first one loaded serves the whole process.  Load torch's first so device
    # buffers and streams from torch and this library share one HIP runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> List[str]:
    return list(_SIGS)


def _check(rc: int) -> None:
    if rc != 0:
        msg = load_library().ehip_last_error().decode()
        raise GpuError(rc, msg)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def _dp(t):
    """Device pointer of a torch tensor (or a raw address; 0/None = null)."""
    if t is None or (isinstance(t, int) and t == 0):
        return _vp(None)
    return _vp(t.data_ptr() if hasattr(t, "data_ptr") else int(t))


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


@dataclass
class TaylorBatchResult:
    """``TaylorBatchResult`` (src/gpu/mod.rs:234-241)."""

    values: np.ndarray
    c1s: np.ndarray
    c2s: np.ndarray


@dataclass
class TaylorKthBatchResult:
    """``TaylorKthBatchResult`` (src/gpu/mod.rs:266-272)."""

    coefficients: List[np.ndarray]
    order: int


def compile_program(data: GpuTapeData, mode: int = 1, fuse: bool = True, f64: Optional[bool] = None,
                    dense: bool = True):
    """Compile without a device.

    Modes 0/1 (forward/gradient) return (words[n,8], side table, info);
    modes 2/3 return (windowed stream words, None, info); adding 4 to a
    gradient mode compiles the dual interpreter's HVP program.  info["pool"]
    holds the DENSE pool words.
    """
    lib = load_library()
    f64 = (np.asarray(data.constants).dtype == np.float64) if f64 is None else f64
    consts = np.ascontiguousarray(data.constants, dtype=np.float64)
    outs = _u32(data.output_indices)
    info = (ctypes.c_uint64 * 5)()
    fz = 0 if not fuse else (1 if dense else 2)
    args = [EHIP_F64 if f64 else EHIP_F32, _u32(data.opcodes), _u32(data.arg0), _u32(data.arg1), consts]
    ptrs = [args[0]] + [a.ctypes.data_as(_c_u32p) for a in args[1:4]] + [consts.ctypes.data_as(_c_f64p)]
    rc = lib.ehip_compile_program(*ptrs, data.num_ops, data.num_inputs, data.output_index,
                                  outs.ctypes.data_as(_c_u32p), len(outs), mode, fz,
                                  None, 0, None, 0, None, 0, info)
    nw, T, npool = int(info[2]), int(info[3]), int(info[4])
    nt = max(4, T * (4 if f64 else 3))
    words = np.zeros(max(nw, 8), dtype=np.uint32)
    terms = np.zeros(nt, dtype=np.uint32)
    pool = np.zeros(max(npool, 1), dtype=np.uint32)
    rc = lib.ehip_compile_program(*ptrs, data.num_ops, data.num_inputs, data.output_index,
                                  outs.ctypes.data_as(_c_u32p), len(outs), mode, fz,
                                  words.ctypes.data_as(_c_u32p), len(words), terms.ctypes.data_as(_c_u32p),
                                  len(terms), pool.ctypes.data_as(_c_u32p), len(pool), info)
    _check(rc)
    info2 = {"instructions": int(info[0]), "registers": int(info[1]), "words": nw, "nterms": T,
             "pool": pool[:npool]}
    if mode & 2:
        return words[:nw], None, info2
    return words[:nw].reshape(-1, 8), terms[:nt], info2


def spec_source(data: GpuTapeData, what: int = 1, f64: Optional[bool] = None, arch: Optional[str] = None,
                waves: int = 0):
    """The tape-specialised kernel's source (no device needed).

    With ``arch`` (e.g. "gfx950") the source is also compiled by hipRTC.
    Returns (source text, {"code_bytes", "compile_ms"}).
    """
    lib = load_library()
    f64 = (np.asarray(data.constants).dtype == np.float64) if f64 is None else f64
    consts = np.ascontiguousarray(data.constants, dtype=np.float64)
    outs = _u32(data.output_indices)
    arrs = [_u32(data.opcodes), _u32(data.arg0), _u32(data.arg1)]
    info = (ctypes.c_uint64 * 3)()
    call = lambda a, buf, cap: lib.ehip_spec_source(
        EHIP_F64 if f64 else EHIP_F32, *[x.ctypes.data_as(_c_u32p) for x in arrs], consts.ctypes.data_as(_c_f64p),
        data.num_ops, data.num_inputs, data.output_index, outs.ctypes.data_as(_c_u32p), len(outs), what, waves, a,
        buf, cap, info)
    _check(call(None, None, 0))
    buf = ctypes.create_string_buffer(int(info[0]) + 1)
    _check(call(arch.encode() if arch else None, buf, len(buf)))
    return buf.value.decode(), {"code_bytes": int(info[1]), "compile_ms": info[2] / 1000.0}


class HipTapeBuffers:
    """Uploaded tape handle (``CudaTapeBuffers``, src/gpu/cuda_backend.rs:449-472)."""

    def __init__(self, ctx: "HipContext", handle, f64: bool, num_inputs: int, num_outputs: int):
        self._ctx = ctx
        self.handle = handle
        self.f64 = f64
        self.num_inputs = num_inputs
        self.num_outputs = num_outputs
        self.dtype = np.float64 if f64 else np.float32

    def program_info(self, mode: int = 1) -> dict:
        """Modes 0..2: the forward, gradient and HVP programs; 3: the segmented
        order-3 jet program of an f32 tape (ehip_set_jet_mma); 4: its direct
        plan for device-drawn directions."""
        out = (ctypes.c_uint64 * 4)()
        _check(load_library().ehip_program_info(self.handle, mode, out))
        if mode == 4:
            return {"direct_gemms": int(out[0]), "inputs_dropped": int(out[1]), "segments_skipped": int(out[2])}
        if mode == 3:
            return {"segments": int(out[0]), "gemms": int(out[1]), "gemm_activations": int(out[2]),
                    "gemm_readouts": int(out[3])}
        return {"instructions": int(out[0]), "registers": int(out[1]), "words": int(out[2]),
                "dot_terms": int(out[3])}

    def spec_error(self, what: int) -> str:
        """Why the upload-time specialisation of ``what`` failed ("" if it did not)."""
        return load_library().ehip_spec_error(self.handle, what).decode()

    def free(self) -> None:
        if self.handle:
            if self._ctx.handle:  # a tape must not outlive its context
                load_library().ehip_tape_free(self.handle)
            self.handle = None
            self._ctx._children.discard(self)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HipGraph:
    """A captured hipGraph of repeated evaluations (``ehip_graph``)."""

    def __init__(self, ctx: "HipContext", handle, tape=None):
        self.ctx = ctx
        self.handle = handle
        self.tape = tape  # the recorded launches read the tape's program

    def launch(self, stream: int = 0) -> None:
        _check(load_library().ehip_graph_launch(self.handle, _vp(stream or None)))

    def close(self) -> None:
        if self.handle:
            if self.ctx.handle:
                load_library().ehip_graph_destroy(self.handle)
            self.handle = None
            self.ctx._children.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipContext:
    """MI355X backend context (``CudaContext``, src/gpu/cuda_backend.rs:477-953)."""

    def __init__(self, handle):
        self.handle = handle
        self._children = weakref.WeakSet()  # objects to release before the context

    @staticmethod
    def new(device: int = 0) -> Optional["HipContext"]:
        """``CudaContext::new``: ``None`` when no gfx950 device is present."""
        lib = load_library()
        h = ctypes.c_void_p()
        rc = lib.ehip_ctx_create(device, ctypes.byref(h))
        if rc == GpuError.NO_DEVICE:
            return None
        _check(rc)
        return HipContext(h)

    def close(self) -> None:
        if self.handle:
            # graphs and networks first, then the tapes they use, then the context
            children = list(self._children)
            for child in children:
                if not isinstance(child, HipTapeBuffers):
                    child.close()
            for child in children:
                if isinstance(child, HipTapeBuffers):
                    child.free()
            load_library().ehip_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch(self, wg: int = 0, lds_budget: int = 0) -> None:
        _check(load_library().ehip_set_launch(self.handle, wg, lds_budget))

    def set_fetch(self, mode: int) -> None:
        """Program fetch of untuned launches and the multi-tape queue (0 LDS windows, 1 scalar cache)."""
        _check(load_library().ehip_set_fetch(self.handle, int(mode)))

    def set_dense(self, enable: bool) -> None:
        _check(load_library().ehip_set_dense(self.handle, int(enable)))

    def set_jet_mma(self, enable: bool) -> None:
        """Order-3 f32 jets: DENSE blocks on FP32 MFMA (f32 rounding, not bit-identical)."""
        _check(load_library().ehip_set_jet_mma(self.handle, int(enable)))

    def set_fusion(self, enable: bool) -> None:
        _check(load_library().ehip_set_fusion(self.handle, int(enable)))

    # -- upload --
    def upload_tape(self, data: GpuTapeData, f64: Optional[bool] = None) -> HipTapeBuffers:
        """``upload_tape`` / ``upload_tape_f64``; invalid data raises GpuError."""
        lib = load_library()
        f64 = (np.asarray(data.constants).dtype == np.float64) if f64 is None else f64
        consts = np.ascontiguousarray(data.constants, dtype=np.float64 if f64 else np.float32)
        op, a0, a1 = _u32(data.opcodes), _u32(data.arg0), _u32(data.arg1)
        outs = _u32(data.output_indices)
        h = ctypes.c_void_p()
        rc = lib.ehip_tape_upload(self.handle, EHIP_F64 if f64 else EHIP_F32, op.ctypes.data_as(_c_u32p),
                                  a0.ctypes.data_as(_c_u32p), a1.ctypes.data_as(_c_u32p), _ptr(consts),
                                  data.num_ops, data.num_inputs, data.num_variables, data.output_index,
                                  outs.ctypes.data_as(_c_u32p), len(outs), ctypes.byref(h))
        _check(rc)
        return self._adopt(HipTapeBuffers(self, h, f64, int(lib.ehip_num_inputs(h)), int(lib.ehip_num_outputs(h))))

    def _adopt(self, tape: "HipTapeBuffers") -> "HipTapeBuffers":
        self._children.add(tape)  # freed before the context
        return tape

    upload_tape_f64 = upload_tape

    def upload_tape_json(self, text, f64: bool = True) -> HipTapeBuffers:
        """``serde_json::from_str::<BytecodeTape<F>>`` then ``upload_tape``: one C call."""
        raw = text.encode() if isinstance(text, str) else bytes(text)
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.ehip_tape_upload_json(self.handle, EHIP_F64 if f64 else EHIP_F32, raw, len(raw), ctypes.byref(h)))
        return self._adopt(HipTapeBuffers(self, h, f64, int(lib.ehip_num_inputs(h)), int(lib.ehip_num_outputs(h))))

    def num_outputs(self, tape: HipTapeBuffers) -> int:
        return int(load_library().ehip_num_outputs(tape.handle))

    # -- host-array entry points (reference semantics) --
    def _arr(self, tape: HipTapeBuffers, a, shape) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=tape.dtype)
        if a.size != int(np.prod(shape)):
            raise ValueError(f"expected {int(np.prod(shape))} values, got {a.size}")
        return a

    def forward_batch(self, tape: HipTapeBuffers, inputs, batch_size: int) -> np.ndarray:
        x = self._arr(tape, inputs, (batch_size, tape.num_inputs))
        y = np.zeros(batch_size * tape.num_outputs, dtype=tape.dtype)
        _check(load_library().ehip_forward_batch(self.handle, tape.handle, _ptr(x), batch_size, _ptr(y)))
        return y

    def gradient_batch(self, tape: HipTapeBuffers, inputs, batch_size: int):
        x = self._arr(tape, inputs, (batch_size, tape.num_inputs))
        y = np.zeros(batch_size * tape.num_outputs, dtype=tape.dtype)
        g = np.zeros(batch_size * tape.num_inputs, dtype=tape.dtype)
        _check(load_library().ehip_gradient_batch(self.handle, tape.handle, _ptr(x), batch_size, _ptr(y), _ptr(g)))
        return y, g

    def hvp_batch(self, tape: HipTapeBuffers, x, tangent_dirs, batch_size: int):
        xv = self._arr(tape, x, (tape.num_inputs,))
        v = self._arr(tape, tangent_dirs, (batch_size, tape.num_inputs))
        g = np.zeros(batch_size * tape.num_inputs, dtype=tape.dtype)
        h = np.zeros(batch_size * tape.num_inputs, dtype=tape.dtype)
        _check(load_library().ehip_hvp_batch(self.handle, tape.handle, _ptr(xv), _ptr(v), batch_size, _ptr(g), _ptr(h)))
        return g, h

    def hvp_batch_points(self, tape: HipTapeBuffers, xs, tangent_dirs, batch_size: int):
        xv = self._arr(tape, xs, (batch_size, tape.num_inputs))
        v = self._arr(tape, tangent_dirs, (batch_size, tape.num_inputs))
        g = np.zeros(batch_size * tape.num_inputs, dtype=tape.dtype)
        h = np.zeros(batch_size * tape.num_inputs, dtype=tape.dtype)
        _check(load_library().ehip_hvp_batch_points(self.handle, tape.handle, _ptr(xv), _ptr(v), batch_size,
                                                     _ptr(g), _ptr(h)))
        return g, h

    def taylor_forward_kth_batch(self, tape: HipTapeBuffers, primal_inputs, direction_seeds, batch_size: int,
                                 order: int) -> TaylorKthBatchResult:
        x = self._arr(tape, primal_inputs, (batch_size, tape.num_inputs))
        v = self._arr(tape, direction_seeds, (batch_size, tape.num_inputs))
        total = batch_size * tape.num_outputs
        c = np.zeros(max(order, 1) * total, dtype=tape.dtype)
        _check(load_library().ehip_taylor_kth_batch(self.handle, tape.handle, _ptr(x), _ptr(v), batch_size, order,
                                                     _ptr(c)))
        return TaylorKthBatchResult([c[k * total:(k + 1) * total] for k in range(order)], order)

    def taylor_forward_2nd_batch(self, tape, primal_inputs, direction_seeds, batch_size: int) -> TaylorBatchResult:
        """Default trait method: delegates to order 3 (src/gpu/mod.rs:189-204)."""
        r = self.taylor_forward_kth_batch(tape, primal_inputs, direction_seeds, batch_size, 3)
        return TaylorBatchResult(*r.coefficients)

    forward_batch_f64 = forward_batch
    gradient_batch_f64 = gradient_batch
    hvp_batch_f64 = hvp_batch
    taylor_forward_kth_batch_f64 = taylor_forward_kth_batch
    taylor_forward_2nd_batch_f64 = taylor_forward_2nd_batch

    # -- device-pointer entry points (torch tensors already in HBM) --
    def forward_batch_dev(self, tape, d_inputs, batch: int, d_outputs, stream: int = 0) -> None:
        _check(load_library().ehip_forward_batch_dev(self.handle, tape.handle, _vp(d_inputs.data_ptr()), batch,
                                                      _vp(d_outputs.data_ptr()), _vp(stream or None)))

    def gradient_batch_dev(self, tape, d_inputs, batch: int, d_outputs, d_grads, stream: int = 0) -> None:
        _check(load_library().ehip_gradient_batch_dev(self.handle, tape.handle, _vp(d_inputs.data_ptr()), batch,
                                                       _vp(d_outputs.data_ptr()), _vp(d_grads.data_ptr()),
                                                       _vp(stream or None)))

    def gradient_batch_dev_noalloc(self, tape, d_inputs, batch: int, d_outputs, d_grads, stream: int = 0) -> None:
        """``gradient_batch_dev`` that allocates nothing (fixed per-GPU scratch)."""
        _check(load_library().ehip_gradient_batch_dev_noalloc(self.handle, tape.handle, _vp(d_inputs.data_ptr()),
                                                               batch, _vp(d_outputs.data_ptr()),
                                                               _vp(d_grads.data_ptr()), _vp(stream or None)))

    def multi_batch_dev(self, jobs, wg: int = 0, lds_budget: int = 0, stream: int = 0) -> None:
        """Run many (tape, inputs, outputs) jobs in one persistent launch.

        ``jobs``: sequence of (tape, what, d_x, batch, d_y, d_g) with device
        tensors (what 0 forward, 1 gradient; d_y / d_g may be None).
        """
        arr = (EhipJob * max(len(jobs), 1))()
        for k, (tape, what, d_x, batch, d_y, d_g) in enumerate(jobs):
            arr[k] = EhipJob(tape.handle, what, _dp(d_x), batch, _dp(d_y), _dp(d_g))
        _check(load_library().ehip_multi_batch_dev(self.handle, ctypes.cast(arr, _vp), len(jobs), wg, lds_budget,
                                                   _vp(stream or None)))

    def autotune_dev(self, tape: "HipTapeBuffers", what: int, batch: int, d_x, d_y=0, d_g=0, d_h=0, d_dirs=0,
                     x_stride: Optional[int] = None, order: int = 0) -> dict:
        """Pick the fastest launch shape for ``what`` (TUNE_*) on these buffers."""
        out = (ctypes.c_uint32 * 5)()
        xs = tape.num_inputs if x_stride is None else x_stride
        _check(load_library().ehip_autotune_dev(self.handle, tape.handle, what, order, _dp(d_x), xs, _dp(d_dirs),
                                                batch, _dp(d_y), _dp(d_g), _dp(d_h), out))
        return {"wg": int(out[0]), "lds_budget": int(out[1]), "us": int(out[2]), "fetch": int(out[3]),
                "points_per_lane": int(out[4])}

    def graph_capture(self, tape: "HipTapeBuffers", what: int, batch: int, d_x, d_y=0, d_g=0, d_h=0, d_dirs=0,
                      x_stride: Optional[int] = None, order: int = 0, repeats: int = 1) -> "HipGraph":
        """Record ``repeats`` launches of ``what`` (TUNE_*) into a hipGraph."""
        h = _vp()
        xs = tape.num_inputs if x_stride is None else x_stride
        _check(load_library().ehip_graph_capture(self.handle, tape.handle, what, order, _dp(d_x), xs, _dp(d_dirs),
                                                 batch, _dp(d_y), _dp(d_g), _dp(d_h), repeats, ctypes.byref(h)))
        g = HipGraph(self, h, tape)
        self._children.add(g)
        return g

    def set_specialize(self, max_entries: int, what_mask: int = 0b011) -> None:
        """Specialise later uploads of at most ``max_entries`` entries (0 = off)
        for the evaluations in ``what_mask`` (bit 0 forward, 1 gradient, 2 hvp,
        3..7 Taylor order 1..5)."""
        _check(load_library().ehip_set_specialize(self.handle, max_entries, what_mask))

    @staticmethod
    def specialize(tape: "HipTapeBuffers", what: int = 1, waves: int = 0) -> dict:
        """Build the tape-specialised kernel for ``what`` (0 forward, 1 gradient, 2 hvp,
        3..7 Taylor jets of order 1..5, which the STDE entry points use too;
        8: the gradient kernel of an f32 tape with binary64 adjoint
        accumulation, serving the gradient evaluations in place of 1).

        Untuned evaluations of that kind then run it; ``set_tuned`` (fetch 2
        or an interpreter shape) and ``autotune_dev`` choose explicitly.
        ``waves``: occupancy target 1..8, or 0 for the automatic choice.
        """
        out = (ctypes.c_uint32 * 5)()
        _check(load_library().ehip_tape_specialize(tape.handle, what, waves, out))
        return {"vgprs": int(out[0]), "scratch": int(out[1]), "compile_ms": out[2] / 1000.0,
                "code_bytes": int(out[3]), "waves_per_eu": int(out[4])}

    @staticmethod
    def set_tuned(tape: "HipTapeBuffers", what: int, wg: int, lds_budget: int, order: int = 0,
                  fetch: int = 0, points_per_lane: int = 0) -> None:
        """Pin a launch shape; ``fetch`` 0 stages the program in LDS windows, 1 reads it through the
        scalar cache, 2 runs the tape-specialised kernel (``specialize``), which alone takes
        ``points_per_lane`` above 1."""
        _check(load_library().ehip_set_tuned(tape.handle, what, order, wg, lds_budget, fetch, points_per_lane))

    @staticmethod
    def get_tuned(tape: "HipTapeBuffers", what: int, order: int = 0) -> dict:
        out = (ctypes.c_uint32 * 4)()
        _check(load_library().ehip_get_tuned(tape.handle, what, order, out))
        return {"wg": int(out[0]), "lds_budget": int(out[1]), "fetch": int(out[2]), "points_per_lane": int(out[3])}

    def hvp_batch_dev(self, tape, d_x, x_stride: int, d_dirs, batch: int, d_grads, d_hvps, stream: int = 0) -> None:
        _check(load_library().ehip_hvp_batch_dev(self.handle, tape.handle, _vp(d_x.data_ptr()), x_stride,
                                                  _vp(d_dirs.data_ptr()), batch, _vp(d_grads.data_ptr()),
                                                  _vp(d_hvps.data_ptr()), _vp(stream or None)))

    def taylor_kth_batch_dev(self, tape, d_x, x_stride: int, d_dirs, batch: int, order: int, d_coeffs,
                             stream: int = 0) -> None:
        _check(load_library().ehip_taylor_kth_batch_dev(self.handle, tape.handle, _vp(d_x.data_ptr()), x_stride,
                                                         _vp(d_dirs.data_ptr()), batch, order,
                                                         _vp(d_coeffs.data_ptr()), _vp(stream or None)))
