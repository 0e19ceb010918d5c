"""Learned low-rank control variate for GPU STDE (Python mirror of the C ABI).

The Laplacian estimator averages samples v^T H(x) v over directions v with
E[v v^T] = I.  A network maps x to L(x) in R^{n x r}; each sample has
v^T L L^T v - tr(L L^T) (mean zero for any L) subtracted, so the estimate
stays unbiased whatever the network predicts, and its variance drops by as
much as L L^T explains H.  No counterpart in the reference; the nearest is
``laplacian_with_control`` (src/stde/laplacian.rs:90-129,
src/gpu/stde_gpu.rs:135-200) with a fixed diagonal instead of L L^T.

Everything runs on the device through ``libechidna_hip.so``: tuples
(x, v, Hv) from the HVP kernel, BF16-MFMA forward/backward GEMMs with FP32
accumulation, Adam, and the FP64 per-point reduction.  ``LearnedStde`` adds
the policy: train, measure the variance ratio on held-out points, and fall
back to plain STDE when the ratio is not below 1.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .backend import HipContext, HipTapeBuffers, _check, _dp, _vp, load_library
from .tape import GpuError

DIR_RADEMACHER = 1
DIR_GAUSSIAN = 2


def _torch():
    import torch

    return torch


class LearnedControlVariate:
    """Handle of an ``ehip_cv`` network (n -> hidden -> hidden -> n*rank)."""

    def __init__(self, ctx: HipContext, n: int, rank: int = 16, hidden: int = 256, seed: int = 0,
                 out_scale: float = 1.0):
        h = _vp()
        _check(load_library().ehip_cv_create(ctx.handle, n, rank, hidden, seed, out_scale, ctypes.byref(h)))
        self.handle = h
        self.ctx = ctx
        self.n, self.rank, self.hidden = n, rank, hidden
        self.npos = rank
        ctx._children.add(self)  # released before the context

    def close(self) -> None:
        if self.handle:
            if self.ctx.handle:  # a network must not outlive its context
                load_library().ehip_cv_destroy(self.handle)
            self.handle = None
            self.ctx._children.discard(self)

    @property
    def gain(self) -> float:
        return float(load_library().ehip_cv_gain(self.handle))

    def set_gain(self, gain: float) -> None:
        """L(x) = gain * (W3 h2 + b3)."""
        _check(load_library().ehip_cv_set_gain(self.handle, gain))

    def set_signs(self, npos: int) -> None:
        """Use v^T L S L^T v - tr(L S L^T), S = diag(+1 x npos, -1 x (rank - npos))."""
        _check(load_library().ehip_cv_set_signs(self.handle, npos))
        self.npos = npos

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parameters -------------------------------------------------------
    def info(self) -> dict:
        out = (ctypes.c_uint64 * 5)()
        _check(load_library().ehip_cv_info(self.handle, out))
        return {"n": out[0], "rank": out[1], "hidden": out[2], "params": out[3], "steps": out[4]}

    @property
    def num_params(self) -> int:
        return int(self.info()["params"])

    def _get(self, which: int) -> np.ndarray:
        a = np.empty(self.num_params, dtype=np.float32)
        _check(load_library().ehip_cv_get(self.handle, which, _vp(a.ctypes.data)))
        return a

    def params(self) -> np.ndarray:
        return self._get(0)

    def grads(self) -> np.ndarray:
        return self._get(1)

    def moments(self):
        return self._get(2), self._get(3)

    def set_params(self, flat) -> None:
        a = np.ascontiguousarray(flat, dtype=np.float32)
        if a.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {a.size}")
        _check(load_library().ehip_cv_set_params(self.handle, _vp(a.ctypes.data)))

    def split(self, flat: np.ndarray):
        """Views (W1 [H,n], b1, W2 [H,H], b2, W3 [n*r,H], b3) of a flat vector."""
        n, H, nr = self.n, self.hidden, self.n * self.rank
        sizes = [H * n, H, H * H, H, nr * H, nr]
        shapes = [(H, n), (H,), (H, H), (H,), (nr, H), (nr,)]
        out, o = [], 0
        for sz, sh in zip(sizes, shapes):
            out.append(flat[o:o + sz].reshape(sh))
            o += sz
        return out

    # -- device passes ----------------------------------------------------
    def forward_dev(self, d_x, batch: int, d_L, stream: int = 0) -> None:
        _check(load_library().ehip_cv_forward_dev(self.handle, _dp(d_x), batch, _dp(d_L), _vp(stream or None)))

    def forward(self, d_x):
        torch = _torch()
        B = d_x.shape[0]
        L = torch.empty(B, self.n * self.rank, dtype=torch.float32, device=d_x.device)
        self.forward_dev(d_x, B, L)
        torch.cuda.synchronize()
        return L

    def loss_grad(self, d_x, d_v, d_hv) -> float:
        torch = _torch()
        loss = torch.zeros(1, dtype=torch.float32, device=d_x.device)
        _check(load_library().ehip_cv_loss_grad_dev(self.handle, _dp(d_x), _dp(d_v), _dp(d_hv), d_x.shape[0],
                                                    _dp(loss), _vp(None)))
        torch.cuda.synchronize()
        return float(loss.cpu()[0])

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        _check(load_library().ehip_cv_adam_step(self.handle, lr, beta1, beta2, eps, _vp(None)))

    def train(self, d_x, d_v, d_hv, batch: int = 512, epochs: int = 1, lr: float = 1e-3,
              schedule: str = "constant") -> np.ndarray:
        """Minibatch Adam over the tuples; returns the per-step mean losses.

        ``schedule="cosine"`` sets epoch e's rate to lr * (1 + cos(pi e / epochs)) / 2.
        """
        torch = _torch()
        T = d_x.shape[0]
        steps = (T + batch - 1) // batch
        hist = torch.zeros(epochs * steps, dtype=torch.float32, device=d_x.device)
        if schedule not in ("constant", "cosine"):
            raise ValueError(f"unknown schedule {schedule!r}")
        rates = [lr] if schedule == "constant" else [lr * 0.5 * (1.0 + np.cos(np.pi * e / epochs))
                                                     for e in range(epochs)]
        per = epochs if schedule == "constant" else 1
        for e, rate in enumerate(rates):
            _check(load_library().ehip_cv_train_dev(self.handle, _dp(d_x), _dp(d_v), _dp(d_hv), T, batch, per,
                                                    float(rate), _dp(hist[e * steps:]), _vp(None)))
        torch.cuda.synchronize()
        return hist.cpu().numpy()


def make_tuples(ctx: HipContext, tape: HipTapeBuffers, ntuples: int, seed: int):
    """(x, v, Hv) training tuples as FP32 device tensors [ntuples, n]."""
    torch = _torch()
    n = tape.num_inputs
    x = torch.empty(ntuples, n, dtype=torch.float32, device="cuda")
    v = torch.empty_like(x)
    hv = torch.empty_like(x)
    _check(load_library().ehip_cv_make_tuples_dev(ctx.handle, tape.handle, ntuples, seed, _dp(x), _dp(v), _dp(hv),
                                                  _vp(None)))
    torch.cuda.synchronize()
    return x, v, hv


def laplacian_batch(ctx: HipContext, tape: HipTapeBuffers, d_x, S: int, kind: int = DIR_GAUSSIAN, seed: int = 0,
                    cv: Optional[LearnedControlVariate] = None, with_plain: bool = False):
    """Per-point STDE [P, 5] (value, estimate, variance, SE, count) with the
    learned control variate (``cv``) or plain; ``with_plain`` also returns the
    plain statistics of the same samples."""
    torch = _torch()
    P = d_x.shape[0]
    out = torch.empty(P, 5, dtype=torch.float64, device=d_x.device)
    plain = torch.empty(P, 5, dtype=torch.float64, device=d_x.device) if with_plain else None
    _check(load_library().ehip_laplacian_cv_batch_dev(ctx.handle, tape.handle, cv.handle if cv else None, _dp(d_x),
                                                      P, S, kind, seed, _dp(out), _dp(plain), _vp(None)))
    torch.cuda.synchronize()
    if with_plain:
        return out.cpu().numpy(), plain.cpu().numpy()
    return out.cpu().numpy()


def hessian_trace_batch(ctx: HipContext, tape: HipTapeBuffers, d_x) -> np.ndarray:
    """Exact tr(H(x_p)) per point from n basis-direction jets."""
    torch = _torch()
    P = d_x.shape[0]
    out = torch.empty(P, 5, dtype=torch.float64, device=d_x.device)
    _check(load_library().ehip_hessian_trace_batch_dev(ctx.handle, tape.handle, _dp(d_x), P, _dp(out), _vp(None)))
    torch.cuda.synchronize()
    return out.cpu().numpy()[:, 1].copy()


@dataclass
class LearnedStde:
    """Train a control variate for a tape, validate it, and estimate with it.

    ``fit`` trains on tuples from the tape, then measures the variance ratio
    (mean CV variance / mean plain variance over the same samples) on
    held-out points; ``use_cv`` is False -- plain STDE -- unless that ratio
    is below 1.
    """

    ctx: HipContext
    tape: HipTapeBuffers
    rank: int = 16
    hidden: int = 256
    seed: int = 0
    npos: Optional[int] = None  # signed factor columns with +1; None: rank (L L^T)
    cv: Optional[LearnedControlVariate] = None
    variance_ratio: float = float("nan")
    use_cv: bool = False
    history: List[float] = field(default_factory=list)

    def fit(self, ntuples: int, epochs: int = 4, batch: int = 512, lr: float = 1e-3, val_points: int = 256,
            val_S: int = 16, out_scale: Optional[float] = None, kind: int = DIR_GAUSSIAN,
            schedule: str = "constant") -> "LearnedStde":
        """Train on ``ntuples`` tuples, then validate.

        Scale: the output gain is set to sqrt(s) with s = rms|Hv| / rms|v|
        (= ||H||_F / sqrt(n) on average) measured on the first tuples, so the
        network fits H / s, whose Frobenius norm is about sqrt(n) for every
        tape; the default ``out_scale`` starts L L^T near that size.
        """
        torch = _torch()
        n = self.tape.num_inputs
        if out_scale is None:
            out_scale = (n * self.rank) ** -0.25 / 0.6
        self.cv = LearnedControlVariate(self.ctx, n, self.rank, self.hidden, self.seed, out_scale)
        if self.npos is not None:
            self.cv.set_signs(self.npos)
        x, v, hv = make_tuples(self.ctx, self.tape, ntuples, self.seed + 1)
        k = min(ntuples, 4096)
        hs = hv[:k].double().cpu().numpy()
        vs = v[:k].double().cpu().numpy()
        s = float(np.sqrt(np.mean(hs * hs) / max(np.mean(vs * vs), 1e-300)))
        if np.isfinite(s) and s > 0:
            self.cv.set_gain(float(np.sqrt(s)))
        self.history = list(self.cv.train(x, v, hv, batch, epochs, lr, schedule))
        del x, v, hv
        xv = self.validation_points(val_points)
        cvr, plain = laplacian_batch(self.ctx, self.tape, xv, val_S, kind, self.seed + 3, self.cv, with_plain=True)
        self.variance_ratio = float(np.mean(cvr[:, 2]) / max(np.mean(plain[:, 2]), 1e-300))
        self.use_cv = bool(self.variance_ratio < 1.0)
        torch.cuda.synchronize()
        return self

    def validation_points(self, P: int):
        torch = _torch()
        from . import workloads as wl

        dt = np.float64 if self.tape.f64 else np.float32
        return torch.from_numpy(wl.points_normal(self.seed + 2, P, self.tape.num_inputs).astype(dt)).cuda()

    def estimate(self, d_x, S: int, kind: int = DIR_GAUSSIAN, seed: int = 0) -> np.ndarray:
        if self.cv is None:
            raise GpuError(5, "GPU error: LearnedStde.fit has not run")
        return laplacian_batch(self.ctx, self.tape, d_x, S, kind, seed, self.cv if self.use_cv else None)
