"""Numpy restatement of the learned control variate's network, loss, gradients
and Adam step (test helper; the device code is the product).

``bf=True`` rounds every GEMM operand to BF16 (round to nearest even) as the
device does when it stages tiles into LDS, and accumulates in float64; the
device accumulates in FP32, so the two agree to FP32 rounding plus the rare
BF16 tie flipped by it.  ``bf=False`` is plain float64, for finite
differences.
"""

import numpy as np


def bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def split(flat, n, H, r):
    nr = n * r
    sizes = [H * n, H, H * H, H, nr * H, nr]
    shapes = [(H, n), (H,), (H, H), (H,), (nr, H), (nr,)]
    out, o = [], 0
    for sz, sh in zip(sizes, shapes):
        out.append(np.asarray(flat[o:o + sz], dtype=np.float64).reshape(sh))
        o += sz
    return out


def _mm(a, b, bf):
    """a [M,K] @ b [K,N] with BF16-rounded operands when bf."""
    if bf:
        return bf16(a) @ bf16(b)
    return np.asarray(a, np.float64) @ np.asarray(b, np.float64)


def forward(flat, X, n, H, r, bf=True, gain=1.0):
    W1, b1, W2, b2, W3, b3 = split(flat, n, H, r)
    H1 = np.tanh(_mm(X, W1.T, bf) + b1)
    H2 = np.tanh(_mm(H1, W2.T, bf) + b2)
    Lf = gain * (_mm(H2, W3.T, bf) + b3)
    if bf:  # the device keeps activations in FP32
        H1, H2, Lf = (a.astype(np.float32).astype(np.float64) for a in (H1, H2, Lf))
    return H1, H2, Lf


def loss_grad(flat, X, V, HV, n, H, r, bf=True, gain=1.0, npos=None):
    """Mean over tuples of ||hv - L S L^T v||^2 / n and its flat gradient,
    L = gain * (W3 h2 + b3), S = diag(+1 x npos, -1 x (r - npos))."""
    sgn = np.where(np.arange(r) < (r if npos is None else npos), 1.0, -1.0)
    W1, b1, W2, b2, W3, b3 = split(flat, n, H, r)
    T = X.shape[0]
    H1, H2, Lf = forward(flat, X, n, H, r, bf, gain)
    L = Lf.reshape(T, n, r)
    V = np.asarray(V, np.float64)
    HV = np.asarray(HV, np.float64)
    u = np.einsum("tir,ti->tr", L, V) * sgn
    e = HV - np.einsum("tir,tr->ti", L, u)
    w = np.einsum("tir,ti->tr", L, e) * sgn
    loss = np.mean(np.sum(e * e, axis=1) / n)
    gs = -2.0 * gain / (n * T)
    dL = gs * (np.einsum("ti,tr->tir", e, u) + np.einsum("ti,tr->tir", V, w))
    dLf = dL.reshape(T, n * r)
    if bf:
        dLf = dLf.astype(np.float32).astype(np.float64)
    dW3 = _mm(dLf.T, H2, bf)
    db3 = dLf.sum(axis=0)
    dZ2 = _mm(dLf, W3, bf) * (1.0 - H2 * H2)
    dW2 = _mm(dZ2.T, H1, bf)
    db2 = dZ2.sum(axis=0)
    dZ1 = _mm(dZ2, W2, bf) * (1.0 - H1 * H1)
    dW1 = _mm(dZ1.T, X, bf)
    db1 = dZ1.sum(axis=0)
    grad = np.concatenate([g.reshape(-1) for g in (dW1, db1, dW2, db2, dW3, db3)])
    return loss, grad


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """PyTorch-form Adam step t (1-based); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    bc1 = 1 - b1 ** t
    bc2 = 1 - b2 ** t
    p = p - (lr / bc1) * m / (np.sqrt(v) / np.sqrt(bc2) + eps)
    return p, m, v


def init_params(n, H, r, seed, out_scale):
    """The device initialisation: N(0, 1/fan_in) at each flat index of the
    (seed, STREAM_CV) Philox stream, output layer times out_scale, zero biases."""
    from echidna_amd import philox

    nr = n * r
    sizes = [H * n, H, H * H, H, nr * H, nr]
    off = np.concatenate([[0], np.cumsum(sizes)])
    flat = np.zeros(off[-1])
    sds = [1 / np.sqrt(n), 1 / np.sqrt(H), out_scale / np.sqrt(H)]
    for k in range(3):
        idx = np.arange(off[2 * k], off[2 * k + 1], dtype=np.uint64)
        flat[off[2 * k]:off[2 * k + 1]] = sds[k] * philox.normal(seed, philox.STREAM_CV, idx)
    return flat.astype(np.float32)
