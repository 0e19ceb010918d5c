"""Synthetic workloads: the tapes and point sets the batched evaluator runs on.

Every tape is produced through ``tape.record`` so it carries exactly the
entries the reference recorder would emit for the same closure (per-use
``Const`` entries, folding rules).  Weights, points and directions are drawn
from the Philox streams in ``philox`` keyed by the workload seed.

Workloads
---------
rosenbrock(n)            extended Rosenbrock, written like the reference's
                         ``benches/common/mod.rs:10-20`` closure.
tanh_mlp(n, H, seed)     f(x) = sum_j w2_j tanh(sum_i W1_ji x_i + b_j);
                         W1 ~ N(0, 1/n), b ~ N(0, 0.1), w2 ~ N(0, 1/H)
                         (second argument is the variance).
tanh_mlp2(n, H1, H2, s)  two hidden tanh layers, same conventions.
random_dag(...)          random expression DAG with fan-in 2 to earlier slots.
"""

from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from . import philox
from .tape import BReverse, BytecodeTape, record


def rosenbrock_fn(x: List[BReverse]) -> BReverse:
    s = BReverse.constant(0.0)
    for i in range(len(x) - 1):
        t1 = BReverse.constant(1.0) - x[i]
        t2 = x[i + 1] - x[i] * x[i]
        s = s + t1 * t1 + BReverse.constant(100.0) * t2 * t2
    return s


def rosenbrock(n: int) -> BytecodeTape:
    tape, _ = record(rosenbrock_fn, [0.5] * n)
    return tape


def _layer(x: List[BReverse], W: np.ndarray, b: np.ndarray) -> List[BReverse]:
    out = []
    for j in range(W.shape[0]):
        s = x[0] * float(W[j, 0])
        for i in range(1, W.shape[1]):
            s = s + x[i] * float(W[j, i])
        s = s + float(b[j])
        out.append(s.tanh())
    return out


def _readout(h: List[BReverse], w: np.ndarray) -> BReverse:
    acc = BReverse.constant(0.0)
    for j in range(len(h)):
        acc = acc + h[j] * float(w[j])
    return acc


def mlp_weights(n: int, H: int, seed: int):
    W1 = philox.normal_matrix(seed, philox.STREAM_W1, H, n, std=math.sqrt(1.0 / n))
    b1 = philox.normal_matrix(seed, philox.STREAM_B1, 1, H, std=math.sqrt(0.1))[0]
    w2 = philox.normal_matrix(seed, philox.STREAM_W2, 1, H, std=math.sqrt(1.0 / H))[0]
    return W1, b1, w2


def tanh_mlp(n: int, H: int, seed: int) -> BytecodeTape:
    W1, b1, w2 = mlp_weights(n, H, seed)
    tape, _ = record(lambda x: _readout(_layer(x, W1, b1), w2), [0.0] * n)
    return tape


def mlp2_weights(n: int, H1: int, H2: int, seed: int):
    W1 = philox.normal_matrix(seed, philox.STREAM_W1, H1, n, std=math.sqrt(1.0 / n))
    b1 = philox.normal_matrix(seed, philox.STREAM_B1, 1, H1, std=math.sqrt(0.1))[0]
    W2 = philox.normal_matrix(seed, philox.STREAM_W2, H2, H1, std=math.sqrt(1.0 / H1))
    b2 = philox.normal_matrix(seed, philox.STREAM_B2, 1, H2, std=math.sqrt(0.1))[0]
    w3 = philox.normal_matrix(seed, philox.STREAM_W3, 1, H2, std=math.sqrt(1.0 / H2))[0]
    return W1, b1, W2, b2, w3


def tanh_mlp2(n: int, H1: int, H2: int, seed: int) -> BytecodeTape:
    W1, b1, W2, b2, w3 = mlp2_weights(n, H1, H2, seed)
    tape, _ = record(lambda x: _readout(_layer(_layer(x, W1, b1), W2, b2), w3), [0.0] * n)
    return tape


def softplus_ridge(n: int, H: int, seed: int) -> BytecodeTape:
    """f(x) = (1/H) sum_j softplus(w_j . x + b_j): a convex function whose
    Hessian (1/H) sum_j s_j(x) w_j w_j^T has rank <= H -- a tape a rank-H
    learned control variate can match exactly."""
    W, b, _ = mlp_weights(n, H, seed)

    def f(x):
        acc = BReverse.constant(0.0)
        for j in range(H):
            s = x[0] * float(W[j, 0])
            for i in range(1, n):
                s = s + x[i] * float(W[j, i])
            acc = acc + _softplus(s + float(b[j])) * (1.0 / H)
        return acc

    tape, _ = record(f, [0.0] * n)
    return tape


def points_uniform(seed: int, B: int, n: int, lo: float, hi: float) -> np.ndarray:
    idx = np.arange(B * n, dtype=np.uint64)
    return philox.uniform_range(seed, philox.STREAM_POINTS, idx, lo, hi).reshape(B, n)


def points_normal(seed: int, B: int, n: int, stream: int = philox.STREAM_POINTS) -> np.ndarray:
    idx = np.arange(B * n, dtype=np.uint64)
    return philox.normal(seed, stream, idx).reshape(B, n)


# ---------------------------------------------------------------------------
# Random expression DAGs (uneven multi-tape batching workload)
# ---------------------------------------------------------------------------

_BIN = ("add", "sub", "mul", "div")
_TRANS = ("sin", "cos", "exp", "log", "tanh", "sqrt")
_MISC = ("powi", "abs", "max")


def _softplus(x: BReverse) -> BReverse:
    # log(1 + exp(x)) recorded as Exp, Add(Const 1), Ln -> strictly positive
    return (x.exp() + 1.0).ln()


def random_dag(seed: int, index: int, length: int, n: int) -> BytecodeTape:
    """One random DAG: ``length`` recorded ops, fan-in 2 to random earlier values.

    Op mix: 60% add/sub/mul/div, 30% sin/cos/exp/log/tanh/sqrt (log and sqrt
    arguments pass through softplus), 10% powi/abs/max.  The output is the
    sum of the last few values so every branch stays live.
    """
    key = seed * 1000003 + index
    # Pre-draw enough randomness (4 words per op).
    count = length * 4 + 16
    w = philox._words(key, philox.STREAM_DAG, np.arange(count, dtype=np.uint64))
    words = np.stack(w, axis=1).reshape(-1)
    pos = [0]

    def nxt() -> int:
        v = int(words[pos[0] % len(words)])
        pos[0] += 1
        return v

    def f(x):
        vals: List[BReverse] = list(x)
        for _ in range(length):
            r = nxt() % 100
            a = vals[nxt() % len(vals)]
            b = vals[nxt() % len(vals)]
            sel = nxt()
            if r < 60:
                op = _BIN[sel % 4]
                if op == "add":
                    v = a + b
                elif op == "sub":
                    v = a - b
                elif op == "mul":
                    v = a * b
                else:
                    v = a / (b * b + 1.0)
            elif r < 90:
                op = _TRANS[sel % 6]
                if op == "sin":
                    v = a.sin()
                elif op == "cos":
                    v = a.cos()
                elif op == "exp":
                    v = (a * 0.25).tanh().exp()
                elif op == "log":
                    v = _softplus(a).ln()
                elif op == "tanh":
                    v = a.tanh()
                else:
                    v = _softplus(a).sqrt()
            else:
                op = _MISC[sel % 3]
                if op == "powi":
                    v = a.tanh().powi(2 + sel % 3)
                elif op == "abs":
                    v = a.abs()
                else:
                    v = a.max(b)
            vals.append(v)
        acc = vals[-1]
        for v in vals[-8:-1]:
            acc = acc + v
        return acc

    tape, _ = record(f, [0.1] * n)
    return tape


def random_dag_suite(seed: int = 3, count: int = 64, scale: float = 1.0):
    """The uneven multi-tape workload: ``count`` DAGs with their point counts.

    Length log-uniform in [200, 20000] ops, n uniform in [4, 128],
    B_i log-uniform in [1e3, 1e5]. ``scale`` shrinks lengths and point
    counts for quick tests.
    """
    idx = np.arange(count * 3, dtype=np.uint64)
    u = philox.uniform(seed, philox.STREAM_DAG + 100, idx).reshape(count, 3)
    out = []
    for k in range(count):
        length = int(round(math.exp(math.log(200) + (math.log(20000) - math.log(200)) * u[k, 0]) * scale))
        length = max(length, 8)
        n = 4 + int(u[k, 1] * 125)
        B = int(round(math.exp(math.log(1e3) + (math.log(1e5) - math.log(1e3)) * u[k, 2]) * scale))
        B = max(B, 1)
        out.append((length, n, B))
    return out
