"""GPU sparse Jacobian / Hessian and kink records against the oracle.

Bars: f64 values within 1e-12 relative of the dense f64 oracle (Hessian from
n HVPs, Jacobian from n basis jets); f32 within 2e-5 of it.  Kink records:
switching values and branches bit-identical to the oracle's
forward_nonsmooth when the recorded operands come from exactly rounded
operations (+, -, *, /, abs, max, floor, ...); with transcendental operands,
switching values agree to the value tolerance and branches agree wherever
the switching value is not within that tolerance of the kink.
"""

import numpy as np
import pytest

import pyoracle as po
from echidna_amd import sparse as sp
from echidna_amd import workloads as wl
from echidna_amd.tape import GpuTapeData, record, record_multi
from test_structure import TAPES

pytestmark = pytest.mark.gpu


def dense_from(pattern_rows, pattern_cols, vals, n, sym=True):
    M = np.zeros((n, n))
    M[pattern_rows, pattern_cols] = vals
    if sym:
        M[pattern_cols, pattern_rows] = vals
    return M


@pytest.mark.parametrize("name", ["kitchen_sink", "rosenbrock", "mlp", "dag0"])
@pytest.mark.parametrize("f64", [True, False])
def test_sparse_hessian_matches_dense_oracle(hip, name, f64):
    tape = TAPES[name]()
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    rng = np.random.default_rng(7)
    x = rng.uniform(-0.8, 0.8, tape.num_inputs)
    if name == "kitchen_sink":
        x[3] = abs(x[3]) + 0.1  # powf base
    if not f64:
        x = x.astype(np.float32).astype(np.float64)
    val, grad, pat, vals = sp.sparse_hessian(hip, bufs, x)
    ref_p = sp.detect_sparsity(GpuTapeData.from_tape_f64(tape))
    assert np.array_equal(pat.rows, ref_p.rows) and np.array_equal(pat.cols, ref_p.cols)
    H = po.hessian(tape, x)
    y_ref, g_ref = po.gradient(tape, x[None, :])
    tol = 1e-12 if f64 else 2e-5
    assert np.all(np.isfinite(H)), "choose a point inside every op's domain"
    scale = max(1.0, np.max(np.abs(H)))
    assert np.max(np.abs(vals - H[pat.rows, pat.cols]), initial=0.0) <= tol * scale
    assert np.max(np.abs(grad - g_ref[0])) <= tol * max(1.0, np.max(np.abs(g_ref)))
    assert abs(val - y_ref[0, 0]) <= tol * max(1.0, abs(y_ref[0, 0]))
    # every Hessian entry outside the pattern is zero
    full = dense_from(pat.rows, pat.cols, np.ones(pat.nnz()), tape.num_inputs) > 0
    assert np.all(np.abs(H[~full]) <= 1e-12 * scale)


def vector_map(v):
    return [v[0] * v[1], v[1] * v[2].sin(), v[0] * v[0] + v[3].exp(), v[2] - v[4].tanh(), v[4] * v[3] * v[0]]


@pytest.mark.parametrize("f64", [True, False])
def test_sparse_jacobian_matches_dense_oracle(hip, f64):
    tape, _ = record_multi(vector_map, [0.1, 0.2, 0.3, 0.4, 0.5])
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    x = np.array([0.3, -0.7, 1.1, 0.2, -0.4])
    out, pat, vals = sp.sparse_jacobian(hip, bufs, x)
    n, m = 5, 5
    J = po.taylor(tape, np.tile(x, (n, 1)), np.eye(n), K=2)[:, :, 1].T  # [m, n]
    y = po.forward(tape, x[None, :])[0]
    tol = 1e-12 if f64 else 2e-5
    assert np.max(np.abs(vals - J[pat.rows, pat.cols])) <= tol * max(1.0, np.max(np.abs(J)))
    assert np.max(np.abs(out - y)) <= tol * max(1.0, np.max(np.abs(y)))
    mask = np.zeros((m, n), bool)
    mask[pat.rows, pat.cols] = True
    assert np.all(J[~mask] == 0.0)
    colors, k = sp.column_coloring(pat)
    assert k < n  # compression happened


def test_sparse_paths_without_interactions(hip):
    """A linear function has an empty Hessian pattern: value and gradient only."""
    tape, _ = record(lambda v: v[0] * 2.0 + v[1] - v[2] * 0.5, [1.0, 2.0, 3.0])
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    val, grad, pat, vals = sp.sparse_hessian(hip, bufs, [1.0, 2.0, 3.0])
    assert pat.nnz() == 0 and len(vals) == 0
    assert val == 2.0 + 2.0 - 1.5 and list(grad) == [2.0, 1.0, -0.5]


def exact_kinks(v):
    a = (v[0] - v[1]).abs()
    b = v[1].max(v[2] * 0.5)
    c = v[2].min(v[0])
    d = (v[3] * 3.0).floor() + (v[3] - 0.25).ceil() + (v[0] * 1.5).round() + (v[1] * 2.0).trunc()
    e = (v[2] * 4.0).fract() + v[3].signum()
    dead = (v[0] * v[2]).abs()  # never reaches the output: still recorded
    return a + b * c + d - e


@pytest.mark.parametrize("f64", [True, False])
def test_kinks_bit_identical_on_exact_operands(hip, f64):
    tape, _ = record(exact_kinks, [0.1, 0.2, 0.3, 0.4])
    F = np.float64 if f64 else np.float32
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    rng = np.random.default_rng(3)
    B = 333
    X = rng.uniform(-2, 2, (B, 4))
    X[:40] = np.round(X[:40] * 2) / 2  # operands exactly on kinks and half-integers
    X[40:45, 1] = X[40:45, 0]
    X = X.astype(F)
    res = sp.forward_nonsmooth(hip, bufs, X.reshape(-1), B)
    y, sw, br = po.nonsmooth(tape, X, dtype=F)
    assert len(res.tape_index) == 10 and res.switching.shape == (B, 10)
    ops = [int(tape.opcodes[i]) for i in res.tape_index]
    assert ops == [int(o) for o in res.opcode]
    assert np.array_equal(res.switching.view(np.uint8), sw.view(np.uint8))
    assert np.array_equal(res.branch, br)
    assert np.array_equal(res.outputs.view(np.uint8), y.view(np.uint8))


def test_kinks_on_transcendental_operands(hip):
    def f(v):
        return (v[0].sin() - v[1].cos()).abs() + v[2].exp().max(v[0] * v[1] + 1.0) + v[1].tanh().floor()

    tape, _ = record(f, [0.1, 0.2, 0.3])
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    X = np.random.default_rng(5).uniform(-2, 2, (500, 3))
    res = sp.forward_nonsmooth(hip, bufs, X.reshape(-1), 500)
    _, sw, br = po.nonsmooth(tape, X)
    assert np.max(np.abs(res.switching - sw)) <= 1e-12 * max(1.0, np.max(np.abs(sw)))
    far = np.abs(sw) > 1e-10
    assert np.array_equal(res.branch[far], br[far])


def test_tape_without_kinks(hip):
    tape, _ = record(lambda v: v[0] * v[1], [1.0, 2.0])
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    res = sp.forward_nonsmooth(hip, bufs, [1.0, 2.0, 3.0, 4.0], 2)
    assert res.switching.shape == (2, 0) and list(res.outputs[:, 0]) == [2.0, 12.0]
