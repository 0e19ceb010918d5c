"""The reference's GPU edge-case regressions on the HIP backend.

* Every case of tests/edge_cases.py (ported from tests/phase7_gpu_fixes.rs,
  tests/gpu_cuda_kernel_parity.rs and tests/gpu_stde.rs:1853-2011) runs
  through the C ABI in f32 and f64 and must meet the reference's expected
  answer; f64 results must also equal the CPU oracle's (NaN for NaN, the
  same infinity, finite values within 1e-12 relative).
* The unary-op x edge-input grid of tests/mode_parity_matrix.rs: gradient
  (reverse sweep) and order-3 jets of each of 29 unary ops at
  0, -0, +-1, 0.5, 2, +-Inf, NaN, 1e-300, 1e300.  f64 against the oracle
  with the bar above; f32 (inputs rounded to f32) against the f32 oracle:
  the same NaN and infinity positions and finite values within 64 ulp
  (the reference's per-op budgets are at most 32 ulp; subnormal results may
  differ by their last bits).
"""

import numpy as np
import pytest

import pyoracle as po
from edge_cases import CASES, EDGES, UNARY, empty_outputs_data, unary_tape
from echidna_amd.tape import GpuTapeData
from parity_cases import ulp_diff_f32
from test_edge_cases import run_oracle

pytestmark = pytest.mark.gpu


def run_gpu(hip, kind, bufs, x, v, K, f64):
    dt = np.float64 if f64 else np.float32
    X = np.asarray(x, dt)
    if kind == "forward":
        return hip.forward_batch(bufs, X, 1)
    if kind == "gradient":
        return hip.gradient_batch(bufs, X, 1)
    if kind == "hvp":
        return hip.hvp_batch(bufs, X, np.asarray(v, dt), 1)
    r = hip.taylor_forward_kth_batch(bufs, X, np.asarray(v, dt), 1, K)
    return np.array([c[0] for c in r.coefficients])


def same64(got, ref):
    got = np.asarray(got, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    nan = np.isnan(ref)
    if not np.array_equal(np.isnan(got), nan):
        return False
    inf = np.isinf(ref)
    if not np.array_equal(got[inf], ref[inf]) or np.isinf(got[~inf & ~nan]).any():
        return False
    fin = ~nan & ~inf
    return bool(np.all(np.abs(got[fin] - ref[fin]) <= 1e-12 * np.abs(ref[fin])))


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_gpu_edge_case(hip, case, dt):
    name, build, kind, x, v, K, check = case
    f64 = dt == "f64"
    if build == "empty_outputs":
        tape, data = empty_outputs_data(f64)
        assert len(data.output_indices) == 0
    else:
        tape = build()
        data = GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape)
    bufs = hip.upload_tape(data)
    r = run_gpu(hip, kind, bufs, x, v, K, f64)
    check(r, dt)
    if f64:
        ref = run_oracle(kind, tape, x, v, K, np.float64)
        parts = zip(r, ref) if isinstance(r, tuple) else [(r, ref)]
        for got, want in parts:
            assert same64(got, want), (name, got, want)


def _grid(hip, name, f64):
    tape = unary_tape(name)
    dt = np.float64 if f64 else np.float32
    X = np.asarray(EDGES, dt)[:, None]
    B = len(EDGES)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    y, g = hip.gradient_batch(bufs, X, B)
    jr = hip.taylor_forward_kth_batch(bufs, X, np.ones_like(X), B, 3)
    jets = np.stack(jr.coefficients, axis=1)
    yr, gr = po.gradient(tape, X, dtype=dt)
    jref = po.taylor(tape, X, np.ones_like(X), K=3, dtype=dt)[:, 0, :]
    return (y, yr[:, 0]), (g, gr[:, 0]), (jets, jref)


@pytest.mark.parametrize("name", UNARY)
def test_gpu_unary_edge_grid_f64(hip, name):
    for what, (got, ref) in zip(("value", "gradient", "jet"), _grid(hip, name, True)):
        assert same64(got, ref), (name, what, got, ref)


def _close32(got, ref):
    got = np.asarray(got, np.float32).ravel()
    ref = np.asarray(ref, np.float32).ravel()
    for a, b in zip(got, ref):
        if np.isnan(b) or np.isnan(a):
            if not (np.isnan(a) and np.isnan(b)):
                return False
        elif np.isinf(b) or np.isinf(a):
            if a != b:
                return False
        elif ulp_diff_f32(a, b) > 64:
            return False
    return True


@pytest.mark.parametrize("name", UNARY)
def test_gpu_unary_edge_grid_f32(hip, name):
    for what, (got, ref) in zip(("value", "gradient", "jet"), _grid(hip, name, False)):
        assert _close32(got, ref), (name, what, got, ref)


@pytest.mark.parametrize("spec", [False, True], ids=["interpreter", "specialised"])
def test_gpu_tanh_f64_ulp(hip, spec):
    """The binary64 tanh and 1/cosh^2 (one exp reduction shared by both)
    against the correctly rounded values (200-bit mpmath): within 2 and 3
    ulp over 4,096 arguments in [-20, 20], tiny and large arguments; the
    specialised kernel's called pair gives the interpreter's bits."""
    mpmath = pytest.importorskip("mpmath")
    from echidna_amd.backend import TUNE_GRADIENT, HipContext

    tape = unary_tape("tanh")
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    rng = np.random.default_rng(11)
    X = np.concatenate([rng.uniform(-20, 20, 3500), rng.uniform(-1e-3, 1e-3, 300),
                        np.sign(rng.uniform(-1, 1, 296)) * 10.0 ** rng.uniform(-300, 1.3, 296)])
    if spec:
        hip.specialize(bufs, TUNE_GRADIENT)
        HipContext.set_tuned(bufs, TUNE_GRADIENT, 256, 0, fetch=2)
    y, g = hip.gradient_batch(bufs, X, len(X))
    y = np.asarray(y, np.float64).ravel()
    g = np.asarray(g, np.float64).ravel()
    mpmath.mp.prec = 200
    worst_t = worst_d = 0.0
    for a, t, d in zip(X, y, g):
        T = mpmath.tanh(mpmath.mpf(float(a)))
        D = 1 / mpmath.cosh(mpmath.mpf(float(a))) ** 2
        worst_t = max(worst_t, float(abs(mpmath.mpf(float(t)) - T)) / np.spacing(abs(float(T))))
        worst_d = max(worst_d, float(abs(mpmath.mpf(float(d)) - D)) / np.spacing(abs(float(D))))
    print(f"tanh f64 ({'specialised' if spec else 'interpreter'}): worst {worst_t:.2f} ulp, 1/cosh^2 {worst_d:.2f} ulp")
    assert worst_t <= 2.0 and worst_d <= 3.0
    if spec:
        HipContext.set_tuned(bufs, TUNE_GRADIENT, 256, 0, fetch=0)
        y0, g0 = hip.gradient_batch(bufs, X, len(X))
        assert np.array_equal(np.asarray(y0, np.float64).ravel().view(np.uint64), y.view(np.uint64))
        assert np.array_equal(np.asarray(g0, np.float64).ravel().view(np.uint64), g.view(np.uint64))
