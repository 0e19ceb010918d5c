"""GPU parity for forward-over-reverse HVPs and Taylor jets against the oracle.

Bars: f64 elementwise |gpu - cpu| <= 1e-12 * max(|cpu|, M), M the sum of
the absolute terms behind the entry (tests/parity_bar.py: gradients, H v,
and jet coefficients c0..c2); jet coefficients c3 and c4 (orders 4 and 5)
within 1e-12 * max(|cpu|, the largest magnitude of c0..c2); f32
against the f64 oracle: the same form with 2e-5 for gradients and HVPs and
1e-4 for second-order jet coefficients (c2 carries one more rounding chain
than c1).
"""

import itertools

import numpy as np
import pytest

import pyoracle as po
from echidna_amd import workloads as wl
from echidna_amd.tape import GpuTapeData, record
from parity_cases import CASES
from parity_bar import assert_f64
from test_gpu_parity import F32_TOL, F64_TOL

pytestmark = pytest.mark.gpu

JET_F32_TOL = 1e-4


def check_hvp(tape, X, V, g, h, tol=F64_TOL):
    X = np.asarray(X, np.float64)
    V = np.asarray(V, np.float64)
    g_ref, h_ref = po.hvp(tape, X, V)
    mag = po.magnitudes(tape, X, V)
    assert_f64(np.asarray(g).reshape(g_ref.shape), g_ref, mag["grad"], "gradient", tol)
    assert_f64(np.asarray(h).reshape(h_ref.shape), h_ref, mag["hv"], "hvp", tol)


def check_jets(tape, X, V, got, K, tol=F64_TOL):
    """got [B, K] for output 0; c0..c2 at tol against their magnitudes;
    c3, c4 at tol relative to max(|c_k|, the largest magnitude of c0..c2)."""
    X = np.asarray(X, np.float64)
    V = np.asarray(V, np.float64)
    ref = po.taylor(tape, X, V, K=K)[:, 0, :]
    mag = po.magnitudes(tape, X, V)
    M = np.stack([mag["value"][:, 0], mag["c1"], mag["c2"]], axis=1)
    k = min(K, 3)
    assert_f64(got[:, :k], ref[:, :k], M[:, :k], "c0..c2", tol)
    if K > 3:
        floor = np.repeat(np.max(np.where(np.isfinite(M), M, 0.0), axis=1, keepdims=True), K - 3, axis=1)
        assert_f64(got[:, 3:], ref[:, 3:], floor, "c3..", tol)


def test_hvp_single_point_matches_reference_semantics(hip):
    """hvp_batch: one x, B directions (src/gpu/mod.rs:149-160)."""
    tape = wl.tanh_mlp(32, 64, 1)
    x = wl.points_normal(1, 1, 32)[0]
    V = wl.points_normal(7, 300, 32)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    g, h = hip.hvp_batch(bufs, x, V, 300)
    check_hvp(tape, np.tile(x, (300, 1)), V, g, h)


@pytest.mark.parametrize("fuse", [True, False])
def test_hvp_points_f64(hip, fuse):
    tape = wl.tanh_mlp(32, 64, 1)
    X = wl.points_normal(1, 500, 32)
    V = wl.points_normal(7, 500, 32)
    hip.set_fusion(fuse)
    try:
        bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    finally:
        hip.set_fusion(True)
    g, h = hip.hvp_batch_points(bufs, X, V, 500)
    check_hvp(tape, X, V, g, h)


def test_hvp_points_f32(hip):
    tape = wl.tanh_mlp(32, 64, 1)
    X = wl.points_normal(1, 500, 32).astype(np.float32)
    V = wl.points_normal(7, 500, 32).astype(np.float32)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    g, h = hip.hvp_batch_points(bufs, X, V, 500)
    check_hvp(tape, X, V, g, h, F32_TOL)


def test_hvp_rosenbrock_analytic(hip):
    """Rosenbrock n=2 Hessian in closed form (tests/hessian.rs:247)."""
    tape = wl.rosenbrock(2)
    x = np.array([-1.2, 1.0])
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    _, h = hip.hvp_batch(bufs, x, np.eye(2), 2)
    H = np.array([[1200 * x[0] ** 2 - 400 * x[1] + 2, -400 * x[0]], [-400 * x[0], 200.0]])
    assert np.allclose(h.reshape(2, 2), H, rtol=1e-14, atol=1e-12)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_hvp_parity_table_f64(hip, case):
    name, build, points, _, _ = case
    tape = build()
    X = np.asarray(points, dtype=np.float64)
    V = np.ones_like(X) * 0.75
    V[:, 0] = 1.25
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    g, h = hip.hvp_batch_points(bufs, X, V, X.shape[0])
    check_hvp(tape, X, V, g, h)


# ---------------------------------------------------------------------------
# Taylor jets
# ---------------------------------------------------------------------------

def test_taylor_closed_form_orders_1_to_5(hip):
    """x^2 + y^2 at (3,4) along (1,0): c = [25, 6, 1, 0, 0] (tests/gpu_cuda_tests.rs:327-372)."""
    tape, _ = record(lambda v: v[0] * v[0] + v[1] * v[1], [3.0, 4.0])
    for f64 in (True, False):
        bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
        for order in range(1, 6):
            r = hip.taylor_forward_kth_batch(bufs, [3.0, 4.0], [1.0, 0.0], 1, order)
            got = [float(c[0]) for c in r.coefficients]
            assert got == [25.0, 6.0, 1.0, 0.0, 0.0][:order], (f64, order, got)


def test_taylor2_mlp_f64(hip):
    tape = wl.tanh_mlp(32, 64, 1)
    X = wl.points_normal(1, 400, 32)
    V = wl.points_normal(9, 400, 32)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    r = hip.taylor_forward_2nd_batch(bufs, X, V, 400)
    check_jets(tape, X, V, np.stack([r.values, r.c1s, r.c2s], axis=1), 3)


def test_taylor2_mlp_f32(hip):
    tape = wl.tanh_mlp(32, 64, 1)
    X = wl.points_normal(1, 400, 32).astype(np.float32)
    V = wl.points_normal(9, 400, 32).astype(np.float32)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    r = hip.taylor_forward_2nd_batch(bufs, X, V, 400)
    check_jets(tape, X, V, np.stack([r.values, r.c1s, r.c2s], axis=1), 3, JET_F32_TOL)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_taylor_parity_table_f64(hip, case):
    name, build, points, _, _ = case
    tape = build()
    X = np.asarray(points, dtype=np.float64)
    V = np.full_like(X, 0.5)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    for K in (3, 5):
        r = hip.taylor_forward_kth_batch(bufs, X, V, X.shape[0], K)
        check_jets(tape, X, V, np.stack(r.coefficients, axis=1), K)


def test_laplacian_known_answer_all_rademacher(hip):
    """x^2 y + y^3 at (1,2,3): Laplacian 16 from all 8 Rademacher vectors (tests/stde_core.rs:61-110)."""
    tape, _ = record(lambda v: v[0] * v[0] * v[1] + v[1] * v[1] * v[1], [1.0, 2.0, 3.0])
    dirs = np.array(list(itertools.product([-1.0, 1.0], repeat=3)))
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    r = hip.taylor_forward_2nd_batch(bufs, np.tile([1.0, 2.0, 3.0], (8, 1)), dirs, 8)
    assert np.mean(2 * r.c2s) == 16.0
    assert r.values[0] == 10.0
