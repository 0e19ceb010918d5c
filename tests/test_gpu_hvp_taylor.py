"""GPU parity for forward-over-reverse HVPs and Taylor jets against the oracle.

Bars: f64 |gpu - cpu| <= 1e-12 * row scale (see test_gpu_parity.rowrel);
f32 against the f64 oracle <= 2e-5 * row scale for gradients and HVPs and
<= 1e-4 * row scale for second-order jet coefficients (c2 carries one more
rounding chain than c1).
"""

import itertools

import numpy as np
import pytest

import pyoracle as po
from echidna_amd import workloads as wl
from echidna_amd.tape import GpuTapeData, record
from parity_cases import CASES
from test_gpu_parity import F32_TOL, F64_TOL, rowrel

pytestmark = pytest.mark.gpu

JET_F32_TOL = 1e-4


def test_hvp_single_point_matches_reference_semantics(hip):
    """hvp_batch: one x, B directions (src/gpu/mod.rs:149-160)."""
    tape = wl.tanh_mlp(32, 64, 1)
    x = wl.points_normal(1, 1, 32)[0]
    V = wl.points_normal(7, 300, 32)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    g, h = hip.hvp_batch(bufs, x, V, 300)
    g_ref, h_ref = po.hvp(tape, x, V)
    assert rowrel(g.reshape(300, 32), g_ref) <= F64_TOL
    assert rowrel(h.reshape(300, 32), h_ref) <= F64_TOL


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
    g_ref, h_ref = po.hvp(tape, X, V)
    assert rowrel(g.reshape(500, 32), g_ref) <= F64_TOL
    assert rowrel(h.reshape(500, 32), h_ref) <= F64_TOL


def test_hvp_points_f32(hip):
    tape = wl.tanh_mlp(32, 64, 1)
    X = wl.points_normal(1, 500, 32).astype(np.float32)
    V = wl.points_normal(7, 500, 32).astype(np.float32)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    g, h = hip.hvp_batch_points(bufs, X, V, 500)
    g_ref, h_ref = po.hvp(tape, X.astype(np.float64), V.astype(np.float64))
    assert rowrel(g.reshape(500, 32), g_ref) <= F32_TOL
    assert rowrel(h.reshape(500, 32), h_ref) <= F32_TOL


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
    g_ref, h_ref = po.hvp(tape, X, V)
    for got, ref in ((g.reshape(X.shape), g_ref), (h.reshape(X.shape), h_ref)):
        both_nan = np.isnan(got) & np.isnan(ref)
        same_inf = np.isinf(ref) & (got == ref)
        ok = both_nan | same_inf | (np.abs(got - ref) <= 1e-11 * np.maximum(1.0, np.abs(ref)))
        assert ok.all(), (name, X, got, ref)


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
    ref = po.taylor(tape, X, V, K=3)[:, 0, :]
    got = np.stack([r.values, r.c1s, r.c2s], axis=1)
    assert rowrel(got, ref) <= F64_TOL


def test_taylor2_mlp_f32(hip):
    tape = wl.tanh_mlp(32, 64, 1)
    X = wl.points_normal(1, 400, 32).astype(np.float32)
    V = wl.points_normal(9, 400, 32).astype(np.float32)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    r = hip.taylor_forward_2nd_batch(bufs, X, V, 400)
    ref = po.taylor(tape, X.astype(np.float64), V.astype(np.float64), K=3)[:, 0, :]
    got = np.stack([r.values, r.c1s, r.c2s], axis=1)
    assert rowrel(got, ref) <= JET_F32_TOL


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_taylor_parity_table_f64(hip, case):
    name, build, points, _, _ = case
    tape = build()
    X = np.asarray(points, dtype=np.float64)
    V = np.full_like(X, 0.5)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    for K in (3, 5):
        r = hip.taylor_forward_kth_batch(bufs, X, V, X.shape[0], K)
        got = np.stack(r.coefficients, axis=1)
        ref = po.taylor(tape, X, V, K=K)[:, 0, :]
        both_nan = np.isnan(got) & np.isnan(ref)
        same_inf = np.isinf(ref) & (got == ref)
        ok = both_nan | same_inf | (np.abs(got - ref) <= 1e-11 * np.maximum(1.0, np.abs(ref)))
        assert ok.all(), (name, K, X, got, ref)


def test_laplacian_known_answer_all_rademacher(hip):
    """x^2 y + y^3 at (1,2,3): Laplacian 16 from all 8 Rademacher vectors (tests/stde_core.rs:61-110)."""
    tape, _ = record(lambda v: v[0] * v[0] * v[1] + v[1] * v[1] * v[1], [1.0, 2.0, 3.0])
    dirs = np.array(list(itertools.product([-1.0, 1.0], repeat=3)))
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    r = hip.taylor_forward_2nd_batch(bufs, np.tile([1.0, 2.0, 3.0], (8, 1)), dirs, 8)
    assert np.mean(2 * r.c2s) == 16.0
    assert r.values[0] == 10.0
