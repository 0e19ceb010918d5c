"""GPU STDE against the oracle's estimators (src/stde/laplacian.rs) and known answers.

Bars: f64 estimates and variances within 1e-10 relative of the oracle (the
device reduces in binary64 with a different summation order: lanes run
Welford, then Chan merges); sample counts exact.  Device-drawn directions are
regenerated on the host from the same Philox stream: Rademacher signs are
bit-exact, Gaussian draws agree to a few ulps (device vs host log/cos).
"""

import itertools

import numpy as np
import pytest

import pyoracle as po
from echidna_amd import philox
from echidna_amd import stde_gpu as sg
from echidna_amd import workloads as wl
from echidna_amd.tape import GpuError, GpuTapeData, record, record_multi

pytestmark = pytest.mark.gpu


def close(a, b, rtol=1e-10, atol=1e-12):
    return abs(a - b) <= rtol * max(abs(a), abs(b)) + atol


def test_laplacian_known_answer(hip):
    tape, _ = record(lambda v: v[0] * v[0] * v[1] + v[1] * v[1] * v[1], [1.0, 2.0, 3.0])
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    dirs = np.array(list(itertools.product([-1.0, 1.0], repeat=3)))
    r = sg.laplacian_gpu(hip, bufs, [1.0, 2.0, 3.0], dirs)
    assert r.estimate == 16.0 and r.value == 10.0 and r.num_samples == 8
    val, diag = sg.hessian_diagonal_gpu(hip, bufs, [1.0, 2.0, 3.0])
    assert val == 10.0 and list(diag) == [4.0, 12.0, 0.0]


def test_laplacian_matches_oracle_f64(hip):
    tape = wl.tanh_mlp(16, 12, 21)
    x = wl.points_normal(2, 1, 16)[0]
    dirs = philox.normal(5, philox.STREAM_DIRS, np.arange(200 * 16)).reshape(200, 16)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    r = sg.laplacian_gpu(hip, bufs, x, dirs)
    ref = po.laplacian_with_stats(tape, x, dirs)
    assert r.num_samples == ref["num_samples"] == 200
    assert close(r.value, ref["value"], 1e-13)
    assert close(r.estimate, ref["estimate"])
    assert close(r.sample_variance, ref["sample_variance"])
    assert close(r.standard_error, ref["standard_error"])


def test_laplacian_with_control_matches_oracle(hip):
    tape = wl.tanh_mlp(10, 8, 22)
    x = wl.points_normal(3, 1, 10)[0]
    dirs = philox.rademacher(6, philox.STREAM_DIRS, np.arange(128 * 10)).reshape(128, 10)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    _, diag = sg.hessian_diagonal_gpu(hip, bufs, x)
    _, diag_ref = po.hessian_diagonal(tape, x)
    assert np.allclose(diag, diag_ref, rtol=1e-12, atol=1e-14)
    r = sg.laplacian_with_control_gpu(hip, bufs, x, dirs, diag_ref)
    ref = po.laplacian_with_control(tape, x, dirs, diag_ref)
    assert close(r.estimate, ref["estimate"])
    assert close(r.sample_variance, ref["sample_variance"], 1e-9)
    # Rademacher v_j^2 = 1, so the diagonal control shifts every sample by the
    # same constant: the variance is unchanged and the estimate agrees
    plain = sg.laplacian_gpu(hip, bufs, x, dirs)
    assert close(r.sample_variance, plain.sample_variance, 1e-9)
    assert close(r.estimate, plain.estimate, 1e-9)
    # with Gaussian directions the exact diagonal removes the diagonal part of
    # the variance (2 sum_j H_jj^2 of 2 ||H||_F^2)
    g = philox.normal(6, philox.STREAM_DIRS, np.arange(512 * 10)).reshape(512, 10)
    rg = sg.laplacian_with_control_gpu(hip, bufs, x, g, diag_ref)
    pg = sg.laplacian_gpu(hip, bufs, x, g)
    refg = po.laplacian_with_control(tape, x, g, diag_ref)
    assert close(rg.sample_variance, refg["sample_variance"], 1e-9)
    assert rg.sample_variance < pg.sample_variance


@pytest.mark.parametrize("kind", [sg.DIR_RADEMACHER, sg.DIR_GAUSSIAN])
def test_batched_device_directions(hip, kind):
    tape = wl.tanh_mlp(12, 10, 23)
    P, S, n = 40, 48, 12
    X = wl.points_normal(4, P, n)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    res = sg.laplacian_batch(hip, bufs, X, S, kind=kind, seed=77)
    idx = np.arange(P * S * n, dtype=np.uint64)
    dirs = philox.direction(kind, 77, idx).reshape(P, S, n)
    for p in (0, 13, P - 1):
        ref = po.laplacian_with_stats(tape, X[p], dirs[p])
        tol = 1e-10 if kind == sg.DIR_RADEMACHER else 1e-9
        assert int(res[p, 4]) == S
        assert close(res[p, 1], ref["estimate"], tol), (p, res[p], ref)
        assert close(res[p, 2], ref["sample_variance"], tol)


def test_batched_supplied_directions_f32(hip):
    tape = wl.tanh_mlp(8, 6, 24)
    P, S, n = 10, 32, 8
    X = wl.points_normal(5, P, n).astype(np.float32)
    dirs = philox.rademacher(9, philox.STREAM_DIRS, np.arange(P * S * n)).reshape(P, S, n).astype(np.float32)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    res = sg.laplacian_batch(hip, bufs, X, S, kind=sg.DIR_SUPPLIED, directions=dirs)
    for p in range(P):
        ref = po.laplacian_with_stats(tape, X[p].astype(np.float64), dirs[p].astype(np.float64))
        scale = max(abs(ref["estimate"]), np.sqrt(ref["sample_variance"]), 1e-3)
        assert abs(res[p, 1] - ref["estimate"]) <= 1e-4 * scale


def test_estimator_is_unbiased_on_a_quadratic(hip):
    """sum_i a_i x_i^2 + x_0 x_1: Laplacian = 2 sum a_i exactly; Rademacher samples are exact too."""
    a = [0.5, 1.5, -2.0, 3.0]

    def f(v):
        s = v[0] * v[1]
        for i, ai in enumerate(a):
            s = s + v[i] * v[i] * ai
        return s

    tape, _ = record(f, [0.1] * 4)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    res = sg.laplacian_batch(hip, bufs, np.zeros((3, 4)), 64, kind=sg.DIR_RADEMACHER, seed=5)
    # v^T H v = 2 sum a_i + 2 v0 v1 (Rademacher): mean over samples -> 2 sum a_i
    for p in range(3):
        assert abs(res[p, 1] - 2 * sum(a)) <= 5 * res[p, 3] + 1e-12


def test_stde_errors(hip):
    tape, _ = record_multi(lambda v: [v[0] * v[1], v[0]], [1.0, 2.0])
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    with pytest.raises(GpuError, match="requires a scalar-output tape; got 2 outputs"):
        sg.laplacian_gpu(hip, bufs, [1.0, 2.0], np.ones((2, 2)))
    tape, _ = record(lambda v: v[0] * v[1], [1.0, 2.0])
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    with pytest.raises(GpuError, match="no directions provided"):
        sg.laplacian_gpu(hip, bufs, [1.0, 2.0], [])
