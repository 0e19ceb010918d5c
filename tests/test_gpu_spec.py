"""Tape-specialised kernels (straight-line code per tape, compiled by hipRTC).

They run the interpreters' arithmetic in the interpreters' order, so values
and gradients equal the interpreter's bit for bit, except the sign bit of
NaN results, which IEEE 754 leaves unspecified; and they match the f64
oracle within 1e-12 relative (f64) like every other path.  Checked on
tapes that use every opcode, on the guard tape (zero, infinite and NaN
adjoints), on random DAGs and on a multi-output map; through the autotuner,
hipGraph replays and ragged batches.
"""

import numpy as np
import pytest
import torch

import pyoracle as po
from parity_bar import assert_f64
from echidna_amd import workloads as wl
from echidna_amd.backend import SPEC_GRAD_ACC64, TUNE_FORWARD, TUNE_GRADIENT, TUNE_TAYLOR, HipContext
from echidna_amd.tape import GpuError, GpuTapeData, record, record_multi
from test_structure import TAPES

pytestmark = pytest.mark.gpu


def _bits(a, f64):
    """Bits with the sign of NaNs cleared."""
    U = np.uint64 if f64 else np.uint32
    b = np.asarray(a).view(U).copy()
    b[np.isnan(a)] &= ~(U(1) << U(63 if f64 else 31))
    return b


def _points(name, B, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.2, 1.2, (B, n))
    if name == "guards":
        import dense_cases as dc

        x = dc.points(B, seed)
    x[:3] = 0.0  # zero rows: zero adjoints and kinks
    return x


def _tape(name):
    if name == "guards":
        import dense_cases as dc

        return dc.tape()
    return TAPES[name]()


def _both(hip, bufs, x, B, what):
    """(interpreter, specialised) results of forward (0) or gradient (1)."""
    outs = []
    for fetch in (0, 2):
        if fetch == 2:
            HipContext.set_tuned(bufs, what, 128, 0, fetch=2)
        else:
            HipContext.set_tuned(bufs, what, 256, 24576, fetch=1)
        if what == TUNE_GRADIENT:
            y, g = hip.gradient_batch(bufs, x.reshape(-1), B)
            outs.append((np.asarray(y).copy(), np.asarray(g).copy()))
        else:
            outs.append((np.asarray(hip.forward_batch(bufs, x.reshape(-1), B)).copy(),))
    return outs


@pytest.mark.parametrize("name", ["kitchen_sink", "rosenbrock", "mlp", "dag0", "dag1", "guards"])
@pytest.mark.parametrize("f64", [False, True])
def test_specialised_equals_interpreter_and_oracle(hip, name, f64):
    tape = _tape(name)
    n = tape.num_inputs
    F = np.float64 if f64 else np.float32
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    for what in (TUNE_FORWARD, TUNE_GRADIENT):
        info = HipContext.specialize(bufs, what)
        assert info["vgprs"] > 0 and info["code_bytes"] > 0
    B = 1000  # not a multiple of any workgroup size
    x = _points(name, B, n, 21).astype(F)
    with np.errstate(all="ignore"):
        for what in (TUNE_FORWARD, TUNE_GRADIENT):
            interp, spec = _both(hip, bufs, x, B, what)
            for a, b in zip(interp, spec):
                assert np.array_equal(_bits(a, f64), _bits(b, f64)), (name, what)
        if f64:
            y, g = spec[0].reshape(B), spec[1].reshape(B, n)
            y_ref, g_ref = po.gradient(tape, x)
            for got, ref in ((y, y_ref[:, 0]), (g, g_ref)):
                assert np.array_equal(np.isnan(got), np.isnan(ref))
                fin = np.isfinite(ref)
                assert np.array_equal(got[~fin & ~np.isnan(ref)], ref[~fin & ~np.isnan(ref)])
                assert np.max(np.abs(got[fin] - ref[fin]) / np.maximum(1.0, np.abs(ref[fin])),
                              initial=0.0) <= 1e-12


def vector_map(v):
    return [v[0] * v[1] + v[2].sin(), v[1] * 2.5 - v[0], v[2].exp() * v[0], v[1]]


@pytest.mark.parametrize("f64", [False, True])
def test_specialised_multi_output(hip, f64):
    """All outputs are written; the gradient is output 0's (the reference
    seeds output_indices[0]); an output that is an input passes through."""
    tape, _ = record_multi(vector_map, [0.1, 0.2, 0.3])
    F = np.float64 if f64 else np.float32
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    HipContext.specialize(bufs, TUNE_GRADIENT)
    HipContext.specialize(bufs, TUNE_FORWARD)
    B = 257
    x = np.random.default_rng(4).uniform(-2, 2, (B, 3)).astype(F)
    for what in (TUNE_FORWARD, TUNE_GRADIENT):
        interp, spec = _both(hip, bufs, x, B, what)
        for a, b in zip(interp, spec):
            assert np.array_equal(_bits(a, f64), _bits(b, f64))
    y = spec[0].reshape(B, 4)
    assert np.array_equal(y[:, 3], x[:, 1])
    g = spec[1].reshape(B, 3)
    assert np.array_equal(g[:, 1], x[:, 0].astype(F))  # d(x0 x1 + sin x2)/dx1


def test_specialised_autotune_graph_and_ragged_batches(hip):
    """The autotuner may pick the specialised kernel; whichever it picks, the
    results stay equal to the interpreter's, also replayed from a graph and
    for batches of 1 and 65 points."""
    tape = wl.tanh_mlp(16, 32, 32)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    B = 1 << 14
    xn = wl.points_normal(10, B, 16).astype(np.float32)
    HipContext.set_tuned(bufs, TUNE_GRADIENT, 256, 24576, fetch=1)
    y0, g0 = hip.gradient_batch(bufs, xn.reshape(-1), B)
    info = HipContext.specialize(bufs, TUNE_GRADIENT)
    assert info["scratch"] == 0
    x = torch.from_numpy(xn).cuda()
    y = torch.empty(B, dtype=torch.float32, device="cuda")
    g = torch.empty(B, 16, dtype=torch.float32, device="cuda")
    res = hip.autotune_dev(bufs, TUNE_GRADIENT, B, x, d_y=y, d_g=g)
    assert res["fetch"] in (0, 1, 2)
    HipContext.set_tuned(bufs, TUNE_GRADIENT, 64, 0, fetch=2)
    y.zero_()
    g.zero_()
    graph = hip.graph_capture(bufs, TUNE_GRADIENT, B, x, d_y=y, d_g=g, repeats=2)
    for _ in range(3):
        y.zero_()
        g.zero_()
        torch.cuda.synchronize()
        graph.launch()
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy(), y0) and np.array_equal(g.cpu().numpy().reshape(-1), g0)
    graph.close()
    for b in (1, 65):
        yb, gb = hip.gradient_batch(bufs, xn[:b].reshape(-1), b)
        assert np.array_equal(yb, y0[:b]) and np.array_equal(gb, g0[: b * 16])


def test_specialise_rejects(hip):
    tape = wl.tanh_mlp(4, 4, 33)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    with pytest.raises(GpuError, match="specialised kernels exist"):
        HipContext.specialize(bufs, 9)
    b64 = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    with pytest.raises(GpuError, match="binary64 adjoint accumulation is for f32 tapes"):
        HipContext.specialize(b64, SPEC_GRAD_ACC64)
    with pytest.raises(GpuError, match="no specialised kernel"):
        HipContext.set_tuned(bufs, TUNE_FORWARD, 64, 0, fetch=2)
    HipContext.specialize(bufs, TUNE_FORWARD)
    HipContext.set_tuned(bufs, TUNE_FORWARD, 64, 0, fetch=2)
    with pytest.raises(GpuError, match="no specialised kernel"):
        HipContext.set_tuned(bufs, TUNE_GRADIENT, 64, 0, fetch=2)


@pytest.mark.parametrize("name", ["kitchen_sink", "rosenbrock", "mlp", "dag0", "guards"])
@pytest.mark.parametrize("f64", [False, True])
def test_specialised_hvp_equals_interpreter_and_oracle(hip, name, f64):
    """Gradient and H v from the specialised dual kernel, one x per point and
    one shared x (the reference's hvp_batch), against the dual interpreter
    (bits, NaN signs aside) and the f64 oracle."""
    from echidna_amd import philox
    from echidna_amd.backend import TUNE_HVP

    tape = _tape(name)
    n = tape.num_inputs
    F = np.float64 if f64 else np.float32
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    info = HipContext.specialize(bufs, TUNE_HVP)
    assert info["vgprs"] > 0
    B = 700
    x = _points(name, B, n, 23).astype(F)
    v = philox.normal(16, philox.STREAM_DIRS, np.arange(B * n)).reshape(B, n).astype(F)
    v[5] = 0.0  # a zero direction: H v = 0 exactly
    outs = {}
    with np.errstate(all="ignore"):
        for fetch in (1, 2):
            if fetch == 2:
                HipContext.set_tuned(bufs, TUNE_HVP, 64, 0, fetch=2)
            else:
                HipContext.set_tuned(bufs, TUNE_HVP, 128, 8192, fetch=1)
            per_point = hip.hvp_batch_points(bufs, x.reshape(-1), v.reshape(-1), B)
            shared = hip.hvp_batch(bufs, x[7], v.reshape(-1), B)
            outs[fetch] = [np.asarray(a).copy() for a in (*per_point, *shared)]
        for a, b in zip(outs[1], outs[2]):
            assert np.array_equal(_bits(a, f64), _bits(b, f64)), name
        if f64:
            g, hv = (a.reshape(B, n) for a in outs[2][:2])
            ref_g, ref_hv = po.hvp(tape, x, v)
            mag = po.magnitudes(tape, x, v)
            assert_f64(g, ref_g.reshape(B, n), mag["grad"], name + " gradient")
            assert_f64(hv, ref_hv.reshape(B, n), mag["hv"], name + " hvp")


def test_upload_time_specialisation_and_default_use(hip):
    """ehip_set_specialize builds the kernels at upload for short tapes only;
    a built kernel runs untuned evaluations, and an interpreter shape pinned
    with set_tuned still gives the same results."""
    from echidna_amd.backend import TUNE_HVP

    short = wl.tanh_mlp(8, 6, 40)
    long_ = wl.tanh_mlp(8, 64, 41)
    with pytest.raises(GpuError, match="what_mask"):
        hip.set_specialize(100, 256)
    with pytest.raises(GpuError, match="limit"):
        hip.set_specialize(1 << 20, 1)
    hip.set_specialize(len(short.opcodes), 0b111)
    try:
        bs = hip.upload_tape(GpuTapeData.from_tape(short))
        bl = hip.upload_tape(GpuTapeData.from_tape(long_))
    finally:
        hip.set_specialize(0, 0)
    for what in (TUNE_FORWARD, TUNE_GRADIENT, TUNE_HVP):
        HipContext.set_tuned(bs, what, 256, 0, fetch=2)  # built at upload
        HipContext.set_tuned(bs, what, 0, 0)
        with pytest.raises(GpuError, match="no specialised kernel"):
            HipContext.set_tuned(bl, what, 256, 0, fetch=2)
    B = 333
    x = wl.points_normal(42, B, 8).astype(np.float32)
    y_spec, g_spec = hip.gradient_batch(bs, x.reshape(-1), B)  # untuned: specialised
    HipContext.set_tuned(bs, TUNE_GRADIENT, 128, 4096, fetch=0)
    y_int, g_int = hip.gradient_batch(bs, x.reshape(-1), B)
    assert np.array_equal(y_spec, y_int) and np.array_equal(g_spec, g_int)


@pytest.mark.parametrize("waves", [2, 4])
def test_occupancy_targets_keep_results(hip, waves):
    """Kernels built for a higher occupancy (the compiler spills what does
    not fit) compute the same bits as the interpreters."""
    from echidna_amd import philox
    from echidna_amd.backend import TUNE_HVP

    tape = wl.tanh_mlp(12, 40, 31)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    B = 500
    x = wl.points_normal(44, B, 12)
    v = philox.normal(45, philox.STREAM_DIRS, np.arange(B * 12)).reshape(B, 12)
    HipContext.set_tuned(bufs, TUNE_GRADIENT, 256, 24576, fetch=1)
    HipContext.set_tuned(bufs, TUNE_HVP, 256, 24576, fetch=1)
    ref = [np.asarray(a).copy() for a in (*hip.gradient_batch(bufs, x.reshape(-1), B),
                                          *hip.hvp_batch_points(bufs, x.reshape(-1), v.reshape(-1), B))]
    for what in (TUNE_GRADIENT, TUNE_HVP):
        info = HipContext.specialize(bufs, what, waves=waves)
        assert info["waves_per_eu"] == waves and info["vgprs"] <= 512 // waves
        HipContext.set_tuned(bufs, what, 128, 0, fetch=2)
    got = [np.asarray(a).copy() for a in (*hip.gradient_batch(bufs, x.reshape(-1), B),
                                          *hip.hvp_batch_points(bufs, x.reshape(-1), v.reshape(-1), B))]
    for a, b in zip(ref, got):
        assert np.array_equal(_bits(a, True), _bits(b, True))


def test_automatic_occupancy_policy(hip):
    """The f64 n=32 H=64 MLP gradient needs more than 256 registers at the
    compiler's own choice (one wave per SIMD): the automatic policy rebuilds
    it for two waves."""
    tape = wl.tanh_mlp(32, 64, 1)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    info = HipContext.specialize(bufs, TUNE_GRADIENT)
    assert info["waves_per_eu"] == 2 and info["vgprs"] <= 256
    B = 300
    x = wl.points_normal(46, B, 32)
    y_s, g_s = hip.gradient_batch(bufs, x.reshape(-1), B)
    HipContext.set_tuned(bufs, TUNE_GRADIENT, 256, 24576, fetch=1)
    y_i, g_i = hip.gradient_batch(bufs, x.reshape(-1), B)
    assert np.array_equal(y_s, y_i) and np.array_equal(g_s, g_i)


@pytest.mark.parametrize("name,orders", [("kitchen_sink", (1, 2, 3, 4, 5)), ("rosenbrock", (3,)), ("mlp", (2, 3)),
                                         ("dag0", (3,)), ("guards", (3,))])
@pytest.mark.parametrize("f64", [False, True])
def test_specialised_jets_equal_interpreter_and_oracle(hip, name, orders, f64):
    from echidna_amd import philox

    tape = _tape(name)
    n = tape.num_inputs
    F = np.float64 if f64 else np.float32
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    B = 700
    x = _points(name, B, n, 25).astype(F)
    v = philox.normal(17, philox.STREAM_DIRS, np.arange(B * n)).reshape(B, n).astype(F)
    with np.errstate(all="ignore"):
        for K in orders:
            HipContext.set_tuned(bufs, TUNE_TAYLOR, 128, 8192, order=K, fetch=1)
            ref = hip.taylor_forward_kth_batch(bufs, x.reshape(-1), v.reshape(-1), B, K)
            info = HipContext.specialize(bufs, 2 + K)
            assert info["vgprs"] > 0
            HipContext.set_tuned(bufs, TUNE_TAYLOR, 0, 0, order=K)  # untuned: the specialised kernel
            got = hip.taylor_forward_kth_batch(bufs, x.reshape(-1), v.reshape(-1), B, K)
            for a, b in zip(ref.coefficients, got.coefficients):
                assert np.array_equal(_bits(a, f64), _bits(b, f64)), (name, K)
            if f64 and K == 3:
                m = len(ref.coefficients[0]) // B
                jets = np.stack(got.coefficients, axis=1).reshape(B, m, 3)
                oj = po.taylor(tape, x, v, K=3)
                fin = np.isfinite(oj)
                assert np.array_equal(np.isnan(jets), np.isnan(oj))
                assert np.max(np.abs(jets[fin] - oj[fin]) / np.maximum(1.0, np.abs(oj[fin])), initial=0.0) <= 1e-12


@pytest.mark.parametrize("kind", ["rademacher", "gaussian"])
def test_specialised_jets_drive_stde(hip, kind):
    """The STDE entry points run the specialised order-3 kernel with device
    directions (S lanes per point) and give the interpreter's estimates."""
    from echidna_amd import cv as lcv

    tape = wl.tanh_mlp(12, 16, 61)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    P, S = 40, 64
    X = torch.from_numpy(wl.points_normal(62, P, 12).astype(np.float32)).cuda()
    d = lcv.DIR_GAUSSIAN if kind == "gaussian" else lcv.DIR_RADEMACHER
    HipContext.set_tuned(bufs, TUNE_TAYLOR, 256, 24576, order=3, fetch=1)
    ref = np.asarray(lcv.laplacian_batch(hip, bufs, X, S, d, 63, None)).copy()
    HipContext.specialize(bufs, 5)
    HipContext.set_tuned(bufs, TUNE_TAYLOR, 64, 0, order=3, fetch=2)
    got = np.asarray(lcv.laplacian_batch(hip, bufs, X, S, d, 63, None)).copy()
    assert np.array_equal(ref, got)


@pytest.mark.parametrize("spec", [False, True])
def test_f32_dual_tanh_guard_keeps_reference_nan_pattern(hip, spec):
    """f32 HVPs through tanh at arguments past cosh^2's overflow (|x| > 44.4)
    mixed with ordinary ones in the same waves: the waves with a large
    argument run the reference's dual formulas, so NaN and Inf positions
    equal the f32 oracle's; finite values agree within the f32 bound."""
    from echidna_amd.backend import TUNE_HVP

    tape, _ = record(lambda v: (v[0] * 1.5 + v[1]).tanh() * v[1] + v[0].tanh(), [0.1, 0.2])
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    if spec:
        HipContext.specialize(bufs, TUNE_HVP)
    B = 512
    rng = np.random.default_rng(70)
    x = rng.uniform(-2, 2, (B, 2))
    x[::9, 0] = rng.uniform(-70, 70, len(x[::9]))  # some lanes past the overflow
    x[300:364, 0] = 0.5  # a whole wave of ordinary arguments
    v = rng.normal(size=(B, 2))
    v[::13] = np.inf
    x, v = x.astype(np.float32), v.astype(np.float32)
    with np.errstate(all="ignore"):
        g, hv = hip.hvp_batch_points(bufs, x.reshape(-1), v.reshape(-1), B)
        rg, rhv = po.hvp(tape, x, v, dtype=np.float32)
    for got, ref in ((np.asarray(g).reshape(B, 2), rg.reshape(B, 2)), (np.asarray(hv).reshape(B, 2), rhv.reshape(B, 2))):
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        assert np.array_equal(np.isinf(got), np.isinf(ref))
        fin = np.isfinite(ref)
        assert np.max(np.abs(got[fin] - ref[fin]) / np.maximum(1.0, np.abs(ref[fin])), initial=0.0) <= 2e-5


@pytest.mark.parametrize("ppl", [2, 4, 7])
def test_points_per_lane_keeps_results(hip, ppl):
    """The specialised kernels loop over several points per lane (grid
    stride): the same values, gradients, HVPs and jets, also for a batch that
    is not a multiple of the workgroup size times the points per lane."""
    from echidna_amd import philox
    from echidna_amd.backend import TUNE_HVP

    tape = wl.tanh_mlp(10, 12, 71)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    B = 1000
    x = wl.points_normal(72, B, 10).astype(np.float32)
    v = philox.normal(73, philox.STREAM_DIRS, np.arange(B * 10)).reshape(B, 10).astype(np.float32)

    def run():
        y, g = hip.gradient_batch(bufs, x.reshape(-1), B)
        gh, hv = hip.hvp_batch_points(bufs, x.reshape(-1), v.reshape(-1), B)
        tk = hip.taylor_forward_kth_batch(bufs, x.reshape(-1), v.reshape(-1), B, 3)
        return [np.asarray(a).copy() for a in (y, g, gh, hv, *tk.coefficients)]

    for what in (TUNE_GRADIENT, TUNE_HVP, 5):
        HipContext.specialize(bufs, what)
    ref = run()
    HipContext.set_tuned(bufs, TUNE_GRADIENT, 64, 0, fetch=2, points_per_lane=ppl)
    HipContext.set_tuned(bufs, TUNE_HVP, 128, 0, fetch=2, points_per_lane=ppl)
    HipContext.set_tuned(bufs, TUNE_TAYLOR, 64, 0, order=3, fetch=2, points_per_lane=ppl)
    assert HipContext.get_tuned(bufs, TUNE_GRADIENT)["points_per_lane"] == ppl
    for a, b in zip(ref, run()):
        assert np.array_equal(a, b)
    with pytest.raises(GpuError, match="points_per_lane"):
        HipContext.set_tuned(bufs, TUNE_GRADIENT, 64, 0, fetch=1, points_per_lane=2)


def _rounded(tape):
    """The tape with its constants rounded to FP32, as an f32 upload holds them."""
    import copy

    t = copy.deepcopy(tape)
    t.values = np.asarray(tape.values, dtype=np.float64).astype(np.float32).astype(np.float64)
    return t


def _grad_f32(hip, bufs, x):
    B, n = x.shape
    X = torch.from_numpy(x.astype(np.float32)).cuda()
    Y = torch.empty(B, bufs.num_outputs, dtype=torch.float32, device="cuda")
    G = torch.empty(B, n, dtype=torch.float32, device="cuda")
    hip.gradient_batch_dev(bufs, X, B, Y, G)
    torch.cuda.synchronize()
    return Y.cpu().numpy(), G.cpu().numpy()


def test_binary64_adjoints_keep_cancelling_fan_in(hip):
    """Binary64 adjoint accumulation on an f32 tape (what 8).  x's adjoint
    collects 1e4 cos x, -1e4 cos x and y cos x (from three recorded sin x):
    in FP32 the running sum passes through 1e4 and keeps y cos x only to
    ulp(1e4) ~ 1e-3; in binary64 the two large contributions (the same FP32
    partial times +-1e4, exact in binary64) cancel exactly and y cos x
    keeps its FP32 accuracy.  Bars against the f64 oracle on the tape's
    FP32 constants: binary64 adjoints within 1e-5 of max |gradient|, and at
    least 100x below the FP32-accumulated kernel's worst error on the same
    points; values unchanged (the forward sweep is the same FP32 code)."""
    def f(v):
        x, y = v[0], v[1]
        return (x.sin() * 1e4 + y * 1e4) - (x.sin() * 1e4 + y * 1e4) + x.sin() * y

    tape, _ = record(f, [0.3, 0.7])
    d32 = GpuTapeData.from_tape(tape)
    x = np.random.default_rng(5).uniform(-2.0, 2.0, (4096, 2))
    x32 = x.astype(np.float32).astype(np.float64)
    _, g_ref = po.gradient(_rounded(tape), x32)
    bufs = hip.upload_tape(d32)
    HipContext.specialize(bufs, TUNE_GRADIENT)
    y1, g1 = _grad_f32(hip, bufs, x)
    HipContext.specialize(bufs, SPEC_GRAD_ACC64)
    y8, g8 = _grad_f32(hip, bufs, x)
    assert np.array_equal(_bits(y1, False), _bits(y8, False))
    scale = float(np.max(np.abs(g_ref)))
    e1 = float(np.max(np.abs(g1 - g_ref)))
    e8 = float(np.max(np.abs(g8 - g_ref)))
    assert e8 <= 1e-5 * scale, (e8, scale)
    assert e1 >= 100.0 * e8, (e1, e8)


def test_binary64_adjoints_on_the_mlp(hip):
    """The n=32 H=64 tanh MLP at 2^16 points: binary64 adjoints stay within
    the f32 bar against the f64 oracle (and no worse than FP32 adjoints at
    their worst point), the values bit-identical to the FP32 kernel's."""
    tape = wl.tanh_mlp(32, 64, 1)
    d32 = GpuTapeData.from_tape(tape)
    x = wl.points_normal(1, 1 << 16, 32)
    bufs = hip.upload_tape(d32)
    HipContext.specialize(bufs, TUNE_GRADIENT)
    y1, g1 = _grad_f32(hip, bufs, x)
    HipContext.specialize(bufs, SPEC_GRAD_ACC64)
    y8, g8 = _grad_f32(hip, bufs, x)
    assert np.array_equal(_bits(y1, False), _bits(y8, False))
    idx = np.arange(0, 1 << 16, 61)
    xs = x[idx].astype(np.float32).astype(np.float64)
    _, g_ref = po.gradient(_rounded(tape), xs)
    scale = np.maximum(np.max(np.abs(g_ref), axis=1, keepdims=True), 1e-30)
    r1 = float(np.max(np.abs(g1[idx] - g_ref) / scale))
    r8 = float(np.max(np.abs(g8[idx] - g_ref) / scale))
    assert r8 <= 2e-5 and r8 <= r1, (r8, r1)
