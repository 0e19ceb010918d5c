"""Autotuner and launch shapes: every workgroup size, LDS budget (all
registers in LDS, split LDS/HBM, all in HBM) and program fetch (LDS windows
or the scalar cache) computes bit-identical results, because each lane runs
the same arithmetic in the same order; only where the registers and the
instruction words live changes.  The tuned choice is kept per tape and
evaluation kind.
"""

import numpy as np
import pytest
import torch

from parity_bar import assert_f64
from echidna_amd import philox
from echidna_amd import workloads as wl
from echidna_amd.backend import TUNE_GRADIENT, TUNE_HVP, TUNE_TAYLOR, HipContext
from echidna_amd.tape import GpuError, GpuTapeData

pytestmark = pytest.mark.gpu

SHAPES = [(64, 0, 0), (64, 1024, 1), (128, 8192, 0), (256, 16384, 1), (256, 131072, 0), (128, 65536, 1),
          (256, 0, 1), (64, 65536, 0)]


@pytest.mark.parametrize("f64", [False, True])
def test_launch_shapes_are_bit_identical(hip, f64):
    tape = wl.tanh_mlp(12, 40, 31)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    F = np.float64 if f64 else np.float32
    B = 3000
    x = wl.points_normal(8, B, 12).astype(F)
    v = philox.normal(9, philox.STREAM_DIRS, np.arange(B * 12)).reshape(B, 12).astype(F)
    ref = None
    for wg, lds, fetch in SHAPES:
        HipContext.set_tuned(bufs, TUNE_GRADIENT, wg, lds, fetch=fetch)
        HipContext.set_tuned(bufs, TUNE_HVP, wg, lds, fetch=fetch)
        HipContext.set_tuned(bufs, TUNE_TAYLOR, wg, lds, order=3, fetch=fetch)
        y, g = hip.gradient_batch(bufs, x.reshape(-1), B)
        gh, hv = hip.hvp_batch_points(bufs, x.reshape(-1), v.reshape(-1), B)
        tk = hip.taylor_forward_kth_batch(bufs, x.reshape(-1), v.reshape(-1), B, 3)
        out = [np.asarray(a).copy() for a in (y, g, gh, hv, *tk.coefficients)]
        if ref is None:
            ref = out
        else:
            for a, b in zip(out, ref):
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (wg, lds, fetch)


def test_autotune_keeps_choice_and_results(hip):
    tape = wl.tanh_mlp(16, 32, 32)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    B = 1 << 14
    x = torch.from_numpy(wl.points_normal(10, B, 16).astype(np.float32)).cuda()
    y = torch.empty(B, dtype=torch.float32, device="cuda")
    g = torch.empty(B, 16, dtype=torch.float32, device="cuda")
    y0, g0 = hip.gradient_batch(bufs, x.cpu().numpy().reshape(-1), B)
    res = hip.autotune_dev(bufs, TUNE_GRADIENT, B, x, d_y=y, d_g=g)
    assert res["wg"] in (64, 128, 256) and res["us"] > 0 and res["fetch"] in (0, 1)
    assert HipContext.get_tuned(bufs, TUNE_GRADIENT) == {"wg": res["wg"], "lds_budget": res["lds_budget"],
                                                         "fetch": res["fetch"], "points_per_lane": 1}
    y1, g1 = hip.gradient_batch(bufs, x.cpu().numpy().reshape(-1), B)
    assert np.array_equal(y0, y1) and np.array_equal(g0, g1)
    HipContext.set_tuned(bufs, TUNE_GRADIENT, 0, 0)
    assert HipContext.get_tuned(bufs, TUNE_GRADIENT)["wg"] == 0


def test_autotune_rejects_bad_arguments(hip):
    tape = wl.tanh_mlp(4, 4, 33)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    with pytest.raises(GpuError, match="workgroup size"):
        HipContext.set_tuned(bufs, TUNE_GRADIENT, 96, 0)
    with pytest.raises(GpuError, match="unknown evaluation kind"):
        HipContext.set_tuned(bufs, 7, 64, 0)
    with pytest.raises(GpuError, match="program fetch"):
        HipContext.set_tuned(bufs, TUNE_GRADIENT, 64, 0, fetch=3)
    with pytest.raises(GpuError, match="no specialised kernel"):
        HipContext.set_tuned(bufs, TUNE_GRADIENT, 64, 0, fetch=2)
    with pytest.raises(GpuError, match="program fetch"):
        hip.set_fetch(3)
    x = torch.zeros(4, 4, device="cuda")
    with pytest.raises(GpuError, match="taylor order"):
        hip.autotune_dev(bufs, TUNE_TAYLOR, 4, x, d_y=x, d_dirs=x, order=9)


@pytest.mark.parametrize("f64", [False, True])
def test_dense_grouping_is_bit_identical_on_device(hip, f64):
    """Uploads with and without DENSE grouping give the same bits for values,
    gradients and Taylor jets."""
    tape = wl.tanh_mlp2(20, 24, 18, 35)
    F = np.float64 if f64 else np.float32
    B = 700
    x = wl.points_normal(12, B, 20).astype(F)
    v = philox.normal(13, philox.STREAM_DIRS, np.arange(B * 20)).reshape(B, 20).astype(F)
    outs = []
    for dense in (True, False):
        hip.set_dense(dense)
        try:
            bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
        finally:
            hip.set_dense(True)
        y, g = hip.gradient_batch(bufs, x.reshape(-1), B)
        tk = hip.taylor_forward_kth_batch(bufs, x.reshape(-1), v.reshape(-1), B, 3)
        outs.append([np.asarray(a).copy() for a in (y, g, *tk.coefficients)])
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("f64", [False, True])
def test_dense_guards_are_bit_identical_on_device(hip, f64):
    """Zero, infinite and NaN adjoints and zero weights inside DENSE groups:
    gradients with and without grouping agree bit for bit across both program
    fetch modes, except the sign bit of NaN results, which IEEE 754 leaves
    unspecified (the device gives +NaN on one path and -NaN on the other for
    0 + c * NaN); NaN positions, Inf values and every finite value agree."""
    import dense_cases as dc

    tape = dc.tape()
    F = np.float64 if f64 else np.float32
    B = 500
    x = dc.points(B, 3).astype(F)
    outs = []
    for dense in (True, False):
        hip.set_dense(dense)
        try:
            bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
        finally:
            hip.set_dense(True)
        for fetch in (0, 1):
            HipContext.set_tuned(bufs, TUNE_GRADIENT, 256, 24576, fetch=fetch)
            y, g = hip.gradient_batch(bufs, x.reshape(-1), B)
            outs.append((np.asarray(y).copy(), np.asarray(g).copy()))
    g0 = outs[0][1]
    g0r = g0.reshape(B, -1)
    assert np.isinf(g0r).any() and np.isnan(g0r).any() and np.isfinite(g0r).all(axis=1).any()
    U = np.uint64 if f64 else np.uint32
    sign = U(1) << U(63 if f64 else 31)

    def bits(a):  # NaN sign cleared
        b = a.view(U).copy()
        b[np.isnan(a)] &= ~sign
        return b

    for y, g in outs[1:]:
        assert np.array_equal(bits(y), bits(outs[0][0]))
        assert np.array_equal(bits(g), bits(g0))
    if f64:  # and the same NaN / Inf pattern and values as the oracle
        import pyoracle as po

        with np.errstate(all="ignore"):
            _, ref = po.gradient(tape, x.astype(np.float64))
        assert np.array_equal(np.isnan(g0r), np.isnan(ref))
        fin = np.isfinite(ref)
        assert np.array_equal(g0r[~fin & ~np.isnan(ref)], ref[~fin & ~np.isnan(ref)])
        assert np.max(np.abs(g0r[fin] - ref[fin]) / np.maximum(1.0, np.abs(ref[fin]))) <= 1e-12


def _bits(a, f64):
    """Bits with the sign of NaNs cleared (IEEE 754 leaves it unspecified)."""
    U = np.uint64 if f64 else np.uint32
    b = np.asarray(a).view(U).copy()
    b[np.isnan(a)] &= ~(U(1) << U(63 if f64 else 31))
    return b


@pytest.mark.parametrize("name", ["mlp2", "guards"])
@pytest.mark.parametrize("f64", [False, True])
def test_dense_hvp_matches_dot_program_and_oracle(hip, name, f64):
    """HVPs (one x per direction) with the dual interpreter's DENSE groups
    equal the dot-level program's bit for bit (NaN signs aside), in both
    program fetch modes, and match the f64 oracle."""
    import dense_cases as dc
    import pyoracle as po

    tape = wl.tanh_mlp2(20, 24, 18, 35) if name == "mlp2" else dc.tape()
    n = tape.num_inputs
    F = np.float64 if f64 else np.float32
    B = 400
    x = (wl.points_normal(14, B, n) if name == "mlp2" else dc.points(B, 5)).astype(F)
    v = philox.normal(15, philox.STREAM_DIRS, np.arange(B * n)).reshape(B, n).astype(F)
    outs = []
    for dense in (True, False):
        hip.set_dense(dense)
        try:
            bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
        finally:
            hip.set_dense(True)
        for fetch in (0, 1):
            HipContext.set_tuned(bufs, TUNE_HVP, 128, 8192, fetch=fetch)
            g, hv = hip.hvp_batch_points(bufs, x.reshape(-1), v.reshape(-1), B)
            outs.append((np.asarray(g).copy(), np.asarray(hv).copy()))
    for g, hv in outs[1:]:
        assert np.array_equal(_bits(g, f64), _bits(outs[0][0], f64))
        assert np.array_equal(_bits(hv, f64), _bits(outs[0][1], f64))
    if f64:
        g, hv = (a.reshape(B, n) for a in outs[0])
        with np.errstate(all="ignore"):
            _, ref = po.hvp(tape, x, v)
        ref = ref.reshape(B, n)
        with np.errstate(all="ignore"):
            mag = po.magnitudes(tape, x, v)["hv"]
        assert_f64(hv, ref, mag, "hvp")


@pytest.mark.parametrize("f64", [False, True])
def test_dense_jets_over_inputs_with_nonfinite_points(hip, f64):
    """Order-3 jets through DENSE groups whose operands are tape inputs (the
    seeds' second coefficients are zero and are not read): bit-identical to
    the ungrouped program (NaN signs aside) also when some inputs and
    directions are infinite or NaN, which sends terms through the full Taylor
    product; f64 against the oracle."""
    import pyoracle as po

    tape = wl.tanh_mlp2(20, 24, 18, 35)
    F = np.float64 if f64 else np.float32
    B = 640
    x = wl.points_normal(17, B, 20)
    v = philox.normal(18, philox.STREAM_DIRS, np.arange(B * 20)).reshape(B, 20)
    x[3, 5] = np.inf
    x[70:72, 0] = np.nan
    v[130, 2] = -np.inf
    x[200:264, 7] = np.inf  # a whole wave of lanes
    x, v = x.astype(F), v.astype(F)
    outs = []
    for dense in (True, False):
        hip.set_dense(dense)
        try:
            bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
        finally:
            hip.set_dense(True)
        tk = hip.taylor_forward_kth_batch(bufs, x.reshape(-1), v.reshape(-1), B, 3)
        outs.append([np.asarray(c).copy() for c in tk.coefficients])
    for a, b in zip(*outs):
        assert np.array_equal(_bits(a, f64), _bits(b, f64))
    if f64:
        with np.errstate(all="ignore"):
            ref = po.taylor(tape, x, v, K=3)[:, 0, :]
        got = np.stack(outs[0], axis=1).reshape(B, 3)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        fin = np.isfinite(ref)
        assert np.array_equal(got[~fin & ~np.isnan(ref)], ref[~fin & ~np.isnan(ref)])
        assert np.max(np.abs(got[fin] - ref[fin]) / np.maximum(1.0, np.abs(ref[fin]))) <= 1e-12


@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("row", ["finite", "nonfinite"])
def test_dense_jets_over_a_shared_input_row(hip, f64, row):
    """Order-3 jets with one input row shared by every lane (x_stride 0, as
    the single-point STDE entry points and S = 64 batches run them): the
    DENSE groups over inputs then take operand values from that row, and
    each lane accumulates one member's value coefficient for the whole wave.
    Bit-identical to the ungrouped program (NaN signs aside) for a finite
    row and for a row holding Inf and NaN, with infinite and NaN directions
    in single lanes and in a whole wave, a partial last wave, and the layer
    split into full and partial blocks of members; f64 also against the
    oracle."""
    import pyoracle as po

    n = 20
    tape = wl.tanh_mlp2(n, 37, 18, 41)
    F = np.float64 if f64 else np.float32
    B = 200  # three full waves and a partial one
    x = wl.points_normal(19, 1, n)[0]
    if row == "nonfinite":
        x[4] = np.inf
        x[11] = np.nan
    v = philox.normal(20, philox.STREAM_DIRS, np.arange(B * n)).reshape(B, n)
    v[5, 3] = np.inf
    v[77, 0] = np.nan
    v[128:192, 9] = -np.inf  # a whole wave
    xd = torch.tensor(x.astype(F), device="cuda")
    vd = torch.tensor(v.astype(F), device="cuda")
    outs = []
    for dense in (True, False):
        hip.set_dense(dense)
        try:
            bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
        finally:
            hip.set_dense(True)
        for fetch in (0, 1):
            HipContext.set_tuned(bufs, TUNE_TAYLOR, 64, 4096, order=3, fetch=fetch)
            c = torch.zeros(B * 3, dtype=xd.dtype, device="cuda")
            hip.taylor_kth_batch_dev(bufs, xd, 0, vd, B, 3, c)
            torch.cuda.synchronize()
            outs.append(c.cpu().numpy().reshape(B, 3))
    for o in outs[1:]:
        assert np.array_equal(_bits(o, f64), _bits(outs[0], f64))
    if f64:
        with np.errstate(all="ignore"):
            ref = po.taylor(tape, np.tile(x, (B, 1)), v, K=3)[:, 0, :]
        got = outs[0]
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        fin = np.isfinite(ref)
        assert np.array_equal(got[~fin & ~np.isnan(ref)], ref[~fin & ~np.isnan(ref)])
        assert np.max(np.abs(got[fin] - ref[fin]) / np.maximum(1.0, np.abs(ref[fin])), initial=0.0) <= 1e-12
