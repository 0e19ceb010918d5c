"""Order-3 f32 jets with DENSE blocks as batch GEMMs on FP32 MFMA (ehip_set_jet_mma).

The matrix path sums the same products in another order, fused, so it is
held to the f32 bar against the f64 oracle rather than to bit identity:
|gpu - oracle| <= tol * max(|oracle|, M) with M the sum of the absolute
terms behind the coefficient (pyoracle.magnitudes), tol 2e-5 for c0 and c1
and 1e-4 for c2 (one more rounding chain), elementwise; the exact path is
held to the same bar for comparison.  Lanes whose operands hold an Inf or a
NaN are recomputed with full Taylor products and must equal the exact path
bit for bit; f64 tapes never take the matrix path.  Tapes: one-layer MLPs
whose DENSE reads tape inputs (c2 of the operands is zero), a two-layer MLP
whose second DENSE reads full jets, member counts off the 32-member tile,
and the n=256 H=256 STDE tape; STDE lane groups of 4 to 128 lanes per point,
whose shared primal is formed once per point.
"""

import numpy as np
import pytest
import torch

import pyoracle as po
from echidna_amd import cv as lcv
from echidna_amd import philox
from echidna_amd import stde_gpu as sg
from echidna_amd import workloads as wl
from echidna_amd.backend import TUNE_TAYLOR, HipContext
from echidna_amd.tape import GpuTapeData, record
from parity_bar import f64_mismatch

pytestmark = pytest.mark.gpu

TOL = (2e-5, 2e-5, 1e-4)


@pytest.fixture
def mma(hip):
    hip.set_jet_mma(True)
    yield hip
    hip.set_jet_mma(False)


def _jets(hip, bufs, X, V, stride, dtype=torch.float32):
    B = V.shape[0]
    dX = torch.from_numpy(np.ascontiguousarray(X)).to(dtype).cuda()
    dV = torch.from_numpy(np.ascontiguousarray(V)).to(dtype).cuda()
    out = torch.full((B, 3), float("nan"), dtype=dtype, device="cuda")
    hip.taylor_kth_batch_dev(bufs, dX, stride, dV, B, 3, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _check(got, tape, X, V, what):
    ref = po.taylor(tape, X, V, K=3)[:, 0, :]
    mag = po.magnitudes(tape, X, V)
    M = np.stack([mag["value"][:, 0], mag["c1"], mag["c2"]], axis=1)
    worst = []
    for q in range(3):
        idx, w = f64_mismatch(got[:, q], ref[:, q], M[:, q], TOL[q])
        assert len(idx) == 0, (what, q, w, idx[:5])
        worst.append(w)
    return worst


def _mlp_skip(n=48, H=40, seed=64):
    """tanh MLP plus x_0 x_1 / 2: two inputs are read again after the first
    layer, so the inputs keep all their rows (no direction-only seeding)."""
    W, b, w2 = wl.mlp_weights(n, H, seed)

    def f(x):
        acc = x[0] * x[1] * 0.5
        for j in range(H):
            s = x[0] * float(W[j, 0])
            for i in range(1, n):
                s = s + x[i] * float(W[j, i])
            acc = acc + (s + float(b[j])).tanh() * float(w2[j])
        return acc

    tape, _ = record(f, [0.0] * n)
    return tape


# tanh (and softplus's exp) run in the GEMM's epilogue, the ridge's ln1p on
# the interpreter after it; the skip connection reads two inputs again.
TAPES = [("mlp_n48_h40", lambda: wl.tanh_mlp(48, 40, 61)),
         ("mlp2_n40_h70_h33", lambda: wl.tanh_mlp2(40, 70, 33, 62)),
         ("mlp_n256_h256", lambda: wl.tanh_mlp(256, 256, 2)),
         ("mlp_skip_n48_h40", _mlp_skip),
         ("ridge_n40_h36", lambda: wl.softplus_ridge(40, 36, 63))]


@pytest.mark.parametrize("shape", [(256, 24576, 1), (128, 0, 0), (64, 1 << 16, 1)], ids=["split", "hbm", "lds"])
@pytest.mark.parametrize("shared_row", [False, True])
@pytest.mark.parametrize("name,build", TAPES, ids=[t[0] for t in TAPES])
def test_mma_jets_against_oracle_and_exact(mma, name, build, shape, shared_row):
    hip = mma
    tape = build()
    n = tape.num_inputs
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    wg, lds, fetch = shape
    HipContext.set_tuned(bufs, TUNE_TAYLOR, wg, lds, order=3, fetch=fetch)
    B = 1000  # not a multiple of 64: a partial last wave
    X = wl.points_normal(3, 1 if shared_row else B, n).astype(np.float32).astype(np.float64)
    V = philox.normal(4, philox.STREAM_DIRS, np.arange(B * n)).reshape(B, n).astype(np.float32).astype(np.float64)
    got = _jets(hip, bufs, X, V, 0 if shared_row else n)
    hip.set_jet_mma(False)
    exact = _jets(hip, bufs, X, V, 0 if shared_row else n)
    hip.set_jet_mma(True)
    Xf = np.repeat(X, B, axis=0) if shared_row else X
    w_mma = _check(got, tape, Xf, V, "mma")
    w_exact = _check(exact, tape, Xf, V, "exact")
    assert not np.array_equal(got, exact) or name == "none"  # the matrix path ran
    print(f"{name} {shape} shared={shared_row}: worst ratio mma {np.round(w_mma, 4)}, exact {np.round(w_exact, 4)}")


def test_mma_nonfinite_lanes_take_full_products(mma):
    """A lane whose DENSE operands hold an Inf or a NaN is recomputed term by
    term with full Taylor products: bit for bit the exact path's result
    (0 * Inf stays NaN); the other lanes of its 64-lane tile keep the
    matrix result."""
    hip = mma
    tape = wl.tanh_mlp(48, 40, 61)
    n, B = 48, 256
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    HipContext.set_tuned(bufs, TUNE_TAYLOR, 64, 24576, order=3, fetch=1)
    X = wl.points_normal(5, B, n).astype(np.float32)
    V = philox.normal(6, philox.STREAM_DIRS, np.arange(B * n)).reshape(B, n).astype(np.float32)
    X[70, 3] = np.inf
    V[200, 5] = np.nan
    got = _jets(hip, bufs, X, V, n)
    hip.set_jet_mma(False)
    exact = _jets(hip, bufs, X, V, n)
    hip.set_jet_mma(True)
    for lane in (70, 200):
        assert np.array_equal(got[lane].view(np.uint32), exact[lane].view(np.uint32)), (lane, got[lane], exact[lane])
    assert np.isnan(got[200]).any() and not np.isfinite(got[70]).all()
    assert not np.array_equal(got[:64], exact[:64])  # finite tiles: the matrix path
    Xd, Vd = X.astype(np.float64), V.astype(np.float64)
    fin = np.ones(B, bool)
    fin[[70, 200]] = False
    _check(got[fin], tape, Xd[fin], Vd[fin], "finite lanes")


def test_mma_f64_tapes_unaffected(mma):
    hip = mma
    tape = wl.tanh_mlp(48, 40, 61)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    X = wl.points_normal(7, 300, 48)
    V = wl.points_normal(8, 300, 48)
    got = _jets(hip, bufs, X, V, 48, torch.float64)
    hip.set_jet_mma(False)
    exact = _jets(hip, bufs, X, V, 48, torch.float64)
    hip.set_jet_mma(True)
    assert np.array_equal(got.view(np.uint64), exact.view(np.uint64))


@pytest.mark.parametrize("S", [4, 8, 16, 32, 24, 128])
@pytest.mark.parametrize("name,build", TAPES[:2], ids=[t[0] for t in TAPES[:2]])
def test_mma_stde_shared_primal_groups(mma, name, build, S):
    """STDE lanes of one point share c0 (the primal): with S dividing the
    64-lane tile (4..32) or a multiple of it (128), c0 is formed once per
    point on the vector ALUs and broadcast; S = 24 takes the full matrix
    path.  Every point's estimate and sample variance against the oracle on
    the same directions, 1e-4 of max(|estimate|, SE) and of the variance, and
    the exact path's results within the same bar."""
    hip = mma
    tape = build()
    n = tape.num_inputs
    P, seed = 7, 3  # 7 S lanes: partial tiles for every S
    X = wl.points_normal(9, P, n).astype(np.float32)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    dX = torch.from_numpy(X).cuda()
    res = lcv.laplacian_batch(hip, bufs, dX, S, lcv.DIR_GAUSSIAN, seed)
    hip.set_jet_mma(False)
    exact = lcv.laplacian_batch(hip, bufs, dX, S, lcv.DIR_GAUSSIAN, seed)
    hip.set_jet_mma(True)
    for p in range(P):
        idx = np.arange(p * S * n, (p + 1) * S * n, dtype=np.uint64)
        dirs = philox.direction(philox.DIR_GAUSSIAN, seed, idx).reshape(S, n).astype(np.float32).astype(np.float64)
        ref = po.laplacian_with_stats(tape, X[p].astype(np.float64), dirs)
        scale = max(abs(ref["estimate"]), ref["standard_error"])
        for got in (res, exact):
            assert abs(got[p, 1] - ref["estimate"]) <= 1e-4 * scale, (p, got[p], ref)
            assert abs(got[p, 2] - ref["sample_variance"]) <= 1e-4 * ref["sample_variance"], (p, got[p], ref)
            assert abs(got[p, 0] - ref["value"]) <= 2e-5 * max(1.0, abs(ref["value"])) and got[p, 4] == S
    assert not np.array_equal(res[:, 1], exact[:, 1])  # the matrix path ran


@pytest.mark.parametrize("kind", [lcv.DIR_RADEMACHER, lcv.DIR_GAUSSIAN])
def test_mma_stde_n256_full_batch(mma, kind):
    """Plain STDE on the n=256 H=256 tape (65,536 points x 64 device
    directions) on the matrix path: estimates of three points against the
    oracle on the same directions (1e-4 of max(|estimate|, SE)), and
    z-scores of 256 points against exact traces."""
    hip = mma
    n, P, S, seed = 256, 65536, 64, 2
    tape = wl.tanh_mlp(n, 256, 2)
    X = wl.points_normal(2, P, n).astype(np.float32)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    dX = torch.from_numpy(X).cuda()
    res = lcv.laplacian_batch(hip, bufs, dX, S, kind, seed)
    worst = 0.0
    for p in (0, 31337, P - 1):
        idx = np.arange(p * S * n, (p + 1) * S * n, dtype=np.uint64)
        dirs = philox.direction(kind, seed, idx).reshape(S, n).astype(np.float32).astype(np.float64)
        ref = po.laplacian_with_stats(tape, X[p].astype(np.float64), dirs)
        err = abs(res[p, 1] - ref["estimate"]) / max(abs(ref["estimate"]), ref["standard_error"])
        verr = abs(res[p, 2] - ref["sample_variance"]) / ref["sample_variance"]
        worst = max(worst, err, verr)
        assert err <= 1e-4 and verr <= 1e-4, (p, res[p], ref)
    k = 256
    b64 = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    exact = lcv.hessian_trace_batch(hip, b64, torch.from_numpy(X[:k].astype(np.float64)).cuda())
    z = (res[:k, 1] - exact) / res[:k, 3]
    print(f"mma stde kind {kind}: worst rel err vs oracle {worst:.3g}; mean z {np.mean(z):.3f}, "
          f"max |z| {np.max(np.abs(z)):.2f}")
    assert abs(np.mean(z)) <= 4.0 / np.sqrt(k) and np.max(np.abs(z)) <= 5.0


@pytest.mark.parametrize("name,want", [("mlp_n48_h40", (3, 1, 1, 1)), ("mlp2_n40_h70_h33", (4, 2, 2, 1)),
                                       ("mlp_skip_n48_h40", (3, 1, 1, 0)), ("ridge_n40_h36", (3, 1, 1, 0))])
def test_segment_plan(hip, name, want):
    """The cut: a LOADN segment, the GEMMs (each applying its layer's tanh
    or exp in its epilogue), the read-out.  Where the read-out dot product
    follows the last GEMM at once and takes every member, that GEMM sums it
    per workgroup row of members instead of storing them.  No fusion where
    the ridge's ln1p sits between, nor on the skip-connection tape, whose
    sum records as a chain of two-term dot products (each starting from the
    running sum)."""
    tape = dict(TAPES)[name]()
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    info = bufs.program_info(3)
    got = (info["segments"], info["gemms"], info["gemm_activations"], info["gemm_readouts"])
    assert got == want, info


@pytest.mark.parametrize("kind", [philox.DIR_RADEMACHER, philox.DIR_GAUSSIAN])
@pytest.mark.parametrize("name,want", [("mlp_n48_h40", (1, 48, 1)), ("mlp2_n40_h70_h33", (1, 40, 1)),
                                       ("mlp_skip_n48_h40", (1, 46, 0)), ("ridge_n40_h36", (1, 40, 1)),
                                       ("mlp_n45_h40", (1, 45, 1))])
def test_direct_plan_equals_supplied_directions(mma, kind, name, want):
    """Directions drawn on the device take the direct plan: the GEMM over
    tape inputs reads x and the directions itself (Rademacher signs drawn
    in the GEMM, Gaussians drawn once coordinate-major), and the LOADN loads
    only inputs another instruction reads (the skip connection's two), so
    the interpreter segment of a plain MLP is not launched.  The same
    directions supplied from the host (philox.lane_directions) take the
    register plan.  Estimates, variances and values are bit-identical: the
    GEMM sees the values the LOADN would have stored, and the host mirror
    gives the device's directions.  S = 2 puts c0 on the matrix cores, 16
    and 64 share it per point; lane counts leave partial tiles.  With n = 45
    (not a multiple of 4) the Rademacher signs are drawn coordinate-major
    first, as the Gaussians are."""
    hip = mma
    tape = wl.tanh_mlp(45, 40, 65) if name == "mlp_n45_h40" else dict(TAPES)[name]()
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    info = bufs.program_info(4)
    assert (info["direct_gemms"], info["inputs_dropped"], info["segments_skipped"]) == want, info
    n = tape.num_inputs
    for P, S, seed in ((150, 2, 5), (97, 16, 6), (13, 64, 7)):
        X = wl.points_normal(seed, P, n).astype(np.float32)
        dX = torch.from_numpy(X).cuda()
        V = philox.lane_directions(kind, seed, n, np.arange(P * S)).astype(np.float32)
        dV = torch.from_numpy(V).cuda()
        got = torch.full((P, 5), float("nan"), dtype=torch.float64, device="cuda")
        ref = torch.full((P, 5), float("nan"), dtype=torch.float64, device="cuda")
        sg.laplacian_batch_dev(hip, bufs, dX, P, S, kind, seed, got)
        sg.laplacian_batch_dev(hip, bufs, dX, P, S, sg.DIR_SUPPLIED, 0, ref, d_dirs=dV)
        torch.cuda.synchronize()
        g, r = got.cpu().numpy(), ref.cpu().numpy()
        assert np.all(np.isfinite(r)), (P, S)
        assert np.array_equal(g.view(np.uint64), r.view(np.uint64)), (P, S, np.argwhere(g != r)[:4])


@pytest.mark.parametrize("kind", [lcv.DIR_RADEMACHER, lcv.DIR_GAUSSIAN])
def test_direct_plan_writes_the_directions_the_cv_reads(mma, kind):
    """The learned control variate reads the jets' directions back
    (coordinate-major) to form u = L(x)^T v.  On the direct plan the GEMM
    writes the Rademacher signs it draws and the draw kernel the Gaussians;
    the controlled and plain statistics must agree with the interpreter's
    (register plan, directions written by its LOADN) to f32 rounding: 1e-4
    of max(|estimate|, SE) of the plain estimator."""
    hip = mma
    n = 48
    tape = dict(TAPES)["mlp_n48_h40"]()
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    net = lcv.LearnedControlVariate(hip, n, 4, 32, seed=3, out_scale=1.0)
    P, S, seed = 40, 16, 11
    dX = torch.from_numpy(wl.points_normal(12, P, n).astype(np.float32)).cuda()
    a, ap = lcv.laplacian_batch(hip, bufs, dX, S, kind, seed, net, with_plain=True)
    hip.set_jet_mma(False)
    b, bp = lcv.laplacian_batch(hip, bufs, dX, S, kind, seed, net, with_plain=True)
    hip.set_jet_mma(True)
    scale = np.maximum(np.abs(bp[:, 1]), bp[:, 3])
    assert np.all(np.abs(ap[:, 1] - bp[:, 1]) <= 1e-4 * scale), (ap[:3], bp[:3])
    assert np.all(np.abs(a[:, 1] - b[:, 1]) <= 1e-4 * scale), (a[:3], b[:3])
    assert np.all(np.abs(a[:, 2] - b[:, 2]) <= 1e-4 * bp[:, 2])
    assert not np.allclose(a[:, 1], ap[:, 1])  # the control variate changed the samples
