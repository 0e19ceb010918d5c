"""Learned control variate on the GPU: forward, gradients and Adam against
the numpy restatement (BF16-rounded GEMM operands, tests/cv_ref.py), the
training tuples against Philox and the HVP kernel, exact traces against the
oracle, unbiasedness of the controlled estimator, the variance it removes
after training, and the plain-STDE fallback.

Bars: forward and gradients within 2e-3 of the reference in relative
Frobenius norm per tensor (FP32 accumulation order and the occasional BF16
tie flipped by it; a wrong index or a dropped term gives O(1) errors);
Adam within 1e-6 relative; exact traces (f64 tape) within 1e-12 relative.
"""

import numpy as np
import pytest
import torch

import cv_ref
import pyoracle as po
from echidna_amd import cv as lcv
from echidna_amd import philox
from echidna_amd import workloads as wl
from echidna_amd.tape import GpuError, GpuTapeData

pytestmark = pytest.mark.gpu

N, H, R = 24, 40, 4  # sizes off the 64x64x32 GEMM tile on purpose


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def random_cv(hip, seed=3, out_scale=1.0):
    net = lcv.LearnedControlVariate(hip, N, R, H, seed=seed, out_scale=out_scale)
    return net


def test_init_matches_philox_formula(hip):
    net = random_cv(hip, seed=11, out_scale=0.5)
    assert np.array_equal(net.params(), cv_ref.init_params(N, H, R, 11, 0.5))


def test_context_stream_waits_for_the_default_stream(hip):
    """An input written by a kernel still queued on torch's default stream
    is read complete by a call on the context's stream (a blocking stream,
    ordered after the legacy default stream), with no synchronisation in
    between: the f64 -> f32 conversion of a learned-CV caller's points is
    the case that exposed the race."""
    net = random_cv(hip)
    X = torch.from_numpy(wl.points_normal(3, 37, N)).cuda()
    want = net.forward(X.float())
    busy = torch.full((2048, 2048), 1e-4, device="cuda")
    for _ in range(8):
        busy = busy @ busy  # keeps the default stream busy; entries decay towards 0
    xf = X.float() + busy[0, :1] * 0.0  # written only after the products
    L = torch.empty_like(want)
    net.forward_dev(xf, 37, L)
    torch.cuda.synchronize()
    assert torch.equal(L, want)


def test_forward_matches_reference(hip):
    net = random_cv(hip)
    X = wl.points_normal(3, 37, N).astype(np.float32)
    L = net.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    _, _, ref = cv_ref.forward(net.params(), X, N, H, R, bf=True)
    assert L.shape == (37, N * R)
    assert rel(L, ref) <= 2e-3


def test_gradients_match_reference(hip):
    net = random_cv(hip)
    T = 53
    X = wl.points_normal(4, T, N).astype(np.float32)
    V = philox.normal(5, philox.STREAM_DIRS, np.arange(T * N)).reshape(T, N).astype(np.float32)
    HV = (V * 0.3 + philox.normal(6, 1, np.arange(T * N)).reshape(T, N) * 0.1).astype(np.float32)
    dev = [torch.from_numpy(a).cuda() for a in (X, V, HV)]
    loss = net.loss_grad(*dev)
    ref_loss, ref_g = cv_ref.loss_grad(net.params(), X, V, HV, N, H, R, bf=True)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    g = net.grads()
    for k, (a, b) in enumerate(zip(net.split(g), net.split(ref_g))):
        assert rel(a, b) <= 2e-3, k


def test_signed_factor_gradients_and_unbiasedness(hip):
    """S = diag(+1, +1, -1, -1): gradients against the reference and the
    controlled estimate still centred on the exact trace."""
    net = random_cv(hip, seed=4)
    net.set_signs(2)
    net.set_gain(0.7)
    T = 45
    X = wl.points_normal(11, T, N).astype(np.float32)
    V = philox.normal(12, philox.STREAM_DIRS, np.arange(T * N)).reshape(T, N).astype(np.float32)
    HV = (V * -0.4).astype(np.float32)
    loss = net.loss_grad(*[torch.from_numpy(a).cuda() for a in (X, V, HV)])
    ref_loss, ref_g = cv_ref.loss_grad(net.params(), X, V, HV, N, H, R, bf=True, gain=0.7, npos=2)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    for k, (a, b) in enumerate(zip(net.split(net.grads()), net.split(ref_g))):
        assert rel(a, b) <= 2e-3, k
    tape = wl.tanh_mlp(N, 20, 56)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    P = 48
    Xp = torch.from_numpy(wl.points_normal(26, P, N)).cuda()
    exact = lcv.hessian_trace_batch(hip, bufs, Xp)
    res = lcv.laplacian_batch(hip, bufs, Xp, 512, lcv.DIR_GAUSSIAN, 13, net)
    z = (res[:, 1] - exact) / res[:, 3]
    assert np.all(np.abs(z) < 5.0) and abs(np.mean(z)) < 4.0 / np.sqrt(P)


def test_gradients_chunked_equal_single_pass(hip):
    """Batches above the 2048-tuple chunk accumulate gradients over chunks."""
    net = random_cv(hip)
    T = 2048 + 300
    X = wl.points_normal(7, T, N).astype(np.float32)
    V = philox.normal(8, philox.STREAM_DIRS, np.arange(T * N)).reshape(T, N).astype(np.float32)
    HV = (0.5 * V).astype(np.float32)
    loss = net.loss_grad(*[torch.from_numpy(a).cuda() for a in (X, V, HV)])
    ref_loss, ref_g = cv_ref.loss_grad(net.params(), X, V, HV, N, H, R, bf=True)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    for k, (a, b) in enumerate(zip(net.split(net.grads()), net.split(ref_g))):
        assert rel(a, b) <= 2e-3, k


@pytest.mark.parametrize("r2", [8, 16])
def test_split_k_gradients_match_reference(hip, r2):
    """n r >= 4096 over a 300-tuple batch: the dZ2 product runs split over K
    (few output tiles, long K), and the loss runs its rank-specialised
    kernel with signed columns; gradients still match the reference."""
    n2, h2, T = 512, 64, 300
    net = lcv.LearnedControlVariate(hip, n2, r2, h2, seed=11, out_scale=1.0)
    net.set_signs(r2 // 2)
    X = wl.points_normal(21, T, n2).astype(np.float32)
    V = philox.normal(22, philox.STREAM_DIRS, np.arange(T * n2)).reshape(T, n2).astype(np.float32)
    HV = (0.5 * V).astype(np.float32)
    loss = net.loss_grad(*[torch.from_numpy(a).cuda() for a in (X, V, HV)])
    ref_loss, ref_g = cv_ref.loss_grad(net.params(), X, V, HV, n2, h2, r2, bf=True, npos=r2 // 2)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    for k, (a, b) in enumerate(zip(net.split(net.grads()), net.split(ref_g))):
        assert rel(a, b) <= 2e-3, k


def test_adam_step_matches_reference(hip):
    net = random_cv(hip)
    T = 40
    X = wl.points_normal(9, T, N).astype(np.float32)
    V = philox.normal(10, philox.STREAM_DIRS, np.arange(T * N)).reshape(T, N).astype(np.float32)
    HV = (0.2 * V).astype(np.float32)
    dev = [torch.from_numpy(a).cuda() for a in (X, V, HV)]
    p = net.params().astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t in (1, 2):
        net.loss_grad(*dev)
        g = net.grads().astype(np.float64)
        net.adam_step(1e-2)
        p, m, v = cv_ref.adam(p, g, m, v, t, 1e-2)
        assert np.max(np.abs(net.params() - p)) <= 1e-6 * max(1.0, np.max(np.abs(p)))
    assert net.info()["steps"] == 2


def test_cosine_schedule_is_per_epoch_adam(hip):
    """schedule="cosine" equals epoch-by-epoch training at lr (1 + cos(pi e / E)) / 2."""
    T, E, lr = 96, 3, 1e-2
    X = wl.points_normal(11, T, N).astype(np.float32)
    V = philox.normal(12, philox.STREAM_DIRS, np.arange(T * N)).reshape(T, N).astype(np.float32)
    HV = (0.3 * V).astype(np.float32)
    dev = [torch.from_numpy(a).cuda() for a in (X, V, HV)]
    a, b = random_cv(hip, seed=5), random_cv(hip, seed=5)
    ha = a.train(*dev, batch=32, epochs=E, lr=lr, schedule="cosine")
    hb = np.concatenate([b.train(*dev, batch=32, epochs=1, lr=lr * 0.5 * (1 + np.cos(np.pi * e / E)))
                         for e in range(E)])
    assert len(ha) == E * 3 and np.array_equal(ha, hb)
    assert np.array_equal(a.params(), b.params()) and a.info()["steps"] == E * 3
    with pytest.raises(ValueError, match="unknown schedule"):
        a.train(*dev, schedule="linear")


def test_tuples_match_philox_and_hvp(hip):
    tape = wl.tanh_mlp(10, 12, 50)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    x, v, hv = lcv.make_tuples(hip, bufs, 300, seed=21)
    xr = wl.points_normal(21, 300, 10).astype(np.float32)
    vr = philox.normal(21, philox.STREAM_DIRS, np.arange(3000)).reshape(300, 10).astype(np.float32)
    assert np.allclose(x.cpu().numpy(), xr, rtol=1e-6, atol=1e-7)
    assert np.allclose(v.cpu().numpy(), vr, rtol=1e-6, atol=1e-7)
    _, hv_ref = hip.hvp_batch_points(bufs, x.cpu().numpy().reshape(-1), v.cpu().numpy().reshape(-1), 300)
    assert np.array_equal(hv.cpu().numpy().reshape(-1), np.asarray(hv_ref).reshape(-1))


def test_exact_trace_matches_oracle(hip):
    tape = wl.tanh_mlp(9, 11, 51)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    X = wl.points_normal(22, 17, 9)
    tr = lcv.hessian_trace_batch(hip, bufs, torch.from_numpy(X).cuda())
    for p in range(17):
        _, d = po.hessian_diagonal(tape, X[p])
        ref = float(np.sum(d))
        assert abs(tr[p] - ref) <= 1e-12 * max(1.0, abs(ref)), p


def test_controlled_estimator_is_unbiased(hip):
    """An untrained (badly scaled) network changes the variance, not the mean."""
    tape = wl.tanh_mlp(16, 24, 52)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    net = lcv.LearnedControlVariate(hip, 16, 4, 32, seed=5, out_scale=2.0)
    P = 64
    X = torch.from_numpy(wl.points_normal(23, P, 16)).cuda()
    exact = lcv.hessian_trace_batch(hip, bufs, X)
    res, plain = lcv.laplacian_batch(hip, bufs, X, 512, lcv.DIR_GAUSSIAN, 9, net, with_plain=True)
    z = (res[:, 1] - exact) / res[:, 3]
    assert np.all(np.abs(z) < 5.0)
    assert abs(np.mean(z)) < 4.0 / np.sqrt(P)
    assert np.all(res[:, 4] == 512) and np.all(plain[:, 4] == 512)
    assert np.array_equal(res[:, 0], plain[:, 0])


def _welford(x):
    return np.mean(x), np.var(x, ddof=1) if len(x) > 1 else 0.0


def _bf16(a):
    """FP32 values rounded to bfloat16, to nearest even (finite inputs)."""
    b = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = ((b + np.uint64(0x7FFF) + ((b >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)) << np.uint64(16)
    return b.astype(np.uint32).view(np.float32)


def test_bf16_rounding_helper():
    """Ties go to the even bfloat16; torch's conversion agrees."""
    x = np.array([1.0, 1.00390625, 1.01171875, 1.0000001, -3.14159, 1e-30, 65504.0], np.float32)
    assert _bf16(x)[1] == 1.0 and _bf16(x)[2] == np.float32(1.015625)
    assert np.array_equal(_bf16(x), torch.from_numpy(x).bfloat16().float().numpy())


@pytest.mark.parametrize("f64", [False, True], ids=["f32", "f64"])
@pytest.mark.parametrize("n,rank,npos,S,P", [(40, 12, 5, 20, 10), (64, 16, 8, 16, 10), (36, 24, 24, 37, 10),
                                             (33, 32, 10, 3, 10), (64, 16, 8, 4, 128), (512, 16, 8, 2, 1024)])
def test_controlled_statistics_match_a_host_restatement(hip, f64, n, rank, npos, S, P):
    """Per-point statistics of the controlled samples 2 c2 - (v^T L S L^T v -
    tr(L S L^T)) against numpy in binary64 from the same pieces: the jets of
    the same points and Philox directions (one GPU evaluation), L(x) from the
    network's forward pass rounded to bfloat16 (as the pass stores it), the
    Welford mean and sample variance.  Shapes
    straddle the 16-sample x 16-column matrix tile and the 16-coordinate slab
    (partial tiles, two column blocks, n not a multiple of 4, S below 4);
    P = 128 (n = 64) and 1024 (n = 512) make L's GEMM store whole 64 x 64
    and 128 x 128 tiles through its bfloat16 row staging.
    Bar: the two sides sum the same binary64 products in different orders,
    1e-11 of the samples' magnitude for the mean, 1e-9 relative for the
    variance, widened for f32 tapes by 1e-6 of the jets' part in case the
    separate jet evaluation rounds a coefficient differently; a swapped row
    or column of the tile gives O(1) errors."""
    tape = wl.tanh_mlp(n, 24, 70 + n)
    bufs = hip.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
    net = lcv.LearnedControlVariate(hip, n, rank, 32, seed=n, out_scale=1.0)
    net.set_signs(npos)
    seed = 13
    dt = np.float64 if f64 else np.float32
    tdt = torch.float64 if f64 else torch.float32
    X = wl.points_normal(31, P, n).astype(dt)
    dX = torch.from_numpy(X).cuda()
    res, plain = lcv.laplacian_batch(hip, bufs, dX, S, lcv.DIR_GAUSSIAN, seed, net, with_plain=True)
    V = philox.direction(philox.DIR_GAUSSIAN, seed, np.arange(P * S * n, dtype=np.uint64)).reshape(P * S, n).astype(dt)
    J = np.empty((P * S, 3), dt)
    dJ = torch.empty(P * S, 3, dtype=tdt, device="cuda")
    hip.taylor_kth_batch_dev(bufs, torch.from_numpy(np.repeat(X, S, axis=0)).cuda(), n,
                             torch.from_numpy(V).cuda(), P * S, 3, dJ)
    torch.cuda.synchronize()
    J[:] = dJ.cpu().numpy()
    # the pass keeps L(x) in bfloat16 (the network's FP32 output rounded)
    L = _bf16(net.forward(dX.float()).cpu().numpy()).astype(np.float64).reshape(P, n, rank)
    sign = np.where(np.arange(rank) < npos, 1.0, -1.0)
    for p in range(P):
        v = V[p * S:(p + 1) * S].astype(np.float64)
        u = v @ L[p]
        q = (u * u) @ sign
        lnorm = float(np.sum(L[p] * L[p] * sign))
        base = 2.0 * J[p * S:(p + 1) * S, 2].astype(np.float64)
        samples = base - (q - lnorm)
        mag = float(np.mean(np.abs(base) + np.abs(q) + abs(lnorm)))
        m, var = _welford(samples)
        pm, pvar = _welford(base)
        tm = 1e-11 * mag + (0.0 if f64 else 1e-6 * float(np.mean(np.abs(base))))
        tv = 1e-9 if f64 else 1e-6
        assert res[p, 4] == S and plain[p, 4] == S
        assert abs(res[p, 1] - m) <= tm, (p, res[p], m, mag)
        assert abs(res[p, 2] - var) <= tv * var + 1e-20 * mag * mag, (p, res[p], var)
        assert abs(plain[p, 1] - pm) <= tm and abs(plain[p, 2] - pvar) <= tv * pvar + 1e-20 * mag * mag
        assert res[p, 0] == plain[p, 0] == J[p * S, 0]


def test_training_removes_variance_on_a_low_rank_hessian(hip):
    """softplus ridge with 4 units: H(x) has rank 4 and is PSD, so a rank-8
    L L^T can absorb it; the trained control variate keeps a small fraction
    of the plain variance and the estimate stays on the exact trace."""
    tape = wl.softplus_ridge(16, 4, 53)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    est = lcv.LearnedStde(hip, bufs, rank=8, hidden=64, seed=7)
    est.fit(ntuples=16384, epochs=6, batch=256, lr=3e-3, val_points=128, val_S=32)
    assert est.history[-1] < est.history[0]
    assert est.use_cv and est.variance_ratio < 0.25, est.variance_ratio
    X64 = torch.from_numpy(wl.points_normal(24, 32, 16)).cuda()
    bufs64 = hip.upload_tape(GpuTapeData.from_tape_f64(tape))
    exact = lcv.hessian_trace_batch(hip, bufs64, X64)
    res = est.estimate(X64.float(), 64, seed=12)
    z = (res[:, 1] - exact) / np.maximum(res[:, 3], 1e-12)
    assert abs(np.mean(z)) < 4.0 / np.sqrt(32) + 0.2


def test_fallback_to_plain_when_the_network_hurts(hip):
    tape = wl.tanh_mlp(12, 16, 54)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    est = lcv.LearnedStde(hip, bufs, rank=4, hidden=16, seed=8)
    est.fit(ntuples=256, epochs=0, out_scale=30.0, val_points=32, val_S=16)
    assert est.variance_ratio > 1.0 and not est.use_cv
    X = torch.from_numpy(wl.points_normal(25, 8, 12).astype(np.float32)).cuda()
    a = est.estimate(X, 16, seed=3)
    b = lcv.laplacian_batch(hip, bufs, X, 16, lcv.DIR_GAUSSIAN, 3, None)
    assert np.array_equal(a, b)


def test_cv_errors(hip):
    with pytest.raises(GpuError, match="rank must be 1..=32"):
        lcv.LearnedControlVariate(hip, 4, 33, 8)
    with pytest.raises(GpuError, match="npos must be <= rank"):
        lcv.LearnedControlVariate(hip, 4, 3, 8).set_signs(4)
    tape = wl.tanh_mlp(5, 6, 55)
    bufs = hip.upload_tape(GpuTapeData.from_tape(tape))
    net = lcv.LearnedControlVariate(hip, 6, 2, 8)
    X = torch.zeros(3, 5, device="cuda")
    with pytest.raises(GpuError, match="built for n=6"):
        lcv.laplacian_batch(hip, bufs, X, 4, lcv.DIR_GAUSSIAN, 0, net)
    with pytest.raises(GpuError, match="device-drawn directions"):
        lcv.laplacian_batch(hip, bufs, X, 4, 0, 0, None)
