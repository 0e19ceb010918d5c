"""Pins the CPU oracle against the reference's known answers, closed forms and
invariants (CPU only).  Each check is chosen so a dropped term, a wrong sign or
a transposed operand in the restatement fails it.
"""

import itertools
import math

import numpy as np
import pytest

import pyoracle as po
from echidna_amd import philox
from echidna_amd import workloads as wl
from echidna_amd.tape import BReverse, record

C = BReverse.constant


def test_philox_known_answers():
    """Random123 philox4x32-10 KAT vectors."""
    assert [int(v) for v in philox.philox4x32_10(0, 0, 0, 0, 0, 0)] == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    f = 0xFFFFFFFF
    assert [int(v) for v in philox.philox4x32_10(f, f, f, f, f, f)] == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    r = philox.philox4x32_10(0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344, 0xA4093822, 0x299F31D0)
    assert [int(v) for v in r] == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_philox_distributions():
    u = philox.uniform(0, 1, np.arange(200000))
    assert 0 < u.min() and u.max() < 1 and abs(u.mean() - 0.5) < 0.005
    z = philox.normal(0, 1, np.arange(200000))
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    r = philox.rademacher(0, 1, np.arange(100000))
    assert set(np.unique(r)) == {-1.0, 1.0} and abs(r.mean()) < 0.02


def test_direction_blocks():
    """Device direction components (philox.direction): Rademacher component
    128 b + k is bit k of block b's 128 bits (word k // 32, -1 when set),
    from the KAT-pinned generator; Gaussian components 4k .. 4k + 3 are block
    k's two Box-Muller pairs: each pair's squares sum to -2 ln u_r and its
    angle is 2 pi u_a, with u = (top 23 bits + 1/2) 2^-23 of words 0, 1 and
    2, 3.  Then moments: mean 0, variance 1, pairs and neighbouring signs
    uncorrelated."""
    seed = 21
    for b in (0, 1, 7, 2 ** 33 + 5):
        w = [int(v) for v in philox.philox4x32_10(b & 0xFFFFFFFF, b >> 32, 0, 0, seed, philox.STREAM_DIRS)]
        got = philox.direction(philox.DIR_RADEMACHER, seed, np.arange(128, dtype=np.uint64) + np.uint64(128 * b))
        want = [-1.0 if (w[k // 32] >> (k % 32)) & 1 else 1.0 for k in range(128)]
        assert list(got) == want, b
    k = np.array([0, 1, 5, 2 ** 34 + 3], dtype=np.uint64)
    g = philox.direction(philox.DIR_GAUSSIAN, seed, (4 * k[:, None] + np.arange(4, dtype=np.uint64)).ravel())
    g = g.reshape(-1, 4)
    w = philox._words(seed, philox.STREAM_DIRS, k)
    for j in (0, 1):
        ur = ((w[2 * j] >> 9).astype(np.float64) + 0.5) * 2.0 ** -23
        ua = ((w[2 * j + 1] >> 9).astype(np.float64) + 0.5) * 2.0 ** -23
        a, c = g[:, 2 * j], g[:, 2 * j + 1]
        assert np.allclose(a ** 2 + c ** 2, -2 * np.log(ur), rtol=2e-6)
        assert np.allclose(np.mod(np.arctan2(c, a), 2 * np.pi), 2 * np.pi * ua, rtol=2e-6, atol=2e-6)
    assert np.all(g.astype(np.float32).astype(np.float64) == g)
    z = philox.direction(philox.DIR_GAUSSIAN, 3, np.arange(2000000, dtype=np.uint64))
    assert abs(z.mean()) < 0.004 and abs(z.var() - 1) < 0.005
    assert abs(np.mean(z[0::2] * z[1::2])) < 0.005 and abs(np.mean(z[0::4] * z[2::4])) < 0.007
    r = philox.direction(philox.DIR_RADEMACHER, 3, np.arange(400000, dtype=np.uint64))
    assert set(np.unique(r)) == {-1.0, 1.0} and abs(r.mean()) < 0.008 and abs(np.mean(r[1:] * r[:-1])) < 0.008
    L = philox.lane_directions(philox.DIR_GAUSSIAN, 3, 10, [0, 7])
    assert np.array_equal(L[1], z[70:80]) and np.array_equal(L[0], z[:10])


def test_direction_box_muller_operations():
    """The Gaussian directions' ln and (cos, sin) of 2 pi u, computed from
    basic binary32 operations (philox.dir_log, dir_cos_sin, the device's
    order), over every uniform the draws produce ((m + 1/2) 2^-23): ln
    within 2 ulp of libm's (binary64 log rounded), cos and sin within 1.2e-7
    of libm's; ln 1 = 0 and ln 1/2 = -ln 2 to the last bit; the octant edges;
    and the grid's moments E[-2 ln u] = 2 and E[cos^2] = 1/2 (the Box-Muller
    variance) to 1e-7."""
    m = np.arange(2 ** 23, dtype=np.uint64)
    u = (m.astype(np.float64) + 0.5) * 2.0 ** -23
    u32 = u.astype(np.float32)
    assert np.array_equal(u32.astype(np.float64), u)
    assert np.array_equal(philox._u23((m << np.uint64(9)).astype(np.uint32) | np.uint32(0x1FF)), u32)
    lg = philox.dir_log(u32).astype(np.float64)
    ref = np.log(u)
    assert np.all(np.abs(lg - ref) <= 2 * np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64))
    assert philox.dir_log(np.array([1.0], np.float32))[0] == 0.0
    assert philox.dir_log(np.array([0.5], np.float32))[0] == np.float32(-math.log(2.0))
    c, sn = philox.dir_cos_sin(u32)
    th = 2.0 * np.pi * u
    assert np.max(np.abs(c - np.cos(th))) <= 1.2e-7 and np.max(np.abs(sn - np.sin(th))) <= 1.2e-7
    for q in range(8):  # octant edges: 8u = q exactly
        cq, sq = philox.dir_cos_sin(np.array([q / 8.0], np.float32))
        assert abs(cq[0] - math.cos(q * math.pi / 4)) <= 1.2e-7 and abs(sq[0] - math.sin(q * math.pi / 4)) <= 1.2e-7
    assert abs(np.mean(-2.0 * lg) - 2.0) < 1e-7 and abs(np.mean(c.astype(np.float64) ** 2) - 0.5) < 1e-7
    # the radius a planted ln error would move: sqrt(-2 ln u) against libm's
    rad = np.sqrt(np.float32(-2) * philox.dir_log(u32)).astype(np.float64)
    assert np.max(np.abs(rad - np.sqrt(-2.0 * ref)) / np.sqrt(-2.0 * ref)) < 3e-7


def test_rosenbrock_gradient_closed_form():
    tape = wl.rosenbrock(2)
    pts = np.array([[-1.2, 1.0], [1.0, 1.0], [0.0, 0.0], [0.3, -0.7]])
    y, g = po.gradient(tape, pts)
    for (x0, x1), yy, gg in zip(pts, y[:, 0], g):
        assert math.isclose(yy, (1 - x0) ** 2 + 100 * (x1 - x0 ** 2) ** 2, rel_tol=1e-15, abs_tol=1e-15)
        assert np.allclose(gg, [-2 * (1 - x0) - 400 * x0 * (x1 - x0 ** 2), 200 * (x1 - x0 ** 2)], rtol=1e-14, atol=1e-13)


def test_rosenbrock_hessian_closed_form():
    """tests/hessian.rs:247 analytic Rosenbrock Hessian."""
    tape = wl.rosenbrock(2)
    x0, x1 = -1.2, 1.0
    H = po.hessian(tape, [x0, x1])
    assert np.allclose(H, [[1200 * x0 ** 2 - 400 * x1 + 2, -400 * x0], [-400 * x0, 200.0]], rtol=1e-14)


def test_jet_closed_form():
    """tests/gpu_cuda_tests.rs:327-372: x^2+y^2 at (3,4) along (1,0) -> 25, 6, 1, 0."""
    tape, _ = record(lambda v: v[0] * v[0] + v[1] * v[1], [3.0, 4.0])
    assert list(po.taylor(tape, [3.0, 4.0], [[1.0, 0.0]], K=5)[0, 0]) == [25.0, 6.0, 1.0, 0.0, 0.0]


def test_stde_core_known_answers():
    """tests/stde_core.rs:61-110."""
    tape, _ = record(lambda v: v[0] * v[0] * v[1] + v[1] * v[1] * v[1], [1.0, 2.0, 3.0])
    assert np.array_equal(po.hessian(tape, [1.0, 2.0, 3.0]), [[4, 2, 0], [2, 12, 0], [0, 0, 0]])
    dirs = np.array(list(itertools.product([-1.0, 1.0], repeat=3)))
    r = po.laplacian_with_stats(tape, [1.0, 2.0, 3.0], dirs)
    assert r["estimate"] == 16.0 and r["value"] == 10.0 and r["num_samples"] == 8
    val, diag = po.hessian_diagonal(tape, [1.0, 2.0, 3.0])
    assert val == 10.0 and list(diag) == [4.0, 12.0, 0.0]
    # sum x_i^2 -> Laplacian 2n; xy -> 0
    t2, _ = record(lambda v: v[0] * v[0] + v[1] * v[1] + v[2] * v[2], [1.0, 1.0, 1.0])
    assert po.laplacian_with_stats(t2, [0.3, 0.1, 2.0], dirs)["estimate"] == 6.0
    t3, _ = record(lambda v: v[0] * v[1], [1.0, 1.0])
    d2 = np.array(list(itertools.product([-1.0, 1.0], repeat=2)))
    assert po.laplacian_with_stats(t3, [0.3, 0.1], d2)["estimate"] == 0.0


def test_control_variate_exact_diag_gives_zero_variance_for_diagonal_hessian():
    """laplacian_with_control: with the exact diagonal of a separable f, residuals vanish."""
    tape, _ = record(lambda v: v[0] * v[0] * 3.0 + v[1].sin(), [1.0, 2.0])
    x = [0.4, 1.1]
    _, diag = po.hessian_diagonal(tape, x)
    dirs = philox.normal(3, 7, np.arange(64)).reshape(32, 2)
    r = po.laplacian_with_control(tape, x, dirs, diag)
    assert abs(r["estimate"] - (6.0 - math.sin(1.1))) < 1e-13 and r["sample_variance"] < 1e-25


def test_welford():
    s = np.array([1.0, 2.0, 4.0, np.nan, np.inf, 7.0])
    est, var, se, n = po.welford(s)
    ref = np.array([1.0, 2.0, 4.0, 7.0])
    assert n == 4 and math.isclose(est, ref.mean()) and math.isclose(var, ref.var(ddof=1))
    assert math.isclose(se, math.sqrt(ref.var(ddof=1) / 4))
    est, var, se, n = po.welford(np.array([np.nan]))
    assert n == 0 and math.isnan(est) and var == 0 and se == 0


def test_hvp_matches_finite_differences():
    tape = wl.tanh_mlp(8, 5, 11)
    x = philox.normal(1, 1, np.arange(8))
    v = philox.normal(1, 2, np.arange(8))
    g, hv = po.hvp(tape, x, v)
    eps = 1e-6
    _, gp = po.gradient(tape, (x + eps * v)[None])
    _, gm = po.gradient(tape, (x - eps * v)[None])
    assert np.allclose(hv[0], (gp[0] - gm[0]) / (2 * eps), rtol=1e-6, atol=1e-8)
    assert np.array_equal(g[0], po.gradient(tape, x[None])[1][0])


def test_gradient_matches_finite_differences_dag():
    tape = wl.random_dag(3, 5, 200, 6)
    x = philox.uniform_range(4, 1, np.arange(6), -1, 1)
    _, g = po.gradient(tape, x[None])
    eps = 1e-6
    for i in range(6):
        e = np.zeros(6)
        e[i] = eps
        fd = (po.forward(tape, (x + e)[None])[0, 0] - po.forward(tape, (x - e)[None])[0, 0]) / (2 * eps)
        assert abs(fd - g[0, i]) <= 1e-5 * max(1.0, abs(g[0, i]))


def test_hessian_symmetry_and_jets_agree_with_hvp():
    tape = wl.tanh_mlp(6, 4, 3)
    x = philox.normal(5, 1, np.arange(6))
    H = po.hessian(tape, x)
    assert np.allclose(H, H.T, rtol=1e-13, atol=1e-15)
    v = philox.normal(5, 2, np.arange(6))
    c2 = po.taylor(tape, x, v[None], K=3)[0, 0, 2]
    assert math.isclose(2 * c2, v @ H @ v, rel_tol=1e-12, abs_tol=1e-14)


def test_batch_equals_single_and_parallel_equals_serial():
    """tests/batch.rs:14-36 and tests/parallel.rs."""
    tape = wl.tanh_mlp(8, 6, 2)
    X = philox.normal(3, 1, np.arange(40 * 8)).reshape(40, 8)
    y, g = po.gradient(tape, X)
    y4, g4 = po.gradient(tape, X, nthreads=4)
    assert np.array_equal(y, y4) and np.array_equal(g, g4)
    for b in (0, 17, 39):
        yb, gb = po.gradient(tape, X[b:b + 1])
        assert np.array_equal(yb[0], y[b]) and np.array_equal(gb[0], g[b])


def test_conventions():
    """Domain NaN, zero-multiplier, abs kink and max/min NaN routing (Appendix A)."""
    t, _ = record(lambda v: v[0].ln(), [1.0])
    assert math.isnan(po.gradient(t, [[-2.0]])[1][0, 0])
    assert po.gradient(t, [[0.0]])[1][0, 0] == math.inf
    t, _ = record(lambda v: v[0].abs(), [1.0])
    assert po.gradient(t, [[0.0]])[1][0, 0] == 0.0 and po.gradient(t, [[-0.0]])[1][0, 0] == 0.0
    t, _ = record(lambda v: v[0].max(v[1]), [1.0, 2.0])
    y, g = po.gradient(t, [[1.0, math.nan]])
    assert y[0, 0] == 1.0 and list(g[0]) == [1.0, 0.0]
    # zero multiplier absorbs an Inf adjoint: d/dx [sqrt(x*0)] at x=1
    t, _ = record(lambda v: (v[0] * 0.0).sqrt(), [1.0])
    assert po.gradient(t, [[1.0]])[1][0, 0] == 0.0
    # atanh outside its domain -> NaN partial; acosh(a<1) NaN
    t, _ = record(lambda v: v[0].atanh(), [0.5])
    assert math.isnan(po.gradient(t, [[1.5]])[1][0, 0])
    # signum(+-0) = +-1, fract by truncation, round half away from zero
    t, _ = record(lambda v: v[0].signum(), [1.0])
    assert po.forward(t, [[-0.0]])[0, 0] == -1.0
    t, _ = record(lambda v: v[0].fract(), [1.5])
    assert po.forward(t, [[-1.25]])[0, 0] == -0.25
    t, _ = record(lambda v: v[0].round(), [1.5])
    assert po.forward(t, [[-2.5]])[0, 0] == -3.0


def test_powi_partial_closed_forms():
    """src/opcode.rs:541-560 via the tape: d/dx x^n."""
    for n, x, want in [(3, 2.0, 12.0), (-1, 2.0, -0.25), (2, 0.0, 0.0)]:
        t, _ = record(lambda v, n=n: v[0].powi(n), [1.5])
        assert po.gradient(t, [[x]])[1][0, 0] == want


def test_remainder_known_answer():
    """tests/phase7_gpu_parity.rs:77-96: 5 % 2 = 1, d = (1, -2)."""
    t, _ = record(lambda v: v[0] % v[1], [5.0, 2.0])
    y, g = po.gradient(t, [[5.0, 2.0]])
    assert y[0, 0] == 1.0 and list(g[0]) == [1.0, -2.0]


def test_asinh_acosh_large_argument_derivatives():
    """tests/gpu_cuda_kernel_parity.rs: asinh'/acosh' at 1e20 ~ 1e-20; powf negative base -> 12."""
    t, _ = record(lambda v: v[0].asinh(), [1.0])
    assert math.isclose(po.gradient(t, [[1e20]])[1][0, 0], 1e-20, rel_tol=1e-12)
    t, _ = record(lambda v: v[0].acosh(), [2.0])
    assert math.isclose(po.gradient(t, [[1e20]])[1][0, 0], 1e-20, rel_tol=1e-12)
    t, _ = record(lambda v: v[0].powf(C(3.0)), [2.0])
    assert po.gradient(t, [[-2.0]])[1][0, 0] == 12.0


@pytest.mark.parametrize("K", [1, 2, 3, 4, 5])
def test_taylor_orders_exp_series(K):
    """exp jet along v=1 at 0: c_k = 1/k!."""
    t, _ = record(lambda v: v[0].exp(), [0.0])
    c = po.taylor(t, [0.0], [[1.0]], K=K)[0, 0]
    assert np.allclose(c, [1 / math.factorial(k) for k in range(K)], rtol=1e-15)


def test_magnitudes_closed_form_mlp():
    """The magnitude sweep on f = sum_j w2_j tanh(sum_i W1_ji x_i + b_j):
    gradient terms |w2_j (1 - t_j^2) W1_ji| summed over j; H v terms
    |w2_j| |2 t_j (1 - t_j^2)| |sum_i W1_ji v_i|^abs |W1_ji| with the inner
    sum taken over absolute values; the value bound is at least |f|."""
    tape = wl.tanh_mlp(5, 4, 3)
    W1, b, w2 = wl.mlp_weights(5, 4, 3)
    x = philox.normal(2, 1, np.arange(5))
    v = philox.normal(2, 2, np.arange(5))
    m = po.magnitudes(tape, x[None], v[None])
    a = W1 @ x + b
    t = np.tanh(a)
    d1 = 1 - t * t
    gmag = np.abs(w2 * d1) @ np.abs(W1)
    assert np.allclose(m["grad"][0], gmag, rtol=1e-12)
    tv = np.abs(W1) @ np.abs(v)
    hmag = (np.abs(w2) * np.abs(2 * t * d1) * tv) @ np.abs(W1)
    assert np.allclose(m["hv"][0], hmag, rtol=1e-12)
    assert np.isclose(m["c1"][0], np.dot(np.abs(w2 * d1), tv), rtol=1e-12)
    assert np.isclose(m["c2"][0], 0.5 * np.dot(np.abs(v), hmag), rtol=1e-12)
    y = po.forward(tape, x[None])[0, 0]
    assert m["value"][0, 0] >= abs(y)
    _, g = po.gradient(tape, x[None])
    _, hv = po.hvp(tape, x[None], v[None])
    assert np.all(m["grad"][0] >= np.abs(g[0])) and np.all(m["hv"][0] >= np.abs(hv[0]) * (1 - 1e-12))
