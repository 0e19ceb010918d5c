"""Pins every derivative rule of the CPU oracle (CPU only).

The oracle restates three independent sets of rules from the reference: the
reverse partials (src/opcode.rs:331-483, src/kernels/mod.rs:85-290), the
Dual arithmetic that forward-over-reverse HVPs run on (src/dual.rs) and the
Taylor recurrences (src/taylor_ops.rs).  Here each is checked against
values that share none of them:

* `mp_tape` evaluates the tape's primal values in 60-digit arithmetic and
  mpmath differentiates that numerically: gradients, H v and Taylor
  coefficients c_0..c_4 along v, for all 41 parity cases at random smooth
  points and for a tape that uses every opcode at once;
* the oracle's own paths against each other: c1 = grad . v (reverse sweep
  vs Taylor), 2 c2 = v . Hv (Dual HVP vs Taylor), and the K=3 jet equal to
  the first three coefficients of the K=5 jet;
* the reference's own known answers for the Taylor rules
  (tests/taylor.rs:11-160, 732-975; tests/hypot_leading_zero_parity.rs;
  tests/gpu_cuda_tests.rs:327-372).

Bars: against the 60-digit values, |oracle - exact| <= 1e-9 * scale_k with
scale_k = max(|exact_k|, 1e-6 * max_j |exact_j|) (f64 rounding on these
tapes is ~1e-15; a wrong rule is off by O(1)); between the oracle's own
paths, 1e-12 relative to the sum of the magnitudes of the terms, floored at
1e-4 |v|^k max(1, |f|) where the result cancels to about zero.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import mp_tape
import pyoracle as po
from echidna_amd.tape import BReverse, record
from parity_cases import CASES

C = BReverse.constant

# Ranges where each case is smooth (away from kinks, branch points and
# domain edges); inputs not listed use (-2, 2).
DOMAINS = {
    "sqrt": [(0.2, 3.0)], "cbrt": [(0.3, 3.0)], "one_div_x": [(0.3, 3.0)], "recip": [(0.3, 3.0)],
    "div": [(-3.0, 3.0), (0.3, 3.0)], "ln": [(0.2, 5.0)], "log2": [(0.2, 5.0)], "log10": [(0.2, 5.0)],
    "ln1p": [(-0.7, 3.0)], "tan": [(-1.2, 1.2)], "asin": [(-0.8, 0.8)], "acos": [(-0.8, 0.8)],
    "atanh": [(-0.8, 0.8)], "acosh": [(1.2, 5.0)], "powf": [(0.2, 3.0)], "powf_int": [(-3.0, 3.0)],
    "powf_var": [(0.3, 3.0), (0.3, 3.0)], "rem": [(-5.0, 5.0), (0.5, 3.0)], "exp": [(-3.0, 3.0)],
    "sinh": [(-3.0, 3.0)], "cosh": [(-3.0, 3.0)], "log_sum_exp": [(-3.0, 3.0), (-3.0, 3.0)],
}


def _smooth(name, x):
    """Reject points within reach of a case's kink or singular line."""
    if name in ("abs", "cbrt", "one_div_x", "recip") and abs(x[0]) < 0.1:
        return False
    if name in ("max", "min") and abs(x[0] - x[1]) < 0.1:
        return False
    if name in ("atan2", "hypot") and math.hypot(x[0], x[1]) < 0.3:
        return False
    if name == "rem":
        if x[1] == 0:
            return False
        q = x[0] / x[1]
        return abs(q - round(q)) > 0.05 and abs(q) > 0.05
    return True


def _points(name, n, count, seed):
    rng = np.random.default_rng(seed)
    rngs = DOMAINS.get(name, [(-2.0, 2.0)] * n)
    rngs = rngs + [rngs[-1]] * (n - len(rngs))
    out = []
    while len(out) < count:
        x = np.array([rng.uniform(lo, hi) for lo, hi in rngs[:n]])
        if name in ("cbrt", "one_div_x", "recip") and len(out) % 2:
            x = -x
        if _smooth(name, x):
            out.append(x)
    return np.array(out)


def _close(got, exact, tol=1e-9, rel_floor=1e-6):
    got = np.asarray(got, np.float64)
    exact = np.asarray(exact, np.float64)
    scale = np.maximum(np.abs(exact), rel_floor * np.max(np.abs(exact), initial=0.0))
    scale = np.where(scale == 0, 1.0, scale)
    return np.all(np.abs(got - exact) <= tol * scale), np.max(np.abs(got - exact) / scale)


def kitchen_sink():
    """One scalar tape that runs every opcode but Input/Const/Custom, at
    x0 in (0.3, 0.7), x1 in (1.2, 1.8), x2 in (-0.6, -0.2)."""

    def f(v):
        x0, x1, x2 = v
        t1 = x0.sin() * x1.cos() + x2.tan() - x0 / x1
        t2 = x0.asin() + x0.acos() * x2.atan() + x0.atanh()
        t3 = x1.acosh() + x1.asinh() * x2 + x1.sinh() / x1.cosh() + x2.tanh()
        t4 = x1.ln() + x1.log2() * x0 + x1.log10() + x0.ln_1p() + x2.exp_m1() + x2.exp2() + x2.exp()
        t5 = x1.sqrt() * x2.cbrt() + x0.powi(3) + x1.powi(-2) + x1.powf(x0) + x1.powf(C(2.5))
        t6 = x0.atan2(x2) + x1.hypot(x2) + (x1 % x0) + x0.max(x2) + x2.min(x0) + x2.abs() + x1.recip()
        t7 = x0.signum() * x1 + x1.floor() + x1.ceil() + x1.round() + x1.trunc() + x1.fract() * x0 - (-x2)
        return t1 * t2 + t3 * t4 - t5 / t6 + t7 * x2

    return record(f, [0.5, 1.5, -0.4])[0]


SINK_DOMAIN = [(0.3, 0.7), (1.2, 1.8), (-0.6, -0.2)]

ALL = [(c[0], c[1]) for c in CASES] + [("kitchen_sink", kitchen_sink)]


def test_kitchen_sink_uses_every_opcode():
    ops = {int(o) for o in kitchen_sink().opcodes}
    assert ops == set(range(0, 43)), sorted(set(range(43)) - ops)


@pytest.mark.parametrize("name,build", ALL, ids=[a[0] for a in ALL])
def test_derivatives_match_60_digit_differentiation(name, build):
    """Gradient (reverse), H v (Dual) and Taylor c0..c4 (K=5) of the oracle
    against numerical derivatives of the 60-digit primal evaluator."""
    tape = build()
    n = tape.num_inputs
    if name == "kitchen_sink":
        rng = np.random.default_rng(7)
        X = np.array([[rng.uniform(lo, hi) for lo, hi in SINK_DOMAIN] for _ in range(3)])
    else:
        X = _points(name, n, 3, hash(name) % 2**32)
    V = np.random.default_rng(11).normal(size=X.shape)
    _, G = po.gradient(tape, X)
    _, HV = po.hvp(tape, X, V)
    J = po.taylor(tape, X, V, K=5)[:, 0, :]
    for p in range(X.shape[0]):
        ok, e = _close(G[p], mp_tape.gradient(tape, X[p]))
        assert ok, (name, "gradient", X[p], e)
        ok, e = _close(HV[p], mp_tape.hvp(tape, X[p], V[p]))
        assert ok, (name, "hvp", X[p], e)
        ok, e = _close(J[p], mp_tape.jet(tape, X[p], V[p], 5))
        assert ok, (name, "jet", X[p], J[p], e)


@pytest.mark.parametrize("name,build", ALL, ids=[a[0] for a in ALL])
def test_oracle_paths_agree(name, build):
    """c1 = grad . v (reverse vs Taylor), 2 c2 = v . Hv (Dual vs Taylor),
    and the K=3 jet is the first three coefficients of the K=5 jet, at the
    case's own parity points (finite results only) and random smooth
    points."""
    tape = build()
    n = tape.num_inputs
    # the case's own points, less those on a kink: there the reverse rule
    # (abs'(0) = 0) and the Taylor rule (the sign of the first non-zero
    # coefficient) differ on purpose
    pts = [np.asarray(p, np.float64) for p in dict((c[0], c[2]) for c in CASES).get(name, [])]
    pts = [p for p in pts if _smooth(name, p)]
    if name == "kitchen_sink":
        rng = np.random.default_rng(8)
        pts += [np.array([rng.uniform(lo, hi) for lo, hi in SINK_DOMAIN]) for _ in range(4)]
    else:
        pts += list(_points(name, n, 4, 99))
    X = np.array(pts)
    V = np.random.default_rng(5).normal(size=X.shape)
    _, G = po.gradient(tape, X)
    _, HV = po.hvp(tape, X, V)
    J5 = po.taylor(tape, X, V, K=5)[:, 0, :]
    J3 = po.taylor(tape, X, V, K=3)[:, 0, :]
    for p in range(X.shape[0]):
        terms1 = G[p] * V[p]
        terms2 = V[p] * HV[p]
        if not (np.all(np.isfinite(terms1)) and np.all(np.isfinite(terms2)) and np.all(np.isfinite(J5[p, :3]))):
            continue
        # floor for results that cancel to ~0: 1e-4 of |v|^k max(1, |f|)
        f0 = max(1.0, abs(J5[p, 0]))
        vv = float(np.dot(V[p], V[p]))
        s1 = max(np.abs(terms1).sum(), 1e-4 * math.sqrt(vv) * f0)
        s2 = max(np.abs(terms2).sum(), 1e-4 * vv * f0)
        assert abs(J5[p, 1] - terms1.sum()) <= 1e-12 * s1, (name, X[p])
        assert abs(2 * J5[p, 2] - terms2.sum()) <= 1e-12 * s2, (name, X[p])
    same = (J3 == J5[:, :3]) | (np.isnan(J3) & np.isnan(J5[:, :3]))
    assert same.all(), name


# ---------------------------------------------------------------------------
# the reference's own known answers for the Taylor rules
# ---------------------------------------------------------------------------

def _jet1(f, x0, K=5, v=1.0):
    """Coefficients of f(x0 + v t) for a one-input closure."""
    tape = record(lambda a: f(a[0]), [x0])[0]
    return po.taylor(tape, [x0], [[v]], K=K)[0, 0]


SERIES = [
    # tests/taylor.rs:11-70
    ("exp", lambda x: x.exp(), 0.0, [1, 1, 1 / 2, 1 / 6, 1 / 24]),
    ("sin", lambda x: x.sin(), 0.0, [0, 1, 0, -1 / 6, 0]),
    ("cos", lambda x: x.cos(), 0.0, [1, 0, -1 / 2, 0, 1 / 24]),
    ("ln(1+x)", lambda x: (C(1.0) + x).ln(), 0.0, [0, 1, -1 / 2, 1 / 3, -1 / 4]),
    ("1/(1-x)", lambda x: C(1.0) / (C(1.0) - x), 0.0, [1, 1, 1, 1, 1]),
    # tests/taylor.rs:732-810
    ("tan", lambda x: x.tan(), 0.0, [0, 1, 0, 1 / 3, 0]),
    ("sinh", lambda x: x.sinh(), 0.0, [0, 1, 0, 1 / 6, 0]),
    ("cosh", lambda x: x.cosh(), 0.0, [1, 0, 1 / 2, 0, 1 / 24]),
    ("atan", lambda x: x.atan(), 0.0, [0, 1, 0, -1 / 3, 0]),
    ("asin", lambda x: x.asin(), 0.0, [0, 1, 0, 1 / 6, 0]),
    ("tanh", lambda x: x.tanh(), 0.0, [0, 1, 0, -1 / 3, 0]),
    # powi on negative bases, tests/taylor.rs:822-875
    ("(-2+t)^2", lambda x: x.powi(2), -2.0, [4, -4, 1, 0, 0]),
    ("(-2+t)^3", lambda x: x.powi(3), -2.0, [-8, 12, -6, 1, 0]),
    ("(-3+t)^4", lambda x: x.powi(4), -3.0, [81, -108, 54, -12, 1]),
    ("(-2+t)^-1", lambda x: x.powi(-1), -2.0, [-0.5, -0.25, -0.125, -0.0625, -0.03125]),
]


@pytest.mark.parametrize("name,f,x0,want", SERIES, ids=[s[0] for s in SERIES])
def test_known_taylor_series(name, f, x0, want):
    got = _jet1(f, x0)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12), (name, got)


def test_powi_positive_base_large_exponent():
    """tests/taylor.rs:878-887: (2+t)^10 -> 1024, 5120."""
    got = _jet1(lambda x: x.powi(10), 2.0, K=3)
    assert abs(got[0] - 1024.0) <= 1e-8 and abs(got[1] - 5120.0) <= 1e-6


K2_DUAL = [
    # tests/taylor.rs:72-130: K=2 jets equal (f, f')
    ("exp", lambda x: x.exp(), 1.5, math.exp(1.5), math.exp(1.5)),
    ("sin", lambda x: x.sin(), 0.7, math.sin(0.7), math.cos(0.7)),
    ("ln", lambda x: x.ln(), 2.0, math.log(2.0), 0.5),
    ("sqrt", lambda x: x.sqrt(), 4.0, 2.0, 0.25),
    ("atan", lambda x: x.atan(), 0.5, math.atan(0.5), 1 / 1.25),
    ("tanh", lambda x: x.tanh(), 0.3, math.tanh(0.3), 1 / math.cosh(0.3) ** 2),
]


@pytest.mark.parametrize("name,f,x0,val,der", K2_DUAL, ids=[s[0] for s in K2_DUAL])
def test_k2_jets_match_dual(name, f, x0, val, der):
    got = _jet1(f, x0, K=2)
    assert math.isclose(got[0], val, rel_tol=1e-12) and math.isclose(got[1], der, rel_tol=1e-12)
    tape = record(lambda a: f(a[0]), [x0])[0]
    _, g = po.gradient(tape, [[x0]])
    assert math.isclose(g[0, 0], der, rel_tol=1e-12)


def test_cauchy_product_and_recursive_division():
    """tests/taylor.rs:136-160: (1+t)(1+t) = 1 + 2t + t^2; (1+2t+t^2)/(1+t) = 1+t."""
    got = _jet1(lambda x: (C(1.0) + x) * (C(1.0) + x), 0.0, K=3)
    assert list(got) == [1.0, 2.0, 1.0]
    got = _jet1(lambda x: (C(1.0) + x * C(2.0) + x * x) / (C(1.0) + x), 0.0, K=3)
    assert np.allclose(got, [1.0, 1.0, 0.0], atol=1e-12)


def _hypot_tape(a_of, b_of, x0):
    return record(lambda v: a_of(v[0]).hypot(b_of(v[0])), [x0])[0]


def test_hypot_identically_zero_operands():
    """tests/taylor.rs:900-950: hypot of two zero series (literal zeros or
    x - x) is the zero series, not [0, Inf, ...]; hypot(t^2, 0) = t^2."""
    zero = lambda x: x - x  # noqa: E731  (x - x is never folded when recording)
    tape = _hypot_tape(zero, zero, 3.0)
    assert list(po.taylor(tape, [3.0], [[1.0]], K=5)[0, 0]) == [0.0] * 5
    tape = _hypot_tape(lambda x: x * x, zero, 0.0)
    assert list(po.taylor(tape, [0.0], [[1.0]], K=5)[0, 0]) == [0.0, 0.0, 1.0, 0.0, 0.0]


def test_hypot_nan_leading_coefficient_propagates():
    """tests/taylor.rs:955-975: a NaN primal stays NaN through the zero-scale
    branch, with or without a first-order signal."""
    zero = lambda x: x - x  # noqa: E731
    tape = _hypot_tape(lambda x: x + C(math.nan), zero, 0.0)
    assert np.isnan(po.taylor(tape, [0.0], [[0.0]], K=3)[0, 0]).all()
    assert np.isnan(po.taylor(tape, [0.0], [[1.0]], K=3)[0, 0, 0])


def _poly_tape(coeffs_a, coeffs_b):
    """hypot(a(t), b(t)) with a, b polynomials in the input t (recorded at
    t = 0 so that seeding (0, 1, 0, ...) gives exactly these series)."""

    def poly(t, cs):
        acc = t - t
        p = None
        for k, c in enumerate(cs):
            if c == 0:
                continue
            term = C(float(c)) if k == 0 else (t.powi(k) * C(float(c)) if k > 1 else t * C(float(c)))
            acc = acc + term
            p = 1
        return acc if p else t - t

    return record(lambda v: poly(v[0], coeffs_a).hypot(poly(v[0], coeffs_b)), [0.0])[0]


HYPOT_PEEL = [
    # tests/hypot_leading_zero_parity.rs (finite cases, order 6)
    ([0, 0, 1, 2, 0, 0], [0, 0, 0, 1, 0, 0], [0, 0, 1, 2, 0.5, -1.0]),
    ([0, 0, 0, 0, 0, 3], [0] * 6, [0, 0, 0, 0, 0, 3]),
    ([0] * 6, [0] * 6, [0] * 6),
    ([1, 2, 3, 4, 5, 6], [7, 8, 9, 10, 11, 12],
     [7.0710678118654755, 8.20243866176395, 9.384721199907856, 10.60976955617793, 11.870557917596285,
      13.161059969448717]),
]


@pytest.mark.parametrize("a,b,want", HYPOT_PEEL)
def test_hypot_leading_zero_peel(a, b, want):
    """The reference's expected jets to the bit, evaluated through the tape
    at order 5 (the oracle's highest) — the first five of the six."""
    tape = _poly_tape(a, b)
    got = po.taylor(tape, [0.0], [[1.0]], K=5)[0, 0]
    assert np.allclose(got, want[:5], rtol=2e-16, atol=0), (got, want)


def test_jet_of_x2_plus_y2():
    """tests/gpu_cuda_tests.rs:327-372: x^2+y^2 at (3,4) along (1,0)."""
    tape, _ = record(lambda v: v[0] * v[0] + v[1] * v[1], [3.0, 4.0])
    assert list(po.taylor(tape, [3.0, 4.0], [[1.0, 0.0]], K=3)[0, 0]) == [25.0, 6.0, 1.0]


def test_powf_partial_at_zero_base():
    """powf's reverse partial at a zero base takes b a^(b-1)
    (src/opcode.rs:350-381): d/dx x^1.0 = 1 at x = 0, d/dx x^2.0 = 0, and
    d/db 0^b = 0 for b > 0; the same through the Dual (HVP) path."""
    t1 = record(lambda v: v[0].powf(C(1.0)), [0.5])[0]
    t2 = record(lambda v: v[0].powf(C(2.0)), [0.5])[0]
    tv = record(lambda v: v[0].powf(v[1]), [0.5, 1.0])[0]
    assert po.gradient(t1, [[0.0]])[1][0, 0] == 1.0
    assert po.gradient(t2, [[0.0]])[1][0, 0] == 0.0
    assert list(po.gradient(tv, [[0.0, 1.0]])[1][0]) == [1.0, 0.0]
    g, hv = po.hvp(t2, [[0.0]], [[1.0]])
    assert g[0, 0] == 0.0 and hv[0, 0] == 2.0
