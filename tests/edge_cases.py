"""Edge-case known answers from the reference's GPU regression tests, as one
table used by the CPU oracle tests and the GPU tests.

Sources: tests/phase7_gpu_fixes.rs (atan' at large |a|, div' at a tiny
denominator, hypot with Inf and at large magnitudes, upload with empty
output_indices, expm1/ln1p at small |a|, powi n=1 HVP at 0, negative-base
power HVPs), tests/gpu_cuda_kernel_parity.rs:20-163 (atan2 at large
magnitudes, asinh'/acosh' at 1e20, powf on a negative base, max with a NaN
operand, fract of a negative input), tests/gpu_stde.rs:1853-2011 (Taylor
fract at a negative input, hypot(t^2, 0), powi with a zero base and an
exponent above the order, powf at a zero base) and
tests/tape_optimization.rs:560-650 (x/x, x*0, x-x and x+(+-0) recorded at a
regular point and replayed at a singular one).

Each case: (name, tape builder, kind, x, v, K, check).  `kind` is forward,
gradient, hvp (one x, one direction) or jet (order K along v).  `check`
takes the result (forward: y; gradient: (y, g); hvp: (g, hv); jet: the K
coefficients of output 0) and the scalar type ("f32" or "f64") and asserts
the reference's expectation; the tolerance for f32 is the reference's own.
"""

from __future__ import annotations

import math

import numpy as np

from echidna_amd.tape import BReverse, record

C = BReverse.constant
INF = math.inf
NAN = math.nan


def _t(f, x0):
    return lambda: record(f, x0)[0]


def _rel(got, want):
    return abs(float(got) - want) / abs(want)


def _atan_large(r, dt):
    g = r[1][0]
    assert np.isfinite(g) and g > 0, g
    if dt == "f64":
        assert _rel(g, 1.0 / (1.0 + 1e40)) <= 1e-12


def _atan_3e9(r, dt):
    g = r[1][0]
    assert np.isfinite(g) and g > 0
    assert _rel(g, 1.0 / (1.0 + 9e18)) < (1e-3 if dt == "f32" else 1e-12)


def _div_tiny(r, dt):
    g = r[1]
    assert np.isfinite(g[1]), g
    assert _rel(g[1], -1e30) < (1e-5 if dt == "f32" else 1e-12)
    assert _rel(g[0], 1e20) < (1e-5 if dt == "f32" else 1e-12)


def _pos_inf(r, dt):
    assert np.isinf(r[0]) and r[0] > 0, r


def _hypot_large(r, dt):
    y, g = r
    assert np.isfinite(y[0]) and _rel(y[0], math.sqrt(2.0) * 1e20) < (1e-4 if dt == "f32" else 1e-15)
    assert np.all(np.isfinite(g)) and all(_rel(v, math.sqrt(0.5)) < 1e-6 for v in g)


def _six(r, dt):
    y, g = r
    assert len(y) == 1 and abs(y[0] - 6.0) < 1e-5 and abs(g[0] - 5.0) < 1e-5


def _expm1_small(r, dt):
    assert abs(float(r[0][0]) - math.expm1(1e-6)) < 1e-12


def _ln1p_small(r, dt):
    assert abs(float(r[0][0]) - math.log1p(1e-6)) < 1e-12


def _powi1_hvp(r, dt):
    g, hv = r
    assert g[0] == 1.0 and np.isfinite(hv[0]) and hv[0] == 0.0


def _neg_base_hvp(r, dt):
    g, hv = r
    assert np.isfinite(g[0]) and abs(g[0] - 12.0) < 1e-3
    assert np.isfinite(hv[0]) and abs(hv[0] + 12.0) < 1e-3


def _atan2_large(r, dt):
    g = r[1]
    assert np.all(np.isfinite(g)) and g[0] != 0 and g[1] != 0
    assert _rel(g[0], 0.5e-20) < 1e-5 and _rel(g[1], -0.5e-20) < 1e-5


def _inv_1e20(r, dt):
    g = r[1][0]
    assert np.isfinite(g) and _rel(g, 1e-20) < 1e-5


def _grad12(r, dt):
    g = r[1][0]
    assert np.isfinite(g) and _rel(g, 12.0) < 1e-5


def _max_nan(r, dt):
    g = r[1]
    assert g[0] == 1.0 and g[1] == 0.0


def _fract_neg(r, dt):
    y = float(r[0])
    assert y < 0 and abs(y - (-1.3 - math.trunc(-1.3))) < 1e-6


def _jet_fract_neg(c, dt):
    assert abs(c[0] - (-0.3)) < 1e-5 and abs(c[1] - 1.0) < 1e-6


def _jet_hypot_depth2(c, dt):
    assert all(abs(a - b) < 1e-6 for a, b in zip(c, [0.0, 0.0, 1.0, 0.0, 0.0])), c


def _jet_powi_zero10(c, dt):
    assert all(v == 0.0 for v in c), c


def _jet_powf_zero(c, dt):
    assert all(v == 0.0 for v in c[:3]) and all(np.isinf(v) and v > 0 for v in c[3:]), c


def _nan_value_and_grad(r, dt):
    assert np.isnan(r[0][0]) and np.isnan(r[1][0]), r


def _nan(r, dt):
    assert np.isnan(r[0]), r


def _pos_zero(r, dt):
    assert r[0] == 0.0 and not np.signbit(r[0]), r


def _neg_zero(r, dt):
    assert r[0] == 0.0 and np.signbit(r[0]), r


CASES = [
    ("atan_grad_1e20", _t(lambda v: v[0].atan(), [1.0]), "gradient", [1e20], None, 0, _atan_large),
    ("atan_grad_3e9", _t(lambda v: v[0].atan(), [1.0]), "gradient", [3e9], None, 0, _atan_3e9),
    ("div_grad_tiny_denominator", _t(lambda v: v[0] / v[1], [1.0, 1.0]), "gradient", [1e-10, 1e-20], None, 0,
     _div_tiny),
    ("hypot_inf_1", _t(lambda v: v[0].hypot(v[1]), [1.0, 1.0]), "forward", [INF, 1.0], None, 0, _pos_inf),
    ("hypot_inf_inf", _t(lambda v: v[0].hypot(v[1]), [1.0, 1.0]), "forward", [INF, INF], None, 0, _pos_inf),
    ("hypot_grad_1e20", _t(lambda v: v[0].hypot(v[1]), [1.0, 1.0]), "gradient", [1e20, 1e20], None, 0,
     _hypot_large),
    ("empty_output_indices", "empty_outputs", "gradient", [2.0], None, 0, _six),
    ("expm1_small", _t(lambda v: v[0].exp_m1(), [1.0]), "gradient", [1e-6], None, 0, _expm1_small),
    ("ln1p_small", _t(lambda v: v[0].ln_1p(), [1.0]), "gradient", [1e-6], None, 0, _ln1p_small),
    ("powi1_hvp_at_0", _t(lambda v: v[0].powi(1), [0.0]), "hvp", [0.0], [1.0], 0, _powi1_hvp),
    ("powf3_hvp_negative_base", _t(lambda v: v[0].powf(C(3.0)), [-2.0]), "hvp", [-2.0], [1.0], 0, _neg_base_hvp),
    ("powi3_hvp_negative_base", _t(lambda v: v[0].powi(3), [-2.0]), "hvp", [-2.0], [1.0], 0, _neg_base_hvp),
    ("atan2_grad_1e20", _t(lambda v: v[0].atan2(v[1]), [1.0, 1.0]), "gradient", [1e20, 1e20], None, 0,
     _atan2_large),
    ("asinh_grad_1e20", _t(lambda v: v[0].asinh(), [1.0]), "gradient", [1e20], None, 0, _inv_1e20),
    ("acosh_grad_1e20", _t(lambda v: v[0].acosh(), [2.0]), "gradient", [1e20], None, 0, _inv_1e20),
    ("powf3_grad_negative_base", _t(lambda v: v[0].powf(C(3.0)), [-2.0]), "gradient", [-2.0], None, 0, _grad12),
    ("max_nan_operand", _t(lambda v: v[0].max(v[1]), [1.5, NAN]), "gradient", [1.5, NAN], None, 0, _max_nan),
    ("fract_negative", _t(lambda v: v[0].fract(), [-1.3]), "forward", [-1.3], None, 0, _fract_neg),
    ("jet_fract_negative", _t(lambda v: v[0].fract(), [-2.3]), "jet", [-2.3], [1.0], 3, _jet_fract_neg),
    ("jet_hypot_depth_two", _t(lambda v: (v[0] * v[0]).hypot(v[1]), [0.0, 0.0]), "jet", [0.0, 0.0], [1.0, 0.0], 5,
     _jet_hypot_depth2),
    ("jet_powi_zero_base_10", _t(lambda v: v[0].powi(10), [0.0]), "jet", [0.0], [1.0], 5, _jet_powi_zero10),
    ("jet_powf_zero_base_2_5", _t(lambda v: v[0].powf(C(2.5)), [1.0]), "jet", [0.0], [1.0], 5, _jet_powf_zero),
    # recorded identities replayed at singular points (tests/tape_optimization.rs:560-650)
    ("div_self_replayed_at_0", _t(lambda v: v[0] / v[0], [2.0]), "gradient", [0.0], None, 0, _nan_value_and_grad),
    ("mul_zero_replayed_at_inf", _t(lambda v: v[0] * 0.0, [3.0]), "forward", [INF], None, 0, _nan),
    ("sub_self_replayed_at_inf", _t(lambda v: v[0] - v[0], [3.0]), "forward", [INF], None, 0, _nan),
    ("add_pos_zero_replayed_at_neg_zero", _t(lambda v: v[0] + 0.0, [3.0]), "forward", [-0.0], None, 0, _pos_zero),
    ("add_neg_zero_replayed_at_neg_zero", _t(lambda v: v[0] + (-0.0), [3.0]), "forward", [-0.0], None, 0, _neg_zero),
]


def empty_outputs_data(f64: bool):
    """GpuTapeData of x*x + x with output_indices cleared (phase7 M24): the
    upload must fall back to [output_index]."""
    from echidna_amd.tape import GpuTapeData

    tape = record(lambda v: v[0] * v[0] + v[0], [2.0])[0]
    d = GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape)
    d.output_indices = np.zeros(0, np.uint32)
    return tape, d


# Unary ops x edge inputs (tests/mode_parity_matrix.rs:26-93).
EDGES = [0.0, -0.0, 1.0, -1.0, 0.5, 2.0, INF, -INF, NAN, 1e-300, 1e300]
UNARY = ["sqrt", "cbrt", "exp", "exp2", "exp_m1", "ln", "ln_1p", "log2", "log10", "sin", "cos", "tan", "asin",
         "acos", "atan", "sinh", "cosh", "tanh", "asinh", "acosh", "atanh", "recip", "abs", "signum", "floor", "ceil",
         "round", "trunc", "fract"]


def unary_tape(name: str):
    return record(lambda v: getattr(v[0], name)(), [0.7])[0]
