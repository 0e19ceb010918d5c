"""CPU <-> GPU parity table, ported from the reference's PARITY_CASES
(tests/gpu_cpu_parity.rs:208-644): tape builders, evaluation points and the
per-case ULP budgets (f32 GPU vs f64 CPU cast to f32; f64 GPU vs f64 CPU),
plus the reference's ULP distance (tests/gpu_cpu_parity.rs:646-696).
"""

from __future__ import annotations

import math

import numpy as np

from echidna_amd.tape import BReverse, record

C = BReverse.constant
NAN = math.nan


def _r(f, x0):
    return lambda: record(f, x0)[0]


CASES = [
    ("add", _r(lambda v: v[0] + v[1], [1.0, 1.0]), [[1.0, 1.0], [-3.5, 2.25], [1e10, 1e-10], [0.0, 0.0]], 2, 2),
    ("sub", _r(lambda v: v[0] - v[1], [1.0, 1.0]), [[1.0, 1.0], [-3.5, 2.25], [2.0, -1.5]], 4, 4),
    ("mul", _r(lambda v: v[0] * v[1], [1.0, 1.0]), [[1.0, 1.0], [-2.5, 4.0], [1e10, 1e-10], [0.0, 5.0]], 2, 2),
    ("div", _r(lambda v: v[0] / v[1], [1.0, 1.0]), [[1.0, 2.0], [-6.0, 3.0], [1.0, 1e-3]], 4, 4),
    ("sqrt", _r(lambda v: v[0].sqrt(), [1.0]), [[1.0], [4.0], [0.25], [1e6], [0.0]], 4, 4),
    ("cbrt", _r(lambda v: v[0].cbrt(), [1.0]), [[1.0], [8.0], [-27.0], [0.0]], 8, 8),
    ("one_div_x", _r(lambda v: C(1.0) / v[0], [1.0]), [[1.0], [2.0], [-0.5]], 4, 4),
    ("recip", _r(lambda v: v[0].powi(-1), [2.0]), [[2.0], [-0.5], [0.0]], 4, 4),
    ("neg", _r(lambda v: -v[0], [1.0]), [[1.0], [-3.5], [0.0]], 0, 0),
    ("abs", _r(lambda v: v[0].abs(), [1.0]), [[1.0], [-3.5], [2.0], [0.0]], 0, 0),
    ("exp", _r(lambda v: v[0].exp(), [1.0]), [[0.0], [1.0], [-1.0], [5.0], [1000.0]], 8, 8),
    ("expm1", _r(lambda v: v[0].exp_m1(), [1.0]), [[0.0], [1e-6], [0.5], [-2.0]], 16, 8),
    ("ln", _r(lambda v: v[0].ln(), [1.0]), [[1.0], [2.0], [10.0], [0.5], [0.0]], 8, 8),
    ("ln1p", _r(lambda v: v[0].ln_1p(), [1.0]), [[0.0], [1e-6], [1.0], [-0.5], [-1.0]], 16, 8),
    ("sin", _r(lambda v: v[0].sin(), [1.0]), [[0.0], [0.5], [1.0], [-0.3]], 16, 16),
    ("cos", _r(lambda v: v[0].cos(), [1.0]), [[0.0], [0.5], [1.0], [-0.3]], 16, 16),
    ("tan", _r(lambda v: v[0].tan(), [0.5]), [[0.0], [0.5], [-0.5]], 16, 16),
    ("atan", _r(lambda v: v[0].atan(), [1.0]), [[0.0], [1.0], [-2.5], [100.0]], 8, 8),
    ("atan2", _r(lambda v: v[0].atan2(v[1]), [1.0, 1.0]),
     [[1.0, 1.0], [-1.0, 1.0], [3.0, 4.0], [1e10, 1e10], [0.0, 0.0]], 16, 16),
    ("sinh", _r(lambda v: v[0].sinh(), [1.0]), [[0.0], [1.0], [-2.0], [1000.0]], 16, 16),
    ("cosh", _r(lambda v: v[0].cosh(), [1.0]), [[0.0], [1.0], [-2.0], [1000.0]], 16, 16),
    ("tanh", _r(lambda v: v[0].tanh(), [1.0]), [[0.0], [1.0], [-2.0]], 16, 16),
    ("asinh", _r(lambda v: v[0].asinh(), [1.0]), [[0.0], [1.0], [-3.0], [1e6], [1e20], [-1e20]], 16, 16),
    ("acosh", _r(lambda v: v[0].acosh(), [2.0]), [[1.5], [2.0], [10.0], [1.0], [1e20]], 16, 16),
    ("atanh", _r(lambda v: v[0].atanh(), [0.5]), [[0.0], [0.25], [-0.5], [0.9], [1.0], [-1.0]], 16, 16),
    ("asin", _r(lambda v: v[0].asin(), [0.5]), [[0.0], [0.5], [-0.25], [1.0], [-1.0]], 16, 16),
    ("acos", _r(lambda v: v[0].acos(), [0.5]), [[0.0], [0.5], [-0.25], [1.0], [-1.0]], 16, 16),
    ("exp2", _r(lambda v: v[0].exp2(), [1.0]), [[0.0], [1.0], [-1.0], [3.0]], 8, 8),
    ("log2", _r(lambda v: v[0].log2(), [1.0]), [[1.0], [2.0], [8.0], [0.0]], 8, 8),
    ("log10", _r(lambda v: v[0].log10(), [1.0]), [[1.0], [10.0], [100.0], [0.0]], 8, 8),
    ("powi", _r(lambda v: v[0].powi(3), [2.0]), [[2.0], [-3.0], [0.5], [0.0]], 8, 8),
    ("powf", _r(lambda v: v[0].powf(C(2.5)), [2.0]), [[2.0], [0.5], [10.0]], 32, 16),
    ("powf_int", _r(lambda v: v[0].powf(C(3.0)), [2.0]), [[2.0], [-2.0], [-3.0], [0.0]], 16, 16),
    ("powf_var", _r(lambda v: v[0].powf(v[1]), [2.0, 2.0]), [[2.0, 3.0], [0.0, 0.0], [0.0, 2.0], [0.0, 3.0]], 16, 16),
    ("rem", _r(lambda v: v[0] % v[1], [5.0, 2.0]), [[5.0, 2.0], [7.5, 2.5], [-3.0, 2.0], [5.0, 0.0], [1e6, 3.0]], 4, 4),
    ("hypot", _r(lambda v: v[0].hypot(v[1]), [3.0, 4.0]),
     [[3.0, 4.0], [1e10, 1e10], [1.0, 0.0], [0.0, 1e-6], [0.0, 0.0]], 8, 8),
    ("max", _r(lambda v: v[0].max(v[1]), [1.0, 2.0]),
     [[1.0, 2.0], [-1.0, -2.0], [3.0, 3.0], [NAN, 1.0], [1.0, NAN]], 0, 0),
    ("min", _r(lambda v: v[0].min(v[1]), [1.0, 2.0]),
     [[1.0, 2.0], [-1.0, -2.0], [3.0, 3.0], [NAN, 1.0], [1.0, NAN]], 0, 0),
    ("rosenbrock", _r(lambda v: (C(1.0) - v[0]) * (C(1.0) - v[0]) + C(100.0) * (v[1] - v[0] * v[0]) * (v[1] - v[0] * v[0]),
                      [1.0, 1.0]), [[1.0, 1.0], [0.0, 0.0], [-1.2, 1.0]], 64, 32),
    ("sin_x2_mul_exp_neg_x", _r(lambda v: (v[0] * v[0] + C(1.0)).sin() * (-v[0]).exp(), [1.0]),
     [[0.5], [1.0], [-0.5]], 64, 32),
    ("log_sum_exp", _r(lambda v: (v[0].exp() + v[1].exp()).ln(), [1.0, 0.5]), [[0.0, 0.0], [1.0, -1.0], [3.0, 5.0]], 64, 32),
]


def _rosen_faithful(v):
    # tests/gpu_cpu_parity.rs:188-197 records a, b once and reuses them.
    a = C(1.0) - v[0]
    b = v[1] - v[0] * v[0]
    return a * a + C(100.0) * b * b


CASES = [c if c[0] != "rosenbrock" else ("rosenbrock", _r(_rosen_faithful, [1.0, 1.0]), c[2], c[3], c[4])
         for c in CASES]


def ulp_diff_f32(a: float, b: float) -> int:
    a = np.float32(a)
    b = np.float32(b)
    if not (np.isfinite(a) and np.isfinite(b)):
        if np.isnan(a) and np.isnan(b):
            return 0
        return 0 if a.view(np.uint32) == b.view(np.uint32) else 2**32 - 1
    ab, bb = int(a.view(np.uint32)), int(b.view(np.uint32))
    if (ab ^ bb) & 0x80000000 == 0:
        return abs(ab - bb)
    return min((ab & 0x7FFFFFFF) + (bb & 0x7FFFFFFF), 2**32 - 1)


def ulp_diff_f64(a: float, b: float) -> int:
    a = np.float64(a)
    b = np.float64(b)
    if not (np.isfinite(a) and np.isfinite(b)):
        if np.isnan(a) and np.isnan(b):
            return 0
        return 0 if a.view(np.uint64) == b.view(np.uint64) else 2**64 - 1
    ab, bb = int(a.view(np.uint64)), int(b.view(np.uint64))
    if (ab ^ bb) & 0x8000000000000000 == 0:
        return abs(ab - bb)
    return min((ab & 0x7FFFFFFFFFFFFFFF) + (bb & 0x7FFFFFFFFFFFFFFF), 2**64 - 1)
