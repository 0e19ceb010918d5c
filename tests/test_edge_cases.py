"""The reference's GPU edge-case known answers (tests/edge_cases.py) held by
the CPU oracle, in f64 and f32 (CPU only)."""

import numpy as np
import pytest

import pyoracle as po
from edge_cases import CASES


def run_oracle(kind, tape, x, v, K, dtype):
    X = np.asarray([x], dtype=np.float64)
    if kind == "forward":
        return po.forward(tape, X, dtype=dtype)[0]
    if kind == "gradient":
        y, g = po.gradient(tape, X, dtype=dtype)
        return y[0], g[0]
    if kind == "hvp":
        g, hv = po.hvp(tape, X, np.asarray([v]), dtype=dtype)
        return g[0], hv[0]
    return po.taylor(tape, X, np.asarray([v]), K=K, dtype=dtype)[0, 0]


ORACLE_CASES = [c for c in CASES if c[1] != "empty_outputs"]


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("case", ORACLE_CASES, ids=[c[0] for c in ORACLE_CASES])
def test_oracle_edge_case(case, dt):
    name, build, kind, x, v, K, check = case
    r = run_oracle(kind, build(), x, v, K, np.float64 if dt == "f64" else np.float32)
    check(r, dt)
