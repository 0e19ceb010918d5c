"""Sparsity patterns and colorings (no GPU): the library's host code against
the reference's own expectations (tests/sparse_hessian.rs,
tests/sparse_jacobian.rs, tests/sparse_coloring.rs), bit-exact against the
oracle's independent restatement, and pinned by brute force: every entry a
dense f64 Hessian / Jacobian shows as non-zero at random points is in the
detected pattern, and every coloring separates what it must."""

import numpy as np
import pytest

import pyoracle as po
from echidna_amd import sparse as sp
from echidna_amd import workloads as wl
from echidna_amd.tape import GpuTapeData, record, record_multi


def pattern_of(f, x):
    tape, _ = record(f, x)
    return tape, sp.detect_sparsity(GpuTapeData.from_tape_f64(tape))


# ---- the reference's expectations -----------------------------------------

def test_reference_hessian_patterns():
    _, p = pattern_of(lambda v: v[0] * v[0] + v[1] * v[1], [3.0, 4.0])
    assert p.contains(0, 0) and p.contains(1, 1) and not p.contains(0, 1)
    _, p = pattern_of(lambda v: v[0] * v[1], [2.0, 3.0])
    assert p.contains(0, 1) and p.contains(1, 0)
    _, p = pattern_of(lambda v: v[0] * v[0] + v[1] * v[2], [1.0, 2.0, 3.0])
    assert p.contains(0, 0) and p.contains(1, 2) and not p.contains(0, 1) and not p.contains(0, 2)


def test_reference_tridiagonal_pattern_and_three_colors():
    n = 10

    def f(v):
        s = v[0] - v[0]
        for i in range(n - 1):
            s = s + v[i] * v[i + 1]
        return s

    _, p = pattern_of(f, [0.5 + 0.01 * i for i in range(n)])
    for i in range(n - 1):
        assert p.contains(i, i + 1)
    for i in range(n):
        for j in range(n):
            if abs(i - j) > 1:
                assert not p.contains(i, j)
    assert sp.greedy_coloring(p)[1] == 3


def test_reference_jacobian_patterns():
    tape, _ = record_multi(lambda v: [v[0] * 2.0 + v[1], v[1] + v[2] * 3.0], [1.0, 2.0, 3.0])
    p = sp.detect_jacobian_sparsity(GpuTapeData.from_tape_f64(tape))
    assert p.contains(0, 0) and p.contains(0, 1) and not p.contains(0, 2)
    assert not p.contains(1, 0) and p.contains(1, 1) and p.contains(1, 2)
    tape, _ = record_multi(lambda v: [v[0] * v[1], v[1] * v[2], v[0] * v[0]], [1.0, 2.0, 3.0])
    p = sp.detect_jacobian_sparsity(GpuTapeData.from_tape_f64(tape))
    assert p.contains(2, 0) and not p.contains(2, 1) and not p.contains(2, 2)
    tape, _ = record_multi(lambda v: [v[0], v[1], v[2]], [1.0, 2.0, 3.0])
    assert sp.column_coloring(sp.detect_jacobian_sparsity(GpuTapeData.from_tape_f64(tape)))[1] == 1


U32 = np.uint32


@pytest.mark.parametrize("rows,cols,dim,expect", [
    ([], [], 0, 0),
    ([0, 1, 2, 3], [0, 1, 2, 3], 4, 1),
    ([0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2], 3, 3),
    ([0, 1, 1, 2, 2], [0, 0, 1, 1, 2], 3, 3),
])
def test_reference_greedy_coloring(rows, cols, dim, expect):
    colors, k = sp.greedy_coloring(sp.SparsityPattern(dim, np.array(rows, U32), np.array(cols, U32)))
    assert k == expect and len(colors) == dim and len(set(colors.tolist())) == k


def test_reference_column_and_row_coloring():
    J = sp.JacobianSparsityPattern
    assert sp.column_coloring(J(0, 0, np.array([], U32), np.array([], U32)))[1] == 0
    assert sp.column_coloring(J(3, 3, np.array([0, 1, 2], U32), np.array([0, 1, 2], U32)))[1] == 1
    assert sp.column_coloring(J(2, 3, np.array([0, 0, 0, 1, 1, 1], U32), np.array([0, 1, 2, 0, 1, 2], U32)))[1] == 3
    c, k = sp.column_coloring(J(2, 3, np.array([0, 0, 1, 1], U32), np.array([0, 1, 1, 2], U32)))
    assert k == 2 and c[0] != c[1] and c[1] != c[2] and c[0] == c[2]
    assert sp.row_coloring(J(0, 0, np.array([], U32), np.array([], U32)))[1] == 0
    assert sp.row_coloring(J(3, 3, np.array([0, 1, 2], U32), np.array([0, 1, 2], U32)))[1] == 1
    assert sp.row_coloring(J(3, 2, np.array([0, 0, 1, 1, 2, 2], U32), np.array([0, 1, 0, 1, 0, 1], U32)))[1] == 3


# ---- bit-exact against the oracle, pinned by brute force -------------------

def kitchen_sink(v):
    a = v[0] * v[1] + v[2].sin() * v[3]
    b = (v[1] - v[4]).abs() + v[5].max(v[0]) + (v[2] % 1.5)
    c = v[3].powf(v[4] * v[4] + 1.0) + v[0].atan2(v[5] + 2.0) + v[1].hypot(v[2])
    d = (v[6] / (v[7] * v[7] + 1.0)).exp() + v[6].floor() * v[7] + v[4].powi(3) - v[5].signum()
    return a * b + c - d.tanh() + v[7].fract() * 2.0 + v[0].round() + v[1].trunc() + v[2].ceil()


TAPES = {
    "kitchen_sink": lambda: record(kitchen_sink, [0.3, -0.4, 0.5, 0.6, -0.7, 0.8, 0.9, -1.1])[0],
    "rosenbrock": lambda: wl.rosenbrock(12),
    "mlp": lambda: wl.tanh_mlp(6, 5, 9),
    "dag0": lambda: wl.random_dag(3, 0, 300, 12),
    "dag1": lambda: wl.random_dag(3, 1, 500, 20),
    "dag2": lambda: wl.random_dag(5, 2, 800, 70),
}


@pytest.mark.parametrize("name", list(TAPES))
def test_patterns_and_colorings_match_oracle(name):
    tape = TAPES[name]()
    d = GpuTapeData.from_tape_f64(tape)
    hp = sp.detect_sparsity(d)
    r, c = po.hessian_pattern(tape)
    assert np.array_equal(hp.rows, r) and np.array_equal(hp.cols, c)
    assert np.all(hp.rows >= hp.cols)
    key = hp.rows.astype(np.int64) * 100000 + hp.cols
    assert np.all(np.diff(key) > 0)  # sorted, no duplicates
    colors, k = sp.greedy_coloring(hp)
    oc, ok_ = po.greedy_coloring(tape.num_inputs, r, c)
    assert k == ok_ and np.array_equal(colors, oc)
    jp = sp.detect_jacobian_sparsity(d)
    jr, jc = po.jacobian_pattern(tape)
    assert np.array_equal(jp.rows, jr) and np.array_equal(jp.cols, jc)
    for mine, theirs in ((sp.column_coloring(jp), po.column_coloring(jp.num_outputs, jp.num_inputs, jr, jc)),
                         (sp.row_coloring(jp), po.row_coloring(jp.num_outputs, jp.num_inputs, jr, jc))):
        assert mine[1] == theirs[1] and np.array_equal(mine[0], theirs[0])


@pytest.mark.parametrize("name", ["kitchen_sink", "rosenbrock", "mlp", "dag0"])
def test_patterns_cover_numeric_nonzeros(name):
    tape = TAPES[name]()
    d = GpuTapeData.from_tape_f64(tape)
    hp = sp.detect_sparsity(d)
    n = tape.num_inputs
    inpat = np.zeros((n, n), bool)
    inpat[hp.rows, hp.cols] = True
    inpat |= inpat.T
    rng = np.random.default_rng(4)
    for _ in range(3):
        x = rng.uniform(-0.9, 0.9, n)
        H = po.hessian(tape, x)
        nz = np.abs(H) > 1e-12 * max(1.0, np.max(np.abs(H)))
        assert not np.any(nz & ~inpat), name
    # distance-2 coloring: two columns sharing a row never share a color
    colors, _ = sp.greedy_coloring(hp)
    for i in range(n):
        cs = colors[np.nonzero(inpat[i])[0]]
        assert len(cs) == len(set(cs.tolist())), (name, i)


def test_coloring_beyond_64_colors_uses_the_fallback():
    dim = 70
    rows, cols = np.tril_indices(dim)
    p = sp.SparsityPattern(dim, rows.astype(U32), cols.astype(U32))
    colors, k = sp.greedy_coloring(p)
    oc, ok_ = po.greedy_coloring(dim, rows, cols)
    assert k == dim == ok_ and np.array_equal(colors, oc) and len(set(colors.tolist())) == dim


def test_pattern_rejects_invalid_tapes():
    from echidna_amd.tape import GpuError

    bad = GpuTapeData.from_tape_f64(wl.rosenbrock(3))
    bad.arg0 = np.array(bad.arg0, np.uint32).copy()
    bad.arg0[-1] = len(bad.arg0) + 5
    with pytest.raises(GpuError, match="must reference an earlier slot"):
        sp.detect_sparsity(bad)
