"""The upload-time compiler, checked on the CPU: the compiled register program
(emulated in numpy exactly as the scalar HIP interpreter executes it) must
reproduce the oracle's values and gradients, with and without dot fusion."""

import numpy as np
import pytest

import program_emu as emu
import pyoracle as po
from echidna_amd import workloads as wl
from echidna_amd.backend import compile_program
from echidna_amd.tape import GpuTapeData, record, record_multi

TAPES = {
    "rosenbrock16": lambda: wl.rosenbrock(16),
    "mlp32": lambda: wl.tanh_mlp(32, 64, 1),
    "mlp2": lambda: wl.tanh_mlp2(12, 9, 7, 4),
    "dag0": lambda: wl.random_dag(3, 0, 300, 12),
    "dag1": lambda: wl.random_dag(3, 1, 600, 30),
    "square_same_operand": lambda: record(lambda v: v[0] * v[0] + v[0] * v[1], [1.0, 2.0])[0],
    "output_is_input": lambda: record(lambda v: v[1], [1.0, 2.0])[0],
}


@pytest.mark.parametrize("name", list(TAPES))
@pytest.mark.parametrize("fuse", [True, False])
@pytest.mark.parametrize("f64", [True, False])
def test_program_matches_oracle(name, fuse, f64):
    tape = TAPES[name]()
    d = GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape)
    x = wl.points_normal(5, 9, tape.num_inputs) * 0.7
    F = np.float64 if f64 else np.float32
    x = x.astype(F)
    y_ref, g_ref = po.gradient(tape, x, dtype=F)
    for mode in (0, 1):
        words, terms, info = compile_program(d, mode=mode, fuse=fuse, f64=f64)
        y, g = emu.run(words, terms, x, tape.num_inputs, tape.num_outputs(), info["registers"], f64=f64,
                       nterms=info["nterms"], pool=info["pool"])
        tol = 1e-13 if f64 else 2e-5
        scale = np.maximum(np.abs(y_ref), 1.0)
        assert np.all(np.abs(y - y_ref) <= tol * scale), (name, mode)
        if mode == 1:
            gs = np.maximum(np.max(np.abs(g_ref), axis=1, keepdims=True), 1.0)
            assert np.all(np.abs(g - g_ref) <= tol * gs), (name, mode)


@pytest.mark.parametrize("name", list(TAPES) + ["wide_mlp"])
@pytest.mark.parametrize("f64", [True, False])
def test_windowed_stream_runs_like_the_program(name, f64):
    """The windowed stream (what the kernels execute) decodes into windows
    that hold whole instructions and computes bit-identical results; long dot
    products split into pieces across windows."""
    tape = wl.tanh_mlp(6, 300, 2) if name == "wide_mlp" else TAPES[name]()
    d = GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape)
    F = np.float64 if f64 else np.float32
    x = (wl.points_normal(6, 7, tape.num_inputs) * 0.7).astype(F)
    for mode in (0, 1):
        words, terms, info = compile_program(d, mode=mode, fuse=True, f64=f64)
        stream, _, sinfo = compile_program(d, mode=mode + 2, fuse=True, f64=f64)
        assert sinfo["registers"] == info["registers"] and len(stream) % emu.WIN_WORDS == 0
        y, g = emu.run(words, terms, x, tape.num_inputs, tape.num_outputs(), info["registers"], f64=f64,
                       nterms=info["nterms"], pool=info["pool"])
        ys, gs = emu.run_stream(stream, x, tape.num_inputs, tape.num_outputs(), info["registers"], f64=f64,
                                pool=sinfo["pool"])
        assert np.array_equal(y.view(np.uint8), ys.view(np.uint8)), (name, mode)
        assert np.array_equal(g.view(np.uint8), gs.view(np.uint8)), (name, mode)
    if name == "wide_mlp":
        ins = emu.decode_stream(stream, f64)
        assert len(stream) // emu.WIN_WORDS > 10
        assert any(w[0] & 0xFF == emu.I_DOT and (w[0] >> 8) & emu.F_DOT_CONT for w, _ in ins)


@pytest.mark.parametrize("name", ["mlp32", "mlp2", "ridge"])
@pytest.mark.parametrize("f64", [True, False])
def test_dense_grouping_is_bit_identical(name, f64):
    """DENSE / DENSE_REV replace groups of DOTs sharing one operand list;
    every value and adjoint is computed by the same operations in the same
    order, so the emulated results match the DOT-level program bit for bit."""
    tape = wl.softplus_ridge(16, 6, 53) if name == "ridge" else TAPES[name]()
    d = GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape)
    F = np.float64 if f64 else np.float32
    x = (wl.points_normal(12, 11, tape.num_inputs) * 0.8).astype(F)
    for mode in (0, 1):
        outs = []
        for dense in (True, False):
            words, terms, info = compile_program(d, mode=mode, dense=dense, f64=f64)
            ops = [int(w[0]) & 0xFF for w in words]
            assert (emu.I_DENSE in ops) == dense
            assert (emu.I_DENSE_REV in ops) == (dense and mode == 1)
            outs.append(emu.run(words, terms, x, tape.num_inputs, tape.num_outputs(), info["registers"], f64=f64,
                                nterms=info["nterms"], pool=info["pool"]))
        for a, b in zip(outs[0], outs[1]):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (name, mode)


@pytest.mark.parametrize("f64", [True, False])
def test_dense_guards_are_bit_identical(f64):
    """Zero and infinite adjoints, zero weights and NaN partials inside DENSE
    groups: the grouped program matches the DOT-level one bit for bit."""
    import dense_cases as dc

    tape = dc.tape()
    d = GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape)
    F = np.float64 if f64 else np.float32
    x = dc.points(60).astype(F)
    outs = []
    for dense in (True, False):
        words, terms, info = compile_program(d, mode=1, dense=dense, f64=f64)
        ops = [int(w[0]) & 0xFF for w in words]
        assert (emu.I_DENSE_REV in ops) == dense
        with np.errstate(all="ignore"):
            outs.append(emu.run(words, terms, x, tape.num_inputs, tape.num_outputs(), info["registers"], f64=f64,
                                nterms=info["nterms"], pool=info["pool"]))
    g = outs[0][1]
    assert np.isinf(g).any() and np.isnan(g).any() and np.isfinite(g).all(axis=1).any()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("name", ["mlp32", "mlp2"])
def test_hvp_programs_group_dense_layers(name):
    """The dual (HVP) program groups the same dense layers as the gradient
    program; its DENSE_REV term words carry P_DEEP for operands reached
    through an Add link (every product operand of an MLP layer's chain)."""
    tape = TAPES[name]()
    d = GpuTapeData.from_tape(tape)
    grad_words, _, _ = compile_program(d, mode=1)
    hvp_words, _, hinfo = compile_program(d, mode=5)
    count = lambda words, op: sum(1 for w in words if int(w[0]) & 0xFF == op)
    assert count(hvp_words, emu.I_DENSE) == count(grad_words, emu.I_DENSE) > 0
    assert count(hvp_words, emu.I_DENSE_REV) == count(grad_words, emu.I_DENSE_REV) > 0
    _, _, plain = compile_program(d, mode=5, dense=False)
    assert hinfo["instructions"] < plain["instructions"]
    pool = hinfo["pool"]
    for w in hvp_words:
        if int(w[0]) & 0xFF == emu.I_DENSE_REV:
            T, off = int(w[2]), int(w[3])
            words = pool[off:off + T]
            assert np.all((words >> 30) == 0) and np.all(words & (1 << 28))  # products, all deep


def test_dense_needs_distinct_operands():
    """A dot product that uses one operand twice keeps its DOT form."""
    def f(v):
        s = v[0] * 0.0
        for j in range(6):
            s = s + (v[0] * (0.1 + j) + v[0] * (0.2 * j) + v[1] * 0.3).tanh()
        return s

    tape, _ = record(f, [0.3, 0.4])
    words, _, info = compile_program(GpuTapeData.from_tape_f64(tape), mode=1)
    ops = [int(w[0]) & 0xFF for w in words]
    assert emu.I_DENSE not in ops and emu.I_DENSE_REV not in ops


def test_fusion_collapses_mlp_dot_products():
    tape = wl.tanh_mlp(32, 64, 1)
    d = GpuTapeData.from_tape(tape)
    _, _, plain = compile_program(d, mode=1, fuse=False)
    _, _, fused = compile_program(d, mode=1, fuse=True)
    assert plain["instructions"] > 8000
    assert fused["instructions"] < 800
    assert fused["registers"] <= 100 < plain["registers"]


def test_constants_become_immediates_and_dead_code_is_dropped():
    # a dead branch (v[1].sin()) and per-use constants: only live ops remain
    tape, _ = record(lambda v: (v[1].sin(), v[0] * 2.0 + 3.0)[1], [1.0, 2.0])
    words, _, info = compile_program(GpuTapeData.from_tape_f64(tape), mode=0, fuse=False)
    ops = [int(w[0]) & 0xFF for w in words]
    assert 24 not in ops  # no Sin
    assert ops.count(1) == 0  # no Const instructions
    assert info["registers"] <= 3


def test_register_reuse_bounds_live_set():
    # a long chain keeps O(1) values alive
    def f(v):
        s = v[0]
        for _ in range(500):
            s = (s * 0.999).sin()
        return s
    tape, _ = record(f, [0.3])
    _, _, info = compile_program(GpuTapeData.from_tape_f64(tape), mode=0)
    assert info["registers"] <= 3
    x = np.array([[0.3], [1.1]])
    words, terms, info = compile_program(GpuTapeData.from_tape_f64(tape), mode=1)
    y, g = emu.run(words, terms, x, 1, 1, info["registers"], nterms=info["nterms"], pool=info["pool"])
    y_ref, g_ref = po.gradient(tape, x)
    assert np.allclose(y, y_ref, rtol=1e-13) and np.allclose(g, g_ref, rtol=1e-12)


def test_multi_output_program():
    tape, _ = record_multi(lambda v: [v[0] * v[1], v[0].sin() + v[1], v[1]], [1.0, 2.0])
    words, terms, info = compile_program(GpuTapeData.from_tape_f64(tape), mode=1)
    x = np.array([[0.5, 2.0], [1.5, -1.0]])
    y, g = emu.run(words, terms, x, 2, 3, info["registers"], nterms=info["nterms"], pool=info["pool"])
    y_ref, g_ref = po.gradient(tape, x)
    assert np.allclose(y, y_ref, rtol=1e-14) and np.allclose(g, g_ref, rtol=1e-14)


@pytest.mark.parametrize("which", ["mlp1", "mlp2", "mixed"])
def test_segmented_jet_program_moves_activations_next_to_their_dense(which):
    """The segmented jet program (whole-layer DENSE groups, mode 8): each
    activation whose operand is a DENSE member read by nothing else follows
    its DENSE at once and computes in place (destination = operand = the
    member's register), which is what lets the layer's batch GEMM apply it in
    its epilogue; the read-out DOT pieces left adjacent merge into one DOT.
    Values are unchanged: the emulated program equals the DOT-level one bit
    for bit.  In `mixed` a member also feeds a second op, so its activation
    stays where the tape put it."""
    if which == "mlp1":
        tape = wl.tanh_mlp(40, 36, 7)
    elif which == "mlp2":
        tape = wl.tanh_mlp2(24, 40, 33, 8)
    else:
        W = np.random.default_rng(3).normal(size=(6, 5))

        def f(v):
            hs = [sum(float(W[j, i]) * v[i] for i in range(5)) for j in range(6)]
            acts = [h.tanh() if j != 2 else h.exp() for j, h in enumerate(hs)]
            return sum(acts[1:], acts[0]) + hs[4] * 0.5

        tape, _ = record(f, [0.1 * i for i in range(5)])
    d = GpuTapeData.from_tape(tape)
    words, terms, info = compile_program(d, mode=8, f64=False)
    ops = [int(w[0]) & 0xFF for w in words]
    dense_at = [i for i, o in enumerate(ops) if o == emu.I_DENSE]
    assert dense_at
    pool = info["pool"]
    for i in dense_at:
        _, members, _ = emu.dense_parts(words[i], pool, False)
        members = set(int(r) for r in members)
        k = i + 1
        moved = []
        while k < len(ops) and ops[k] < 64 and int(words[k][1]) == int(words[k][2]) and int(words[k][2]) in members:
            moved.append(int(words[k][2]))
            k += 1
        hidden = 5 if which == "mixed" else len(members)
        assert len(moved) == hidden and len(set(moved)) == hidden, (which, i, moved)
    if which != "mixed":
        assert ops.count(emu.I_DOT) == 1  # the read-out, one instruction
    x = (wl.points_normal(12, 9, tape.num_inputs) * 0.8).astype(np.float32)
    ref_words, ref_terms, ref_info = compile_program(d, mode=0, f64=False, dense=False)
    got = emu.run(words, terms, x, tape.num_inputs, 1, info["registers"], f64=False, mode=0, nterms=info["nterms"],
                  pool=pool)
    ref = emu.run(ref_words, ref_terms, x, tape.num_inputs, 1, ref_info["registers"], f64=False, mode=0,
                  nterms=ref_info["nterms"], pool=ref_info["pool"])
    for a, b in zip(got, ref):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
