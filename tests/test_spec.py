"""Tape-specialised kernel source (no device): generation and hipRTC
compilation for gfx950 on the host, literal constants exact to the bit, and
the reverse sweep's shape (zero tests only where a partial is not a fixed
finite constant)."""

import re

import numpy as np
import pytest

from echidna_amd import backend as be
from echidna_amd.tape import GpuError, GpuTapeData, record, record_multi
from test_structure import TAPES


def _data(tape, f64):
    return GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape)


@pytest.mark.parametrize("name", ["kitchen_sink", "rosenbrock", "mlp", "dag0", "guards"])
@pytest.mark.parametrize("f64", [False, True])
def test_source_compiles_for_gfx950(name, f64):
    if name == "guards":
        import dense_cases as dc

        tape = dc.tape()
    else:
        tape = TAPES[name]()
    for what in (0, 1, 2, 5):  # 5: Taylor jets of order 3
        src, info = be.spec_source(_data(tape, f64), what, f64=f64, arch="gfx950")
        assert info["code_bytes"] > 0
        assert ("typedef double F;" if f64 else "typedef float F;") in src
        stores = re.findall(r"^gp\[(\d+)\] = ", src, re.M)
        hstores = re.findall(r"^hp\[(\d+)\] = ", src, re.M)
        assert len(stores) == (tape.num_inputs if what in (1, 2) else 0)
        assert len(hstores) == (tape.num_inputs if what == 2 else 0)


def test_constants_are_exact_literals():
    tape, _ = record(lambda v: v[0] * 0.1 + v[1] * -0.0 + v[1] * float("inf"), [1.0, 2.0])
    src32, _ = be.spec_source(_data(tape, False), 1, f64=False)
    src64, _ = be.spec_source(_data(tape, True), 1, f64=True)
    assert "__builtin_bit_cast(float, 0x3dcccccdu)" in src32  # 0.1f
    assert "__builtin_bit_cast(double, 0x3fb999999999999aull)" in src64  # 0.1
    assert "__builtin_bit_cast(double, 0x8000000000000000ull)" in src64  # -0.0
    assert "__builtin_bit_cast(float, 0x7f800000u)" in src32  # inf
    # multiplication by -0.0 has a zero partial: no update from it; by inf a
    # guarded one (0 * inf would be NaN)
    assert "0x80000000u) * g" not in src32
    assert "if (lv && da != F(0))" in src32


def test_reverse_sweep_shape():
    """Adds, subs and products with constants update unconditionally; other
    ops keep the reference's zero tests; entries off the path to output 0
    and constant subexpressions get no adjoint."""
    def f(v):
        dead = v[2].exp()  # recorded, never reaches the output
        c = 3.0
        return (v[0] * 2.0 - v[1]).tanh() + v[0] * v[1]

    tape, _ = record(f, [0.5, 0.25, 1.0])
    src, _ = be.spec_source(_data(tape, True), 1, f64=True)
    rev = src.split("F g", 1)[1]
    # the variable product; binary64 tanh reuses the derivative its forward
    # pair made (s_tanh_dtanh), the f32 source calls s_partials for it too
    assert rev.count("s_partials") == 1 and "s_tanh_dtanh" in src
    assert "const F da = t" in rev
    src32, _ = be.spec_source(_data(tape, False), 1, f64=False)
    assert src32.split("F g", 1)[1].count("s_partials") == 2 and "s_tanh_dtanh" not in src32
    assert "gp[2] = F(0);" in src  # x2 only feeds the dead exp
    assert "s_partials<F>(17u" not in src  # no reverse for exp


def test_source_rejects_bad_input():
    tape, _ = record(lambda v: v[0] * v[1], [1.0, 2.0])
    with pytest.raises(GpuError, match="bad argument"):
        be.spec_source(_data(tape, False), 9, f64=False)
    with pytest.raises(GpuError, match="bad argument"):  # binary64 adjoints: f32 tapes only
        be.spec_source(_data(tape, True), be.SPEC_GRAD_ACC64, f64=True)
    d = _data(tape, False)
    d.arg0 = np.array(d.arg0, dtype=np.uint32)
    d.arg0[-1] = 99
    with pytest.raises(GpuError):
        be.spec_source(d, 1, f64=False)


def test_dual_source_shape():
    """The HVP kernel seeds inputs with (x, v), keeps the reference's Dual
    zero tests for runtime partials and writes SUB's partial as d_neg(one) =
    {-1, -0}."""
    tape, _ = record(lambda v: (v[0] - v[1]).sin() * 3.0, [0.5, 0.25])
    src, _ = be.spec_source(_data(tape, True), 2, f64=True)
    assert "const D v0 = {x[0], v[0]};" in src
    assert "!ehip::d_zero(" in src and "ehip::d_partials<F>(24u" in src
    assert "D{__builtin_bit_cast(double, 0xbff0000000000000ull), __builtin_bit_cast(double, 0x8000000000000000ull)}" in src
    multi, _ = record_multi(lambda v: [v[0] * v[1], v[1]], [1.0, 2.0])
    with pytest.raises(GpuError, match="scalar-output"):
        be.spec_source(_data(multi, False), 2, f64=False)


def test_occupancy_target_in_source():
    tape, _ = record(lambda v: v[0] * v[1] + v[0].sin(), [0.5, 0.25])
    src0, _ = be.spec_source(_data(tape, True), 1, f64=True)
    src2, info = be.spec_source(_data(tape, True), 1, f64=True, arch="gfx950", waves=2)
    assert "#define EHIP_OCC\n" in src0 and "amdgpu_waves_per_eu(2)" in src2 and info["code_bytes"] > 0
    with pytest.raises(GpuError, match="bad argument"):
        be.spec_source(_data(tape, True), 1, f64=True, waves=9)


@pytest.mark.parametrize("order", [1, 2, 4, 5])
def test_jet_source_compiles_for_every_order(order):
    """Jets of every order on the tape that uses every opcode; inputs are
    seeded (x, v, 0, ..) with v read or drawn like the interpreters."""
    tape = TAPES["kitchen_sink"]()
    src, info = be.spec_source(_data(tape, False), 2 + order, f64=False, arch="gfx950")
    assert info["code_bytes"] > 0 and f"typedef ehip::Jet<F, {order}> J;" in src
    assert ("ehip::lane_direction" in src) == (order > 1)
    assert src.count("y[") == len(_data(tape, False).output_indices) * order


@pytest.mark.parametrize("name", ["kitchen_sink", "mlp", "guards"])
def test_binary64_adjoint_source(name):
    """what 8 on an f32 tape: the same FP32 forward sweep and FP32 partials as
    the gradient kernel, every adjoint declared binary64 with each
    contribution widened before it is added, one rounding per stored
    gradient; it compiles for gfx950."""
    if name == "guards":
        import dense_cases as dc

        tape = dc.tape()
    else:
        tape = TAPES[name]()
    d = _data(tape, False)
    src1, _ = be.spec_source(d, 1, f64=False)
    src8, info = be.spec_source(d, be.SPEC_GRAD_ACC64, f64=False, arch="gfx950")
    assert info["code_bytes"] > 0 and "binary64 adjoints" in src8
    fwd1, rev1 = src1.split("F g", 1)
    fwd8, rev8 = src8.split("double g", 1)
    assert fwd1.replace(", value and gradient of output 0\n", ", value and gradient of output 0, binary64 adjoints\n") \
        == fwd8
    assert not re.search(r"^F g\d+ = ", src8, re.M)
    # every update widens its contribution; gradients are rounded once
    updates = re.findall(r"(g\d+) = \1 \+ (.*?) \* g\d+;", rev8)
    assert updates and all(u[1].startswith("(double)") for u in updates)
    assert len(re.findall(r"^gp\[\d+\] = \(F\)g\d+;$", src8, re.M)) + src8.count("] = F(0);") == tape.num_inputs
