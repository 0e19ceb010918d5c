"""The arithmetic cost model behind the roofline, on hand-counted tapes."""

from echidna_amd import costmodel as cm
from echidna_amd import workloads as wl
from echidna_amd.tape import record


def test_hand_counted_tape():
    """f = x0 x1 + sin x0.  Forward: mul, sin, add = 3.  Reverse: add passes
    its adjoint on (partials +1: no multiply); sin: cos (1) and a multiply
    (1); mul: one multiply per operand (partials are the other operand);
    x0 receives two contributions: one add.  5 in all."""
    tape, _ = record(lambda v: v[0] * v[1] + v[0].sin(), [0.3, 0.4])
    assert cm.sweep_counts(tape) == {"forward": 3, "reverse": 5}
    assert cm.fwdgrad_flops(tape) == 8
    # jets, K = 3: a product of two jets 9, sin 18, add 3
    assert cm.jet_flops(tape, 3) == 30
    # seeded (x, v, 0): x0 x1 has pairs (0,0) | (0,1),(1,0) | (1,1): 1 | 3 | 1;
    # sin 6 per coefficient; the add 1 per coefficient
    assert cm.jet_flops_split(tape, 3) == (1 + 6 + 1, 4 + 12 + 2)


def test_jet_split_on_the_stde_mlp():
    """n=256 H=256 tanh MLP: the layer's 256 x 256 products and 255 x 256
    sums have c0 (per point) and c1 (per direction) only, the bias adds c0
    only (c1 is a copy); tanh costs 6 per coefficient, c0..c2; the
    read-out's 256 products and 255 sums take all three coefficients, and
    its first add (to the constant 0) c0 alone."""
    tape = wl.tanh_mlp(256, 256, 2)
    layer = 256 * 256 + 255 * 256
    point = layer + 256 + 6 * 256 + (256 + 256)
    lane = layer + 12 * 256 + 2 * (256 + 255)
    assert cm.jet_flops_split(tape, 3) == (point, lane)


def test_mlp_counts():
    """tanh MLP n=32 H=64: 2048 products and 2048 sums in the layer, 64 tanh,
    64 readout products and 64 readout sums forward (4,288); reverse 2,048
    layer products, 63 extra contributions to each of 32 inputs, 64 tanh
    partials with their multiplies and 64 readout multiplies (4,256)."""
    tape = wl.tanh_mlp(32, 64, 1)
    assert cm.sweep_counts(tape) == {"forward": 4288, "reverse": 2048 + 63 * 32 + 2 * 64 + 64}
    assert cm.compulsory_bytes(tape, "gradient", 4) == (32 + 1 + 32) * 4
