"""An arbitrary-precision evaluator of a tape's primal values (test
infrastructure).

Each opcode is its mathematical function on mpmath reals, as
``eval_forward`` (src/opcode.rs:228-318) states it: Rust's `%` is the
truncated remainder (the sign of the dividend), `round` rounds half away from zero, `signum`
of a signed zero is +-1, `max`/`min` return the first operand on ties and
when the second is NaN, `cbrt` is the real cube root, and `powi` decodes
its exponent from the raw u32 argument (src/opcode.rs:487-499).

It holds no derivative rule.  Derivatives taken from it by mpmath's
high-precision numerical differentiation (`mpmath.diff`, `mpmath.taylor`)
are therefore independent of every derivative rule the oracle restates:
the reverse partials, the Dual (forward-over-reverse) arithmetic and the
Taylor recurrences.  A dropped term, a wrong sign or a transposed operand
in any of those rules changes a derivative by O(1) and fails a comparison
against these values.
"""

from __future__ import annotations

import mpmath as mp

from echidna_amd.tape import UNUSED, OpCode, powi_exp_decode_raw

DPS = 60


class NotReal(Exception):
    """The tape left the reals (or hit an undefined value) at this point."""


def _real(v):
    if isinstance(v, mp.mpc):
        if v.imag != 0:
            raise NotReal()
        v = v.real
    if not mp.isfinite(v):
        raise NotReal()
    return v


def _cbrt(a):
    return mp.cbrt(a) if a >= 0 else -mp.cbrt(-a)


def _round(a):
    return mp.sign(a) * mp.floor(abs(a) + mp.mpf("0.5"))


def _signum(a):
    return mp.mpf(1) if a >= 0 else mp.mpf(-1)


def _op(op: int, a, b, raw_b: int):
    O = OpCode
    if op == O.Add:
        return a + b
    if op == O.Sub:
        return a - b
    if op == O.Mul:
        return a * b
    if op == O.Div:
        return a / b
    if op == O.Rem:
        q = a / b
        return a - b * (mp.floor(q) if q >= 0 else mp.ceil(q))
    if op == O.Powf:
        return mp.power(a, b)
    if op == O.Atan2:
        return mp.atan2(a, b)
    if op == O.Hypot:
        return mp.hypot(a, b)
    if op == O.Max:
        return a if a >= b else b
    if op == O.Min:
        return a if a <= b else b
    if op == O.Neg:
        return -a
    if op == O.Recip:
        return 1 / a
    if op == O.Sqrt:
        return mp.sqrt(a)
    if op == O.Cbrt:
        return _cbrt(a)
    if op == O.Powi:
        return mp.power(a, powi_exp_decode_raw(raw_b))
    if op == O.Exp:
        return mp.exp(a)
    if op == O.Exp2:
        return mp.power(2, a)
    if op == O.ExpM1:
        return mp.expm1(a)
    if op == O.Ln:
        return mp.log(a)
    if op == O.Log2:
        return mp.log(a, 2)
    if op == O.Log10:
        return mp.log10(a)
    if op == O.Ln1p:
        return mp.log1p(a)
    if op == O.Sin:
        return mp.sin(a)
    if op == O.Cos:
        return mp.cos(a)
    if op == O.Tan:
        return mp.tan(a)
    if op == O.Asin:
        return mp.asin(a)
    if op == O.Acos:
        return mp.acos(a)
    if op == O.Atan:
        return mp.atan(a)
    if op == O.Sinh:
        return mp.sinh(a)
    if op == O.Cosh:
        return mp.cosh(a)
    if op == O.Tanh:
        return mp.tanh(a)
    if op == O.Asinh:
        return mp.asinh(a)
    if op == O.Acosh:
        return mp.acosh(a)
    if op == O.Atanh:
        return mp.atanh(a)
    if op == O.Abs:
        return abs(a)
    if op == O.Signum:
        return _signum(a)
    if op == O.Floor:
        return mp.floor(a)
    if op == O.Ceil:
        return mp.ceil(a)
    if op == O.Round:
        return _round(a)
    if op == O.Trunc:
        return mp.floor(a) if a >= 0 else mp.ceil(a)
    if op == O.Fract:
        return a - (mp.floor(a) if a >= 0 else mp.ceil(a))
    raise ValueError(f"opcode {op}")


def forward(tape, x):
    """All outputs of the tape at the point x (mpmath reals)."""
    vals = []
    ni = tape.num_inputs
    for i, (op, a, b) in enumerate(zip(tape.opcodes, tape.arg0, tape.arg1)):
        op = int(op)
        if op == OpCode.Input:
            vals.append(mp.mpf(x[i]))
        elif op == OpCode.Const:
            vals.append(mp.mpf(float(tape.values[i])))
        else:
            av = vals[a]
            bv = vals[b] if (b != UNUSED and op != OpCode.Powi) else None
            vals.append(_real(_op(op, av, bv, b)))
    assert len(x) == ni
    return [vals[o] for o in tape.all_output_indices()]


def jet(tape, x, v, K: int, out: int = 0):
    """Taylor coefficients c_k = g^(k)(0) / k! of g(t) = f(x + t v), k < K."""
    with mp.workdps(DPS):
        xs = [mp.mpf(float(t)) for t in x]
        vs = [mp.mpf(float(t)) for t in v]

        def g(t):
            return forward(tape, [a + t * d for a, d in zip(xs, vs)])[out]

        return [float(c) for c in mp.taylor(g, 0, K - 1)]


def gradient(tape, x, out: int = 0):
    with mp.workdps(DPS):
        xs = [mp.mpf(float(t)) for t in x]
        n = len(xs)
        res = []
        for i in range(n):
            def g(t, i=i):
                y = list(xs)
                y[i] = y[i] + t
                return forward(tape, y)[out]
            res.append(float(mp.diff(g, 0)))
        return res


def hvp(tape, x, v, out: int = 0):
    """(H v)_i = d/ds d/dt f(x + s e_i + t v) at s = t = 0."""
    with mp.workdps(DPS):
        xs = [mp.mpf(float(t)) for t in x]
        vs = [mp.mpf(float(t)) for t in v]
        res = []
        for i in range(len(xs)):
            def g(s, t, i=i):
                y = [a + t * d for a, d in zip(xs, vs)]
                y[i] = y[i] + s
                return forward(tape, y)[out]
            res.append(float(mp.diff(g, (0, 0), (1, 1))))
        return res
