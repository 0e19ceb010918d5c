"""Bytecode tape recording and the flattened upload format.

Host-side mirror of the reference's graph-mode recorder and GPU upload
structure, kept here so tapes can be built for the HIP backend without the
Rust crate:

* ``OpCode`` discriminants pin ``src/opcode.rs:21-130`` (Input=0 ... Fract=42,
  Custom=43), exactly as ``GpuTapeData::validate`` pins them
  (``src/gpu/mod.rs:427-472``).
* ``BytecodeTape.push_const / push_op / push_powi`` follow the recording rules
  of ``src/bytecode_tape/mod.rs:249-382``: all-constant folding, the four
  IEEE-exact aliases (``x*1``, ``x/1``, ``x+(-0)``, ``x-(+0)``), and the powi
  folds (``x^0`` only when the value is 1, ``x^1``, ``x^-1`` -> Recip).
* ``BReverse`` follows ``src/breverse.rs:22-30`` and
  ``src/traits/breverse_ops.rs:16-42``: a constant operand is promoted to a
  fresh ``Const`` entry at every use (no deduplication).
* ``record`` follows ``src/api.rs:197-216``.
* ``GpuTapeData.from_tape`` / ``from_tape_f64_lossy`` / ``validate`` follow
  ``src/gpu/mod.rs:313-565``.

Values are carried as Python floats (binary64) during recording; they matter
only for the constants stored on the tape and for the folding decisions.
"""

from __future__ import annotations

import enum
import math
import struct
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

UNUSED = 0xFFFFFFFF
CONSTANT = 0xFFFFFFFF


class OpCode(enum.IntEnum):
    """Opcode discriminants, pinned to ``src/opcode.rs:21-130``."""

    Input = 0
    Const = 1
    Add = 2
    Sub = 3
    Mul = 4
    Div = 5
    Rem = 6
    Powf = 7
    Atan2 = 8
    Hypot = 9
    Max = 10
    Min = 11
    Neg = 12
    Recip = 13
    Sqrt = 14
    Cbrt = 15
    Powi = 16
    Exp = 17
    Exp2 = 18
    ExpM1 = 19
    Ln = 20
    Log2 = 21
    Log10 = 22
    Ln1p = 23
    Sin = 24
    Cos = 25
    Tan = 26
    Asin = 27
    Acos = 28
    Atan = 29
    Sinh = 30
    Cosh = 31
    Tanh = 32
    Asinh = 33
    Acosh = 34
    Atanh = 35
    Abs = 36
    Signum = 37
    Floor = 38
    Ceil = 39
    Round = 40
    Trunc = 41
    Fract = 42
    Custom = 43


BINARY_OPS = frozenset(range(int(OpCode.Add), int(OpCode.Min) + 1))


class GpuError(Exception):
    """Mirror of ``GpuError`` (``src/gpu/mod.rs:276-303``).

    ``code`` is the C-ABI status: 1 NoDevice, 2 ShaderCompilation,
    3 OutOfMemory, 4 CustomOpsNotSupported, 5 Other.
    """

    NO_DEVICE = 1
    SHADER_COMPILATION = 2
    OUT_OF_MEMORY = 3
    CUSTOM_OPS_NOT_SUPPORTED = 4
    OTHER = 5

    def __init__(self, code: int, msg: str = ""):
        super().__init__(msg or f"GpuError({code})")
        self.code = code


def powi_exp_encode(exp: int) -> int:
    """Bit-preserving ``i32 as u32`` (``src/opcode.rs:520``)."""
    return exp & 0xFFFFFFFF


def powi_exp_decode_raw(raw: int) -> int:
    """Bit-preserving ``u32 as i32`` (``src/opcode.rs:491``)."""
    raw &= 0xFFFFFFFF
    return raw - (1 << 32) if raw & 0x80000000 else raw


def _is_neg_zero(x: float) -> bool:
    return x == 0.0 and math.copysign(1.0, x) < 0


def _is_pos_zero(x: float) -> bool:
    return x == 0.0 and math.copysign(1.0, x) > 0


class BytecodeTape:
    """Structure-of-arrays tape (``src/bytecode_tape/mod.rs:123-142``)."""

    def __init__(self) -> None:
        self.opcodes: List[int] = []
        self.arg0: List[int] = []
        self.arg1: List[int] = []
        self.values: List[float] = []
        self.num_inputs = 0
        self.output_index = 0
        self.output_indices: List[int] = []

    # -- appends (src/bytecode_tape/mod.rs:226-251) --
    def _push_entry(self, op: int, a: int, b: int, value: float) -> int:
        idx = len(self.opcodes)
        self.opcodes.append(int(op))
        self.arg0.append(a)
        self.arg1.append(b)
        self.values.append(value)
        return idx

    def new_input(self, value: float) -> int:
        self.num_inputs += 1
        return self._push_entry(OpCode.Input, UNUSED, UNUSED, value)

    def push_const(self, value: float) -> int:
        return self._push_entry(OpCode.Const, UNUSED, UNUSED, value)

    def push_op(self, op: int, a: int, b: int, value: float) -> int:
        """Record an op with folding (``src/bytecode_tape/mod.rs:268-286``)."""
        a_const = self.opcodes[a] == OpCode.Const
        b_const = b == UNUSED or self.opcodes[b] == OpCode.Const
        if a_const and b_const:
            return self.push_const(value)
        if (a_const or b_const) and b != UNUSED:
            alias = self._try_algebraic_simplify(op, a, b, a_const, b_const)
            if alias is not None:
                return alias
        return self._push_entry(op, a, b, value)

    def _try_algebraic_simplify(self, op, a, b, a_const, b_const) -> Optional[int]:
        """IEEE-exact aliases only (``src/bytecode_tape/mod.rs:307-354``)."""
        if op == OpCode.Add:
            if b_const and _is_neg_zero(self.values[b]):
                return a
            if a_const and _is_neg_zero(self.values[a]):
                return b
        elif op == OpCode.Sub:
            if b_const and _is_pos_zero(self.values[b]):
                return a
        elif op == OpCode.Mul:
            if b_const and self.values[b] == 1.0:
                return a
            if a_const and self.values[a] == 1.0:
                return b
        elif op == OpCode.Div:
            if b_const and self.values[b] == 1.0:
                return a
        return None

    def push_powi(self, a: int, exp: int, value: float) -> int:
        """Powi recording (``src/bytecode_tape/mod.rs:363-382``)."""
        if self.opcodes[a] == OpCode.Const:
            return self.push_const(value)
        if exp == 0 and value == 1.0:
            return self.push_const(1.0)
        if exp == 1:
            return a
        if exp == -1:
            return self.push_op(OpCode.Recip, a, UNUSED, value)
        return self._push_entry(OpCode.Powi, a, powi_exp_encode(exp), value)

    def set_output(self, index: int) -> None:
        self.output_index = index

    def set_outputs(self, indices: Sequence[int]) -> None:
        n = len(self.values)
        for i, idx in enumerate(indices):
            if not 0 <= idx < n:
                raise IndexError(f"set_outputs: indices[{i}] = {idx} is out of range")
        self.output_indices = list(indices)
        if indices:
            self.output_index = indices[0]

    def all_output_indices(self) -> List[int]:
        return list(self.output_indices) if self.output_indices else [self.output_index]

    def num_outputs(self) -> int:
        return len(self.output_indices) if self.output_indices else 1

    def num_ops(self) -> int:
        return len(self.opcodes)

    def counts(self) -> dict:
        """Entry counts used by the traffic model: n, N (computed), C, m."""
        ops = np.asarray(self.opcodes, dtype=np.int64)
        n_const = int(np.count_nonzero(ops == OpCode.Const))
        return {
            "n": self.num_inputs,
            "N": len(self.opcodes) - self.num_inputs - n_const,
            "C": n_const,
            "m": self.num_outputs(),
            "entries": len(self.opcodes),
        }


# ---------------------------------------------------------------------------
# BReverse: recording variable (src/breverse.rs, src/traits/breverse_ops.rs)
# ---------------------------------------------------------------------------

_ACTIVE: List[BytecodeTape] = []


def _tape() -> BytecodeTape:
    if not _ACTIVE:
        raise RuntimeError("no active recording")
    return _ACTIVE[-1]


def _rust_signum(x: float) -> float:
    if math.isnan(x):
        return x
    return math.copysign(1.0, x)


def _rust_round(x: float) -> float:
    # Round half away from zero (Rust f64::round).
    if not math.isfinite(x):
        return x
    r = math.floor(abs(x) + 0.5)
    # guard for values where abs(x)+0.5 rounds up incorrectly
    if r - abs(x) > 0.5:
        r -= 1.0
    return math.copysign(r, x)


def _rust_rem(a: float, b: float) -> float:
    try:
        return math.fmod(a, b)
    except ValueError:
        return math.nan


def _safe(f, *args):
    try:
        return f(*args)
    except (ValueError, OverflowError, ZeroDivisionError):
        return _ieee_fallback(f, *args)


def _ieee_fallback(f, *args):
    with np.errstate(all="ignore"):
        name = getattr(f, "__name__", "")
        table = {
            "exp": np.exp, "log": np.log, "sqrt": np.sqrt, "sin": np.sin,
            "cos": np.cos, "tan": np.tan, "asin": np.arcsin, "acos": np.arccos,
            "sinh": np.sinh, "cosh": np.cosh, "acosh": np.arccosh,
            "atanh": np.arctanh, "log2": np.log2, "log10": np.log10,
            "log1p": np.log1p, "expm1": np.expm1, "pow": np.power,
        }
        g = table.get(name)
        if g is None:
            return math.nan
        return float(g(*[np.float64(a) for a in args]))


class BReverse:
    """A recorded scalar: value plus tape index (CONSTANT when untracked)."""

    __slots__ = ("value", "index")

    def __init__(self, value: float, index: int = CONSTANT):
        self.value = float(value)
        self.index = index

    @staticmethod
    def constant(value: float) -> "BReverse":
        return BReverse(value, CONSTANT)

    def _on_tape(self, t: BytecodeTape) -> int:
        if self.index == CONSTANT:
            return t.push_const(self.value)
        return self.index

    @staticmethod
    def _lift(x) -> "BReverse":
        return x if isinstance(x, BReverse) else BReverse.constant(float(x))

    def _bin(self, other, op: int, value: float, swap: bool = False) -> "BReverse":
        o = BReverse._lift(other)
        lhs, rhs = (o, self) if swap else (self, o)
        t = _tape()
        li = lhs._on_tape(t)
        ri = rhs._on_tape(t)
        return BReverse(value, t.push_op(op, li, ri, value))

    def _un(self, op: int, value: float) -> "BReverse":
        t = _tape()
        xi = self._on_tape(t)
        return BReverse(value, t.push_op(op, xi, UNUSED, value))

    # arithmetic
    def __add__(self, o):
        return self._bin(o, OpCode.Add, self.value + BReverse._lift(o).value)

    def __radd__(self, o):
        return self._bin(o, OpCode.Add, BReverse._lift(o).value + self.value, swap=True)

    def __sub__(self, o):
        return self._bin(o, OpCode.Sub, self.value - BReverse._lift(o).value)

    def __rsub__(self, o):
        return self._bin(o, OpCode.Sub, BReverse._lift(o).value - self.value, swap=True)

    def __mul__(self, o):
        return self._bin(o, OpCode.Mul, self.value * BReverse._lift(o).value)

    def __rmul__(self, o):
        return self._bin(o, OpCode.Mul, BReverse._lift(o).value * self.value, swap=True)

    def __truediv__(self, o):
        b = BReverse._lift(o).value
        return self._bin(o, OpCode.Div, _div(self.value, b))

    def __rtruediv__(self, o):
        a = BReverse._lift(o).value
        return self._bin(o, OpCode.Div, _div(a, self.value), swap=True)

    def __mod__(self, o):
        return self._bin(o, OpCode.Rem, _rust_rem(self.value, BReverse._lift(o).value))

    def __neg__(self):
        return self._un(OpCode.Neg, -self.value)

    # elementals (src/traits/breverse_num_traits.rs:255-420)
    def powi(self, n: int) -> "BReverse":
        with np.errstate(all="ignore"):
            val = 1.0 if n == 0 else float(np.power(np.float64(self.value), float(n)))
        t = _tape()
        xi = self._on_tape(t)
        return BReverse(val, t.push_powi(xi, int(n), val))

    def powf(self, o) -> "BReverse":
        return self._bin(o, OpCode.Powf, _safe(math.pow, self.value, BReverse._lift(o).value))

    def atan2(self, o) -> "BReverse":
        return self._bin(o, OpCode.Atan2, math.atan2(self.value, BReverse._lift(o).value))

    def hypot(self, o) -> "BReverse":
        return self._bin(o, OpCode.Hypot, math.hypot(self.value, BReverse._lift(o).value))

    def max(self, o) -> "BReverse":
        b = BReverse._lift(o).value
        v = self.value if (self.value >= b or math.isnan(b)) else b
        return self._bin(o, OpCode.Max, v)

    def min(self, o) -> "BReverse":
        b = BReverse._lift(o).value
        v = self.value if (self.value <= b or math.isnan(b)) else b
        return self._bin(o, OpCode.Min, v)

    def recip(self):
        return self._un(OpCode.Recip, _div(1.0, self.value))

    def sqrt(self):
        return self._un(OpCode.Sqrt, _safe(math.sqrt, self.value))

    def cbrt(self):
        return self._un(OpCode.Cbrt, float(np.cbrt(self.value)))

    def exp(self):
        return self._un(OpCode.Exp, _safe(math.exp, self.value))

    def exp2(self):
        return self._un(OpCode.Exp2, float(np.exp2(self.value)))

    def exp_m1(self):
        return self._un(OpCode.ExpM1, _safe(math.expm1, self.value))

    def ln(self):
        return self._un(OpCode.Ln, _safe(math.log, self.value))

    def log2(self):
        return self._un(OpCode.Log2, _safe(math.log2, self.value))

    def log10(self):
        return self._un(OpCode.Log10, _safe(math.log10, self.value))

    def ln_1p(self):
        return self._un(OpCode.Ln1p, _safe(math.log1p, self.value))

    def sin(self):
        return self._un(OpCode.Sin, _safe(math.sin, self.value))

    def cos(self):
        return self._un(OpCode.Cos, _safe(math.cos, self.value))

    def tan(self):
        return self._un(OpCode.Tan, _safe(math.tan, self.value))

    def asin(self):
        return self._un(OpCode.Asin, _safe(math.asin, self.value))

    def acos(self):
        return self._un(OpCode.Acos, _safe(math.acos, self.value))

    def atan(self):
        return self._un(OpCode.Atan, math.atan(self.value))

    def sinh(self):
        return self._un(OpCode.Sinh, _safe(math.sinh, self.value))

    def cosh(self):
        return self._un(OpCode.Cosh, _safe(math.cosh, self.value))

    def tanh(self):
        return self._un(OpCode.Tanh, math.tanh(self.value))

    def asinh(self):
        return self._un(OpCode.Asinh, math.asinh(self.value))

    def acosh(self):
        return self._un(OpCode.Acosh, _safe(math.acosh, self.value))

    def atanh(self):
        return self._un(OpCode.Atanh, _safe(math.atanh, self.value))

    def abs(self):
        return self._un(OpCode.Abs, abs(self.value))

    def signum(self):
        return self._un(OpCode.Signum, _rust_signum(self.value))

    def floor(self):
        return self._un(OpCode.Floor, float(np.floor(self.value)))

    def ceil(self):
        return self._un(OpCode.Ceil, float(np.ceil(self.value)))

    def round(self):
        return self._un(OpCode.Round, _rust_round(self.value))

    def trunc(self):
        return self._un(OpCode.Trunc, float(np.trunc(self.value)))

    def fract(self):
        return self._un(OpCode.Fract, self.value - float(np.trunc(self.value)))


def _div(a: float, b: float) -> float:
    with np.errstate(all="ignore"):
        return float(np.float64(a) / np.float64(b))


def record(f: Callable[[List[BReverse]], BReverse], x: Sequence[float]):
    """Record ``f`` at ``x`` (``src/api.rs:197-216``). Returns ``(tape, value)``."""
    tape = BytecodeTape()
    inputs = [BReverse(v, tape.new_input(float(v))) for v in x]
    _ACTIVE.append(tape)
    try:
        out = f(inputs)
    finally:
        _ACTIVE.pop()
    out = BReverse._lift(out)
    idx = out._on_tape(tape)
    tape.set_output(idx)
    return tape, out.value


def record_multi(f: Callable[[List[BReverse]], Sequence[BReverse]], x: Sequence[float]):
    """Multi-output recording (``src/api.rs:218-260``)."""
    tape = BytecodeTape()
    inputs = [BReverse(v, tape.new_input(float(v))) for v in x]
    _ACTIVE.append(tape)
    try:
        outs = list(f(inputs))
    finally:
        _ACTIVE.pop()
    if not outs:
        raise ValueError("record_multi: f returned no outputs")
    idxs = [BReverse._lift(o)._on_tape(tape) for o in outs]
    tape.set_outputs(idxs)
    return tape, [BReverse._lift(o).value for o in outs]


# ---------------------------------------------------------------------------
# Flattened upload format (src/gpu/mod.rs:313-565)
# ---------------------------------------------------------------------------


@dataclass
class GpuTapeData:
    """Flattened tape for upload; arrays all have ``num_ops`` entries."""

    opcodes: np.ndarray
    arg0: np.ndarray
    arg1: np.ndarray
    constants: np.ndarray  # float32 (f32 upload) or float64 (f64 upload)
    num_ops: int
    num_inputs: int
    num_variables: int
    output_index: int
    output_indices: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))

    @staticmethod
    def _build(tape: BytecodeTape, constants: np.ndarray) -> "GpuTapeData":
        n = len(tape.opcodes)
        return GpuTapeData(
            opcodes=np.asarray(tape.opcodes, dtype=np.uint32),
            arg0=np.asarray(tape.arg0, dtype=np.uint32),
            arg1=np.asarray(tape.arg1, dtype=np.uint32),
            constants=constants,
            num_ops=n,
            num_inputs=tape.num_inputs,
            num_variables=n,
            output_index=tape.output_index,
            output_indices=np.asarray(tape.all_output_indices(), dtype=np.uint32),
        )

    @staticmethod
    def from_tape(tape: BytecodeTape) -> "GpuTapeData":
        """f32 upload data (``from_tape``, ``src/gpu/mod.rs:365-370``)."""
        with np.errstate(over="ignore"):
            return GpuTapeData._build(tape, np.asarray(tape.values, dtype=np.float64).astype(np.float32))

    from_tape_f64_lossy = from_tape

    @staticmethod
    def from_tape_f64(tape: BytecodeTape) -> "GpuTapeData":
        """f64 upload data (the CUDA backend's ``upload_tape_f64``, ``src/gpu/cuda_backend.rs:566-611``)."""
        return GpuTapeData._build(tape, np.asarray(tape.values, dtype=np.float64))

    def validate(self) -> None:
        """Structural checks of ``GpuTapeData::validate`` (``src/gpu/mod.rs:397-564``).

        Raises ``GpuError(OTHER)`` with the reference's message shapes.
        """
        def err(msg):
            raise GpuError(GpuError.OTHER, f"invalid GpuTapeData: {msg}")

        n = self.num_ops
        if len(self.opcodes) != n:
            err(f"opcodes.len() ({len(self.opcodes)}) != num_ops ({n})")
        if len(self.arg0) != n:
            err(f"arg0.len() ({len(self.arg0)}) != num_ops ({n})")
        if len(self.arg1) != n:
            err(f"arg1.len() ({len(self.arg1)}) != num_ops ({n})")
        if self.num_ops != self.num_variables:
            err(f"num_ops ({self.num_ops}) != num_variables ({self.num_variables})")
        if len(self.constants) != self.num_variables:
            err(f"constants.len() ({len(self.constants)}) != num_variables ({self.num_variables})")
        if self.num_inputs > self.num_variables:
            err(f"num_inputs ({self.num_inputs}) > num_variables ({self.num_variables})")
        ni = self.num_inputs
        for i in range(n):
            op = int(self.opcodes[i])
            if op > OpCode.Fract:
                err(f"opcodes[{i}] ({op}) is not a GPU-executable opcode")
            if (op == OpCode.Input) != (i < ni):
                if i < ni:
                    err(f"opcodes[{i}] ({op}) should be Input")
                err(f"opcodes[{i}] is Input, but only the first {ni} entries may be inputs")
        for i in range(ni, n):
            op = int(self.opcodes[i])
            if op == OpCode.Const:
                continue
            a = int(self.arg0[i])
            if a >= i:
                err(f"arg0[{i}] ({a}) must reference an earlier slot")
            b = int(self.arg1[i])
            if op == OpCode.Powi:
                continue
            if op in BINARY_OPS:
                if b >= i:
                    err(f"arg1[{i}] ({b}) must reference an earlier slot")
            elif b != UNUSED:
                err(f"arg1[{i}] ({b}) should be the UNUSED sentinel for unary opcode {op}")
        if self.output_index >= self.num_variables:
            err(f"output_index ({self.output_index}) >= num_variables ({self.num_variables})")
        for j, oi in enumerate(self.output_indices):
            if int(oi) >= self.num_variables:
                err(f"output_indices[{j}] ({int(oi)}) >= num_variables ({self.num_variables})")

    def counts(self) -> dict:
        ops = np.asarray(self.opcodes)
        n_const = int(np.count_nonzero(ops == OpCode.Const))
        m = len(self.output_indices) if len(self.output_indices) else 1
        return {"n": self.num_inputs, "N": self.num_ops - self.num_inputs - n_const,
                "C": n_const, "m": m, "entries": self.num_ops}


def f32_bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]
