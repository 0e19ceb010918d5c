"""Algorithmic arithmetic per point of a tape's sweeps (the roofline's work).

The tape-specialised kernels keep every value in registers, so their HBM
traffic is only the inputs in and the results out and their bound is FP32
(or FP64) vector issue.  The work they are priced by is the arithmetic the
sweep cannot avoid, counted here from the tape's own opcode arrays, one
floating-point operation per add, subtract, multiply, divide or elementary
function (a transcendental counts as one, so the figure is a lower bound on
the instructions any implementation issues):

* forward: one operation per computed entry (opcode not Input/Const);
* reverse (entries the output depends on, in the reference's descending
  order, src/bytecode_tape/reverse.rs:32-105): per variable operand, the
  partial's own arithmetic (none for add, sub, neg and mul, whose partials
  are +-1 or the other operand; one for each unary function; div 1 and 2;
  powf and atan2/hypot 2), one multiply by the adjoint unless the partial
  is the constant +-1, and one add for every contribution to an adjoint
  after its first (the first is a copy);
* HVP (forward-over-reverse in dual numbers): each real operation above
  becomes its dual form: add 2, multiply by a constant 2, multiply of two
  duals 3 + 1, unary function 3 (value, derivative, tangent product);
* order-K Taylor jet along one direction: add K, multiply by a constant K,
  multiply of two jets K(K+1)/2 products and K(K-1)/2 adds, unary function
  about 2K^2 (its recurrence), per computed entry.

`jet_flops_split` counts the same jet arithmetic where it has to happen
when S directions share one point (STDE): the primal coefficient c0 is a
function of x alone, once per point, and the higher coefficients once per
direction; coefficients that are zero by construction (an input's jet is
(x, v, 0, ..), a constant's (c, 0, ..)) and their products are not counted.

`compulsory_bytes` is the HBM traffic no implementation avoids: the inputs
read once and the results written once.
"""

from __future__ import annotations

from typing import Dict

import numpy as np

from .tape import UNUSED, OpCode

O = OpCode
_LINEAR = {int(O.Add), int(O.Sub)}
_BINARY = set(range(int(O.Add), int(O.Min) + 1))
_PIECEWISE = {int(O.Max), int(O.Min), int(O.Abs), int(O.Signum), int(O.Floor), int(O.Ceil), int(O.Round),
              int(O.Trunc), int(O.Fract), int(O.Neg)}


def _arrays(tape):
    ops = np.asarray(tape.opcodes, dtype=np.int64)
    a0 = np.asarray(tape.arg0, dtype=np.int64)
    a1 = np.asarray(tape.arg1, dtype=np.int64)
    outs = list(tape.all_output_indices()) if hasattr(tape, "all_output_indices") else list(tape.output_indices)
    return ops, a0, a1, outs


def _variable(ops):
    return (ops != int(O.Input)) & (ops != int(O.Const))


def _partial_cost(op: int, which: int) -> int:
    if op in _LINEAR or op == int(O.Mul) or op in _PIECEWISE or op == int(O.Rem):
        return 0
    if op == int(O.Div):
        return 1 if which == 0 else 2
    if op in (int(O.Powf), int(O.Atan2), int(O.Hypot)):
        return 2
    return 1  # unary functions, powi


def _unit_partial(op: int, which: int) -> bool:
    """The partial is the constant +-1 (no multiply by the adjoint)."""
    return op in _LINEAR or op in (int(O.Neg), int(O.Fract), int(O.Max), int(O.Min)) or (
        op == int(O.Rem) and which == 0)


def _operands(ops, a0, a1, i):
    op = int(ops[i])
    out = [(0, int(a0[i]))]
    if op in _BINARY and a1[i] != UNUSED:
        out.append((1, int(a1[i])))
    return op, out


def sweep_counts(tape) -> Dict[str, int]:
    """Operation counts per point: forward, reverse (gradient of output 0)."""
    ops, a0, a1, outs = _arrays(tape)
    var = _variable(ops)
    N = len(ops)
    fwd = int(np.count_nonzero(var))
    live = np.zeros(N, dtype=bool)
    live[outs[0]] = True
    rev = 0
    contrib = np.zeros(N, dtype=np.int64)
    for i in range(N - 1, -1, -1):
        if not (live[i] and var[i]):
            continue
        op, operands = _operands(ops, a0, a1, i)
        for which, a in operands:
            if ops[a] == int(O.Const):
                continue
            if op in (int(O.Signum), int(O.Floor), int(O.Ceil), int(O.Round), int(O.Trunc)):
                continue  # zero partial: nothing flows
            live[a] = True
            rev += _partial_cost(op, which) + (0 if _unit_partial(op, which) else 1)
            contrib[a] += 1
    rev += int(np.sum(np.maximum(contrib - 1, 0)))
    return {"forward": fwd, "reverse": rev}


def fwdgrad_flops(tape) -> int:
    c = sweep_counts(tape)
    return c["forward"] + c["reverse"]


def hvp_flops(tape) -> int:
    """Dual-number forward and reverse: about twice the real sweep for linear
    operations and four times for products of two varying operands."""
    ops, a0, a1, _ = _arrays(tape)
    var = _variable(ops)
    total = 0
    for i in np.nonzero(var)[0]:
        op, operands = _operands(ops, a0, a1, int(i))
        nvar = sum(1 for _, a in operands if ops[a] != int(O.Const))
        if op in _LINEAR:
            total += 2
        elif op == int(O.Mul):
            total += 4 if nvar == 2 else 2
        else:
            total += 3
    c = sweep_counts(tape)
    return total + 2 * c["reverse"]


def jet_flops(tape, K: int = 3) -> int:
    """One order-K Taylor jet (one point, one direction) of every computed entry."""
    ops, a0, a1, _ = _arrays(tape)
    var = _variable(ops)
    total = 0
    for i in np.nonzero(var)[0]:
        op, operands = _operands(ops, a0, a1, int(i))
        nvar = sum(1 for _, a in operands if ops[a] != int(O.Const))
        if op in _LINEAR or op == int(O.Neg):
            total += K
        elif op == int(O.Mul):
            total += K * K if nvar == 2 else K
        else:
            total += 2 * K * K
    return total


def jet_flops_split(tape, K: int = 3):
    """(per point, per direction) arithmetic of order-K jets seeded (x, v, 0, ..):
    per coefficient k that can be non-zero, a multiply by a constant or a
    negation costs 1, and an add 1 where both of its operands can be
    non-zero (else it copies); a product of two jets costs 2p - 1 for its p
    non-zero pairs a_j b_(k-j); an elementary function (and a division or
    binary function of two jets) costs 2K per coefficient, as `jet_flops`
    charges 2K^2 per entry.  Coefficient 0 goes to the point, the rest to
    the direction.  An elementary function's c_k is non-zero when one of its
    operand's c_1 .. c_k is."""
    ops, a0, a1, _ = _arrays(tape)
    full = (1 << K) - 1
    nz = np.zeros(len(ops), dtype=np.int64)
    point = lane = 0
    for i in range(len(ops)):
        op = int(ops[i])
        if op == int(O.Input):
            nz[i] = 0b11 & full
            continue
        if op == int(O.Const):
            nz[i] = 1
            continue
        _, operands = _operands(ops, a0, a1, i)
        ms = [int(nz[a]) for _, a in operands]
        var = [ops[a] != int(O.Const) for _, a in operands]
        cost = [0] * K
        if op in _LINEAR and len(ms) == 2:
            nz[i] = ms[0] | ms[1]
            for k in range(K):  # an add where both coefficients can be non-zero, else a copy
                if (ms[0] >> k) & (ms[1] >> k) & 1:
                    cost[k] = 1
        elif op in _LINEAR or op == int(O.Neg) or (op in (int(O.Mul), int(O.Div)) and not all(var)):
            m = 0
            for x, v in zip(ms, var):
                if v:
                    m |= x
            nz[i] = m | 1
            for k in range(K):
                if (m >> k) & 1:
                    cost[k] = 1
        elif op == int(O.Mul):
            m = 0
            for k in range(K):
                pairs = sum(1 for j in range(k + 1) if (ms[0] >> j) & 1 and (ms[1] >> (k - j)) & 1)
                if pairs:
                    m |= 1 << k
                    cost[k] = 2 * pairs - 1
            nz[i] = m
        else:
            m = 1
            anyv = 0
            for x in ms:
                anyv |= x
            for k in range(1, K):
                if anyv & (((1 << (k + 1)) - 1) & ~1):
                    m |= 1 << k
            nz[i] = m
            for k in range(K):
                if (m >> k) & 1:
                    cost[k] = 2 * K
        point += cost[0]
        lane += sum(cost[1:])
    return point, lane


def compulsory_bytes(tape, what: str, scalar_bytes: int) -> int:
    """HBM bytes per point: inputs read once, results written once."""
    n = tape.num_inputs
    m = len(tape.all_output_indices()) if hasattr(tape, "all_output_indices") else len(tape.output_indices)
    per = {"forward": n + m, "gradient": n + m + n, "hvp": 2 * n + 2 * n}[what]
    return per * scalar_bytes
