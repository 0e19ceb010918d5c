"""Numpy emulator of the scalar register-program interpreter (test helper).

Executes the words produced by ``ehip_compile_program`` exactly as the HIP
scalar interpreter does — same register numbers, same FIRST/ACC flags, same
DOT term order — vectorised over points.  Comparing it with the CPU oracle
checks the upload-time compiler (pruning, liveness, register reuse, fusion,
adjoint order) without a GPU.  Arithmetic is numpy's, so transcendentals may
differ from the oracle by an ulp; structure errors show up as gross
mismatches.
"""

from __future__ import annotations

import numpy as np

I_LOAD, I_STORE_OUT, I_SEED, I_STORE_GRAD, I_DOT, I_DOT_REV, I_LOADN, I_REV = 64, 65, 66, 67, 68, 69, 70, 128
F_DOT_CONT = 1024
F_A_IMM, F_B_IMM, F_FIRST_A, F_FIRST_B, F_ACC_A, F_ACC_B, F_HAS_B, F_RD_A, F_RD_B, F_RD_R = (
    1, 2, 4, 8, 16, 32, 64, 128, 256, 512)
T_PROD, T_REG, T_IMM = 0, 1, 2


def _imm(lo, hi, f64):
    if f64:
        return np.array([int(lo) | (int(hi) << 32)], dtype=np.uint64).view(np.float64)[0]
    return np.array([int(lo)], dtype=np.uint32).view(np.float32)[0]


def _i32(v):
    v = int(v) & 0xFFFFFFFF
    return v - (1 << 32) if v & 0x80000000 else v


def _powi(a, n, F):
    r = np.ones_like(a)
    recip = n < 0
    e = n
    b = a.copy()
    while True:
        if e & 1:
            r = r * b
        e = int(e / 2)
        if e == 0:
            break
        b = b * b
    return F(1) / r if recip else r


def _eval(op, a, b, F):
    one = F(1)
    return {
        2: lambda: a + b, 3: lambda: a - b, 4: lambda: a * b, 5: lambda: a / b, 6: lambda: np.fmod(a, b),
        7: lambda: np.power(a, b), 8: lambda: np.arctan2(a, b), 9: lambda: np.hypot(a, b),
        10: lambda: np.where((a >= b) | np.isnan(b), a, b), 11: lambda: np.where((a <= b) | np.isnan(b), a, b),
        12: lambda: -a, 13: lambda: one / a, 14: lambda: np.sqrt(a), 15: lambda: np.cbrt(a),
        17: lambda: np.exp(a), 18: lambda: np.exp2(a), 19: lambda: np.expm1(a), 20: lambda: np.log(a),
        21: lambda: np.log2(a), 22: lambda: np.log10(a), 23: lambda: np.log1p(a), 24: lambda: np.sin(a),
        25: lambda: np.cos(a), 26: lambda: np.tan(a), 27: lambda: np.arcsin(a), 28: lambda: np.arccos(a),
        29: lambda: np.arctan(a), 30: lambda: np.sinh(a), 31: lambda: np.cosh(a), 32: lambda: np.tanh(a),
        33: lambda: np.arcsinh(a), 34: lambda: np.arccosh(a), 35: lambda: np.arctanh(a), 36: lambda: np.abs(a),
        37: lambda: np.where(np.isnan(a), a, np.copysign(one, a)), 38: lambda: np.floor(a),
        39: lambda: np.ceil(a), 40: lambda: np.where(a >= 0, np.floor(a + F(0.5)), np.ceil(a - F(0.5))),
        41: lambda: np.trunc(a), 42: lambda: a - np.trunc(a),
    }[op]()


def _partials(op, a, b, r, F):
    z = np.zeros_like(r) if hasattr(r, "shape") else F(0)
    one = F(1)
    nan = F(np.nan)
    if op == 2:
        return one + z, one + z
    if op == 3:
        return one + z, -one + z
    if op == 4:
        return b + z, a + z
    if op == 5:
        inv = one / b
        return inv + z, -r * inv
    if op == 6:
        return one + z, -np.trunc(a / b)
    if op == 12:
        return -one + z, z
    if op == 13:
        inv = one / a
        return -inv * inv, z
    if op == 14:
        return one / ((one + one) * r), z
    if op == 17:
        return r + z, z
    if op == 20:
        return np.where(a >= 0, one / a, nan), z
    if op == 24:
        return np.cos(a), z
    if op == 25:
        return -np.sin(a), z
    if op == 32:
        c = np.cosh(a)
        return one / (c * c), z
    if op == 36:
        return np.where(a == 0, F(0), np.copysign(one, a)), z
    if op in (10, 11):
        win = ((a >= b) if op == 10 else (a <= b)) | np.isnan(b)
        return np.where(win, one, F(0)) + z, np.where(win, F(0), one) + z
    if op in (37, 38, 39, 40, 41):
        return z, z
    if op == 42:
        return one + z, z
    raise NotImplementedError(f"emulator partial for opcode {op}")


def split_terms(terms, nterms, f64):
    """meta[T] | reg[T] | imm[T*IW] -> per-term tuples (meta, reg, lo, hi)."""
    t = np.asarray(terms, dtype=np.uint32).reshape(-1)
    meta = t[:nterms]
    reg = t[nterms:2 * nterms]
    iw = 2 if f64 else 1
    imm = t[2 * nterms:2 * nterms + iw * nterms]
    out = []
    for k in range(nterms):
        lo = int(imm[k * iw])
        hi = int(imm[k * iw + 1]) if f64 else 0
        out.append((int(meta[k]), int(reg[k]), lo, hi))
    return out


WIN_WORDS = 512
I_WEND, I_END = 71, 72
I_DENSE, I_DENSE_REV = 73, 74
P_KIND_SHIFT, P_FIRST, P_CLEFT, P_REG_MASK = 30, 1 << 29, 1 << 28, (1 << 28) - 1


def decode_stream(stream, f64):
    """Windowed stream -> [(w[8], inline terms)], checking the window layout.

    Every instruction lies inside one window, a window ends with I_WEND (or
    the final I_END) followed by zero padding, and the length field of each
    header matches its contents.
    """
    st = [int(v) for v in np.asarray(stream, dtype=np.uint32)]
    assert len(st) % WIN_WORDS == 0 and st, "stream is not whole windows"
    tww = 4 if f64 else 3
    out = []
    nwin = len(st) // WIN_WORDS
    for wi in range(nwin):
        win = st[wi * WIN_WORDS:(wi + 1) * WIN_WORDS]
        pos = 0
        while True:
            assert pos < WIN_WORDS, "instruction runs past its window"
            w0 = win[pos]
            op = w0 & 0xFF
            if op in (I_WEND, I_END):
                assert (op == I_END) == (wi == nwin - 1), "misplaced stream end"
                assert all(v == 0 for v in win[pos + 1:]), "window padding is not zero"
                break
            ln = w0 >> 24
            assert ln >= 3 and pos + ln < WIN_WORDS, "instruction crosses its window"
            ins = win[pos:pos + ln]
            head = [ins[0] & 0x00FFFFFF] + ins[1:min(ln, 8)] + [0] * max(0, 8 - ln)
            tl = []
            if op in (I_DOT, I_DOT_REV, I_LOADN):
                k = ins[2]
                assert ln == 3 + k * tww, "term count does not match length"
                for j in range(k):
                    t = ins[3 + j * tww:3 + (j + 1) * tww]
                    tl.append((t[0], t[1], t[2], t[3] if f64 else 0))
                head = [ins[0] & 0x00FFFFFF, ins[1], k, 0, 0, 0, 0, 0]
            else:
                assert ln == 8, "ordinary instructions are 8 words"
            out.append((head, tl))
            pos += ln
    return out


def run(words, terms, x, nin, nout, nregs, f64=True, mode=1, nterms=None, pool=None):
    """Run the program over points ``x`` [B, nin]; returns (y [B,nout], g [B,nin])."""
    tl = split_terms(terms, nterms, f64)
    instrs = []
    for w in words:
        w = [int(v) for v in w]
        op = w[0] & 0xFF
        if op in (I_DOT, I_DOT_REV, I_LOADN):
            instrs.append((w, tl[w[3]:w[3] + w[2]]))
        else:
            instrs.append((w, []))
    return execute(instrs, x, nin, nout, nregs, f64, pool)


def run_stream(stream, x, nin, nout, nregs, f64=True, pool=None):
    """Run a windowed stream (``ehip_compile_program`` modes 2/3)."""
    return execute(decode_stream(stream, f64), x, nin, nout, nregs, f64, pool)


def dense_parts(w, pool, f64):
    """(term words, member registers, weights) of a DENSE / DENSE_REV."""
    pool = np.asarray(pool, dtype=np.uint32)
    m, T, o_r, o_w = w[1], w[2], w[3], w[4]
    tw = [int(v) for v in pool[o_r:o_r + T]]
    mem = [int(v) for v in pool[o_r + T:o_r + T + m]]
    if f64:
        W = pool[o_w:o_w + 2 * m * T].view(np.float64)
    else:
        W = pool[o_w:o_w + m * T].view(np.float32)
    return tw, mem, W


def execute(instrs, x, nin, nout, nregs, f64=True, pool=None):
    F = np.float64 if f64 else np.float32
    B = x.shape[0]
    regs = np.full((nregs, B), np.nan, dtype=F)
    y = np.zeros((B, nout), F)
    g = np.zeros((B, nin), F)
    with np.errstate(all="ignore"):
        for w, terms in instrs:
            op, fl = w[0] & 0xFF, w[0] >> 8

            def val(lo, hi, imm_flag):
                return F(_imm(lo, hi, f64)) if fl & imm_flag else regs[lo].copy()

            if op < 64:
                a = val(w[2], w[3], F_A_IMM)
                b = val(w[4], w[5], F_B_IMM) if fl & F_HAS_B else F(0)
                r = _powi(np.broadcast_to(a, (B,)).astype(F), _i32(w[7]), F) if op == 16 else _eval(op, a, b, F)
                regs[w[1]] = r
            elif op >= I_REV:
                top = op - I_REV
                gg = regs[w[1]].copy()
                live = gg != 0
                a = F(_imm(w[2], w[3], f64)) if fl & F_A_IMM else (regs[w[2]].copy() if fl & F_RD_A else np.zeros(B, F))
                b = F(_imm(w[4], w[5], f64)) if fl & F_B_IMM else (regs[w[4]].copy() if fl & F_RD_B else np.zeros(B, F))
                r = regs[w[6]].copy() if fl & F_RD_R else np.zeros(B, F)
                if top == 16:
                    e = _i32(w[7])
                    aa = np.broadcast_to(a, (B,)).astype(F)
                    da = np.zeros(B, F) if e == 0 else F(e) * _powi(aa, e - 1, F)
                    db = np.zeros(B, F)
                else:
                    da, db = _partials(top, a, b, r, F)
                    da = np.broadcast_to(da, (B,)).astype(F)
                    db = np.broadcast_to(db, (B,)).astype(F)
                if fl & F_ACC_A:
                    cur = np.zeros(B, F) if fl & F_FIRST_A else regs[w[3]].copy()
                    m = live & (da != 0)
                    cur[m] = cur[m] + da[m] * gg[m]
                    regs[w[3]] = cur
                if fl & F_ACC_B:
                    cur = np.zeros(B, F) if fl & F_FIRST_B else regs[w[5]].copy()
                    m = live & (db != 0)
                    cur[m] = cur[m] + db[m] * gg[m]
                    regs[w[5]] = cur
            elif op == I_LOADN:
                for k in range(w[2]):
                    t0, t1, t2, t3 = terms[k]
                    regs[t1] = x[:, t2]
            elif op == I_DOT:
                acc = regs[w[1]].copy() if fl & F_DOT_CONT else None
                for k in range(w[2]):
                    t0, t1, t2, t3 = terms[k]
                    c = F(_imm(t2, t3, f64))
                    kind = t0 & 0x3
                    term = (regs[t1] * c if kind == T_PROD else regs[t1].copy()) if kind != T_IMM else np.full(B, c, F)
                    acc = term if acc is None else acc + term
                regs[w[1]] = acc
            elif op == I_DOT_REV:
                gg = regs[w[1]].copy()
                live = gg != 0
                for k in range(w[2]):
                    t0, t1, t2, t3 = terms[k]
                    c = F(_imm(t2, t3, f64))
                    kind = t0 & 0x3
                    cur = np.zeros(B, F) if (t0 >> 2) & 1 else regs[t1].copy()
                    if kind == T_PROD:
                        if c != 0:
                            cur[live] = cur[live] + c * gg[live]
                    else:
                        cur[live] = cur[live] + F(1) * gg[live]
                    regs[t1] = cur
            elif op == I_DENSE:
                tw, mem, W = dense_parts(w, pool, f64)
                m, T = len(mem), len(tw)
                W = W.reshape(T, m)
                for j in range(m):
                    acc = None
                    for p, word in enumerate(tw):
                        kind, r = word >> P_KIND_SHIFT, word & P_REG_MASK
                        c = F(W[p, j])
                        term = regs[r] * c if kind == T_PROD else (regs[r].copy() if kind == T_REG
                                                                   else np.full(B, c, F))
                        acc = term if acc is None else acc + term
                    regs[mem[j]] = acc
            elif op == I_DENSE_REV:
                tw, mem, W = dense_parts(w, pool, f64)
                m, T = len(mem), len(tw)
                W = W.reshape(m, T)
                for i, word in enumerate(tw):
                    kind, r = word >> P_KIND_SHIFT, word & P_REG_MASK
                    cur = np.zeros(B, F) if word & P_FIRST else regs[r].copy()
                    for j in range(m):
                        gg = regs[mem[j]]
                        live = gg != 0
                        c = F(W[j, i])
                        if kind == T_PROD:
                            if c != 0:
                                cur[live] = cur[live] + c * gg[live]
                        else:
                            cur[live] = cur[live] + F(1) * gg[live]
                    regs[r] = cur
            elif op == I_LOAD:
                regs[w[1]] = x[:, w[2]]
            elif op == I_STORE_OUT:
                y[:, w[1]] = val(w[2], w[3], F_A_IMM)
            elif op == I_SEED:
                regs[w[1]] = F(1)
            elif op == I_STORE_GRAD:
                g[:, w[1]] = val(w[2], w[3], F_A_IMM)
            else:
                raise ValueError(f"unknown instruction {op}")
    return y, g
