"""Philox4x32-10 counter-based generator (host side, numpy).

Algorithm: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
1, 2, 3" (SC'11), Philox4x32 with 10 rounds: multipliers 0xD2511F53 and
0xCD9E8D57, Weyl key increments 0x9E3779B9 and 0xBB67AE85.  The device copy
in ``csrc/philox.h`` implements the same function, so points, weights and
STDE directions drawn on either side agree bit for bit in their u32 words.

Streams used by this package (key = (seed, stream)):
  counter = (block_lo, block_hi, 0, 0); one Philox block per scalar draw.
Uniform in (0, 1): 53-bit mantissa from words 0,1 with a half-ulp offset.
Gaussian: Box-Muller in binary64 from two such uniforms (words 0,1 and 2,3),
keeping the cosine branch.  Rademacher: sign from bit 0 of word 0.

Direction components (``direction``: the STDE and jet directions drawn on
the device, stream STREAM_DIRS) share blocks: Rademacher component idx is
bit idx mod 128 of block idx // 128 (bit b of word b // 32, set = -1);
Gaussian components 4k .. 4k + 3 are the cosine and sine branches of block
k's two Box-Muller pairs (radius from word 0 or 2, angle from word 1 or 3,
each uniform (top 23 bits + 1/2) 2^-23), computed from binary32 basic
operations (``dir_log``, ``dir_cos_sin``) in the device's order, so both
sides give the same bits.
"""

from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)

# Stream tags for the synthetic workloads.
STREAM_POINTS = 1
STREAM_W1 = 2
STREAM_B1 = 3
STREAM_W2 = 4
STREAM_B2 = 5
STREAM_W3 = 6
STREAM_DIRS = 7
STREAM_DAG = 8
STREAM_CV = 9


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10. All inputs broadcastable uint32 arrays."""
    c0 = np.asarray(c0, dtype=np.uint32).astype(np.uint64)
    c1 = np.asarray(c1, dtype=np.uint32).astype(np.uint64)
    c2 = np.asarray(c2, dtype=np.uint32).astype(np.uint64)
    c3 = np.asarray(c3, dtype=np.uint32).astype(np.uint64)
    k0 = np.asarray(k0, dtype=np.uint32).astype(np.uint64)
    k1 = np.asarray(k1, dtype=np.uint32).astype(np.uint64)
    for r in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        n0 = hi1 ^ c1 ^ k0
        n1 = lo1
        n2 = hi0 ^ c3 ^ k1
        n3 = lo0
        c0, c1, c2, c3 = n0, n1, n2, n3
        if r < 9:
            k0 = (k0 + np.uint64(W0)) & MASK
            k1 = (k1 + np.uint64(W1)) & MASK
    return (c0.astype(np.uint32), c1.astype(np.uint32),
            c2.astype(np.uint32), c3.astype(np.uint32))


def _words(seed: int, stream: int, idx: np.ndarray):
    idx = np.asarray(idx, dtype=np.uint64)
    lo = (idx & MASK).astype(np.uint32)
    hi = (idx >> np.uint64(32)).astype(np.uint32)
    z = np.zeros_like(lo)
    return philox4x32_10(lo, hi, z, z, np.uint32(seed & 0xFFFFFFFF), np.uint32(stream & 0xFFFFFFFF))


_TWO_M53 = 1.0 / 9007199254740992.0  # 2^-53


def _u53(wa, wb):
    """Uniform in (0,1) from two u32 words: 53 random bits + half-ulp offset."""
    hi = wa.astype(np.uint64) >> np.uint64(6)  # 26 bits
    lo = wb.astype(np.uint64) >> np.uint64(5)  # 27 bits
    m = (hi << np.uint64(27)) | lo
    return (m.astype(np.float64) + 0.5) * _TWO_M53


def uniform(seed: int, stream: int, idx) -> np.ndarray:
    w0, w1, _, _ = _words(seed, stream, idx)
    return _u53(w0, w1)


def uniform_range(seed: int, stream: int, idx, lo: float, hi: float) -> np.ndarray:
    return lo + (hi - lo) * uniform(seed, stream, idx)


def normal(seed: int, stream: int, idx) -> np.ndarray:
    w0, w1, w2, w3 = _words(seed, stream, idx)
    u1 = _u53(w0, w1)
    u2 = _u53(w2, w3)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def rademacher(seed: int, stream: int, idx) -> np.ndarray:
    w0, _, _, _ = _words(seed, stream, idx)
    return np.where((w0 & np.uint32(1)) == 0, 1.0, -1.0)


DIR_RADEMACHER, DIR_GAUSSIAN = 1, 2


_F = np.float32


def _u23(w):
    """Uniform in (0, 1) from the top 23 bits of a u32 word with a half-ulp
    offset, exact in binary32 (``csrc/philox.h`` philox_u23)."""
    return ((w >> np.uint32(9)).astype(np.float32) + _F(0.5)) * _F(1.1920928955078125e-07)


def dir_log(u: np.ndarray) -> np.ndarray:
    """ln u for u in (0, 1], from binary32 basic operations in the order of
    ``csrc/philox.h`` dir_log: u = 2^e f, f in [sqrt(1/2), sqrt(2)),
    s = (f - 1)/(f + 1), ln f = 2 atanh s by its series to s^11, and a
    Cody-Waite split of ln 2."""
    u = np.asarray(u, dtype=np.float32)
    b = u.view(np.uint32)
    e = ((b >> np.uint32(23)) & np.uint32(0xFF)).astype(np.int32) - 127
    f = ((b & np.uint32(0x007FFFFF)) | np.uint32(0x3F800000)).view(np.float32)
    big = f > _F(1.41421354)
    f = np.where(big, f * _F(0.5), f)
    e = np.where(big, e + 1, e)
    s = (f - _F(1)) / (f + _F(1))
    t = s * s
    p = np.full_like(s, _F(1) / _F(11))
    for k in (9, 7, 5, 3):
        p = p * t + _F(1) / _F(k)
    lnf = _F(2) * s + (_F(2) * s) * (t * p)
    de = e.astype(np.float32)
    return de * _F(0.693145751953125) + (lnf + de * _F(1.42860677e-06))


# Taylor coefficients, highest order first, each 1/k! rounded to binary32
_SIN = ((1, 39916800), (-1, 362880), (1, 5040), (-1, 120), (1, 6))
_COS = ((1, 479001600), (-1, 3628800), (1, 40320), (-1, 720), (1, 24))


def dir_cos_sin(u: np.ndarray):
    """cos and sin of 2 pi u in binary32, in the order of ``csrc/philox.h``
    dir_cos_sin: exact octant reduction of 8u, Taylor series within the
    octant, the quadrant's rotation."""
    u = np.asarray(u, dtype=np.float32)
    t8 = u * _F(8)
    q = t8.astype(np.int32)
    r = t8 - q.astype(np.float32)
    odd = (q & 1) == 1
    x = np.where(odd, _F(1) - r, r) * _F(0.785398185)
    x2 = x * x

    def horner(coefs):
        sg, d = coefs[0]
        acc = np.full_like(x, _F(1) / _F(d) if sg > 0 else -(_F(1) / _F(d)))
        for sg, d in coefs[1:]:
            acc = acc * x2 + _F(1) / _F(d) if sg > 0 else acc * x2 - _F(1) / _F(d)
        return acc

    ps = horner(_SIN)
    sx = x - (x * x2) * ps
    pc = horner(_COS)
    cx = (_F(1) - _F(0.5) * x2) + (x2 * x2) * pc
    S = np.where(odd, cx, sx)
    C = np.where(odd, sx, cx)
    k = (q >> 1) & 3
    cos = np.select([k == 0, k == 1, k == 2], [C, -S, -C], S)
    sin = np.select([k == 0, k == 1, k == 2], [S, C, -S], -C)
    return cos.astype(np.float32), sin.astype(np.float32)


def dir_pair(wr, wa):
    """One binary32 Box-Muller pair (cosine, sine branch): radius from word
    ``wr``, angle from word ``wa`` (``csrc/philox.h`` dir_pair)."""
    rad = np.sqrt(_F(-2) * dir_log(_u23(wr)))
    c, s = dir_cos_sin(_u23(wa))
    return rad * c, rad * s


def direction(kind: int, seed: int, idx) -> np.ndarray:
    """Direction components ``idx`` of the (seed, STREAM_DIRS) stream, as the
    device draws them (``csrc/philox.h`` philox_direction), as float64."""
    idx = np.asarray(idx, dtype=np.uint64)
    if kind == DIR_RADEMACHER:
        w = np.stack(_words(seed, STREAM_DIRS, idx >> np.uint64(7)))
        b = (idx & np.uint64(127)).astype(np.int64)
        word = np.take_along_axis(w, (b >> 5)[None, ...], axis=0)[0]
        bit = (word >> (b & 31).astype(np.uint32)) & np.uint32(1)
        return np.where(bit == 0, 1.0, -1.0)
    if kind == DIR_GAUSSIAN:
        w0, w1, w2, w3 = _words(seed, STREAM_DIRS, idx >> np.uint64(2))
        g0, g1 = dir_pair(w0, w1)
        g2, g3 = dir_pair(w2, w3)
        k = (idx & np.uint64(3)).astype(np.int64)
        return np.select([k == 0, k == 1, k == 2], [g0, g1, g2], g3).astype(np.float64)
    raise ValueError(f"direction kind {kind}: 1 Rademacher or 2 Gaussian")


def lane_directions(kind: int, seed: int, n: int, lanes) -> np.ndarray:
    """Directions of jet lanes ``lanes`` (lane p = point * S + sample) over
    ``n`` inputs: rows of components p * n .. p * n + n - 1."""
    lanes = np.asarray(lanes, dtype=np.uint64).reshape(-1, 1)
    return direction(kind, seed, lanes * np.uint64(n) + np.arange(n, dtype=np.uint64)[None, :])


def normal_matrix(seed: int, stream: int, rows: int, cols: int, std: float = 1.0) -> np.ndarray:
    idx = np.arange(rows * cols, dtype=np.uint64)
    return (std * normal(seed, stream, idx)).reshape(rows, cols)
