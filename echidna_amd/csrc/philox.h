// Philox4x32-10 (Salmon et al., SC'11) for device and host, plus the draws the
// package uses.  Same definition as echidna_amd/philox.py: key = (seed, stream),
// counter = (block_lo, block_hi, 0, 0).  Scalar draws (philox_draw) take one
// block each: uniforms take 53 bits from words 0,1 with a half-ulp offset;
// Gaussians are Box-Muller in binary64 (cosine branch) from words (0,1) and
// (2,3); Rademacher signs come from bit 0 of word 0.  Direction components
// (philox_direction, the STDE and jet directions) share blocks: Rademacher
// component idx is bit idx mod 128 of block idx / 128 (bit b of word b / 32),
// Gaussian components idx = 4k .. 4k + 3 are the cosine and sine branches of
// block k's two binary32 Box-Muller pairs, (radius word 0, angle word 1) and
// (word 2, word 3) (dir_log, dir_cos_sin: basic operations only).
#pragma once
// Also compiled at run time by hipRTC (tape-specialised jet kernels), which
// supplies the HIP runtime declarations and fixed-width integers itself.
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace ehip {

struct Philox4 { uint32_t v[4]; };

__host__ __device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                 uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return Philox4{{c0, c1, c2, c3}};
}

__host__ __device__ __forceinline__ double philox_u53(uint32_t wa, uint32_t wb) {
  uint64_t m = ((uint64_t)(wa >> 6) << 27) | (uint64_t)(wb >> 5);
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

// Direction sources: read from memory, Rademacher, Gaussian, or the basis
// vectors (sample s of a point is e_s; needs S == n).
enum : uint32_t { RNG_NONE = 0, RNG_RADEMACHER = 1, RNG_GAUSSIAN = 2, RNG_BASIS = 3 };
constexpr uint32_t STREAM_POINTS = 1, STREAM_DIRS = 7, STREAM_CV = 9;

// Draw `idx` of the (seed, stream) sequence: a Rademacher sign or a Gaussian.
__host__ __device__ __forceinline__ double philox_draw(uint32_t kind, uint64_t seed, uint32_t stream, uint64_t idx) {
  Philox4 w = philox4x32_10((uint32_t)idx, (uint32_t)(idx >> 32), 0u, 0u, (uint32_t)seed, stream);
  if (kind == RNG_RADEMACHER) return (w.v[0] & 1u) ? -1.0 : 1.0;
  double u1 = philox_u53(w.v[0], w.v[1]);
  double u2 = philox_u53(w.v[2], w.v[3]);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2);
}

// Direction components come in blocks: 128 Rademacher signs or 4 Gaussians.
__host__ __device__ __forceinline__ uint64_t direction_block(uint32_t kind, uint64_t idx) {
  return kind == RNG_RADEMACHER ? idx >> 7 : idx >> 2;
}

// Uniform in (0, 1) from the top 23 bits of a word with a half-ulp offset:
// (m + 1/2) 2^-23, exact in binary32.
__host__ __device__ __forceinline__ float philox_u23(uint32_t w) {
  return ((float)(w >> 9) + 0.5f) * 1.1920928955078125e-07f;
}

// Box-Muller for the direction draws from IEEE binary32 basic operations
// only (+ - * /, sqrt, exact bit manipulation; no contraction), so that the
// host mirror (echidna_amd/philox.py, numpy float32) gives the same bits.
//
// ln u for u in (0, 1]: u = 2^e f with f in [sqrt(1/2), sqrt(2)); with
// s = (f - 1) / (f + 1), ln f = 2 atanh s = 2 (s + s^3/3 + .. + s^11/11)
// (|s| <= 0.1716: the next term is below 2^-30 of the sum); ln u = e ln2_hi
// + (ln f + e ln2_lo) with ln 2 split Cody-Waite style (ln2_hi has 15
// significant bits, so e ln2_hi is exact for the exponents of u).
__host__ __device__ __forceinline__ float dir_log(float u) {
  uint32_t b;
  __builtin_memcpy(&b, &u, 4);
  int e = (int)((b >> 23) & 0xFFu) - 127;
  b = (b & 0x007FFFFFu) | 0x3F800000u;  // f in [1, 2)
  float f;
  __builtin_memcpy(&f, &b, 4);
  if (f > 1.41421354f) {
    f = f * 0.5f;
    e = e + 1;
  }
  const float s = (f - 1.0f) / (f + 1.0f), t = s * s;
  float p = 1.0f / 11.0f;
  p = p * t + 1.0f / 9.0f;
  p = p * t + 1.0f / 7.0f;
  p = p * t + 1.0f / 5.0f;
  p = p * t + 1.0f / 3.0f;
  const float lnf = 2.0f * s + (2.0f * s) * (t * p);
  const float de = (float)e;
  return de * 0.693145751953125f + (lnf + de * 1.42860677e-06f);
}

// cos and sin of 2 pi u, u in (0, 1) an odd multiple of 2^-24: 8u = q + r
// exactly (octant q, r in [0, 1)); the angle within the octant's quadrant is
// r pi/4 (even q) or (1 - r) pi/4 (odd q, 1 - r exact), whose sine and
// cosine come from their Taylor series to x^11 and x^12 (|x| <= pi/4: the
// next terms are below 2^-36), then the quadrant's rotation.
__host__ __device__ __forceinline__ void dir_cos_sin(float u, float& c, float& sn) {
  const float t8 = u * 8.0f;
  const int q = (int)t8;
  const float r = t8 - (float)q;
  const float x = ((q & 1) ? (1.0f - r) : r) * 0.785398185f, x2 = x * x;
  float ps = 1.0f / 39916800.0f;  // sin: 1/11!, -1/9!, ...
  ps = ps * x2 - 1.0f / 362880.0f;
  ps = ps * x2 + 1.0f / 5040.0f;
  ps = ps * x2 - 1.0f / 120.0f;
  ps = ps * x2 + 1.0f / 6.0f;
  const float sx = x - (x * x2) * ps;
  float pc = 1.0f / 479001600.0f;  // cos: 1/12!, -1/10!, ...
  pc = pc * x2 - 1.0f / 3628800.0f;
  pc = pc * x2 + 1.0f / 40320.0f;
  pc = pc * x2 - 1.0f / 720.0f;
  pc = pc * x2 + 1.0f / 24.0f;
  const float cx = (1.0f - 0.5f * x2) + (x2 * x2) * pc;
  const float S = (q & 1) ? cx : sx, C = (q & 1) ? sx : cx;  // sine, cosine within the quadrant
  switch ((q >> 1) & 3) {
    case 0: c = C, sn = S; break;
    case 1: c = -S, sn = C; break;
    case 2: c = -C, sn = -S; break;
    default: c = S, sn = -C; break;
  }
}

// One Box-Muller pair: radius from word wr, angle from word wa.
__host__ __device__ __forceinline__ void dir_pair(uint32_t wr, uint32_t wa, float& g0, float& g1) {
  const float rad = sqrtf(-2.0f * dir_log(philox_u23(wr)));
  float c, sn;
  dir_cos_sin(philox_u23(wa), c, sn);
  g0 = rad * c;
  g1 = rad * sn;
}

// One block's direction values, kept while consecutive components of a lane
// come from it.  Rademacher: the block's four words; Gaussian: the two pairs'
// cosine and sine branches.
struct DirBlock {  // scalar fields: an array here would be indexed through scratch memory
  uint64_t blk = ~(uint64_t)0;
  uint32_t w0 = 0u, w1 = 0u, w2 = 0u, w3 = 0u;
  float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, g3 = 0.0f;
};

__host__ __device__ __forceinline__ void direction_fill(uint32_t kind, uint64_t seed, uint64_t blk, DirBlock& d) {
  Philox4 r = philox4x32_10((uint32_t)blk, (uint32_t)(blk >> 32), 0u, 0u, (uint32_t)seed, STREAM_DIRS);
  d.blk = blk;
  if (kind == RNG_RADEMACHER) {
    d.w0 = r.v[0], d.w1 = r.v[1], d.w2 = r.v[2], d.w3 = r.v[3];
    return;
  }
  dir_pair(r.v[0], r.v[1], d.g0, d.g1);
  dir_pair(r.v[2], r.v[3], d.g2, d.g3);
}

// Direction component `idx` of the (seed, STREAM_DIRS) stream, through a
// block cache `d` (refilled when idx leaves its block).
__host__ __device__ __forceinline__ double philox_direction(uint32_t kind, uint64_t seed, uint64_t idx, DirBlock& d) {
  const uint64_t blk = direction_block(kind, idx);
  if (blk != d.blk) direction_fill(kind, seed, blk, d);
  if (kind == RNG_RADEMACHER) {
    const uint32_t b = (uint32_t)idx & 127u;
    const uint32_t lo = (b & 32u) ? d.w1 : d.w0, hi = (b & 32u) ? d.w3 : d.w2;
    return ((((b & 64u) ? hi : lo) >> (b & 31u)) & 1u) ? -1.0 : 1.0;
  }
  const uint32_t k = (uint32_t)idx & 3u;
  return (double)((k & 2u) ? ((k & 1u) ? d.g3 : d.g2) : ((k & 1u) ? d.g1 : d.g0));
}

__host__ __device__ __forceinline__ double philox_direction(uint32_t kind, uint64_t seed, uint64_t idx) {
  DirBlock d;
  return philox_direction(kind, seed, idx, d);
}

// Component i of the direction of lane p (p = point * S + sample) with n
// inputs: drawn at index p * n + i, or the basis vector e_(p mod S).  `d`
// caches the block of the previous component read through it.
__host__ __device__ __forceinline__ double lane_direction(uint32_t kind, uint64_t seed, uint32_t S, uint64_t p, uint32_t n,
                                                 uint32_t i, DirBlock& d) {
  if (kind == RNG_BASIS) return (p % S) == i ? 1.0 : 0.0;
  return philox_direction(kind, seed, p * n + i, d);
}

__host__ __device__ __forceinline__ double lane_direction(uint32_t kind, uint64_t seed, uint32_t S, uint64_t p, uint32_t n,
                                                 uint32_t i) {
  DirBlock d;
  return lane_direction(kind, seed, S, p, n, i, d);
}

}  // namespace ehip
