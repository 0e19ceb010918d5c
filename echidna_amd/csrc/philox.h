// Philox4x32-10 (Salmon et al., SC'11) for device and host, plus the draws the
// package uses.  Same definition as echidna_amd/philox.py: key = (seed, stream),
// counter = (index_lo, index_hi, 0, 0), one block per scalar draw; uniforms
// take 53 bits from words 0,1 with a half-ulp offset; Gaussians are Box-Muller
// in binary64 (cosine branch) from words (0,1) and (2,3); Rademacher signs
// come from bit 0 of word 0.
#pragma once
// Also compiled at run time by hipRTC (tape-specialised jet kernels), which
// supplies the HIP runtime declarations and fixed-width integers itself.
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace ehip {

struct Philox4 { uint32_t v[4]; };

__host__ __device__ inline Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
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

__host__ __device__ inline double philox_u53(uint32_t wa, uint32_t wb) {
  uint64_t m = ((uint64_t)(wa >> 6) << 27) | (uint64_t)(wb >> 5);
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

// Direction sources: read from memory, Rademacher, Gaussian, or the basis
// vectors (sample s of a point is e_s; needs S == n).
enum : uint32_t { RNG_NONE = 0, RNG_RADEMACHER = 1, RNG_GAUSSIAN = 2, RNG_BASIS = 3 };
constexpr uint32_t STREAM_POINTS = 1, STREAM_DIRS = 7, STREAM_CV = 9;

// Draw `idx` of the (seed, stream) sequence: a Rademacher sign or a Gaussian.
__host__ __device__ inline double philox_draw(uint32_t kind, uint64_t seed, uint32_t stream, uint64_t idx) {
  Philox4 w = philox4x32_10((uint32_t)idx, (uint32_t)(idx >> 32), 0u, 0u, (uint32_t)seed, stream);
  if (kind == RNG_RADEMACHER) return (w.v[0] & 1u) ? -1.0 : 1.0;
  double u1 = philox_u53(w.v[0], w.v[1]);
  double u2 = philox_u53(w.v[2], w.v[3]);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2);
}

// Direction component `idx` of the (seed, STREAM_DIRS) stream.
__host__ __device__ inline double philox_direction(uint32_t kind, uint64_t seed, uint64_t idx) {
  return philox_draw(kind, seed, STREAM_DIRS, idx);
}

// Component i of the direction of lane p (p = point * S + sample) with n
// inputs: drawn at index p * n + i, or the basis vector e_(p mod S).
__host__ __device__ inline double lane_direction(uint32_t kind, uint64_t seed, uint32_t S, uint64_t p, uint32_t n,
                                                 uint32_t i) {
  if (kind == RNG_BASIS) return (p % S) == i ? 1.0 : 0.0;
  return philox_direction(kind, seed, p * n + i);
}

}  // namespace ehip
