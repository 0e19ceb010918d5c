// Philox4x32-10 (Salmon et al., SC'11) for device and host, plus the draws the
// package uses.  Same definition as echidna_amd/philox.py: key = (seed, stream),
// counter = (index_lo, index_hi, 0, 0), one block per scalar draw; uniforms
// take 53 bits from words 0,1 with a half-ulp offset; Gaussians are Box-Muller
// in binary64 (cosine branch) from words (0,1) and (2,3); Rademacher signs
// come from bit 0 of word 0.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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

enum : uint32_t { RNG_NONE = 0, RNG_RADEMACHER = 1, RNG_GAUSSIAN = 2 };
constexpr uint32_t STREAM_DIRS = 7;

// Direction component `idx` of the (seed, STREAM_DIRS) stream.
__host__ __device__ inline double philox_direction(uint32_t kind, uint64_t seed, uint64_t idx) {
  Philox4 w = philox4x32_10((uint32_t)idx, (uint32_t)(idx >> 32), 0u, 0u, (uint32_t)seed, STREAM_DIRS);
  if (kind == RNG_RADEMACHER) return (w.v[0] & 1u) ? -1.0 : 1.0;
  double u1 = philox_u53(w.v[0], w.v[1]);
  double u2 = philox_u53(w.v[2], w.v[3]);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.141592653589793 * u2);
}

}  // namespace ehip
