// Order-3 f32 Taylor-jet DENSE blocks as batch GEMMs on the matrix cores
// (ehip_set_jet_mma).
//
// The segmented jet program (segment_jet_program) leaves every register of
// every lane in HBM, one row of `lanes` floats per (register, coefficient).
// A DENSE's members then are, per coefficient q,
//     C_q[j][lane] = sum_p W[p][j] X_q[p][lane]
// (members x terms times terms x lanes), with X_q[p] the row of term p's
// operand register.  A workgroup owns 64 lanes and 128 members: it stages
// 32 terms x 64 lanes of every coefficient in LDS, and each of its four
// waves multiplies its 32 members' weights into them with
// v_mfma_f32_32x32x2_f32 (two terms a step, FP32 accumulation; lane l
// supplies A[j = l % 32][k = l / 32] and B[k = l / 32][n = l % 32], and holds
// the results of rows 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32).
// Immediate terms add their weight to c0; operands that are tape inputs
// (D_INPUT_OPERANDS) carry +0 from order 2 on, and so do the results.
//
// These are the products of the interpreter's exact path (0 + x c per term
// when x is finite) summed in another order and fused, so results agree
// with it to f32 rounding.  A lane whose operand coefficients include an Inf
// or a NaN is recomputed term by term with full Taylor products, the
// reference's arithmetic, so 0 * Inf stays NaN there.
#include <hip/hip_runtime.h>

#include "ehip_kernels.h"
#include "ehip_math.h"
#include "ehip_program.h"

namespace ehip {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t NC = 3;    // coefficients per register (order 3)
constexpr uint32_t MT = 128;  // members per workgroup: four waves of 32
constexpr uint32_t NT = 64;   // lanes per workgroup
constexpr uint32_t KC = 32;   // terms per LDS stage

__device__ __forceinline__ float* row_of(const DenseGemm& G, uint32_t reg, uint32_t q) {
  return G.regs + ((uint64_t)reg * NC + q) * G.lanes;
}

template <uint32_t NZ>
__global__ __launch_bounds__(256) void k_dense_gemm(DenseGemm G) {
  __shared__ float Bs[NZ][KC][NT];
  __shared__ uint32_t bad[2];  // tile lanes (0..31, 32..63) with a non-finite operand coefficient
  const uint32_t tid = threadIdx.x, l = tid & 63u, w = tid >> 6, hi = l >> 5, n = l & 31u;
  const uint64_t lane0 = (uint64_t)blockIdx.x * NT;
  const uint32_t m = G.m, T = G.T;
  const uint32_t* PR = G.pool + G.off_regs;
  const float* PW = reinterpret_cast<const float*>(G.pool + G.off_w);
  const uint32_t jw = blockIdx.y * MT + w * 32u;  // this wave's first member
  const bool live = jw < m;
  if (tid < 2) bad[tid] = 0u;
  f32x16 acc[NZ][2];
#pragma unroll
  for (uint32_t q = 0; q < NZ; ++q) acc[q][0] = acc[q][1] = f32x16{};
  for (uint32_t k0 = 0; k0 < T; k0 += KC) {
    __syncthreads();
    // wave w stages terms w, w + 4, ..: each a row segment of 64 lanes
    for (uint32_t i = w; i < KC; i += 4) {
      const uint32_t p = k0 + i;
      float v[NZ];
#pragma unroll
      for (uint32_t q = 0; q < NZ; ++q) v[q] = 0.f;
      if (p < T) {
        const uint32_t kr = PR[p];
        if ((kr >> P_KIND_SHIFT) != T_IMM) {
          bool fin = true;
#pragma unroll
          for (uint32_t q = 0; q < NZ; ++q) {
            v[q] = row_of(G, kr & P_REG_MASK, q)[lane0 + l];
            fin = fin && isfinite(v[q]);
          }
          if (!fin) atomicOr(&bad[hi], 1u << n);
        }
      }
#pragma unroll
      for (uint32_t q = 0; q < NZ; ++q) Bs[q][i][l] = v[q];
    }
    __syncthreads();
    if (live) {
      const uint32_t j = jw + n;
#pragma unroll 4
      for (uint32_t kk = 0; kk < KC / 2; ++kk) {
        const uint32_t k = 2 * kk + hi, p = k0 + k;
        const float a = (p < T && j < m) ? PW[(uint64_t)p * m + j] : 0.f;
#pragma unroll
        for (uint32_t q = 0; q < NZ; ++q) {
          acc[q][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, Bs[q][k][n], acc[q][0], 0, 0, 0);
          acc[q][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, Bs[q][k][32 + n], acc[q][1], 0, 0, 0);
        }
      }
    }
  }
  __syncthreads();  // bad[] complete
  const uint32_t bad0 = bad[0], bad1 = bad[1];
  if (live) {
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const uint32_t row = jw + 8u * (v >> 2) + 4u * hi + (v & 3);
      if (row >= m) continue;
      float c0[2] = {acc[0][0][v], acc[0][1][v]};
      for (uint32_t i = 0; i < G.nimm; ++i) {
        const float wi = PW[(uint64_t)G.imm[i] * m + row];
        c0[0] = c0[0] + wi;
        c0[1] = c0[1] + wi;
      }
      const uint32_t dst = PR[T + row];
#pragma unroll
      for (uint32_t t = 0; t < 2; ++t) {
        if (((t ? bad1 : bad0) >> n) & 1u) continue;  // recomputed below
        const uint64_t ln = lane0 + t * 32u + n;
        row_of(G, dst, 0)[ln] = c0[t];
#pragma unroll
        for (uint32_t q = 1; q < NC; ++q) row_of(G, dst, q)[ln] = q < NZ ? acc[q][t][v] : 0.f;
      }
    }
  }
  if ((bad0 | bad1) == 0u) return;
  // Lanes with a non-finite operand: every member of this workgroup term by
  // term with full Taylor products, in program order, as the interpreter's
  // exact path computes them.
  const uint32_t jend = min(m, (blockIdx.y + 1) * MT);
  for (uint32_t ln = 0; ln < NT; ++ln) {
    if (!(((ln < 32 ? bad0 : bad1) >> (ln & 31u)) & 1u)) continue;
    for (uint32_t j = blockIdx.y * MT + tid; j < jend; j += blockDim.x) {
      float s[NC] = {0.f, 0.f, 0.f};
      for (uint32_t p = 0; p < T; ++p) {
        const uint32_t kr = PR[p];
        const uint32_t kind = kr >> P_KIND_SHIFT;
        const float c[NC] = {PW[(uint64_t)p * m + j], 0.f, 0.f};
        float x[NC] = {0.f, 0.f, 0.f}, term[NC];
        if (kind != T_IMM)
#pragma unroll
          for (uint32_t q = 0; q < NZ; ++q) x[q] = row_of(G, kr & P_REG_MASK, q)[lane0 + ln];
        if (kind == T_PROD) {
          if (kr & P_CLEFT) jt::mul<float, NC>(c, x, term);
          else jt::mul<float, NC>(x, c, term);
        } else if (kind == T_REG) {
#pragma unroll
          for (uint32_t q = 0; q < NC; ++q) term[q] = x[q];
        } else {
#pragma unroll
          for (uint32_t q = 0; q < NC; ++q) term[q] = c[q];
        }
#pragma unroll
        for (uint32_t q = 0; q < NC; ++q) s[q] = (p == 0) ? term[q] : s[q] + term[q];
      }
      const uint32_t dst = PR[T + j];
#pragma unroll
      for (uint32_t q = 0; q < NC; ++q) row_of(G, dst, q)[lane0 + ln] = s[q];
    }
  }
}

}  // namespace

hipError_t launch_dense_gemm(const DenseGemm& G, hipStream_t s) {
  if (G.m == 0 || G.lanes == 0) return hipSuccess;
  if (G.lanes % NT != 0) return hipErrorInvalidValue;
  const dim3 grid((uint32_t)(G.lanes / NT), (G.m + MT - 1) / MT);
  void* args[] = {const_cast<DenseGemm*>(&G)};
  const void* fn = (G.flags & D_INPUT_OPERANDS) ? (const void*)k_dense_gemm<2> : (const void*)k_dense_gemm<3>;
  return hipLaunchKernel(fn, grid, dim3(256), args, 0, s);
}

}  // namespace ehip
