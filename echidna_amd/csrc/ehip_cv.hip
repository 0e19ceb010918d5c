// Kernels of the learned control variate (ehip_cv.h).
//
// GEMM: a (64 T) x (64 T) output tile per 256-thread workgroup (T = 1 or
// 2), four waves in a 2x2 grid, each wave T x T blocks of 32x32 from
// v_mfma_f32_32x32x16_bf16.  Operand layout of that instruction (checked on
// the device): lane l holds A[m = l%32][k = 8*(l/32) + i] and
// B[k = 8*(l/32) + i][n = l%32] for i < 8, and C[m][n] for n = l%32,
// m = 8*(j/4) + 4*(l/32) + j%4, j < 16.  Tiles of A and B go global ->
// registers -> LDS as BF16 with k contiguous (row pitch 40 BF16 = 80 bytes
// keeps every 8-element operand read 16-byte aligned); the next K step's
// global loads are in flight while the current step's MFMAs run.  The
// 128x128 tile reads half the operand bytes per product of the 64x64 one
// and is used where it still gives the chip enough workgroups.
#include <hip/hip_runtime.h>

#include "ehip_cv.h"
#include "philox.h"

namespace ehip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Two floats as one word of round-to-nearest-even bfloat16s (lo first).
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  const __bf16 a = (__bf16)lo, b = (__bf16)hi;
  uint16_t ua, ub;
  __builtin_memcpy(&ua, &a, 2);
  __builtin_memcpy(&ub, &b, 2);
  return (uint32_t)ua | ((uint32_t)ub << 16);
}

constexpr int BK = 32, PITCH = BK + 8;

// Eight consecutive elements p[0..7] (stride 1) if all lie below `lim`
// elements from the run start, else the in-range ones and zeros.
__device__ __forceinline__ void load8(const float* p, int64_t avail, float* o) {
  if (avail >= 8 && ((reinterpret_cast<uintptr_t>(p) & 15u) == 0)) {
    const float4 a = reinterpret_cast<const float4*>(p)[0];
    const float4 b = reinterpret_cast<const float4*>(p)[1];
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = (j < avail) ? p[j] : 0.0f;
  }
}

// Stage one operand tile (64 T rows of the non-k dimension x 32 k) into
// registers, eight elements per thread per 64 rows.  KC: k is the
// contiguous dimension of this operand.
template <int T, bool KC>
__device__ __forceinline__ void fetch_tile(const float* base, int64_t s_row, int64_t s_k, int rows, int K, int r0,
                                           int k0, float* o) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < T; ++q) {
    float* oq = o + 8 * q;
    if (KC) {
      const int row = (t >> 2) + 64 * q, kq = (t & 3) * 8;
      const int gr = r0 + row, gk = k0 + kq;
      if (gr < rows && gk < K) load8(base + (int64_t)gr * s_row + gk, K - gk, oq);
      else {
#pragma unroll
        for (int j = 0; j < 8; ++j) oq[j] = 0.0f;
      }
    } else {
      const int kr = t >> 3, rq = (t & 7) * 8 + 64 * q;
      const int gk = k0 + kr, gr = r0 + rq;
      if (gk < K && gr < rows) load8(base + (int64_t)gk * s_k + gr, rows - gr, oq);
      else {
#pragma unroll
        for (int j = 0; j < 8; ++j) oq[j] = 0.0f;
      }
    }
  }
}

template <int T, bool KC>
__device__ __forceinline__ void store_tile(__bf16* lds, const float* o) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < T; ++q) {
    const float* oq = o + 8 * q;
    if (KC) {
      const int row = (t >> 2) + 64 * q, kq = (t & 3) * 8;
      bf16x8 v;
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = (__bf16)oq[j];
      *reinterpret_cast<bf16x8*>(lds + row * PITCH + kq) = v;
    } else {
      const int kr = t >> 3, rq = (t & 7) * 8 + 64 * q;
#pragma unroll
      for (int j = 0; j < 8; ++j) lds[(rq + j) * PITCH + kr] = (__bf16)oq[j];
    }
  }
}

// The epilogue's value of output element (m, n) from its sum v.
__device__ __forceinline__ float epi_value(const Gemm& g, int m, int n, float v) {
  switch (g.epi) {
    case EPI_BIAS: return g.alpha * (v + g.bias[n]);
    case EPI_BIAS_TANH: return tanhf(v + g.bias[n]);
    case EPI_DTANH: {
      const float h = g.aux[(int64_t)m * g.ldaux + n];
      return v * (1.0f - h * h);
    }
    default: return v;
  }
}

// Epilogue of one output element (shared by the direct and split-K paths).
__device__ __forceinline__ void gemm_store(const Gemm& g, int m, int n, float v) {
  float* c = g.C + (int64_t)m * g.ldc + n;
  __bf16* ch = reinterpret_cast<__bf16*>(g.C) + (int64_t)m * g.ldc + n;
  switch (g.epi) {
    case EPI_BIAS: v = g.alpha * (v + g.bias[n]); break;
    case EPI_BIAS_TANH: v = tanhf(v + g.bias[n]); break;
    case EPI_DTANH: {
      const float h = g.aux[(int64_t)m * g.ldaux + n];
      v = v * (1.0f - h * h);
      break;
    }
    default: break;
  }
  if (g.out_bf16) *ch = (__bf16)v;
  else *c = g.accumulate ? *c + v : v;
}

// blockIdx.z = K slice [kz * kc, (kz + 1) * kc) (kc = K when not split).
// Split: raw sums to part[kz][m][n]; otherwise the epilogue.
// After the K loop the operand LDS holds each wave's 32x32 output blocks in
// turn (row pitch 33 floats), so full tiles leave as 16-byte row pieces
// instead of single floats; edge tiles and split-K slices store directly.
constexpr int STAGE_PITCH = 33;

template <int T, bool AK, bool BKC>
__global__ __launch_bounds__(256, 3) void k_gemm(Gemm g, int kc) {
  constexpr int BM = 64 * T, BN = 64 * T;
  constexpr int SPB = 32 * T + 1;  // bfloat16 output: a wave's 32 x 32T strip, row pitch SPB floats
  constexpr int OPB = (BM + BN) * PITCH * 2, STB = 4 * 32 * SPB * 4;
  __shared__ __attribute__((aligned(16))) unsigned char smem[OPB > STB ? OPB : STB];
  __bf16* As = reinterpret_cast<__bf16*>(smem);
  __bf16* Bs = As + BM * PITCH;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32 * T, wn = (wave & 1) * 32 * T;
  const int kb = blockIdx.z * kc, ke = min(g.K, kb + kc);
  // A(m,k): rows = M, row stride sam, k stride sak; B(k,n): rows = N.
  float ra[8 * T], rb[8 * T];
  fetch_tile<T, AK>(g.A, g.sam, g.sak, g.M, ke, m0, kb, ra);
  fetch_tile<T, BKC>(g.B, g.sbn, g.sbk, g.N, ke, n0, kb, rb);
  f32x16 acc[T][T];
#pragma unroll
  for (int i = 0; i < T; ++i)
#pragma unroll
    for (int j = 0; j < T; ++j) acc[i][j] = f32x16{};
  for (int k0 = kb; k0 < ke; k0 += BK) {
    __syncthreads();
    store_tile<T, AK>(As, ra);
    store_tile<T, BKC>(Bs, rb);
    __syncthreads();
    if (k0 + BK < ke) {
      fetch_tile<T, AK>(g.A, g.sam, g.sak, g.M, ke, m0, k0 + BK, ra);
      fetch_tile<T, BKC>(g.B, g.sbn, g.sbk, g.N, ke, n0, k0 + BK, rb);
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int ko = kk + 8 * (lane >> 5);
      bf16x8 a[T], b[T];
#pragma unroll
      for (int i = 0; i < T; ++i) {
        a[i] = *reinterpret_cast<const bf16x8*>(As + (wm + 32 * i + (lane & 31)) * PITCH + ko);
        b[i] = *reinterpret_cast<const bf16x8*>(Bs + (wn + 32 * i + (lane & 31)) * PITCH + ko);
      }
#pragma unroll
      for (int i = 0; i < T; ++i)
#pragma unroll
        for (int j = 0; j < T; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
  const bool full = gridDim.z == 1 && m0 + BM <= g.M && n0 + BN <= g.N && (g.ldc & 3) == 0 &&
                    (reinterpret_cast<uintptr_t>(g.C) & 15u) == 0;
  if (full && g.out_bf16) {
    // bfloat16 C: each wave stages the whole 32 x 32T strip of a row block
    // and stores it as rows of 64T bytes (whole 128-byte lines for T = 2:
    // half lines written at different times leave HBM masked writes)
    __syncthreads();  // every wave is done reading the operand tiles
    float* st = reinterpret_cast<float*>(smem) + wave * (32 * SPB);
#pragma unroll
    for (int bi = 0; bi < T; ++bi) {
      const int mb = m0 + wm + 32 * bi;
#pragma unroll
      for (int bj = 0; bj < T; ++bj) {
        const int nl = lane & 31, nb = n0 + wn + 32 * bj;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int ml = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
          st[ml * SPB + 32 * bj + nl] = epi_value(g, mb + ml, nb + nl, acc[bi][bj][j]);
        }
      }
      __builtin_amdgcn_wave_barrier();
      constexpr int LPR = 8 * T, RPP = 64 / LPR;  // lanes per row, rows per pass
#pragma unroll
      for (int q = 0; q < 32 / RPP; ++q) {
        const int row = q * RPP + lane / LPR, c4 = (lane % LPR) * 4;
        const float* sr = st + row * SPB + c4;
        const u32x2 w = {pack_bf16(sr[0], sr[1]), pack_bf16(sr[2], sr[3])};
        u32x2* dh = reinterpret_cast<u32x2*>(reinterpret_cast<__bf16*>(g.C) + (int64_t)(mb + row) * g.ldc + n0 + wn + c4);
        if (g.stream_out) __builtin_nontemporal_store(w, dh);
        else *dh = w;
      }
      __builtin_amdgcn_wave_barrier();
    }
    return;
  }
  if (full) {
    __syncthreads();  // every wave is done reading the operand tiles
    float* stage = reinterpret_cast<float*>(smem) + wave * (32 * STAGE_PITCH);
#pragma unroll
    for (int bi = 0; bi < T; ++bi)
#pragma unroll
      for (int bj = 0; bj < T; ++bj) {
        const int nl = lane & 31, nb = n0 + wn + 32 * bj, mb = m0 + wm + 32 * bi;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int ml = 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
          float v = acc[bi][bj][j];
          switch (g.epi) {
            case EPI_BIAS: v = g.alpha * (v + g.bias[nb + nl]); break;
            case EPI_BIAS_TANH: v = tanhf(v + g.bias[nb + nl]); break;
            case EPI_DTANH: {
              const float h = g.aux[(int64_t)(mb + ml) * g.ldaux + nb + nl];
              v = v * (1.0f - h * h);
              break;
            }
            default: break;
          }
          stage[ml * STAGE_PITCH + nl] = v;
        }
        __builtin_amdgcn_wave_barrier();
        // rows of the block as eight 16-byte pieces each, eight rows a pass
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = q * 8 + (lane >> 3), c4 = (lane & 7) * 4;
          const float* sr = stage + row * STAGE_PITCH + c4;
          float4 v = make_float4(sr[0], sr[1], sr[2], sr[3]);
          float4* dst = reinterpret_cast<float4*>(g.C + (int64_t)(mb + row) * g.ldc + nb + c4);
          if (g.accumulate) {
            const float4 o = *dst;
            v = make_float4(o.x + v.x, o.y + v.y, o.z + v.z, o.w + v.w);
          }
          if (g.stream_out) {
            const f32x4 w = {v.x, v.y, v.z, v.w};
            __builtin_nontemporal_store(w, reinterpret_cast<f32x4*>(dst));
          } else {
            *dst = v;
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
    return;
  }
#pragma unroll
  for (int bj = 0; bj < T; ++bj) {
    const int n = n0 + wn + 32 * bj + (lane & 31);
    if (n >= g.N) continue;
#pragma unroll
    for (int bi = 0; bi < T; ++bi)
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int m = m0 + wm + 32 * bi + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        if (m >= g.M) continue;
        if (gridDim.z > 1) g.part[((int64_t)blockIdx.z * g.M + m) * g.N + n] = acc[bi][bj][j];
        else gemm_store(g, m, n, acc[bi][bj][j]);
      }
  }
}

// Split-K second pass: the slices' sums added in slice order, then the epilogue.
__global__ __launch_bounds__(256) void k_gemm_reduce(Gemm g, int nsplit) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t MN = (int64_t)g.M * g.N;
  if (i >= MN) return;
  float v = g.part[i];
  for (int z = 1; z < nsplit; ++z) v += g.part[z * MN + i];
  gemm_store(g, (int)(i / g.N), (int)(i % g.N), v);
}

template <int T>
const void* gemm_fn(bool ak, bool bk) {
  return ak ? (bk ? (const void*)k_gemm<T, true, true> : (const void*)k_gemm<T, true, false>)
            : (bk ? (const void*)k_gemm<T, false, true> : (const void*)k_gemm<T, false, false>);
}

hipError_t launch_gemm(const Gemm& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  if (g.out_bf16 && g.accumulate) return hipErrorInvalidValue;
  const bool ak = g.sak == 1, bk = g.sbk == 1;
  if ((!ak && g.sam != 1) || (!bk && g.sbn != 1)) return hipErrorInvalidValue;
  // 128x128 tiles when they still make at least two workgroups per CU (256 CUs)
  const int64_t t128 = (int64_t)((g.M + 127) / 128) * ((g.N + 127) / 128);
  const int T = t128 >= 512 ? 2 : 1;
  const int tm = (g.M + 64 * T - 1) / (64 * T), tn = (g.N + 64 * T - 1) / (64 * T);
  // a long K over too few tiles to fill the chip: split it
  int split = 1;
  const int64_t MN = (int64_t)g.M * g.N;
  while (split < 16 && (int64_t)tm * tn * split < 512 && g.K / (split * 2) >= 512 &&
         g.part && g.part_cap >= MN * split * 2)
    split *= 2;
  const int kc = split > 1 ? ((g.K + split - 1) / split + BK - 1) / BK * BK : g.K;
  const dim3 grid(tn, tm, split);
  const void* fn = T == 2 ? gemm_fn<2>(ak, bk) : gemm_fn<1>(ak, bk);
  int kcv = kc;
  void* args[] = {const_cast<Gemm*>(&g), &kcv};
  hipError_t e = hipLaunchKernel(fn, grid, dim3(256), args, 0, s);
  if (e != hipSuccess || split == 1) return e;
  int sv = split;
  void* rargs[] = {const_cast<Gemm*>(&g), &sv};
  return hipLaunchKernel((const void*)k_gemm_reduce, dim3((uint32_t)((MN + 255) / 256)), dim3(256), rargs, 0, s);
}

// Column sums of a row-major [M][N] block (bias gradients): a workgroup owns
// 64 columns; its four waves take every fourth row, each lane keeping eight
// independent partial sums (eight loads in flight, rows coalesced across
// the wave), then the four waves' sums meet in LDS.
__global__ __launch_bounds__(256) void k_colsum(const float* X, int64_t ld, int64_t M, int64_t N, float scale,
                                                float* out, int accumulate) {
  __shared__ float part[4][64];
  const int c = threadIdx.x & 63, rg = threadIdx.x >> 6;
  const int64_t n = (int64_t)blockIdx.x * 64 + c;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (n < N) {
    int64_t m = rg;
    for (; m + 28 < M; m += 32) {
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] += X[(m + 4 * u) * ld + n];
    }
    for (; m < M; m += 4) acc[0] += X[m * ld + n];
  }
  part[rg][c] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  __syncthreads();
  if (rg == 0 && n < N) {
    const float t = (((part[0][c] + part[1][c]) + part[2][c]) + part[3][c]) * scale;
    out[n] = accumulate ? out[n] + t : t;
  }
}

hipError_t launch_colsum(const float* X, int64_t ld, int64_t M, int64_t N, float scale, float* out, int accumulate,
                         hipStream_t s) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_colsum, dim3((uint32_t)((N + 63) / 64)), dim3(256), 0, s, X, ld, M, N, scale, out,
                     accumulate);
  return hipGetLastError();
}

// Block-wide sum of `cnt` per-thread values (cnt <= 34), result in every thread.
template <int CNT>
__device__ __forceinline__ void block_sum(float* v, float* red /* [4][CNT] */) {
#pragma unroll
  for (int j = 0; j < CNT; ++j) {
    float x = v[j];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    v[j] = x;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < CNT; ++j) red[wave * CNT + j] = v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < CNT; ++j) v[j] = red[j] + red[CNT + j] + red[2 * CNT + j] + red[3 * CNT + j];
}

constexpr int RMAX = 32;

__global__ __launch_bounds__(256) void k_cv_loss(CvLoss a) {
  __shared__ float red[4 * (RMAX + 2)];
  const uint64_t t = blockIdx.x;
  if (t >= a.T) return;
  const uint32_t n = a.n, r = a.r;
  const float* L = a.L + t * (uint64_t)n * r;
  const float* v = a.v + t * (uint64_t)n;
  const float* hv = a.hv + t * (uint64_t)n;
  // pass 1: u = L^T v, |L|_F^2, v.hv
  float acc[RMAX + 2];
#pragma unroll
  for (int j = 0; j < RMAX + 2; ++j) acc[j] = 0.0f;
  for (uint32_t i = threadIdx.x; i < n; i += 256) {
    const float vi = v[i];
#pragma unroll
    for (int j = 0; j < RMAX; ++j)
      if (j < (int)r) {
        const float l = L[(uint64_t)i * r + j];
        acc[j] += l * vi;
        acc[RMAX] += (j < (int)a.npos) ? l * l : -(l * l);
      }
    acc[RMAX + 1] += vi * hv[i];
  }
  block_sum<RMAX + 2>(acc, red);
  // su = S u with S = diag(+1 for the first npos columns, -1 after)
  float su[RMAX];
  float uu = 0.0f;
#pragma unroll
  for (int j = 0; j < RMAX; ++j) {
    const float u = (j < (int)r) ? acc[j] : 0.0f;
    su[j] = (j < (int)a.npos) ? u : -u;
    uu += su[j] * u;
  }
  const float lnorm = acc[RMAX], vhv = acc[RMAX + 1];
  // pass 2: e = hv - L S u, |e|^2, w = L^T e
  float acc2[RMAX + 2];
#pragma unroll
  for (int j = 0; j < RMAX + 2; ++j) acc2[j] = 0.0f;
  for (uint32_t i = threadIdx.x; i < n; i += 256) {
    float lu = 0.0f;
#pragma unroll
    for (int j = 0; j < RMAX; ++j)
      if (j < (int)r) lu += L[(uint64_t)i * r + j] * su[j];
    const float e = hv[i] - lu;
    acc2[RMAX] += e * e;
#pragma unroll
    for (int j = 0; j < RMAX; ++j)
      if (j < (int)r) acc2[j] += L[(uint64_t)i * r + j] * e;
  }
  __syncthreads();
  block_sum<RMAX + 2>(acc2, red);
  if (threadIdx.x == 0) {
    if (a.loss) a.loss[t] = acc2[RMAX] / (float)n;
    if (a.vhv) a.vhv[t] = vhv;
    if (a.cv) a.cv[t] = uu - lnorm;
  }
  if (!a.dL) return;
  // pass 3: dL = gscale (e (S u)^T + v (S w)^T)
  float* dL = a.dL + t * (uint64_t)n * r;
  for (uint32_t i = threadIdx.x; i < n; i += 256) {
    float lu = 0.0f;
#pragma unroll
    for (int j = 0; j < RMAX; ++j)
      if (j < (int)r) lu += L[(uint64_t)i * r + j] * su[j];
    const float e = hv[i] - lu, vi = v[i];
#pragma unroll
    for (int j = 0; j < RMAX; ++j)
      if (j < (int)r) {
        const float sw = (j < (int)a.npos) ? acc2[j] : -acc2[j];
        dL[(uint64_t)i * r + j] = a.gscale * (e * su[j] + vi * sw);
      }
  }
}

// The same three passes for a compile-time rank R (a multiple of 4): each
// lane loads its rows of L as float4s (a wave reads contiguous rows) and
// both inner loops of a pass reuse the loaded row.
template <int R>
__global__ __launch_bounds__(256) void k_cv_loss_r(CvLoss a) {
  static_assert(R % 4 == 0 && R <= RMAX, "rank");
  typedef float f4 __attribute__((ext_vector_type(4)));
  __shared__ float red[4 * (R + 2)];
  const uint64_t t = blockIdx.x;
  if (t >= a.T) return;
  const uint32_t n = a.n, npos = a.npos;
  const f4* L = reinterpret_cast<const f4*>(a.L + t * (uint64_t)n * R);
  const float* v = a.v + t * (uint64_t)n;
  const float* hv = a.hv + t * (uint64_t)n;
  auto row = [&](uint32_t i, float* l) {
#pragma unroll
    for (int q = 0; q < R / 4; ++q) {
      const f4 x = L[(uint64_t)i * (R / 4) + q];
      l[4 * q] = x.x; l[4 * q + 1] = x.y; l[4 * q + 2] = x.z; l[4 * q + 3] = x.w;
    }
  };
  // pass 1: u = L^T v, |L|_F^2 with the signs of S, v.hv
  float acc[R + 2];
#pragma unroll
  for (int j = 0; j < R + 2; ++j) acc[j] = 0.0f;
  for (uint32_t i = threadIdx.x; i < n; i += 256) {
    float l[R];
    row(i, l);
    const float vi = v[i];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      acc[j] += l[j] * vi;
      acc[R] += (j < (int)npos) ? l[j] * l[j] : -(l[j] * l[j]);
    }
    acc[R + 1] += vi * hv[i];
  }
  block_sum<R + 2>(acc, red);
  float su[R];
  float uu = 0.0f;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    su[j] = (j < (int)npos) ? acc[j] : -acc[j];
    uu += su[j] * acc[j];
  }
  const float lnorm = acc[R], vhv = acc[R + 1];
  // pass 2: e = hv - L S u, |e|^2, w = L^T e
  float acc2[R + 2];
#pragma unroll
  for (int j = 0; j < R + 2; ++j) acc2[j] = 0.0f;
  for (uint32_t i = threadIdx.x; i < n; i += 256) {
    float l[R];
    row(i, l);
    float lu = 0.0f;
#pragma unroll
    for (int j = 0; j < R; ++j) lu += l[j] * su[j];
    const float e = hv[i] - lu;
    acc2[R] += e * e;
#pragma unroll
    for (int j = 0; j < R; ++j) acc2[j] += l[j] * e;
  }
  __syncthreads();
  block_sum<R + 2>(acc2, red);
  if (threadIdx.x == 0) {
    if (a.loss) a.loss[t] = acc2[R] / (float)n;
    if (a.vhv) a.vhv[t] = vhv;
    if (a.cv) a.cv[t] = uu - lnorm;
  }
  if (!a.dL) return;
  // pass 3: dL = gscale (e (S u)^T + v (S w)^T)
  f4* dL = reinterpret_cast<f4*>(a.dL + t * (uint64_t)n * R);
  float sw[R];
#pragma unroll
  for (int j = 0; j < R; ++j) sw[j] = (j < (int)npos) ? acc2[j] : -acc2[j];
  for (uint32_t i = threadIdx.x; i < n; i += 256) {
    float l[R];
    row(i, l);
    float lu = 0.0f;
#pragma unroll
    for (int j = 0; j < R; ++j) lu += l[j] * su[j];
    const float e = hv[i] - lu, vi = v[i];
#pragma unroll
    for (int q = 0; q < R / 4; ++q) {
      f4 o;
      o.x = a.gscale * (e * su[4 * q] + vi * sw[4 * q]);
      o.y = a.gscale * (e * su[4 * q + 1] + vi * sw[4 * q + 1]);
      o.z = a.gscale * (e * su[4 * q + 2] + vi * sw[4 * q + 2]);
      o.w = a.gscale * (e * su[4 * q + 3] + vi * sw[4 * q + 3]);
      dL[(uint64_t)i * (R / 4) + q] = o;
    }
  }
}

hipError_t launch_cv_loss(const CvLoss& a, hipStream_t s) {
  if (a.T == 0) return hipSuccess;
  if (a.r == 0 || a.r > (uint32_t)RMAX) return hipErrorInvalidValue;
  // float4 rows need 16-byte aligned tuples
  const bool al = ((reinterpret_cast<uintptr_t>(a.L) | reinterpret_cast<uintptr_t>(a.dL)) & 15u) == 0;
  const dim3 grid((uint32_t)a.T), block(256);
  if (al && a.r == 8) hipLaunchKernelGGL(k_cv_loss_r<8>, grid, block, 0, s, a);
  else if (al && a.r == 16) hipLaunchKernelGGL(k_cv_loss_r<16>, grid, block, 0, s, a);
  else if (al && a.r == 32) hipLaunchKernelGGL(k_cv_loss_r<32>, grid, block, 0, s, a);
  else hipLaunchKernelGGL(k_cv_loss, grid, block, 0, s, a);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_adam(float* p, const float* g, float* m, float* v, uint64_t count, float lr_c,
                                              float b1, float b2, float bc2s, float eps) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] = p[i] - lr_c * mi / (sqrtf(vi) / bc2s + eps);
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, uint64_t count, float lr_c, float b1, float b2,
                       float bc2s, float eps, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adam, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0, s, p, g, m, v, count, lr_c, b1,
                     b2, bc2s, eps);
  return hipGetLastError();
}

template <class F>
__global__ __launch_bounds__(256) void k_philox_fill(F* out, uint64_t count, uint32_t kind, uint64_t seed,
                                                     uint32_t stream, uint64_t offset) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (F)philox_draw(kind, seed, stream, offset + i);
}

hipError_t launch_philox_fill(void* out, bool f64, uint64_t count, uint32_t kind, uint64_t seed, uint32_t stream,
                              uint64_t offset, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const dim3 grid((uint32_t)((count + 255) / 256));
  if (f64) hipLaunchKernelGGL(k_philox_fill<double>, grid, dim3(256), 0, s, (double*)out, count, kind, seed, stream, offset);
  else hipLaunchKernelGGL(k_philox_fill<float>, grid, dim3(256), 0, s, (float*)out, count, kind, seed, stream, offset);
  return hipGetLastError();
}

template <class F>
__global__ __launch_bounds__(256) void k_to_f32(const F* src, uint64_t count, float* dst) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = (float)src[i];
}

hipError_t launch_to_f32(const void* src, bool f64, uint64_t count, float* dst, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const dim3 grid((uint32_t)((count + 255) / 256));
  if (f64) hipLaunchKernelGGL(k_to_f32<double>, grid, dim3(256), 0, s, (const double*)src, count, dst);
  else hipLaunchKernelGGL(k_to_f32<float>, grid, dim3(256), 0, s, (const float*)src, count, dst);
  return hipGetLastError();
}

}  // namespace ehip
