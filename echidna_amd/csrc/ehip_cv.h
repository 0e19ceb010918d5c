// Learned low-rank control variate: device kernels for the network
// x -> L(x) in R^{n x r} (n -> H -> H -> n*r, tanh hidden layers), its
// training loss ||Hv - L L^T v||^2 / n and Adam.  GEMMs run on BF16 MFMA with
// FP32 accumulation; all tensors are FP32 in HBM and are rounded to BF16
// (round to nearest even) as they are staged into LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ehip {

enum : int { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_TANH = 2, EPI_DTANH = 3 };

// C[m][n] = epi(sum_k A(m,k) B(k,n)); A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn].  One of sam/sak and one of sbk/sbn must be 1.
//   EPI_NONE       C = acc            (accumulate: C += acc)
//   EPI_BIAS       C = alpha * (acc + bias[n])
//   EPI_BIAS_TANH  C = tanh(acc + bias[n])
//   EPI_DTANH      C = acc * (1 - aux[m*ldaux + n]^2)
struct Gemm {
  const float* A;
  int64_t sam, sak;
  const float* B;
  int64_t sbk, sbn;
  float* C;
  int64_t ldc;
  int32_t M, N, K;
  int32_t epi;
  const float* bias;
  const float* aux;
  int64_t ldaux;
  int32_t accumulate;
  float alpha;
  // Full tiles store C with the non-temporal hint: an output far larger than
  // L2 that no later step of this kernel reads then does not evict the
  // operand tiles every workgroup re-reads (set for L of a whole batch).
  int32_t stream_out;
  // C holds bfloat16 (round to nearest even after the epilogue; ldc in
  // elements); not with accumulate.
  int32_t out_bf16;
  // Split-K workspace (optional): with room for ksplit * M * N floats the
  // launcher may split a long K over `ksplit` slices of the grid, write each
  // slice's sums there and add them in slice order before the epilogue.
  float* part;
  int64_t part_cap;  // floats
};

hipError_t launch_gemm(const Gemm& g, hipStream_t s);

// out[n] (+)= scale * sum_m X[m*ld + n], n < N.
hipError_t launch_colsum(const float* X, int64_t ld, int64_t M, int64_t N, float scale, float* out, int accumulate,
                         hipStream_t s);

// Per tuple t, with S = diag(+1 x npos, -1 x (r - npos)):
//   u = L^T v, e = hv - L S u, w = L^T e;
//   loss[t] = |e|^2 / n, vhv[t] = v.hv, cv[t] = u^T S u - tr(L S L^T),
//   dL = gscale * (e (S u)^T + v (S w)^T)   (gscale = -2 / (n * T) for the
//   batch mean of |hv - L S L^T v|^2 / n).
struct CvLoss {
  const float* L;  // [T][n*r]
  const float* v;  // [T][n]
  const float* hv; // [T][n]
  float* dL;       // [T][n*r] or null
  float* loss;     // [T] or null
  float* vhv;      // [T] or null
  float* cv;       // [T] or null
  uint32_t n, r, npos;
  uint64_t T;
  float gscale;
};
hipError_t launch_cv_loss(const CvLoss& a, hipStream_t s);

// PyTorch-form Adam: m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// p -= lr_c * m / (sqrt(v) / bc2s + eps), lr_c = lr / (1 - b1^t), bc2s = sqrt(1 - b2^t).
hipError_t launch_adam(float* p, const float* g, float* m, float* v, uint64_t count, float lr_c, float b1, float b2,
                       float bc2s, float eps, hipStream_t s);

// out[i] = draw(kind, seed, stream, offset + i) as float or double.
hipError_t launch_philox_fill(void* out, bool f64, uint64_t count, uint32_t kind, uint64_t seed, uint32_t stream,
                              uint64_t offset, hipStream_t s);

// dst[i] = (float) src[i]; src float or double.
hipError_t launch_to_f32(const void* src, bool f64, uint64_t count, float* dst, hipStream_t s);

}  // namespace ehip
