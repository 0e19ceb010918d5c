// Per-point STDE statistics on the device, in binary64.
//
// The reference reduces the Taylor samples on the host in f32 with
// WelfordAccumulator (src/stde/types.rs:44-113, src/gpu/stde_gpu.rs:84-97,
// 135-200).  Here one wavefront owns one point: each lane runs Welford over
// its strided share of the S samples (skipping non-finite samples, as the
// reference does), then the 64 partial (count, mean, M2) triples merge with
// Chan's pairwise update through cross-lane shuffles.  Finalisation follows
// WelfordAccumulator::finalize: NaN estimate when nothing contributed,
// variance clamped at zero, standard error sqrt(var / n).
//
// Samples: plain STDE uses 2*c2 (= v^T H v); the diagonal control variate
// uses the residual 2*c2 - sum_j d_j v_j^2 with the exact tr(D) added to the
// mean afterwards (laplacian_with_control_gpu's form).  The learned control
// variate subtracts v^T L S L^T v - tr(L S L^T) (S = diag(+-1); S = I is
// |L^T v|^2 - |L|_F^2), which has mean zero for any L when E[v v^T] = I, so
// the estimate stays unbiased; the plain statistics of the same samples can
// be written alongside.
#include <hip/hip_runtime.h>

#include "ehip_kernels.h"
#include "philox.h"

namespace ehip {

struct Wf {
  double n, mean, m2;
};

__device__ __forceinline__ Wf chan_merge(Wf a, Wf b) {
  if (b.n == 0.0) return a;
  if (a.n == 0.0) return b;
  double n = a.n + b.n;
  double delta = b.mean - a.mean;
  Wf r;
  r.n = n;
  r.mean = a.mean + delta * (b.n / n);
  r.m2 = a.m2 + b.m2 + delta * delta * (a.n * b.n / n);
  return r;
}

__device__ __forceinline__ void welford_add(Wf& w, double x) {
  if (!isfinite(x)) return;
  w.n += 1.0;
  double delta = x - w.mean;
  w.mean += delta / w.n;
  w.m2 += delta * (x - w.mean);
}

__device__ __forceinline__ Wf wave_merge(Wf w) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    Wf o;
    o.n = __shfl_xor(w.n, off, 64);
    o.mean = __shfl_xor(w.mean, off, 64);
    o.m2 = __shfl_xor(w.m2, off, 64);
    w = chan_merge(w, o);
  }
  return w;
}

// WelfordAccumulator::finalize (src/stde/types.rs:96-113) plus an offset.
__device__ __forceinline__ void finalize(const Wf& w, double value, double offset, double* out) {
  out[0] = value;
  if (w.n == 0.0) {
    out[1] = __builtin_nan("");
    out[2] = 0.0;
    out[3] = 0.0;
  } else if (w.n > 1.0) {
    double var = w.m2 / (w.n - 1.0);
    if (!(var > 0.0)) var = 0.0;
    out[1] = w.mean + offset;
    out[2] = var;
    out[3] = sqrt(var / w.n);
  } else {
    out[1] = w.mean + offset;
    out[2] = 0.0;
    out[3] = 0.0;
  }
  out[4] = w.n;
}

constexpr int CV_RMAX = 32;

template <class F>
__global__ __launch_bounds__(64) void k_stde_reduce(StdeReduce R) {
  const uint64_t pt = blockIdx.x;
  if (pt >= R.npoints) return;
  const uint32_t lane = threadIdx.x;
  const F* J = reinterpret_cast<const F*>(R.jets);
  const F* D = reinterpret_cast<const F*>(R.dirs);
  const F* DG = reinterpret_cast<const F*>(R.diag);
  const float* LC = R.lcv ? reinterpret_cast<const float*>(R.lcv) + pt * (uint64_t)R.nin * R.rank : nullptr;
  // |L|_F^2 of this point, summed over the wave
  double lnorm = 0.0;
  if (LC) {
    for (uint64_t k = lane; k < (uint64_t)R.nin * R.rank; k += 64) {
      const double l2 = (double)LC[k] * (double)LC[k];
      lnorm += (k % R.rank) < R.npos ? l2 : -l2;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) lnorm += __shfl_xor(lnorm, off, 64);
  }
  Wf w{0.0, 0.0, 0.0}, wp{0.0, 0.0, 0.0};
  // With a learned control variate each sample needs u = L^T v over all n
  // coordinates.  G lanes share one sample (G = 64 / samples per pass, a power
  // of two): lane part g sums coordinates g, g + G, .., and the G partial sums
  // meet by shuffles inside the lane group.  Without one, a lane takes whole
  // samples.
  uint32_t spp = 64, G = 1;  // samples per pass, lanes per sample
  if (LC)
    while (spp / 2 >= R.S && spp > 1) spp /= 2;
  G = 64 / spp;
  const uint32_t part = lane % G;
  for (uint32_t s0 = 0; s0 < R.S; s0 += spp) {
    const uint32_t s = s0 + lane / G;
    const bool has = s < R.S;
    const uint64_t row = pt * R.S + (has ? s : 0);
    double sample = 2.0 * (double)J[row * 3 + 2];
    if (R.sample_scale != 0.0) sample *= R.sample_scale;
    if (LC) {
      double u[CV_RMAX];
#pragma unroll
      for (int j = 0; j < CV_RMAX; ++j) u[j] = 0.0;
      DirBlock db;
      if (has)
        for (uint32_t i = part; i < R.nin; i += G) {
          const double v = R.rng ? lane_direction(R.rng, R.seed, R.S, row, R.nin, i, db) : (double)D[row * R.nin + i];
          const float* li = LC + (uint64_t)i * R.rank;
#pragma unroll
          for (int j = 0; j < CV_RMAX; ++j)
            if (j < (int)R.rank) u[j] += (double)li[j] * v;
        }
      for (uint32_t off = G / 2; off >= 1; off >>= 1)
#pragma unroll
        for (int j = 0; j < CV_RMAX; ++j)
          if (j < (int)R.rank) u[j] += __shfl_xor(u[j], (int)off, 64);
      double q = 0.0;
#pragma unroll
      for (int j = 0; j < CV_RMAX; ++j) q += (j < (int)R.npos) ? u[j] * u[j] : -(u[j] * u[j]);
      if (!has || part != 0) continue;
      welford_add(wp, sample);
      sample -= q - lnorm;
    } else {
      if (!has) continue;
      welford_add(wp, sample);
    }
    if (DG) {
      const F* d = DG + pt * R.diag_stride;
      double cv = 0.0;
      DirBlock db;
      for (uint32_t j = 0; j < R.nin; ++j) {
        double v = R.rng ? lane_direction(R.rng, R.seed, R.S, row, R.nin, j, db) : (double)D[row * R.nin + j];
        cv += (double)d[j] * v * v;
      }
      sample -= cv;
    }
    welford_add(w, sample);
  }
  w = wave_merge(w);
  if (R.out_plain) wp = wave_merge(wp);
  if (lane == 0) {
    double tr = 0.0;
    if (DG) {
      const F* d = DG + pt * R.diag_stride;
      for (uint32_t j = 0; j < R.nin; ++j) tr += (double)d[j];
    }
    const double value = (double)J[(pt * R.S) * 3 + 0];
    finalize(w, value, tr, R.out + pt * 5);
    if (R.out_plain) finalize(wp, value, 0.0, R.out_plain + pt * 5);
  }
}

// The learned control variate with the directions in memory: per point,
// U = V L (S samples x n coordinates times n x r) on the FP64 matrix cores,
// 16 samples x 16 columns per v_mfma_f64_16x16x4f64 tile.  Lane l supplies
// A[i = l % 16][k = l / 16] and B[k = l / 16][j = l % 16] and holds
// D[l / 16 + 4 v][l % 16], v = 0..3, so the 16 lanes of a lane group hold
// one row's columns and q_s = sum_j +-U[s][j]^2 is a shuffle sum inside the
// group.  The directions are coordinate-major (sample p of the pass at
// c * npoints * S + p, as the jets wrote them), so the 16 lanes of a group
// read 16 consecutive samples of one coordinate; step t of a 16-coordinate
// slab gives lane group g coordinate 4 t + g, so the wave reads four
// consecutive rows of L.  Four slabs' operands are loaded before their
// MFMAs run.  L is bfloat16 (half the bytes of the pass's largest read
// after the directions); the products of its values and the FP32 (or FP64)
// directions are formed and summed in binary64.  One wave per point, four
// points per workgroup; |L|_F^2 is summed from the same loads, so the
// control variate's mean is exactly that of the rounded L.
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <class F, uint32_t NRB>
__global__ __launch_bounds__(256) void k_stde_reduce_lcv(StdeReduce R) {
  constexpr uint32_t U = 4;  // slabs of 16 coordinates per batch of loads
  const uint32_t l = threadIdx.x & 63u, g = l >> 4, i16 = l & 15u;
  const uint64_t pt = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6);
  if (pt >= R.npoints) return;  // whole waves: nothing below synchronises the workgroup
  const F* J = reinterpret_cast<const F*>(R.jets);
  const F* D = reinterpret_cast<const F*>(R.dirs_cmajor);
  const uint64_t NL = R.npoints * (uint64_t)R.S;  // coordinate stride of the directions
  const __bf16* LC = reinterpret_cast<const __bf16*>(R.lcv) + pt * (uint64_t)R.nin * R.rank;
  const uint32_t n = R.nin, r = R.rank;
  double lnorm = 0.0;
  Wf w{0.0, 0.0, 0.0}, wp{0.0, 0.0, 0.0};
  // Lanes past the rank or the sample count read a valid element (column 0,
  // sample s0) and zero it: no divergent loads in the loop.
  bool jok[NRB];
  const __bf16* pb0[NRB];
#pragma unroll
  for (uint32_t rb = 0; rb < NRB; ++rb) {
    const uint32_t j = rb * 16u + i16;
    jok[rb] = j < r;
    pb0[rb] = LC + (uint64_t)g * r + (jok[rb] ? j : 0u);
  }
  // one slab (16 coordinates, four MFMA steps) from coordinate k0 + g on
  auto slab = [&](f64x4* acc, const F*& qa, const __bf16** qb, bool first, bool sok, uint32_t k0) {
    F a[4];
    float b[NRB][4];
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) {
      const bool cok = k0 + 4u * t + g < n;
      a[t] = cok ? *qa : F(0);
      qa += 4 * NL;
#pragma unroll
      for (uint32_t rb = 0; rb < NRB; ++rb) {
        b[rb][t] = cok ? (float)*qb[rb] : 0.0f;
        qb[rb] += 4 * r;
      }
    }
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t)
#pragma unroll
      for (uint32_t rb = 0; rb < NRB; ++rb) {
        const double bd = jok[rb] ? (double)b[rb][t] : 0.0;
        if (first) lnorm += rb * 16u + i16 < R.npos ? bd * bd : -(bd * bd);
        acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(sok ? (double)a[t] : 0.0, bd, acc[rb], 0, 0, 0);
      }
  };
  for (uint32_t s0 = 0; s0 < R.S; s0 += 16u) {
    const uint32_t sa = s0 + i16;  // the A row this lane supplies
    const bool sok = sa < R.S, first = s0 == 0;
    const F* qa = D + (uint64_t)g * NL + pt * R.S + (sok ? sa : s0);
    const __bf16* qb[NRB];
#pragma unroll
    for (uint32_t rb = 0; rb < NRB; ++rb) qb[rb] = pb0[rb];
    f64x4 acc[NRB];
#pragma unroll
    for (uint32_t rb = 0; rb < NRB; ++rb) acc[rb] = f64x4{0.0, 0.0, 0.0, 0.0};
    uint32_t k0 = 0;
    // whole batches of U slabs: every coordinate in range, all loads first
    for (; k0 + 16u * U <= n; k0 += 16u * U) {
      F a[U * 4];
      float b[NRB][U * 4];
#pragma unroll
      for (uint32_t q = 0; q < U * 4; ++q) {
        a[q] = *qa;
        qa += 4 * NL;
#pragma unroll
        for (uint32_t rb = 0; rb < NRB; ++rb) {
          b[rb][q] = (float)*qb[rb];
          qb[rb] += 4 * r;
        }
      }
#pragma unroll
      for (uint32_t q = 0; q < U * 4; ++q)
#pragma unroll
        for (uint32_t rb = 0; rb < NRB; ++rb) {
          const double bd = jok[rb] ? (double)b[rb][q] : 0.0;
          if (first) lnorm += rb * 16u + i16 < R.npos ? bd * bd : -(bd * bd);
          acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(sok ? (double)a[q] : 0.0, bd, acc[rb], 0, 0, 0);
        }
    }
    for (; k0 < n; k0 += 16u) slab(acc, qa, qb, first, sok, k0);
    // q for samples s0 + g + 4 v: signed squares summed over the 16 columns of the group
    double qv[4];
#pragma unroll
    for (uint32_t v = 0; v < 4; ++v) {
      double e = 0.0;
#pragma unroll
      for (uint32_t rb = 0; rb < NRB; ++rb) {
        const uint32_t j = rb * 16u + i16;
        const double uu = acc[rb][v];
        e += (j < R.npos) ? uu * uu : -(uu * uu);
      }
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) e += __shfl_xor(e, off, 64);
      qv[v] = e;
    }
    if (s0 == 0)
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) lnorm += __shfl_xor(lnorm, off, 64);
    // lane (g, v) of the group's first four lanes takes sample s0 + g + 4 v
    const uint32_t s = s0 + g + 4u * i16;
    if (i16 < 4u && s < R.S) {
      const double q = i16 == 0 ? qv[0] : (i16 == 1 ? qv[1] : (i16 == 2 ? qv[2] : qv[3]));
      const uint64_t row = pt * R.S + s;
      double sample = 2.0 * (double)J[row * 3 + 2];
      if (R.sample_scale != 0.0) sample *= R.sample_scale;
      welford_add(wp, sample);
      welford_add(w, sample - (q - lnorm));
    }
  }
  w = wave_merge(w);
  if (R.out_plain) wp = wave_merge(wp);
  if (l == 0) {
    const double value = (double)J[(pt * R.S) * 3 + 0];
    finalize(w, value, 0.0, R.out + pt * 5);
    if (R.out_plain) finalize(wp, value, 0.0, R.out_plain + pt * 5);
  }
}

hipError_t launch_stde_reduce(bool f64, const StdeReduce& R, hipStream_t s) {
  if (R.npoints == 0) return hipSuccess;
  void* args[] = {const_cast<StdeReduce*>(&R)};
  if (R.dirs_cmajor) {
    // the jets' own directions: only the learned control variate reads them
    if (!R.lcv || R.diag || R.rng != RNG_NONE || R.rank > 32u) return hipErrorInvalidValue;
    const bool one = R.rank <= 16u;
    const void* fn = f64 ? (one ? (const void*)k_stde_reduce_lcv<double, 1> : (const void*)k_stde_reduce_lcv<double, 2>)
                         : (one ? (const void*)k_stde_reduce_lcv<float, 1> : (const void*)k_stde_reduce_lcv<float, 2>);
    return hipLaunchKernel(fn, dim3((uint32_t)((R.npoints + 3) / 4)), dim3(256), args, 0, s);
  }
  const void* fn = f64 ? (const void*)k_stde_reduce<double> : (const void*)k_stde_reduce<float>;
  return hipLaunchKernel(fn, dim3((uint32_t)R.npoints), dim3(64), args, 0, s);
}

}  // namespace ehip
