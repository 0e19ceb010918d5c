// Wave-uniform interpreters for the compiled register program.
//
// One lane evaluates one point (or one point-direction pair).  Every lane of
// a workgroup runs the same instruction stream, so instruction words are
// fetched with scalar loads into SGPRs and every branch on them is uniform;
// only the data-dependent conventions (zero-adjoint skip, zero-multiplier
// guard, domain checks) are per-lane selects.
//
// Register file: registers [0, R_lds) live in LDS as [reg][component][lane]
// (slot-major, lane-minor: a register access is one conflict-free LDS
// transaction per wave); registers >= R_lds live in HBM as
// [reg][component][point] (slot-major, point-minor: one coalesced
// transaction per wave).  The host sizes R_lds from the LDS budget.
#include <hip/hip_runtime.h>

#include "ehip_kernels.h"
#include "ehip_math.h"
#include "ehip_program.h"
#include "philox.h"

namespace ehip {

template <class F> __device__ __forceinline__ F imm_of(uint32_t lo, uint32_t hi);
template <> __device__ __forceinline__ float imm_of<float>(uint32_t lo, uint32_t) { return __uint_as_float(lo); }
template <> __device__ __forceinline__ double imm_of<double>(uint32_t lo, uint32_t hi) {
  return __hiloint2double((int)hi, (int)lo);
}

template <class F, int NC>
struct Regs {
  F* lds;
  F* gl;
  uint32_t R_lds, wg, tid;
  uint64_t gstride, p;
  __device__ __forceinline__ void rd(uint32_t r, F* o) const {
    if (r < R_lds) {
#pragma unroll
      for (int c = 0; c < NC; ++c) o[c] = lds[((size_t)r * NC + c) * wg + tid];
    } else {
      size_t base = (size_t)(r - R_lds) * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) o[c] = gl[(base + c) * gstride + p];
    }
  }
  __device__ __forceinline__ void wr(uint32_t r, const F* v) const {
    if (r < R_lds) {
#pragma unroll
      for (int c = 0; c < NC; ++c) lds[((size_t)r * NC + c) * wg + tid] = v[c];
    } else {
      size_t base = (size_t)(r - R_lds) * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) gl[(base + c) * gstride + p] = v[c];
    }
  }
};

template <class F, int NC>
__device__ __forceinline__ Regs<F, NC> make_regs(const Launch& L) {
  extern __shared__ unsigned char smem_raw[];
  Regs<F, NC> R;
  R.lds = reinterpret_cast<F*>(smem_raw);
  R.gl = reinterpret_cast<F*>(L.gregs);
  R.R_lds = L.R_lds;
  R.wg = blockDim.x;
  R.tid = threadIdx.x;
  R.gstride = (uint64_t)gridDim.x * blockDim.x;
  R.p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  return R;
}

// ---------------------------------------------------------------------------
// Scalar interpreter: forward values, outputs, reverse adjoint sweep.
// ---------------------------------------------------------------------------
template <class F>
__global__ __launch_bounds__(256) void k_scalar(Launch L) {
  const Regs<F, 1> R = make_regs<F, 1>(L);
  const uint64_t p = R.p;
  const bool act = p < L.B;
  const F* X = reinterpret_cast<const F*>(L.x);
  F* Y = reinterpret_cast<F*>(L.y);
  F* G = reinterpret_cast<F*>(L.g);
  for (uint32_t pc = 0; pc < L.ninstr; ++pc) {
    const uint32_t* w = L.prog + (size_t)pc * NW;
    const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4], w5 = w[5], w6 = w[6], w7 = w[7];
    const uint32_t op = w0 & 0xFFu, fl = w0 >> 8;
    if (op < 64u) {
      F a, b = F(0), r;
      if (fl & F_A_IMM) a = imm_of<F>(w2, w3); else R.rd(w2, &a);
      if (fl & F_HAS_B) { if (fl & F_B_IMM) b = imm_of<F>(w4, w5); else R.rd(w4, &b); }
      r = (op == OP_POWI) ? s_powi(a, (int32_t)w7) : s_eval<F>(op, a, b);
      R.wr(w1, &r);
    } else if (op >= I_REV_BASE) {
      const uint32_t top = op - I_REV_BASE;
      F g;
      R.rd(w1, &g);
      const bool live = g != F(0);
      F a = F(0), b = F(0), r = F(0);
      if (fl & F_A_IMM) a = imm_of<F>(w2, w3); else if (fl & F_RD_A) R.rd(w2, &a);
      if (fl & F_B_IMM) b = imm_of<F>(w4, w5); else if (fl & F_RD_B) R.rd(w4, &b);
      if (fl & F_RD_R) R.rd(w6, &r);
      F da, db;
      if (top == OP_POWI) { da = s_powi_partial(( int32_t)w7, a, r); db = F(0); }
      else s_partials<F>(top, a, b, r, da, db);
      if (fl & F_ACC_A) {
        F cur = F(0);
        if (!(fl & F_FIRST_A)) R.rd(w3, &cur);
        if (live && da != F(0)) cur = cur + da * g;
        R.wr(w3, &cur);
      }
      if (fl & F_ACC_B) {
        F cur = F(0);
        if (!(fl & F_FIRST_B)) R.rd(w5, &cur);
        if (live && db != F(0)) cur = cur + db * g;
        R.wr(w5, &cur);
      }
    } else if (op == I_DOT) {
      const uint32_t* tp = L.terms + (size_t)w3 * TW;
      F acc = F(0);
      for (uint32_t k = 0; k < w2; ++k) {
        const uint32_t t0 = tp[k * TW], t1 = tp[k * TW + 1];
        const F c = imm_of<F>(tp[k * TW + 2], tp[k * TW + 3]);
        const uint32_t kind = t0 & 0xFFu;
        F term = c;
        if (kind != T_IMM) {
          F v;
          R.rd(t1, &v);
          term = (kind == T_PROD) ? v * c : v;
        }
        acc = (k == 0) ? term : acc + term;
      }
      R.wr(w1, &acc);
    } else if (op == I_DOT_REV) {
      const uint32_t* tp = L.terms + (size_t)w3 * TW;
      F g;
      R.rd(w1, &g);
      const bool live = g != F(0);
      for (uint32_t k = 0; k < w2; ++k) {
        const uint32_t t0 = tp[k * TW], t1 = tp[k * TW + 1];
        const F c = imm_of<F>(tp[k * TW + 2], tp[k * TW + 3]);
        const uint32_t kind = t0 & 0xFFu;
        F cur = F(0);
        if (!((t0 >> 8) & 1u)) R.rd(t1, &cur);
        if (kind == T_PROD) {
          if (live && c != F(0)) cur = cur + c * g;
        } else {
          if (live) cur = cur + F(1) * g;
        }
        R.wr(t1, &cur);
      }
    } else if (op == I_LOAD) {
      F v = act ? X[p * L.xstride + w2] : F(0);
      R.wr(w1, &v);
    } else if (op == I_STORE_OUT) {
      F v;
      if (fl & F_A_IMM) v = imm_of<F>(w2, w3); else R.rd(w2, &v);
      if (act && Y) Y[p * L.nout + w1] = v;
    } else if (op == I_SEED) {
      F one = F(1);
      R.wr(w1, &one);
    } else if (op == I_STORE_GRAD) {
      F v;
      if (fl & F_A_IMM) v = imm_of<F>(w2, w3); else R.rd(w2, &v);
      if (act && G) G[p * L.nin + w1] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Dual interpreter: forward-over-reverse Hessian-vector products.
// ---------------------------------------------------------------------------
template <class F>
__device__ __forceinline__ Dual<F> dual_norm(Dual<F> x) {
  // 0 + 1 * x: the adjoint a single Add link hands to each operand.
  return d_add(dc(F(0)), d_mul(dc(F(1)), x));
}

template <class F>
__global__ __launch_bounds__(256) void k_dual(Launch L) {
  const Regs<F, 2> R = make_regs<F, 2>(L);
  const uint64_t p = R.p;
  const bool act = p < L.B;
  const F* X = reinterpret_cast<const F*>(L.x);
  const F* V = reinterpret_cast<const F*>(L.dirs);
  F* G = reinterpret_cast<F*>(L.g);
  F* H = reinterpret_cast<F*>(L.h);
  auto rdd = [&](uint32_t r) { Dual<F> d; F t[2]; R.rd(r, t); d.re = t[0]; d.eps = t[1]; return d; };
  auto wrd = [&](uint32_t r, Dual<F> d) { F t[2] = {d.re, d.eps}; R.wr(r, t); };
  for (uint32_t pc = 0; pc < L.ninstr; ++pc) {
    const uint32_t* w = L.prog + (size_t)pc * NW;
    const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4], w5 = w[5], w6 = w[6], w7 = w[7];
    const uint32_t op = w0 & 0xFFu, fl = w0 >> 8;
    if (op < 64u) {
      Dual<F> a, b = dc(F(0));
      a = (fl & F_A_IMM) ? dc(imm_of<F>(w2, w3)) : rdd(w2);
      if (fl & F_HAS_B) b = (fl & F_B_IMM) ? dc(imm_of<F>(w4, w5)) : rdd(w4);
      Dual<F> r = (op == OP_POWI) ? d_powi(a, (int32_t)w7) : d_eval<F>(op, a, b);
      wrd(w1, r);
    } else if (op >= I_REV_BASE) {
      const uint32_t top = op - I_REV_BASE;
      const Dual<F> g = rdd(w1);
      const bool live = !d_zero(g);
      Dual<F> a = dc(F(0)), b = dc(F(0)), r = dc(F(0));
      if (fl & F_A_IMM) a = dc(imm_of<F>(w2, w3)); else if (fl & F_RD_A) a = rdd(w2);
      if (fl & F_B_IMM) b = dc(imm_of<F>(w4, w5)); else if (fl & F_RD_B) b = rdd(w4);
      if (fl & F_RD_R) r = rdd(w6);
      Dual<F> da, db;
      if (top == OP_POWI) { da = d_powi_partial((int32_t)w7, a, r); db = dc(F(0)); }
      else d_partials<F>(top, a, b, r, da, db);
      if (fl & F_ACC_A) {
        Dual<F> cur = (fl & F_FIRST_A) ? dc(F(0)) : rdd(w3);
        if (live && !d_zero(da)) cur = d_add(cur, d_mul(da, g));
        wrd(w3, cur);
      }
      if (fl & F_ACC_B) {
        Dual<F> cur = (fl & F_FIRST_B) ? dc(F(0)) : rdd(w5);
        if (live && !d_zero(db)) cur = d_add(cur, d_mul(db, g));
        wrd(w5, cur);
      }
    } else if (op == I_DOT) {
      const uint32_t* tp = L.terms + (size_t)w3 * TW;
      Dual<F> acc = dc(F(0));
      for (uint32_t k = 0; k < w2; ++k) {
        const uint32_t t0 = tp[k * TW], t1 = tp[k * TW + 1];
        const Dual<F> c = dc(imm_of<F>(tp[k * TW + 2], tp[k * TW + 3]));
        const uint32_t kind = t0 & 0xFFu;
        Dual<F> term = c;
        if (kind != T_IMM) {
          Dual<F> v = rdd(t1);
          term = (kind == T_PROD) ? (((t0 >> 9) & 1u) ? d_mul(c, v) : d_mul(v, c)) : v;
        }
        acc = (k == 0) ? term : d_add(acc, term);
      }
      wrd(w1, acc);
    } else if (op == I_DOT_REV) {
      const uint32_t* tp = L.terms + (size_t)w3 * TW;
      const Dual<F> g = rdd(w1);
      const bool live = !d_zero(g);
      Dual<F> gd = g;
      uint32_t depth = 0;
      for (uint32_t k = 0; k < w2; ++k) {
        const uint32_t t0 = tp[k * TW], t1 = tp[k * TW + 1];
        const F c = imm_of<F>(tp[k * TW + 2], tp[k * TW + 3]);
        const uint32_t kind = t0 & 0xFFu, tdepth = t0 >> 16;
        while (depth < tdepth) { gd = dual_norm(gd); ++depth; }
        Dual<F> cur = ((t0 >> 8) & 1u) ? dc(F(0)) : rdd(t1);
        if (live) {
          if (kind == T_PROD) {
            if (c != F(0)) cur = d_add(cur, d_mul(dc(c), dual_norm(gd)));
          } else {
            cur = d_add(cur, d_mul(dc(F(1)), gd));
          }
        }
        wrd(t1, cur);
      }
    } else if (op == I_LOAD) {
      Dual<F> v = dc(F(0));
      if (act) {
        const uint64_t row = L.S ? p / L.S : p;
        v.re = X[row * L.xstride + w2];
        v.eps = L.rng ? (F)philox_direction(L.rng, L.seed, p * L.nin + w2) : V[p * L.nin + w2];
      }
      wrd(w1, v);
    } else if (op == I_SEED) {
      wrd(w1, dc(F(1)));
    } else if (op == I_STORE_GRAD) {
      Dual<F> v = (fl & F_A_IMM) ? dc(imm_of<F>(w2, w3)) : rdd(w2);
      if (act) { G[p * L.nin + w1] = v.re; H[p * L.nin + w1] = v.eps; }
    }
    // I_STORE_OUT: the HVP entry points return gradients and products only.
  }
}

// ---------------------------------------------------------------------------
// Jet interpreter: K-coefficient Taylor forward propagation.
// ---------------------------------------------------------------------------
template <class F, int K>
__global__ __launch_bounds__(256) void k_jet(Launch L) {
  const Regs<F, K> R = make_regs<F, K>(L);
  const uint64_t p = R.p;
  const bool act = p < L.B;
  const F* X = reinterpret_cast<const F*>(L.x);
  const F* V = reinterpret_cast<const F*>(L.dirs);
  F* Y = reinterpret_cast<F*>(L.y);
  for (uint32_t pc = 0; pc < L.ninstr; ++pc) {
    const uint32_t* w = L.prog + (size_t)pc * NW;
    const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4], w5 = w[5], w7 = w[7];
    const uint32_t op = w0 & 0xFFu, fl = w0 >> 8;
    if (op < 64u) {
      Jet<F, K> a, b = jc<F, K>(F(0));
      if (fl & F_A_IMM) a = jc<F, K>(imm_of<F>(w2, w3)); else R.rd(w2, a.c);
      if (fl & F_HAS_B) { if (fl & F_B_IMM) b = jc<F, K>(imm_of<F>(w4, w5)); else R.rd(w4, b.c); }
      Jet<F, K> r = (op == OP_POWI) ? j_powi<F, K>(a, (int32_t)w7) : j_eval<F, K>(op, a, b);
      R.wr(w1, r.c);
    } else if (op == I_DOT) {
      const uint32_t* tp = L.terms + (size_t)w3 * TW;
      Jet<F, K> acc = jc<F, K>(F(0));
      for (uint32_t k = 0; k < w2; ++k) {
        const uint32_t t0 = tp[k * TW], t1 = tp[k * TW + 1];
        const Jet<F, K> c = jc<F, K>(imm_of<F>(tp[k * TW + 2], tp[k * TW + 3]));
        const uint32_t kind = t0 & 0xFFu;
        Jet<F, K> term = c;
        if (kind != T_IMM) {
          Jet<F, K> v;
          R.rd(t1, v.c);
          if (kind == T_PROD) {
            if ((t0 >> 9) & 1u) jt::mul<F, K>(c.c, v.c, term.c);
            else jt::mul<F, K>(v.c, c.c, term.c);
          } else {
            term = v;
          }
        }
        if (k == 0) acc = term;
        else {
#pragma unroll
          for (int q = 0; q < K; ++q) acc.c[q] = acc.c[q] + term.c[q];
        }
      }
      R.wr(w1, acc.c);
    } else if (op == I_LOAD) {
      Jet<F, K> v = jc<F, K>(F(0));
      if (act) {
        const uint64_t row = L.S ? p / L.S : p;
        v.c[0] = X[row * L.xstride + w2];
        if (K > 1)
          v.c[K > 1 ? 1 : 0] = L.rng ? (F)philox_direction(L.rng, L.seed, p * L.nin + w2) : V[p * L.nin + w2];
      }
      R.wr(w1, v.c);
    } else if (op == I_STORE_OUT) {
      Jet<F, K> v;
      if (fl & F_A_IMM) v = jc<F, K>(imm_of<F>(w2, w3)); else R.rd(w2, v.c);
      if (act) {
#pragma unroll
        for (int q = 0; q < K; ++q) Y[(p * L.nout + w1) * K + q] = v.c[q];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static hipError_t launch_common(const void* fn, const Launch& L, uint32_t wg, size_t lds, hipStream_t s) {
  uint64_t grid = (L.B + wg - 1) / wg;
  if (grid == 0) grid = 1;
  void* args[] = {const_cast<Launch*>(&L)};
  return hipLaunchKernel(fn, dim3((uint32_t)grid), dim3(wg), args, lds, s);
}

hipError_t launch_scalar(bool f64, const Launch& L, uint32_t wg, size_t lds, hipStream_t s) {
  return f64 ? launch_common((const void*)k_scalar<double>, L, wg, lds, s)
             : launch_common((const void*)k_scalar<float>, L, wg, lds, s);
}

hipError_t launch_dual(bool f64, const Launch& L, uint32_t wg, size_t lds, hipStream_t s) {
  return f64 ? launch_common((const void*)k_dual<double>, L, wg, lds, s)
             : launch_common((const void*)k_dual<float>, L, wg, lds, s);
}

hipError_t launch_jet(bool f64, int K, const Launch& L, uint32_t wg, size_t lds, hipStream_t s) {
#define EHIP_JET(KK)                                                                   \
  case KK:                                                                             \
    return f64 ? launch_common((const void*)k_jet<double, KK>, L, wg, lds, s)          \
               : launch_common((const void*)k_jet<float, KK>, L, wg, lds, s);
  switch (K) {
    EHIP_JET(1)
    EHIP_JET(2)
    EHIP_JET(3)
    EHIP_JET(4)
    EHIP_JET(5)
    default: return hipErrorInvalidValue;
  }
#undef EHIP_JET
}

}  // namespace ehip
