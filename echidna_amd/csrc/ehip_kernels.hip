// Wave-uniform interpreters for the compiled register program.
//
// One lane evaluates one point (or one point-direction pair).  Every lane of
// a workgroup runs the same instruction stream, so instruction words are
// fetched with scalar loads into SGPRs (the next instruction is requested
// while the current one executes) and every branch on them is uniform; only
// the data-dependent conventions (zero-adjoint skip, zero-multiplier guard,
// domain checks) are per-lane selects.
//
// Register file: registers [0, R_lds) live in LDS as [reg][component][lane]
// (slot-major, lane-minor: one conflict-free LDS transaction per wave);
// registers >= R_lds live in HBM as [reg][component][point] (slot-major,
// point-minor: one coalesced transaction per wave).  The LDS-only variant
// (every register fits) has no HBM path at all; the split variant keeps the
// two paths in distinct address spaces so neither degrades to flat accesses.
//
// Fused dot products read their term table (meta | reg | imm arrays) in
// blocks of eight: eight register indices and eight immediates arrive in two
// scalar loads, eight LDS reads issue back to back, then the products are
// added in term order.
#include <hip/hip_runtime.h>

#include "ehip_kernels.h"
#include "ehip_math.h"
#include "ehip_program.h"
#include "philox.h"

namespace ehip {

extern __shared__ __attribute__((aligned(16))) unsigned char g_smem[];

#define AS3 __attribute__((address_space(3)))
#define AS1 __attribute__((address_space(1)))

template <class F> __device__ __forceinline__ F imm_of(uint32_t lo, uint32_t hi);
template <> __device__ __forceinline__ float imm_of<float>(uint32_t lo, uint32_t) { return __uint_as_float(lo); }
template <> __device__ __forceinline__ double imm_of<double>(uint32_t lo, uint32_t hi) {
  return __hiloint2double((int)hi, (int)lo);
}
template <class F> __device__ __forceinline__ F imm_at(const uint32_t* immw, uint32_t k);
template <> __device__ __forceinline__ float imm_at<float>(const uint32_t* immw, uint32_t k) {
  return __uint_as_float(immw[k]);
}
template <> __device__ __forceinline__ double imm_at<double>(const uint32_t* immw, uint32_t k) {
  return __hiloint2double((int)immw[2 * k + 1], (int)immw[2 * k]);
}

template <class F, int NC, bool LDSONLY>
struct Regs {
  AS3 F* lds;
  AS1 F* gl;
  uint32_t R_lds, wg, tid;
  uint64_t gstride, p;
  __device__ __forceinline__ void rd(uint32_t r, F* o) const {
    if (LDSONLY || r < R_lds) {
#pragma unroll
      for (int c = 0; c < NC; ++c) o[c] = lds[(r * NC + c) * wg + tid];
    } else {
      const uint64_t base = (uint64_t)(r - R_lds) * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) o[c] = gl[(base + c) * gstride + p];
    }
  }
  __device__ __forceinline__ void wr(uint32_t r, const F* v) const {
    if (LDSONLY || r < R_lds) {
#pragma unroll
      for (int c = 0; c < NC; ++c) lds[(r * NC + c) * wg + tid] = v[c];
    } else {
      const uint64_t base = (uint64_t)(r - R_lds) * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) gl[(base + c) * gstride + p] = v[c];
    }
  }
};

template <class F, int NC, bool LDSONLY>
__device__ __forceinline__ Regs<F, NC, LDSONLY> make_regs(const Launch& L) {
  Regs<F, NC, LDSONLY> R;
  R.lds = (AS3 F*)(g_smem);
  R.gl = (AS1 F*)(L.gregs);
  R.R_lds = L.R_lds;
  R.wg = blockDim.x;
  R.tid = threadIdx.x;
  R.gstride = (uint64_t)gridDim.x * blockDim.x;
  R.p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  return R;
}

// Side-table views for instruction-relative access.
template <class F>
struct Terms {
  const uint32_t* meta;
  const uint32_t* reg;
  const uint32_t* immw;
  __device__ __forceinline__ Terms(const Launch& L, uint32_t off) {
    constexpr uint32_t IW = sizeof(F) / 4;
    meta = L.terms + off;
    reg = L.terms + L.nterms + off;
    immw = L.terms + 2 * L.nterms + (size_t)off * IW;
  }
  __device__ __forceinline__ F imm(uint32_t k) const { return imm_at<F>(immw, k); }
  __device__ __forceinline__ uint32_t kind(uint32_t k) const { return meta[k] & 3u; }
  __device__ __forceinline__ uint32_t u32(uint32_t k) const { return immw[(size_t)k * (sizeof(F) / 4)]; }
  __device__ __forceinline__ uint32_t run(uint32_t k) const { return (meta[k] >> M_RUN_SHIFT) & 0xFFu; }
};

struct Ins {
  uint32_t w[NW];
};

__device__ __forceinline__ Ins fetch(const uint32_t* prog, uint32_t pc) {
  Ins i;
  const uint32_t* p = prog + (size_t)pc * NW;
#pragma unroll
  for (int k = 0; k < (int)NW; ++k) i.w[k] = p[k];
  return i;
}

// Batched input load: eight loads in flight, then eight register writes.
template <class F, int NC, bool LDSONLY, class MakeVal>
__device__ __forceinline__ void load_inputs(const Regs<F, NC, LDSONLY>& R, const Launch& L, uint32_t n, uint32_t off,
                                            MakeVal make) {
  Terms<F> T(L, off);
  uint32_t k = 0;
  for (; k + 8 <= n; k += 8) {
    F v[8][NC];
#pragma unroll
    for (int j = 0; j < 8; ++j) make(T.u32(k + j), v[j]);
#pragma unroll
    for (int j = 0; j < 8; ++j) R.wr(T.reg[k + j], v[j]);
  }
  for (; k < n; ++k) {
    F v[NC];
    make(T.u32(k), v);
    R.wr(T.reg[k], v);
  }
}

// ---------------------------------------------------------------------------
// Scalar interpreter: forward values, outputs, reverse adjoint sweep.
// ---------------------------------------------------------------------------
template <class F, bool LDSONLY>
__global__ __launch_bounds__(256) void k_scalar(Launch L) {
  const Regs<F, 1, LDSONLY> R = make_regs<F, 1, LDSONLY>(L);
  const uint64_t p = R.p;
  const bool act = p < L.B;
  const F* X = reinterpret_cast<const F*>(L.x);
  F* Y = reinterpret_cast<F*>(L.y);
  F* G = reinterpret_cast<F*>(L.g);
  Ins nxt = fetch(L.prog, 0);
  for (uint32_t pc = 0; pc < L.ninstr; ++pc) {
    const Ins cur = nxt;
    if (pc + 1 < L.ninstr) nxt = fetch(L.prog, pc + 1);
    const uint32_t w0 = cur.w[0], w1 = cur.w[1], w2 = cur.w[2], w3 = cur.w[3], w4 = cur.w[4], w5 = cur.w[5],
                   w6 = cur.w[6], w7 = cur.w[7];
    const uint32_t op = w0 & 0xFFu, fl = w0 >> 8;
    if (op < 64u) {
      F a, b = F(0), r;
      if (fl & F_A_IMM) a = imm_of<F>(w2, w3); else R.rd(w2, &a);
      if (fl & F_HAS_B) { if (fl & F_B_IMM) b = imm_of<F>(w4, w5); else R.rd(w4, &b); }
      r = (op == OP_POWI) ? s_powi(a, (int32_t)w7) : s_eval<F>(op, a, b);
      R.wr(w1, &r);
    } else if (op == I_DOT) {
      const Terms<F> T(L, w3);
      F acc = F(0);
      bool have = (fl & F_DOT_CONT) != 0;
      if (have) R.rd(w1, &acc);
      uint32_t k = 0;
      while (k < w2) {
        if (T.run(k) >= 8) {
          uint32_t rr[8];
          F cc[8], v[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) { rr[j] = T.reg[k + j]; cc[j] = T.imm(k + j); }
#pragma unroll
          for (int j = 0; j < 8; ++j) R.rd(rr[j], &v[j]);
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const F t = v[j] * cc[j];
            acc = (j == 0 && !have) ? t : acc + t;
          }
          have = true;
          k += 8;
        } else {
          const uint32_t kind = T.kind(k);
          const F c = T.imm(k);
          F term = c;
          if (kind != T_IMM) {
            F v;
            R.rd(T.reg[k], &v);
            term = (kind == T_PROD) ? v * c : v;
          }
          acc = have ? acc + term : term;
          have = true;
          ++k;
        }
      }
      R.wr(w1, &acc);
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
      if (top == OP_POWI) { da = s_powi_partial((int32_t)w7, a, r); db = F(0); }
      else s_partials<F>(top, a, b, r, da, db);
      if (fl & F_ACC_A) {
        F c = F(0);
        if (!(fl & F_FIRST_A)) R.rd(w3, &c);
        if (live && da != F(0)) c = c + da * g;
        R.wr(w3, &c);
      }
      if (fl & F_ACC_B) {
        F c = F(0);
        if (!(fl & F_FIRST_B)) R.rd(w5, &c);
        if (live && db != F(0)) c = c + db * g;
        R.wr(w5, &c);
      }
    } else if (op == I_DOT_REV) {
      const Terms<F> T(L, w3);
      F g;
      R.rd(w1, &g);
      const bool live = g != F(0);
      uint32_t k = 0;
      while (k < w2) {
        if (T.run(k) >= 8) {
          // eight product terms with distinct adjoint registers; a first
          // contribution reads its register anyway and discards the value
          uint32_t mm[8], rr[8];
          F cc[8], cur[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) { mm[j] = T.meta[k + j]; rr[j] = T.reg[k + j]; cc[j] = T.imm(k + j); }
#pragma unroll
          for (int j = 0; j < 8; ++j) R.rd(rr[j], &cur[j]);
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            F c0 = (mm[j] & M_FIRST) ? F(0) : cur[j];
            if (live && cc[j] != F(0)) c0 = c0 + cc[j] * g;
            R.wr(rr[j], &c0);
          }
          k += 8;
        } else {
          const uint32_t m = T.meta[k];
          const F c = T.imm(k);
          F cur = F(0);
          if (!(m & M_FIRST)) R.rd(T.reg[k], &cur);
          if ((m & 3u) == T_PROD) {
            if (live && c != F(0)) cur = cur + c * g;
          } else if (live) {
            cur = cur + F(1) * g;
          }
          R.wr(T.reg[k], &cur);
          ++k;
        }
      }
    } else if (op == I_LOADN) {
      load_inputs<F, 1, LDSONLY>(R, L, w2, w3, [&](uint32_t i, F* v) { v[0] = act ? X[p * L.xstride + i] : F(0); });
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

template <class F, bool LDSONLY>
__global__ __launch_bounds__(256) void k_dual(Launch L) {
  const Regs<F, 2, LDSONLY> R = make_regs<F, 2, LDSONLY>(L);
  const uint64_t p = R.p;
  const bool act = p < L.B;
  const F* X = reinterpret_cast<const F*>(L.x);
  const F* V = reinterpret_cast<const F*>(L.dirs);
  F* G = reinterpret_cast<F*>(L.g);
  F* H = reinterpret_cast<F*>(L.h);
  auto rdd = [&](uint32_t r) { Dual<F> d; F t[2]; R.rd(r, t); d.re = t[0]; d.eps = t[1]; return d; };
  auto wrd = [&](uint32_t r, Dual<F> d) { F t[2] = {d.re, d.eps}; R.wr(r, t); };
  Ins nxt = fetch(L.prog, 0);
  for (uint32_t pc = 0; pc < L.ninstr; ++pc) {
    const Ins cur = nxt;
    if (pc + 1 < L.ninstr) nxt = fetch(L.prog, pc + 1);
    const uint32_t w0 = cur.w[0], w1 = cur.w[1], w2 = cur.w[2], w3 = cur.w[3], w4 = cur.w[4], w5 = cur.w[5],
                   w6 = cur.w[6], w7 = cur.w[7];
    const uint32_t op = w0 & 0xFFu, fl = w0 >> 8;
    if (op < 64u) {
      Dual<F> a, b = dc(F(0));
      a = (fl & F_A_IMM) ? dc(imm_of<F>(w2, w3)) : rdd(w2);
      if (fl & F_HAS_B) b = (fl & F_B_IMM) ? dc(imm_of<F>(w4, w5)) : rdd(w4);
      Dual<F> r = (op == OP_POWI) ? d_powi(a, (int32_t)w7) : d_eval<F>(op, a, b);
      wrd(w1, r);
    } else if (op == I_DOT) {
      const Terms<F> T(L, w3);
      Dual<F> acc = dc(F(0));
      bool have = (fl & F_DOT_CONT) != 0;
      if (have) acc = rdd(w1);
      for (uint32_t k = 0; k < w2; ++k) {
        const uint32_t m = T.meta[k], kind = m & 3u;
        const Dual<F> c = dc(T.imm(k));
        Dual<F> term = c;
        if (kind != T_IMM) {
          Dual<F> v = rdd(T.reg[k]);
          term = (kind == T_PROD) ? ((m & M_CONST_LEFT) ? d_mul(c, v) : d_mul(v, c)) : v;
        }
        acc = have ? d_add(acc, term) : term;
        have = true;
      }
      wrd(w1, acc);
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
        Dual<F> c = (fl & F_FIRST_A) ? dc(F(0)) : rdd(w3);
        if (live && !d_zero(da)) c = d_add(c, d_mul(da, g));
        wrd(w3, c);
      }
      if (fl & F_ACC_B) {
        Dual<F> c = (fl & F_FIRST_B) ? dc(F(0)) : rdd(w5);
        if (live && !d_zero(db)) c = d_add(c, d_mul(db, g));
        wrd(w5, c);
      }
    } else if (op == I_DOT_REV) {
      const Terms<F> T(L, w3);
      const Dual<F> g = rdd(w1);
      const bool live = !d_zero(g);
      Dual<F> gd = g;
      uint32_t depth = 0;
      for (uint32_t k = 0; k < w2; ++k) {
        const uint32_t m = T.meta[k], kind = m & 3u, tdepth = m >> M_DEPTH_SHIFT;
        const F c = T.imm(k);
        while (depth < tdepth) { gd = dual_norm(gd); ++depth; }
        Dual<F> cur = (m & M_FIRST) ? dc(F(0)) : rdd(T.reg[k]);
        if (live) {
          if (kind == T_PROD) {
            if (c != F(0)) cur = d_add(cur, d_mul(dc(c), dual_norm(gd)));
          } else {
            cur = d_add(cur, d_mul(dc(F(1)), gd));
          }
        }
        wrd(T.reg[k], cur);
      }
    } else if (op == I_LOADN) {
      load_inputs<F, 2, LDSONLY>(R, L, w2, w3, [&](uint32_t i, F* v) {
        v[0] = F(0);
        v[1] = F(0);
        if (act) {
          const uint64_t row = L.S ? p / L.S : p;
          v[0] = X[row * L.xstride + i];
          v[1] = L.rng ? (F)philox_direction(L.rng, L.seed, p * L.nin + i) : V[p * L.nin + i];
        }
      });
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
template <class F, int K, bool LDSONLY>
__global__ __launch_bounds__(256) void k_jet(Launch L) {
  const Regs<F, K, LDSONLY> R = make_regs<F, K, LDSONLY>(L);
  const uint64_t p = R.p;
  const bool act = p < L.B;
  const F* X = reinterpret_cast<const F*>(L.x);
  const F* V = reinterpret_cast<const F*>(L.dirs);
  F* Y = reinterpret_cast<F*>(L.y);
  Ins nxt = fetch(L.prog, 0);
  for (uint32_t pc = 0; pc < L.ninstr; ++pc) {
    const Ins cur = nxt;
    if (pc + 1 < L.ninstr) nxt = fetch(L.prog, pc + 1);
    const uint32_t w0 = cur.w[0], w1 = cur.w[1], w2 = cur.w[2], w3 = cur.w[3], w4 = cur.w[4], w5 = cur.w[5],
                   w7 = cur.w[7];
    const uint32_t op = w0 & 0xFFu, fl = w0 >> 8;
    if (op < 64u) {
      Jet<F, K> a, b = jc<F, K>(F(0));
      if (fl & F_A_IMM) a = jc<F, K>(imm_of<F>(w2, w3)); else R.rd(w2, a.c);
      if (fl & F_HAS_B) { if (fl & F_B_IMM) b = jc<F, K>(imm_of<F>(w4, w5)); else R.rd(w4, b.c); }
      Jet<F, K> r = (op == OP_POWI) ? j_powi<F, K>(a, (int32_t)w7) : j_eval<F, K>(op, a, b);
      R.wr(w1, r.c);
    } else if (op == I_DOT) {
      const Terms<F> T(L, w3);
      Jet<F, K> acc = jc<F, K>(F(0));
      bool have = (fl & F_DOT_CONT) != 0;
      if (have) R.rd(w1, acc.c);
      uint32_t k = 0;
      while (k < w2) {
        const uint32_t blk = (T.run(k) >= 4) ? 4u : 1u;
        Jet<F, K> v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if ((uint32_t)j < blk && T.kind(k + j) != T_IMM) R.rd(T.reg[k + j], v[j].c);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if ((uint32_t)j >= blk) break;
          const uint32_t m = T.meta[k + j], kind = m & 3u;
          const Jet<F, K> c = jc<F, K>(T.imm(k + j));
          Jet<F, K> term = c;
          if (kind == T_PROD) {
            if (m & M_CONST_LEFT) jt::mul<F, K>(c.c, v[j].c, term.c);
            else jt::mul<F, K>(v[j].c, c.c, term.c);
          } else if (kind == T_REG) {
            term = v[j];
          }
          if (have) {
#pragma unroll
            for (int q = 0; q < K; ++q) acc.c[q] = acc.c[q] + term.c[q];
          } else {
            acc = term;
          }
          have = true;
        }
        k += blk;
      }
      R.wr(w1, acc.c);
    } else if (op == I_LOADN) {
      load_inputs<F, K, LDSONLY>(R, L, w2, w3, [&](uint32_t i, F* v) {
#pragma unroll
        for (int q = 0; q < K; ++q) v[q] = F(0);
        if (act) {
          const uint64_t row = L.S ? p / L.S : p;
          v[0] = X[row * L.xstride + i];
          if (K > 1)
            v[K > 1 ? 1 : 0] = L.rng ? (F)philox_direction(L.rng, L.seed, p * L.nin + i) : V[p * L.nin + i];
        }
      });
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

#define EHIP_PICK(KERNEL, ...)                                                                     \
  (f64 ? (all_lds ? (const void*)KERNEL<double, ##__VA_ARGS__, true>                              \
                  : (const void*)KERNEL<double, ##__VA_ARGS__, false>)                            \
       : (all_lds ? (const void*)KERNEL<float, ##__VA_ARGS__, true>                               \
                  : (const void*)KERNEL<float, ##__VA_ARGS__, false>))

hipError_t launch_scalar(bool f64, const Launch& L, uint32_t wg, size_t lds, hipStream_t s) {
  const bool all_lds = L.gregs == nullptr;
  return launch_common(EHIP_PICK(k_scalar), L, wg, lds, s);
}

hipError_t launch_dual(bool f64, const Launch& L, uint32_t wg, size_t lds, hipStream_t s) {
  const bool all_lds = L.gregs == nullptr;
  return launch_common(EHIP_PICK(k_dual), L, wg, lds, s);
}

hipError_t launch_jet(bool f64, int K, const Launch& L, uint32_t wg, size_t lds, hipStream_t s) {
  const bool all_lds = L.gregs == nullptr;
  switch (K) {
    case 1: return launch_common(EHIP_PICK(k_jet, 1), L, wg, lds, s);
    case 2: return launch_common(EHIP_PICK(k_jet, 2), L, wg, lds, s);
    case 3: return launch_common(EHIP_PICK(k_jet, 3), L, wg, lds, s);
    case 4: return launch_common(EHIP_PICK(k_jet, 4), L, wg, lds, s);
    case 5: return launch_common(EHIP_PICK(k_jet, 5), L, wg, lds, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ehip
