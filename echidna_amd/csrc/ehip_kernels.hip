// Wave-uniform interpreters for the compiled register program.
//
// One lane evaluates one point (or one point-direction pair); every lane of a
// workgroup runs the same instruction stream.  The stream (ehip_program.h,
// encode_stream) is cut into windows that the workgroup stages in LDS with
// coalesced vector loads: two window buffers, the next window arriving while
// the current one executes, one barrier per window.  Instruction words are
// read from LDS by every lane at the same address (broadcast) and made
// wave-uniform with readfirstlane where they steer control flow; register
// numbers and immediates are used straight from VGPRs.  Only the reference's
// data-dependent conventions (zero-adjoint skip, zero-multiplier guard,
// domain checks, NaN routing) are per-lane selects.
//
// Register file: registers [0, R_lds) live in LDS as [reg][component][lane]
// (slot-major, lane-minor: one conflict-free LDS transaction per wave);
// registers >= R_lds live in HBM as [reg][component][point] (slot-major,
// point-minor: one coalesced transaction per wave).  The LDS-only variant
// has no HBM path; the split variant keeps the two paths in distinct address
// spaces so neither degrades to flat accesses.
//
// Fused dot products run their product terms in blocks of eight: eight LDS
// reads issue back to back, then the products are added in term order (no
// FMA, the reference's rounding sequence).
#include <hip/hip_runtime.h>

#include "ehip_kernels.h"
#include "ehip_math.h"
#include "ehip_program.h"
#include "philox.h"

namespace ehip {
#ifndef EHIP_NS
#define EHIP_NS lds_fetch
#endif
namespace EHIP_NS {

extern __shared__ __attribute__((aligned(16))) unsigned char g_smem[];

#define AS3 __attribute__((address_space(3)))
#define AS1 __attribute__((address_space(1)))
// Constant address space: read-only for the whole launch, so wave-uniform
// loads through it become scalar loads (SGPR results, scalar cache).
#define AS4 __attribute__((address_space(4)))

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

template <class F> __device__ __forceinline__ F imm_of(uint32_t lo, uint32_t hi);
template <> __device__ __forceinline__ float imm_of<float>(uint32_t lo, uint32_t) { return __uint_as_float(lo); }
template <> __device__ __forceinline__ double imm_of<double>(uint32_t lo, uint32_t hi) {
  return __hiloint2double((int)hi, (int)lo);
}

// Register file.  Register numbers may arrive in VGPRs (wave-uniform values);
// the split variant makes them scalar before choosing LDS or HBM.
template <class F, int NC, bool LDSONLY>
struct Regs {
  AS3 F* lane;  // this lane's column of the LDS register file
  AS1 F* gl;
  uint32_t R_lds, lsh;  // lsh = log2(workgroup size)
  uint64_t gstride, p;
  __device__ __forceinline__ void rd(uint32_t r, F* o) const {
    if (LDSONLY) {
#pragma unroll
      for (int c = 0; c < NC; ++c) o[c] = lane[(r * NC + c) << lsh];
      return;
    }
    r = uni(r);
    if (r < R_lds) {
#pragma unroll
      for (int c = 0; c < NC; ++c) o[c] = lane[(r * NC + c) << lsh];
    } else {
      const uint64_t base = (uint64_t)(r - R_lds) * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) o[c] = gl[(base + c) * gstride + p];
    }
  }
  // component C only
  template <int C>
  __device__ __forceinline__ F rd_c(uint32_t r) const {
    static_assert(C < NC, "component");
    if (LDSONLY) return lane[(r * NC + C) << lsh];
    r = uni(r);
    if (r < R_lds) return lane[(r * NC + C) << lsh];
    return gl[((uint64_t)(r - R_lds) * NC + C) * gstride + p];
  }
  // the first N components only
  template <int N>
  __device__ __forceinline__ void rd_n(uint32_t r, F* o) const {
    static_assert(N <= NC, "components");
    if (LDSONLY) {
#pragma unroll
      for (int c = 0; c < N; ++c) o[c] = lane[(r * NC + c) << lsh];
      return;
    }
    r = uni(r);
    if (r < R_lds) {
#pragma unroll
      for (int c = 0; c < N; ++c) o[c] = lane[(r * NC + c) << lsh];
    } else {
      const uint64_t base = (uint64_t)(r - R_lds) * NC;
#pragma unroll
      for (int c = 0; c < N; ++c) o[c] = gl[(base + c) * gstride + p];
    }
  }
  __device__ __forceinline__ void wr(uint32_t r, const F* v) const {
    if (LDSONLY) {
#pragma unroll
      for (int c = 0; c < NC; ++c) lane[(r * NC + c) << lsh] = v[c];
      return;
    }
    r = uni(r);
    if (r < R_lds) {
#pragma unroll
      for (int c = 0; c < NC; ++c) lane[(r * NC + c) << lsh] = v[c];
    } else {
      const uint64_t base = (uint64_t)(r - R_lds) * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) gl[(base + c) * gstride + p] = v[c];
    }
  }
};

// DENSE members per block of the order-3 jet interpreter: each block reads
// every operand jet once, so larger blocks read the operands less often and
// hold more accumulators.  On the n=256 H=256 MLP (f32, 65,536 points x 64
// directions) 8 gave 39.7 M jets/s, 16 gave 52.2 M, and 12, 20, 24 and 32
// gave 34, 29, 28 and 17 M (more VGPRs, fewer waves, SGPR spills).
template <class F> constexpr int jet3_block() { return sizeof(F) == 4 ? 16 : 8; }

// LDS layout: [window buffer 0 | window buffer 1 (LDS fetch only) | register file].
constexpr uint32_t WIN_BYTES = WIN_WORDS * 4;

#ifndef EHIP_SFETCH
#define EHIP_SFETCH 0
#endif
// Program fetch.  LDS: each window is staged once per workgroup in LDS
// (double-buffered) and instruction words are read from there.  Scalar: the
// stream is read in place through the constant address space, so words
// arrive in SGPRs through the scalar cache with no LDS round trip and no
// readfirstlane.
constexpr bool SFETCH = EHIP_SFETCH != 0;
// LDS before the register file: the two instruction windows (LDS fetch).
constexpr uint32_t WIN_LDS = SFETCH ? 0u : 2u * WIN_BYTES;

template <class F, int NC, bool LDSONLY>
__device__ __forceinline__ Regs<F, NC, LDSONLY> make_regs(const Launch& L) {
  Regs<F, NC, LDSONLY> R;
  R.lane = (AS3 F*)(g_smem + WIN_LDS) + threadIdx.x;
  R.gl = (AS1 F*)(L.gregs);
  R.R_lds = L.R_lds;
  R.lsh = __builtin_ctz(blockDim.x);
  R.gstride = (uint64_t)gridDim.x * blockDim.x;
  R.p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  return R;
}



// Streams the program window by window and calls `exec(window, pos, w0)`
// for every instruction.
template <class Exec>
__device__ __forceinline__ void run_stream(const Launch& L, Exec exec) {
  if constexpr (SFETCH) {
    AS4 const uint32_t* prog = (AS4 const uint32_t*)L.prog;
    const uint32_t nwin = L.nwin;
    for (uint32_t w = 0; w < nwin; ++w) {
      AS4 const uint32_t* win = prog + (size_t)w * WIN_WORDS;
      uint32_t pos = 0;
      for (;;) {
        const uint32_t w0 = win[pos];
        const uint32_t op = w0 & 0xFFu;
        if (op == I_WEND || op == I_END) break;
        exec(win, pos, w0);
        pos += w0 >> 24;
      }
    }
    return;
  }
  AS3 uint32_t* buf = (AS3 uint32_t*)g_smem;
  const uint32_t tid = threadIdx.x, wg = blockDim.x;
  const u32x4* src = reinterpret_cast<const u32x4*>(L.prog);
  constexpr uint32_t Q = WIN_WORDS / 4;  // 16-byte vectors per window
  constexpr uint32_t PER = (Q + 63) / 64;  // vectors per lane at 64 lanes
  for (uint32_t i = tid; i < Q; i += wg) ((AS3 u32x4*)buf)[i] = src[i];
  __syncthreads();
  const uint32_t nwin = L.nwin;
  for (uint32_t w = 0; w < nwin; ++w) {
    AS3 const uint32_t* win = buf + (w & 1u) * WIN_WORDS;
    // request the next window now; store it once this one has executed
    u32x4 pre[PER];
    const bool more = w + 1 < nwin;
    if (more) {
#pragma unroll
      for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t i = tid + k * wg;
        if (i < Q) pre[k] = src[(size_t)(w + 1) * Q + i];
      }
    }
    uint32_t pos = 0;
    for (;;) {
      const uint32_t w0 = uni(win[pos]);
      const uint32_t op = w0 & 0xFFu;
      if (op == I_WEND || op == I_END) break;
      exec(win, pos, w0);
      pos += w0 >> 24;
    }
    if (more) {
      AS3 u32x4* nb = (AS3 u32x4*)(buf + ((w + 1) & 1u) * WIN_WORDS);
#pragma unroll
      for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t i = tid + k * wg;
        if (i < Q) nb[i] = pre[k];
      }
    }
    __syncthreads();
  }
}

// Word accessors for the current instruction.
template <class F, class P>
struct Ins {
  P p;
  __device__ __forceinline__ uint32_t u(uint32_t k) const { return uni(p[k]); }  // uniform
  __device__ __forceinline__ uint32_t v(uint32_t k) const { return p[k]; }       // VGPR copy
  __device__ __forceinline__ F imm(uint32_t k) const { return imm_of<F>(p[k], sizeof(F) == 8 ? p[k + 1] : 0u); }
};

// Inline term k of a DOT / DOT_REV / LOADN: meta, reg, imm.
template <class F, class P>
struct Term {
  P p;
  __device__ __forceinline__ Term(P ins, uint32_t k) : p(ins + 3 + k * (2 + sizeof(F) / 4)) {}
  __device__ __forceinline__ uint32_t meta() const { return p[0]; }
  __device__ __forceinline__ uint32_t reg() const { return p[1]; }
  __device__ __forceinline__ F imm() const { return imm_of<F>(p[2], sizeof(F) == 8 ? p[3] : 0u); }
  __device__ __forceinline__ uint32_t u32() const { return p[2]; }
};

template <class F, class P>
__device__ __forceinline__ Term<F, P> term(P ins, uint32_t k) {
  return Term<F, P>(ins, k);
}

// Batched input load: eight global loads in flight, then eight register writes.
template <class F, int NC, bool LDSONLY, class P, class MakeVal>
__device__ __forceinline__ void load_inputs(const Regs<F, NC, LDSONLY>& R, P ins, uint32_t n,
                                            MakeVal make) {
  uint32_t k = 0;
  for (; k + 8 <= n; k += 8) {
    F v[8][NC];
#pragma unroll
    for (int j = 0; j < 8; ++j) make(term<F>(ins, k + j).u32(), v[j]);
#pragma unroll
    for (int j = 0; j < 8; ++j) R.wr(term<F>(ins, k + j).reg(), v[j]);
  }
  for (; k < n; ++k) {
    F v[NC];
    make(term<F>(ins, k).u32(), v);
    R.wr(term<F>(ins, k).reg(), v);
  }
}

// ---------------------------------------------------------------------------
// DENSE / DENSE_REV (ehip_program.h): m dot products over one operand list.
// Each member's accumulator sees its terms in program order, with the same
// operations as its DOT (term 0 initialises, then acc + term), so results
// are bit-identical; an operand register is read once per block of JB
// members.  Operand lists and weights come from the pool through
// wave-uniform (scalar) loads.
// ---------------------------------------------------------------------------
// Value programs.  One term of a block of JB members (FULL: exactly JB, no
// per-member bounds checks): FIRST initialises the accumulators with it,
// otherwise it is added.  Pool words and weights arrive through scalar loads.
// Every kind runs the same straight-line x * w: the pool holds weight 1 for a
// plain register term, and an immediate term uses x = 1; x * 1 and 1 * c are
// exact, so each member sees the bits its DOT would produce, and the
// accumulators stay in place across terms.
template <class F, class Regs_>
__device__ __forceinline__ F dense_operand(const Regs_& R, uint32_t kr) {
  F x = F(1);
  if ((kr >> P_KIND_SHIFT) != T_IMM) R.rd(kr & P_REG_MASK, &x);
  return x;
}

template <class F, int JB, bool FULL, bool FIRST>
__device__ __forceinline__ void dense_fwd_term(F x, AS4 const F* PW, uint32_t m, uint32_t p, uint32_t j0, uint32_t jb,
                                               F (&acc)[JB]) {
  AS4 const F* w = PW + (uint64_t)p * m + j0;
#pragma unroll
  for (int j = 0; j < JB; ++j) {
    if (!(FULL || (uint32_t)j < jb)) continue;
    const F term = x * w[j];
    acc[j] = FIRST ? term : acc[j] + term;
  }
}

// The first term peeled, so the loop body is a plain acc + term; each
// term's operand is read one term ahead of its products, so that its
// register-file load overlaps the previous term's arithmetic.  A DENSE has
// at least one term.
template <class F, int JB, bool FULL, class Regs_>
__device__ __forceinline__ void dense_fwd_block(const Regs_& R, AS4 const uint32_t* PR, AS4 const F* PW, uint32_t m,
                                                uint32_t T, uint32_t j0, uint32_t jb) {
  F acc[JB];
  F x = dense_operand<F>(R, PR[0]);
  F xn = T > 1 ? dense_operand<F>(R, PR[1]) : F(0);
  dense_fwd_term<F, JB, FULL, true>(x, PW, m, 0, j0, jb, acc);
  for (uint32_t p = 1; p < T; ++p) {
    x = xn;
    if (p + 1 < T) xn = dense_operand<F>(R, PR[p + 1]);
    dense_fwd_term<F, JB, FULL, false>(x, PW, m, p, j0, jb, acc);
  }
#pragma unroll
  for (int j = 0; j < JB; ++j) {
    if (!(FULL || (uint32_t)j < jb)) continue;
    R.wr(PR[T + j0 + j], &acc[j]);
  }
}

// Jet programs: products follow the Taylor product with a constant jet
// (jt::mul, the operation the jet interpreter's DOT performs); one loop with
// a per-term select (the peeled value-program form raises the jet
// interpreter's register use: 115 instead of 96 VGPRs at order 3 in f32),
// pool words and weights through scalar loads.
// NZ: how many leading coefficients of the operands may be non-zero.  NZ <
// NC when every operand is a tape input (D_INPUT_OPERANDS): jets seed inputs
// as (x, v, 0, ..), so coefficients NZ.. are +0 and are not read.  With
// finite weights, an accumulator of such an order then holds +0, or NaN
// once a non-finite operand went through the full product, and every
// further term adds +0 to it, which changes neither; so those accumulators
// are set once and only the full product touches them.
// ROW (with NZ < NC): XR is the wave's common input row; operand values
// then come from XR[IX[p]] by scalar loads, and only the first-order
// coefficients from the register file.  The members' value coefficients are
// then the same in every lane, so lane l computes only member l mod JB's,
// term by term with the operations and the order every lane would use, and
// the JB results are broadcast by shuffles at the end of the block.
template <class F, int NC, int JB, bool FULL, int NZ, bool ROW, class Regs_>
__device__ __forceinline__ void dense_jet_block(const Regs_& R, AS4 const uint32_t* PR, AS4 const F* PW, uint32_t m,
                                                uint32_t T, uint32_t j0, uint32_t jb, AS4 const F* XR = nullptr,
                                                AS4 const uint32_t* IX = nullptr) {
  static_assert(NZ >= 1 && NZ <= NC, "coefficients");
  static_assert(NZ == NC || NZ == 2, "input operands carry (x, v)");
  static_assert(!ROW || NZ < NC, "a common row holds input operands only");
  static_assert((JB & (JB - 1)) == 0 && JB <= 64, "members per lane group");
  constexpr int Q0 = ROW ? 1 : 0;  // first coefficient kept per member
  F acc[JB][NC];
  // ROW: this lane's member, its weight for the current and the next term,
  // and its value-coefficient accumulator
  const uint32_t lane_m = (uint32_t)(threadIdx.x & (JB - 1));
  const uint32_t mj = (FULL || lane_m < jb) ? lane_m : jb - 1;
  F cl = F(0), cln = F(0), a0 = F(0);
  // true once a product term has entered the sums: a Taylor product's
  // coefficients are sums that start at +0, so from then on no accumulator
  // can be -0 (x + y is -0 only when both are), and acc + (0 + t) == acc + t
  bool absorbed = false;
  // A term's operand coefficients are read one term ahead of its arithmetic:
  // the register-file load (HBM past the LDS budget) and the scalar loads of
  // its input-row value then overlap the previous term's products instead of
  // stalling the wave once per term.  Immediate terms read nothing.
  auto fetch = [&](uint32_t p, uint32_t kr, F* x) {
    if ((kr >> P_KIND_SHIFT) == T_IMM) {
#pragma unroll
      for (int q = 0; q < NZ; ++q) x[q] = F(0);
      return;
    }
    if constexpr (ROW) {
      x[0] = XR[IX[p]];
      x[1] = R.template rd_c<1>(kr & P_REG_MASK);
    } else {
      R.template rd_n<NZ>(kr & P_REG_MASK, x);
    }
  };
  uint32_t knext = 0;
  F xn[NZ];
  if (T > 0) {
    knext = PR[0];
    fetch(0, knext, xn);
    if constexpr (ROW) cln = PW[j0 + mj];
  }
  if (NZ < NC) {
#pragma unroll
    for (int j = 0; j < JB; ++j)
#pragma unroll
      for (int q = NZ; q < NC; ++q) acc[j][q] = F(0);
  }
  for (uint32_t p = 0; p < T; ++p) {
    const uint32_t kr = knext;
    const uint32_t kind = kr >> P_KIND_SHIFT;
    F x[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) x[q] = q < NZ ? xn[q] : F(0);
    if constexpr (ROW) cl = cln;
    if (p + 1 < T) {
      knext = PR[p + 1];
      fetch(p + 1, knext, xn);
      if constexpr (ROW) cln = PW[(uint64_t)(p + 1) * m + j0 + mj];
    }
    AS4 const F* w = PW + (uint64_t)p * m + j0;
    if (kind == T_PROD) {
      const bool cleft = (kr & P_CLEFT) != 0;
      bool fin = true;
#pragma unroll
      for (int q = 0; q < NZ; ++q) fin = fin && isfinite(x[q]);
      if (__all(fin)) {
        // With every coefficient of x finite the Taylor product with the
        // constant jet [c, 0, ..] reduces exactly to 0 + x_k * c: its other
        // products are signed zeros, and a sum that starts at +0 absorbs
        // them (either operand order).
        if (absorbed) {
          if constexpr (ROW) a0 = a0 + x[0] * cl;
#pragma unroll
          for (int j = 0; j < JB; ++j) {
            if (!(FULL || (uint32_t)j < jb)) continue;
            const F c = w[j];
#pragma unroll
            for (int q = Q0; q < NZ; ++q) acc[j][q] = acc[j][q] + x[q] * c;
          }
        } else {
          if constexpr (ROW) {
            const F term = F(0) + x[0] * cl;
            a0 = (p == 0) ? term : a0 + term;
          }
#pragma unroll
          for (int j = 0; j < JB; ++j) {
            if (!(FULL || (uint32_t)j < jb)) continue;
            const F c = w[j];
#pragma unroll
            for (int q = Q0; q < NZ; ++q) {
              const F term = F(0) + x[q] * c;
              acc[j][q] = (p == 0) ? term : acc[j][q] + term;
            }
          }
        }
      } else {
        if constexpr (ROW) {
          // the value coefficient of the full Taylor product: 0 + a0 * b0
          const F term = cleft ? F(0) + cl * x[0] : F(0) + x[0] * cl;
          a0 = (p == 0) ? term : a0 + term;
        }
#pragma unroll
        for (int j = 0; j < JB; ++j) {
          if (!(FULL || (uint32_t)j < jb)) continue;
          const F c = w[j];
          F term[NC], cj[NC];
#pragma unroll
          for (int q = 0; q < NC; ++q) cj[q] = (q == 0) ? c : F(0);
          if (cleft) jt::mul<F, NC>(cj, x, term);
          else jt::mul<F, NC>(x, cj, term);
#pragma unroll
          for (int q = Q0; q < NC; ++q) acc[j][q] = (p == 0) ? term[q] : acc[j][q] + term[q];
        }
      }
      absorbed = true;
    } else if (kind == T_REG) {
      if constexpr (ROW) a0 = (p == 0) ? x[0] : a0 + x[0];
#pragma unroll
      for (int j = 0; j < JB; ++j) {
        if (!(FULL || (uint32_t)j < jb)) continue;
#pragma unroll
        for (int q = Q0; q < NZ; ++q) acc[j][q] = (p == 0) ? x[q] : acc[j][q] + x[q];
      }
    } else {
      if constexpr (ROW) a0 = (p == 0) ? cl : a0 + cl;
#pragma unroll
      for (int j = 0; j < JB; ++j) {
        if (!(FULL || (uint32_t)j < jb)) continue;
        const F c = w[j];
#pragma unroll
        for (int q = Q0; q < NZ; ++q) {
          const F term = (q == 0) ? c : F(0);
          acc[j][q] = (p == 0) ? term : acc[j][q] + term;
        }
      }
    }
  }
  if constexpr (ROW) {
    // member j's value coefficient from lane j (every lane of the wave
    // runs this block: the row is common to the whole wave)
#pragma unroll
    for (int j = 0; j < JB; ++j) acc[j][0] = __shfl(a0, j, 64);
  }
#pragma unroll
  for (int j = 0; j < JB; ++j) {
    if (!(FULL || (uint32_t)j < jb)) continue;
    R.wr(PR[T + j0 + j], acc[j]);
  }
}

template <class F, int NC, int JB, bool JET, class Regs_>
__device__ __forceinline__ void dense_fwd(const Regs_& R, const uint32_t* pool, uint32_t m, uint32_t T,
                                          uint32_t off_regs, uint32_t off_w, uint32_t dflags = 0,
                                          uint32_t off_ix = 0, AS4 const F* XR = nullptr) {
  uint32_t j0 = 0;
  if constexpr (JET) {
    AS4 const uint32_t* PR = (AS4 const uint32_t*)pool + off_regs;
    AS4 const F* PW = reinterpret_cast<AS4 const F*>((AS4 const uint32_t*)pool + off_w);
    constexpr int NZI = NC >= 3 ? 2 : NC;  // input operands: (x, v, 0, ..)
    if constexpr (NZI < NC) {
      if (dflags & D_INPUT_OPERANDS) {
        // operand values from the wave's common input row when it has one
        AS4 const uint32_t* IX = (AS4 const uint32_t*)pool + off_ix;
        if (XR) {
          for (; j0 + JB <= m; j0 += JB) dense_jet_block<F, NC, JB, true, NZI, true>(R, PR, PW, m, T, j0, JB, XR, IX);
          if (j0 < m) dense_jet_block<F, NC, JB, false, NZI, true>(R, PR, PW, m, T, j0, m - j0, XR, IX);
          return;
        }
        for (; j0 + JB <= m; j0 += JB) dense_jet_block<F, NC, JB, true, NZI, false>(R, PR, PW, m, T, j0, JB);
        if (j0 < m) dense_jet_block<F, NC, JB, false, NZI, false>(R, PR, PW, m, T, j0, m - j0);
        return;
      }
    }
    for (; j0 + JB <= m; j0 += JB) dense_jet_block<F, NC, JB, true, NC, false>(R, PR, PW, m, T, j0, JB);
    if (j0 < m) dense_jet_block<F, NC, JB, false, NC, false>(R, PR, PW, m, T, j0, m - j0);
  } else {
    static_assert(NC == 1, "value programs carry one component");
    AS4 const uint32_t* PR = (AS4 const uint32_t*)pool + off_regs;
    AS4 const F* PW = reinterpret_cast<AS4 const F*>((AS4 const uint32_t*)pool + off_w);
    for (; j0 + JB <= m; j0 += JB) dense_fwd_block<F, JB, true>(R, PR, PW, m, T, j0, JB);
    if (j0 < m) dense_fwd_block<F, JB, false>(R, PR, PW, m, T, j0, m - j0);
  }
}

// One block of IB adjoints.  The reference skips an op whose adjoint is
// zero and a contribution whose partial (here the weight; 1 for a plain
// register term) is zero; every other contribution is cur + c * g, Inf and
// NaN adjoints included.  Fast path, taken per member when the compiler
// found every weight non-zero (D_ALL_PROD_NZ) and the member's adjoint is
// non-zero in every lane: then no guard can trigger.  Otherwise each
// contribution is a select between cur + c * g and cur.  The next member's
// register index is loaded one member ahead.
template <class F, int IB, bool FULL, bool LDSONLY>
__device__ __forceinline__ void dense_rev_block(const Regs<F, 1, LDSONLY>& R, AS4 const uint32_t* PR, AS4 const F* PW,
                                                uint32_t m, uint32_t T, uint32_t i0, uint32_t ib, bool fastw) {
  F cur[IB];
#pragma unroll
  for (int i = 0; i < IB; ++i) {
    cur[i] = F(0);
    if (!(FULL || (uint32_t)i < ib)) continue;
    const uint32_t kr = PR[i0 + i];
    if (!(kr & P_FIRST)) R.rd(kr & P_REG_MASK, &cur[i]);
  }
  // each member's adjoint is read one member ahead of its contributions
  F gn;
  R.rd(PR[T], &gn);
  for (uint32_t j = 0; j < m; ++j) {
    const F g = gn;
    if (j + 1 < m) R.rd(PR[T + j + 1], &gn);
    AS4 const F* w = PW + (uint64_t)j * T + i0;
    const bool live = g != F(0);
    if (fastw && __all(live)) {
#pragma unroll
      for (int i = 0; i < IB; ++i) {
        if (!(FULL || (uint32_t)i < ib)) continue;
        cur[i] = cur[i] + w[i] * g;
      }
    } else {
#pragma unroll
      for (int i = 0; i < IB; ++i) {
        if (!(FULL || (uint32_t)i < ib)) continue;
        const F c = w[i];
        const F t = cur[i] + c * g;
        cur[i] = (live && c != F(0)) ? t : cur[i];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < IB; ++i) {
    if (!(FULL || (uint32_t)i < ib)) continue;
    R.wr(PR[i0 + i] & P_REG_MASK, &cur[i]);
  }
}

template <class F, bool LDSONLY>
__device__ __forceinline__ void dense_rev(const Regs<F, 1, LDSONLY>& R, const uint32_t* pool, uint32_t m, uint32_t T,
                                          uint32_t off_regs, uint32_t off_w, uint32_t dflags) {
  constexpr int IB = sizeof(F) == 8 ? 4 : 8;
  AS4 const uint32_t* PR = (AS4 const uint32_t*)pool + off_regs;
  AS4 const F* PW = reinterpret_cast<AS4 const F*>((AS4 const uint32_t*)pool + off_w);
  const bool fastw = (dflags & D_ALL_PROD_NZ) != 0;
  uint32_t i0 = 0;
  for (; i0 + IB <= T; i0 += IB) dense_rev_block<F, IB, true, LDSONLY>(R, PR, PW, m, T, i0, IB, fastw);
  if (i0 < T) dense_rev_block<F, IB, false, LDSONLY>(R, PR, PW, m, T, i0, T - i0, fastw);
}

// ---------------------------------------------------------------------------
// Scalar interpreter: forward values, outputs, reverse adjoint sweep.
// ---------------------------------------------------------------------------
// `p` is the point this lane evaluates; R.p the lane's slot in the HBM
// register file (the same in a one-shot launch, different in the queue).
template <class F, bool LDSONLY>
__device__ __forceinline__ void scalar_body(const Launch& L, const Regs<F, 1, LDSONLY>& R, uint64_t p) {
  const bool act = p < L.B;
  const F* X = reinterpret_cast<const F*>(L.x);
  F* Y = reinterpret_cast<F*>(L.y);
  F* G = reinterpret_cast<F*>(L.g);
  run_stream(L, [&](auto win, uint32_t pos, uint32_t w0) {
    const Ins<F, decltype(win)> I{win + pos};
    const auto ip = win + pos;
    const uint32_t op = w0 & 0xFFu, fl = (w0 >> 8) & 0xFFFFu;
    if (op < 64u) {
      F a, b = F(0), r;
      if (fl & F_A_IMM) a = I.imm(2); else R.rd(I.v(2), &a);
      if (fl & F_HAS_B) { if (fl & F_B_IMM) b = I.imm(4); else R.rd(I.v(4), &b); }
      r = (op == OP_POWI) ? s_powi(a, (int32_t)I.u(7)) : s_eval<F>(op, a, b);
      R.wr(I.v(1), &r);
    } else if (op == I_DOT) {
      const uint32_t n = I.u(2), dst = I.v(1);
      F acc = F(0);
      bool have = (fl & F_DOT_CONT) != 0;
      if (have) R.rd(dst, &acc);
      uint32_t k = 0;
      while (k < n) {
        const uint32_t m = uni(term<F>(ip, k).meta());
        if (((m >> M_RUN_SHIFT) & 0xFFu) >= 8) {
          F v[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) R.rd(term<F>(ip, k + j).reg(), &v[j]);
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const F t = v[j] * term<F>(ip, k + j).imm();
            acc = (j == 0 && !have) ? t : acc + t;
          }
          have = true;
          k += 8;
        } else {
          const auto T = term<F>(ip, k);
          const uint32_t kind = m & 3u;
          const F c = T.imm();
          F term = c;
          if (kind != T_IMM) {
            F v;
            R.rd(T.reg(), &v);
            term = (kind == T_PROD) ? v * c : v;
          }
          acc = have ? acc + term : term;
          have = true;
          ++k;
        }
      }
      R.wr(dst, &acc);
    } else if (op >= I_REV_BASE) {
      const uint32_t top = op - I_REV_BASE;
      F g;
      R.rd(I.v(1), &g);
      const bool live = g != F(0);
      F a = F(0), b = F(0), r = F(0);
      if (fl & F_A_IMM) a = I.imm(2); else if (fl & F_RD_A) R.rd(I.v(2), &a);
      if (fl & F_B_IMM) b = I.imm(4); else if (fl & F_RD_B) R.rd(I.v(4), &b);
      if (fl & F_RD_R) R.rd(I.v(6), &r);
      F da, db;
      if (top == OP_POWI) { da = s_powi_partial((int32_t)I.u(7), a, r); db = F(0); }
      else s_partials<F>(top, a, b, r, da, db);
      if (fl & F_ACC_A) {
        F c = F(0);
        if (!(fl & F_FIRST_A)) R.rd(I.v(3), &c);
        if (live && da != F(0)) c = c + da * g;
        R.wr(I.v(3), &c);
      }
      if (fl & F_ACC_B) {
        F c = F(0);
        if (!(fl & F_FIRST_B)) R.rd(I.v(5), &c);
        if (live && db != F(0)) c = c + db * g;
        R.wr(I.v(5), &c);
      }
    } else if (op == I_DOT_REV) {
      const uint32_t n = I.u(2);
      F g;
      R.rd(I.v(1), &g);
      const bool live = g != F(0);
      uint32_t k = 0;
      while (k < n) {
        const uint32_t m0 = uni(term<F>(ip, k).meta());
        if (((m0 >> M_RUN_SHIFT) & 0xFFu) >= 8) {
          // eight product terms with distinct adjoint registers; a first
          // contribution reads its register anyway and discards the value
          // all term words are read before the first register write, which
          // the compiler must otherwise assume may alias the window
          F cur[8], cj[8];
          uint32_t rj[8], mj[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const auto T = term<F>(ip, k + j);
            rj[j] = T.reg();
            mj[j] = T.meta();
            cj[j] = T.imm();
          }
#pragma unroll
          for (int j = 0; j < 8; ++j) R.rd(rj[j], &cur[j]);
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            F c0 = (mj[j] & M_FIRST) ? F(0) : cur[j];
            if (live && cj[j] != F(0)) c0 = c0 + cj[j] * g;
            R.wr(rj[j], &c0);
          }
          k += 8;
        } else {
          const auto T = term<F>(ip, k);
          const F c = T.imm();
          F cur = F(0);
          if (!(m0 & M_FIRST)) R.rd(T.reg(), &cur);
          if ((m0 & 3u) == T_PROD) {
            if (live && c != F(0)) cur = cur + c * g;
          } else if (live) {
            cur = cur + F(1) * g;
          }
          R.wr(T.reg(), &cur);
          ++k;
        }
      }
    } else if (op == I_KINK) {
      // KinkEntry of forward_nonsmooth (src/bytecode_tape/forward.rs:87-187)
      F a, b = F(0);
      if (fl & F_A_IMM) a = I.imm(2); else R.rd(I.v(2), &a);
      if (fl & F_HAS_B) { if (fl & F_B_IMM) b = I.imm(4); else R.rd(I.v(4), &b); }
      const uint32_t kop = I.u(6);
      F sw;
      bool up;
      if (kop == OP_ABS || kop == OP_SIGNUM) {
        sw = a;
        up = a >= F(0);
      } else if (kop == OP_MAX) {
        sw = a - b;
        up = a >= b || b != b;
      } else if (kop == OP_MIN) {
        sw = a - b;
        up = a <= b || b != b;
      } else if (kop == OP_ROUND) {
        const F sh = a + F(0.5);
        sw = sh - round(sh);
        up = a - floor(a) < F(0.5);
      } else if (kop == OP_FRACT) {
        sw = a - round(a);
        up = a - trunc(a) >= F(0);
      } else {  // floor, ceil, trunc
        sw = a - round(a);
        up = a - floor(a) < F(0.5);
      }
      if (act) {
        const uint64_t o = p * L.nkinks + I.u(1);
        reinterpret_cast<F*>(L.ksw)[o] = sw;
        L.kbr[o] = up ? int8_t(1) : int8_t(-1);
      }
    } else if (op == I_DENSE) {
      dense_fwd<F, 1, 16, false>(R, L.pool, I.u(1), I.u(2), I.u(3), I.u(4));
    } else if (op == I_DENSE_REV) {
      dense_rev<F, LDSONLY>(R, L.pool, I.u(1), I.u(2), I.u(3), I.u(4), I.u(5));
    } else if (op == I_LOADN) {
      load_inputs<F, 1, LDSONLY>(R, ip, I.u(2), [&](uint32_t i, F* v) {
        v[0] = act ? X[p * L.xstride + i] : F(0);
      });
    } else if (op == I_STORE_OUT) {
      F v;
      if (fl & F_A_IMM) v = I.imm(2); else R.rd(I.v(2), &v);
      if (act && Y) Y[p * L.nout + I.u(1)] = v;
    } else if (op == I_SEED) {
      F one = F(1);
      R.wr(I.v(1), &one);
    } else if (op == I_STORE_GRAD) {
      F v;
      if (fl & F_A_IMM) v = I.imm(2); else R.rd(I.v(2), &v);
      if (act && G) G[p * L.nin + I.u(1)] = v;
    }
  });
}

template <class F, bool LDSONLY>
__global__ __launch_bounds__(256) void k_scalar(Launch L) {
  const Regs<F, 1, LDSONLY> R = make_regs<F, 1, LDSONLY>(L);
  scalar_body<F, LDSONLY>(L, R, R.p);
}

// Many tapes in one persistent launch.  Each workgroup takes (job, tile)
// items from a global counter until the queue is empty; the host orders the
// tiles longest program first.  A workgroup keeps its slot in the HBM
// register file for every tile it runs.
template <class F>
__global__ __launch_bounds__(256) void k_queue_scalar(Queue Q) {
  __shared__ uint32_t item;
  for (;;) {
    if (threadIdx.x == 0) item = atomicAdd(Q.counter, 1u);
    __syncthreads();
    const uint32_t it = uni(item);
    // every wave has read `item` before thread 0 may overwrite it for the
    // next tile (the scalar-fetch body has no barrier of its own)
    __syncthreads();
    if (it >= Q.ntiles) break;
    const QueueTile t = Q.tiles[it];
    const Launch L = Q.jobs[t.job];
    const Regs<F, 1, false> R = make_regs<F, 1, false>(L);
    scalar_body<F, false>(L, R, t.first + threadIdx.x);
  }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Dual interpreter: forward-over-reverse Hessian-vector products.
// ---------------------------------------------------------------------------
template <class F>
__device__ __forceinline__ Dual<F> dual_norm(Dual<F> x) {
  // 0 + 1 * x: the adjoint a single Add link hands to each operand.
  return d_add(dc(F(0)), d_mul(dc(F(1)), x));
}

// Dual programs.  DENSE: each member's DOT in dual arithmetic (the product
// with the constant dual (c, 0) and d_add), term by term in program order,
// so every member gets the bits its DOT gives; an operand is read once per
// block of JB members.
template <class F, int JB, bool FULL, bool LDSONLY>
__device__ __forceinline__ void dense_dual_block(const Regs<F, 2, LDSONLY>& R, AS4 const uint32_t* PR,
                                                 AS4 const F* PW, uint32_t m, uint32_t T, uint32_t j0, uint32_t jb) {
  Dual<F> acc[JB];
  // a term's operand is read one term ahead of its products
  auto fetch = [&](uint32_t kr, F* t) {
    if ((kr >> P_KIND_SHIFT) != T_IMM) {
      R.rd(kr & P_REG_MASK, t);
    } else {
      t[0] = F(0);
      t[1] = F(0);
    }
  };
  uint32_t knext = 0;
  F tn[2] = {F(0), F(0)};
  if (T > 0) {
    knext = PR[0];
    fetch(knext, tn);
  }
  for (uint32_t p = 0; p < T; ++p) {
    const uint32_t kr = knext;
    const uint32_t kind = kr >> P_KIND_SHIFT;
    AS4 const F* w = PW + (uint64_t)p * m + j0;
    Dual<F> x = dc(F(0));
    if (kind != T_IMM) x = Dual<F>{tn[0], tn[1]};
    if (p + 1 < T) {
      knext = PR[p + 1];
      fetch(knext, tn);
    }
    const bool cleft = (kr & P_CLEFT) != 0;
#pragma unroll
    for (int j = 0; j < JB; ++j) {
      if (!(FULL || (uint32_t)j < jb)) continue;
      const Dual<F> c = dc(w[j]);
      Dual<F> term = c;
      if (kind == T_PROD) term = cleft ? d_mul(c, x) : d_mul(x, c);
      else if (kind == T_REG) term = x;
      acc[j] = (p == 0) ? term : d_add(acc[j], term);
    }
  }
#pragma unroll
  for (int j = 0; j < JB; ++j) {
    if (!(FULL || (uint32_t)j < jb)) continue;
    F t[2] = {acc[j].re, acc[j].eps};
    R.wr(PR[T + j0 + j], t);
  }
}

// DENSE_REV: the members' DOT_REVs applied to each operand adjoint in their
// original order.  A DOT_REV hands a term at depth d the adjoint normalised
// d times by 0 + 1 * g (one Add link each), and a product term normalises
// once more.  One normalisation N(g) already gives every later one's value,
// so a product contributes c * N(g) and a plain term g (depth 0, P_DEEP
// clear) or N(g).  Skips as in the reference: an all-zero adjoint (per lane)
// and a zero weight (uniform).
template <class F, int IB, bool FULL, bool LDSONLY>
__device__ __forceinline__ void dense_dual_rev_block(const Regs<F, 2, LDSONLY>& R, AS4 const uint32_t* PR,
                                                     AS4 const F* PW, uint32_t m, uint32_t T, uint32_t i0,
                                                     uint32_t ib) {
  static_assert(IB <= 32, "one flag bit per adjoint");
  Dual<F> cur[IB];
  uint32_t prod = 0, deep = 0;
#pragma unroll
  for (int i = 0; i < IB; ++i) {
    cur[i] = dc(F(0));
    if (!(FULL || (uint32_t)i < ib)) continue;
    const uint32_t kr = PR[i0 + i];
    if ((kr >> P_KIND_SHIFT) == T_PROD) prod |= 1u << i;
    if (kr & P_DEEP) deep |= 1u << i;
    if (!(kr & P_FIRST)) {
      F t[2];
      R.rd(kr & P_REG_MASK, t);
      cur[i] = Dual<F>{t[0], t[1]};
    }
  }
  // each member's adjoint is read one member ahead of its contributions
  F tn[2];
  R.rd(PR[T], tn);
  for (uint32_t j = 0; j < m; ++j) {
    const Dual<F> g{tn[0], tn[1]};
    if (j + 1 < m) R.rd(PR[T + j + 1], tn);
    const bool live = !d_zero(g);
    const Dual<F> n1 = dual_norm(g);
    AS4 const F* w = PW + (uint64_t)j * T + i0;
#pragma unroll
    for (int i = 0; i < IB; ++i) {
      if (!(FULL || (uint32_t)i < ib)) continue;
      Dual<F> nx;
      if (prod & (1u << i)) {
        const F c = w[i];
        if (c == F(0)) continue;
        nx = d_add(cur[i], d_mul(dc(c), n1));
      } else {
        nx = d_add(cur[i], d_mul(dc(F(1)), (deep & (1u << i)) ? n1 : g));
      }
      cur[i].re = live ? nx.re : cur[i].re;
      cur[i].eps = live ? nx.eps : cur[i].eps;
    }
  }
#pragma unroll
  for (int i = 0; i < IB; ++i) {
    if (!(FULL || (uint32_t)i < ib)) continue;
    F t[2] = {cur[i].re, cur[i].eps};
    R.wr(PR[i0 + i] & P_REG_MASK, t);
  }
}

template <class F, bool LDSONLY>
__device__ __forceinline__ void dense_dual(const Regs<F, 2, LDSONLY>& R, const uint32_t* pool, uint32_t m, uint32_t T,
                                           uint32_t off_regs, uint32_t off_w, bool rev) {
  constexpr int B = 8;
  AS4 const uint32_t* PR = (AS4 const uint32_t*)pool + off_regs;
  AS4 const F* PW = reinterpret_cast<AS4 const F*>((AS4 const uint32_t*)pool + off_w);
  const uint32_t n = rev ? T : m;
  uint32_t k0 = 0;
  for (; k0 + B <= n; k0 += B) {
    if (rev) dense_dual_rev_block<F, B, true, LDSONLY>(R, PR, PW, m, T, k0, B);
    else dense_dual_block<F, B, true, LDSONLY>(R, PR, PW, m, T, k0, B);
  }
  if (k0 < n) {
    if (rev) dense_dual_rev_block<F, B, false, LDSONLY>(R, PR, PW, m, T, k0, n - k0);
    else dense_dual_block<F, B, false, LDSONLY>(R, PR, PW, m, T, k0, n - k0);
  }
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
  run_stream(L, [&](auto win, uint32_t pos, uint32_t w0) {
    const Ins<F, decltype(win)> I{win + pos};
    const auto ip = win + pos;
    const uint32_t op = w0 & 0xFFu, fl = (w0 >> 8) & 0xFFFFu;
    if (op < 64u) {
      Dual<F> a, b = dc(F(0));
      a = (fl & F_A_IMM) ? dc(I.imm(2)) : rdd(I.v(2));
      if (fl & F_HAS_B) b = (fl & F_B_IMM) ? dc(I.imm(4)) : rdd(I.v(4));
      Dual<F> r = (op == OP_POWI) ? d_powi(a, (int32_t)I.u(7)) : d_eval<F>(op, a, b);
      wrd(I.v(1), r);
    } else if (op == I_DOT) {
      const uint32_t n = I.u(2), dst = I.v(1);
      Dual<F> acc = dc(F(0));
      bool have = (fl & F_DOT_CONT) != 0;
      if (have) acc = rdd(dst);
      for (uint32_t k = 0; k < n; ++k) {
        const auto T = term<F>(ip, k);
        const uint32_t m = uni(T.meta()), kind = m & 3u;
        const Dual<F> c = dc(T.imm());
        Dual<F> term = c;
        if (kind != T_IMM) {
          Dual<F> v = rdd(T.reg());
          term = (kind == T_PROD) ? ((m & M_CONST_LEFT) ? d_mul(c, v) : d_mul(v, c)) : v;
        }
        acc = have ? d_add(acc, term) : term;
        have = true;
      }
      wrd(dst, acc);
    } else if (op == I_DENSE || op == I_DENSE_REV) {
      dense_dual<F, LDSONLY>(R, L.pool, I.u(1), I.u(2), I.u(3), I.u(4), op == I_DENSE_REV);
    } else if (op >= I_REV_BASE) {
      const uint32_t top = op - I_REV_BASE;
      const Dual<F> g = rdd(I.v(1));
      const bool live = !d_zero(g);
      Dual<F> a = dc(F(0)), b = dc(F(0)), r = dc(F(0));
      if (fl & F_A_IMM) a = dc(I.imm(2)); else if (fl & F_RD_A) a = rdd(I.v(2));
      if (fl & F_B_IMM) b = dc(I.imm(4)); else if (fl & F_RD_B) b = rdd(I.v(4));
      if (fl & F_RD_R) r = rdd(I.v(6));
      Dual<F> da, db;
      if (top == OP_POWI) { da = d_powi_partial((int32_t)I.u(7), a, r); db = dc(F(0)); }
      else d_partials<F>(top, a, b, r, da, db);
      if (fl & F_ACC_A) {
        Dual<F> c = (fl & F_FIRST_A) ? dc(F(0)) : rdd(I.v(3));
        if (live && !d_zero(da)) c = d_add(c, d_mul(da, g));
        wrd(I.v(3), c);
      }
      if (fl & F_ACC_B) {
        Dual<F> c = (fl & F_FIRST_B) ? dc(F(0)) : rdd(I.v(5));
        if (live && !d_zero(db)) c = d_add(c, d_mul(db, g));
        wrd(I.v(5), c);
      }
    } else if (op == I_DOT_REV) {
      const uint32_t n = I.u(2);
      const Dual<F> g = rdd(I.v(1));
      const bool live = !d_zero(g);
      Dual<F> gd = g;
      uint32_t depth = 0;
      for (uint32_t k = 0; k < n; ++k) {
        const auto T = term<F>(ip, k);
        const uint32_t m = uni(T.meta()), kind = m & 3u, tdepth = m >> M_DEPTH_SHIFT;
        const F c = T.imm();
        if (tdepth < depth) { gd = g; depth = 0; }
        while (depth < tdepth) { gd = dual_norm(gd); ++depth; }
        Dual<F> cur = (m & M_FIRST) ? dc(F(0)) : rdd(T.reg());
        if (live) {
          if (kind == T_PROD) {
            if (c != F(0)) cur = d_add(cur, d_mul(dc(c), dual_norm(gd)));
          } else {
            cur = d_add(cur, d_mul(dc(F(1)), gd));
          }
        }
        wrd(T.reg(), cur);
      }
    } else if (op == I_LOADN) {
      DirBlock db;  // consecutive inputs share a direction block
      load_inputs<F, 2, LDSONLY>(R, ip, I.u(2), [&](uint32_t i, F* v) {
        v[0] = F(0);
        v[1] = F(0);
        if (act) {
          const uint64_t row = L.S ? p / L.S : p;
          v[0] = X[row * L.xstride + i];
          v[1] = L.rng ? (F)lane_direction(L.rng, L.seed, L.S, p, L.nin, i, db) : V[p * L.nin + i];
        }
      });
    } else if (op == I_SEED) {
      wrd(I.v(1), dc(F(1)));
    } else if (op == I_STORE_GRAD) {
      Dual<F> v = (fl & F_A_IMM) ? dc(I.imm(2)) : rdd(I.v(2));
      if (act) { G[p * L.nin + I.u(1)] = v.re; H[p * L.nin + I.u(1)] = v.eps; }
    }
    // I_STORE_OUT: the HVP entry points return gradients and products only.
  });
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
  // The wave's common input row, when every lane reads the same one: x
  // shared by all points, or S lanes per point with S a multiple of 64 (a
  // wave is then one point's samples) in a wave inside the batch.  DENSE
  // over inputs then reads operand values from it by scalar loads.
  const uint64_t p0 = p - (threadIdx.x & 63u);
  const bool one_row = uni((L.xstride == 0 || (L.S >= 64 && (L.S & 63u) == 0 && p0 < L.B)) ? 1u : 0u) != 0;
  const uint64_t roff = (one_row && L.xstride) ? (L.S ? p0 / L.S : p0) * L.xstride : 0;
  const uint64_t roff_u = ((uint64_t)uni((uint32_t)(roff >> 32)) << 32) | uni((uint32_t)roff);
  AS4 const F* XR = one_row ? (AS4 const F*)L.x + roff_u : nullptr;
  const F* V = reinterpret_cast<const F*>(L.dirs);
  F* Y = reinterpret_cast<F*>(L.y);
  run_stream(L, [&](auto win, uint32_t pos, uint32_t w0) {
    const Ins<F, decltype(win)> I{win + pos};
    const auto ip = win + pos;
    const uint32_t op = w0 & 0xFFu, fl = (w0 >> 8) & 0xFFFFu;
    if (op < 64u) {
      Jet<F, K> a, b = jc<F, K>(F(0));
      if (fl & F_A_IMM) a = jc<F, K>(I.imm(2)); else R.rd(I.v(2), a.c);
      if (fl & F_HAS_B) { if (fl & F_B_IMM) b = jc<F, K>(I.imm(4)); else R.rd(I.v(4), b.c); }
      Jet<F, K> r = (op == OP_POWI) ? j_powi<F, K>(a, (int32_t)I.u(7)) : j_eval<F, K>(op, a, b);
      R.wr(I.v(1), r.c);
    } else if (op == I_DOT) {
      const uint32_t n = I.u(2), dst = I.v(1);
      Jet<F, K> acc = jc<F, K>(F(0));
      bool have = (fl & F_DOT_CONT) != 0;
      if (have) R.rd(dst, acc.c);
      uint32_t k = 0;
      while (k < n) {
        const uint32_t m0 = uni(term<F>(ip, k).meta());
        const uint32_t blk = (((m0 >> M_RUN_SHIFT) & 0xFFu) >= 4) ? 4u : 1u;
        Jet<F, K> v[4];
        if (blk == 4) {
#pragma unroll
          for (int j = 0; j < 4; ++j) R.rd(term<F>(ip, k + j).reg(), v[j].c);
        } else if ((m0 & 3u) != T_IMM) {
          R.rd(term<F>(ip, k).reg(), v[0].c);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if ((uint32_t)j >= blk) break;
          const auto T = term<F>(ip, k + j);
          const uint32_t m = uni(T.meta()), kind = m & 3u;
          const Jet<F, K> c = jc<F, K>(T.imm());
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
      R.wr(dst, acc.c);
    } else if (op == I_DENSE) {
      dense_fwd<F, K, (K <= 2 ? 16 : (K == 3 ? jet3_block<F>() : 4)), true>(R, L.pool, I.u(1), I.u(2), I.u(3), I.u(4),
                                                                          I.u(5), I.u(6), XR);
    } else if (op == I_LOADN) {
      DirBlock db;  // consecutive inputs share a direction block
      load_inputs<F, K, LDSONLY>(R, ip, I.u(2), [&](uint32_t i, F* v) {
#pragma unroll
        for (int q = 0; q < K; ++q) v[q] = F(0);
        if (act) {
          const uint64_t row = L.S ? p / L.S : p;
          v[0] = X[row * L.xstride + i];
          if (K > 1) {
            if (L.rng) {
              const F d = (F)lane_direction(L.rng, L.seed, L.S, p, L.nin, i, db);
              v[1] = d;
              if (L.dirs_out) reinterpret_cast<F*>(L.dirs_out)[(uint64_t)i * L.B + p] = d;
            } else {
              v[1] = V[p * L.nin + i];
            }
          }
        }
      });
    } else if (op == I_STORE_OUT) {
      Jet<F, K> v;
      if (fl & F_A_IMM) v = jc<F, K>(I.imm(2)); else R.rd(I.v(2), v.c);
      if (act) {
        const uint32_t j = I.u(1);
#pragma unroll
        for (int q = 0; q < K; ++q) Y[(p * L.nout + j) * K + q] = v.c[q];
      }
    }
  });
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static hipError_t launch_common(const void* fn, const Launch& L, uint32_t wg, size_t lds, hipStream_t s) {
  uint64_t grid = (L.B + wg - 1) / wg;
  if (grid == 0) grid = 1;
  void* args[] = {const_cast<Launch*>(&L)};
  return hipLaunchKernel(fn, dim3((uint32_t)grid), dim3(wg), args, lds + WIN_LDS, s);
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

const void* queue_kernel(bool f64) {
  return f64 ? (const void*)k_queue_scalar<double> : (const void*)k_queue_scalar<float>;
}

hipError_t launch_queue(bool f64, const Queue& Q, uint32_t grid, uint32_t wg, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, Q.counter, 1u, 0u);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  void* args[] = {const_cast<Queue*>(&Q)};
  return hipLaunchKernel(queue_kernel(f64), dim3(grid), dim3(wg), args, lds + WIN_LDS, s);
}

size_t queue_lds_bytes(size_t reg_lds) { return reg_lds + WIN_LDS; }

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

}  // namespace EHIP_NS
}  // namespace ehip
