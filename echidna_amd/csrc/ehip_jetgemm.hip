// Order-3 f32 Taylor-jet DENSE blocks as batch GEMMs on the matrix cores
// (ehip_set_jet_mma).
//
// The segmented jet program (segment_jet_program) leaves every register of
// every lane in HBM, one row of `lanes` floats per (register, coefficient).
// A DENSE's members then are, per coefficient q,
//     C_q[j][lane] = sum_p W[p][j] X_q[p][lane]
// (members x terms times terms x lanes), with X_q[p] the row of term p's
// operand register.  A workgroup owns NT lanes and 128 members: it stages
// 16 terms x NT lanes of every coefficient in LDS, and each of its four
// waves multiplies its 32 members' weights into them with
// v_mfma_f32_32x32x2_f32 (two terms a step, FP32 accumulation; lane l
// supplies A[j = l % 32][k = l / 32] and B[k = l / 32][n = l % 32], and holds
// the results of rows 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32), one
// product per 32-lane column block.  NT is 64, or 128 when the matrix cores
// form one coefficient only (c1 of tape inputs, c0 apart): every weight read
// from LDS then feeds four products, as with two coefficients at 64 lanes.  The
// next stage's operands (and the term descriptors of the one after) are
// loaded into registers while the matrix cores work on the current one.  Immediate terms add their weight to c0;
// operands that are tape inputs (D_INPUT_OPERANDS) carry +0 from order 2 on,
// and so do the results.
//
// c0 is the primal, equal on every lane that reads the same x row.  When
// the lanes of a 64-lane tile fall into PT such groups (`share` lanes each,
// aligned to the tile), c0 is a PT-column product on the vector ALUs, fused
// into the same pass, and is broadcast to the group's lanes; the matrix
// cores then form only c1 (and c2).
//
// An activation the program computes in place right after the DENSE (the
// next layer's tanh, say) is applied here, with the interpreter's jet rule,
// before the member is stored: the member's pre-activation is never written
// and the interpreter segment after the GEMM skips those instructions.
//
// Direct operands (DenseGemm::direct): when every operand is a tape input
// and the directions are drawn on the device, the operand jets (x_i, v_i, 0)
// come from their sources instead of the input registers: x from its row,
// v as the Rademacher sign drawn here when the operands are the inputs in
// order (DenseGemm::nseq: a lane's four signs of a stage are one 4-bit
// field of its Philox block, which a cache keeps for 128 signs) or from
// directions drawn coordinate-major beforehand (k_draw_dirs).  The values
// are the ones the interpreter's LOADN writes, so results do not change;
// the program then skips writing inputs only GEMMs read.
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
#include "philox.h"

namespace ehip {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t NC = 3;    // coefficients per register (order 3)
constexpr uint32_t MT = JET_GEMM_MT;  // members per workgroup: four waves of 32
constexpr uint32_t KC = JET_GEMM_KC;  // terms per stage
constexpr uint32_t RPT = KC / 4;  // stage rows per wave

__device__ __forceinline__ float* row_of(const DenseGemm& G, uint32_t reg, uint32_t q) {
  return G.regs + ((uint64_t)reg * NC + q) * G.lanes;
}

// The member's activation when the program computes it in place right
// after the DENSE (jet_gemm_epilogue_op): the interpreter's own jet rules
// for these ops (j_eval), applied to the member's jet before it is stored.
// The primal part -- tanh, exp, sin and cos of c0 -- depends only on c0;
// where c0 is shared by a group of lanes it is formed once per (member,
// group) and passed in as p0, p1 (sin, cos), the same values the rule would
// compute per lane.
__device__ __forceinline__ void epilogue_primal(uint32_t op, float c0, float& p0, float& p1) {
  switch (op) {
    case OP_TANH: p0 = tanhf(c0); break;
    case OP_EXP: p0 = expf(c0); break;
    case OP_SIN:
    case OP_COS: p0 = sinf(c0), p1 = cosf(c0); break;
    default: break;
  }
}
__device__ __forceinline__ void epilogue_tail(uint32_t op, float p0, float p1, float* c) {
  float r[NC];
  switch (op) {
    case OP_TANH: jt::tanh_tail<float, NC>(c, p0, r); break;
    case OP_EXP: jt::exp_tail<float, NC>(c, p0, r); break;
    case OP_SIN: {
      float co[NC];
      jt::sin_cos_tail<float, NC>(c, p0, p1, r, co);
      break;
    }
    case OP_COS: {
      float sn[NC];
      jt::sin_cos_tail<float, NC>(c, p0, p1, sn, r);
      break;
    }
    case OP_NEG:
#pragma unroll
      for (uint32_t q = 0; q < NC; ++q) r[q] = -c[q];
      break;
    default: return;
  }
#pragma unroll
  for (uint32_t q = 0; q < NC; ++q) c[q] = r[q];
}
__device__ __forceinline__ void epilogue_op(uint32_t op, float* c) {
  float p0 = 0.f, p1 = 0.f;
  epilogue_primal(op, c[0], p0, p1);
  epilogue_tail(op, p0, p1, c);
}

// NZ: operand coefficients that can be non-zero (2 when the operands are
// tape inputs).  PT: c0 groups per tile (0: c0 on the matrix cores too).
// DM: operand source (DenseGemm::direct; NZ == 2 only).  NT: lanes per
// workgroup, 64 or 128 (one coefficient on the matrix cores).
// Three waves per SIMD: accumulators in VGPRs, no spills (four spill).
template <uint32_t NZ, uint32_t PT, uint32_t DM, uint32_t NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_dense_gemm(DenseGemm G) {
  static_assert(DM == 0 || NZ == 2, "direct operands are tape inputs");
  static_assert(NT == 64 || (NT == 128 && PT != 0 && NZ == 2), "128 lanes: one coefficient on the matrix cores");
  constexpr uint32_t NB = NT / 64;   // lanes each thread stages per operand row
  constexpr uint32_t NTC = NT / 32;  // 32-lane column blocks: products per wave and step
  constexpr uint32_t Q0 = PT ? 1u : 0u;  // first coefficient on the matrix cores
  constexpr uint32_t NQ = NZ - Q0;
  constexpr uint32_t PTS = PT ? PT : 1u;
  constexpr uint32_t LPP = NT / PTS;  // lanes per c0 group inside the tile
  __shared__ float Bs[NQ][KC][NT];
  __shared__ float As[4][KC][32];  // each wave's weights: [term][member]
  __shared__ float C0s[KC][PTS];   // distinct c0 operands of the stage
  __shared__ float C0r[MT][PTS];   // c0 results
  __shared__ float P0r[PT ? MT : 1][PTS], P1r[PT ? MT : 1][PTS];  // their activation's primal part
  __shared__ float RO[4][NT][NC];  // read-out: each wave's partial sums per lane
  __shared__ float ROB[MT][NC];    // read-out of a recomputed lane: its members in order
  __shared__ uint32_t bad[NTC];    // tile lanes (32 t .. 32 t + 31) with a non-finite operand coefficient
  const uint32_t tid = threadIdx.x, l = tid & 63u, hi = l >> 5, n = l & 31u;
  const uint32_t w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  // Workgroup -> (lane tile, member row).  Workgroups go to the eight XCDs
  // in turn (id mod 8), each XCD with its own L2: ids 8 apart share one, so
  // the member rows of a lane tile are placed 8 ids apart and read the
  // tile's operand rows from the same L2 at about the same time.
  const uint32_t nrow = (G.m + MT - 1) / MT, grp = blockIdx.x / (8u * nrow), rem = blockIdx.x % (8u * nrow);
  const uint32_t by = rem / 8u;                            // member row
  const uint64_t tile = (uint64_t)grp * 8u + rem % 8u;     // lane tile
  if (tile * NT >= G.lanes) return;  // padding of the last group: the whole workgroup leaves
  const uint64_t lane0 = tile * NT;
  const uint32_t m = G.m, T = G.T;
  const uint32_t* PR = G.pool + G.off_regs;
  const float* PW = reinterpret_cast<const float*>(G.pool + G.off_w);
  const uint32_t* PE = G.off_epi != ~0u ? G.pool + G.off_epi : nullptr;  // members' activations
  const float* PRO = G.off_ro != ~0u ? reinterpret_cast<const float*>(G.pool + G.off_ro) : nullptr;
  const uint32_t* PI = G.pool + (DM ? G.off_in : 0u);  // direct: the terms' input indices
  const uint32_t nst = (T + KC - 1) / KC;  // stages
  const uint32_t jw = by * MT + w * 32u;  // this wave's first member
  const uint32_t j = jw + n;                      // the member whose weights this lane supplies
  const bool live = jw < m;
  if (tid < NTC) bad[tid] = 0u;
  f32x16 acc[NQ][NTC];
#pragma unroll
  for (uint32_t q = 0; q < NQ; ++q)
#pragma unroll
    for (uint32_t t = 0; t < NTC; ++t) acc[q][t] = f32x16{};
  float c0acc[PTS];
#pragma unroll
  for (uint32_t t = 0; t < PTS; ++t) c0acc[t] = 0.f;
  // Direct operands: the x rows and draw bases of the lanes this thread
  // stages (l, l + 64), and the x row of the c0 group it loads (lane l <
  // PT), fixed for the whole product.
  bool dact[NB];
  uint64_t dx[NB], dbase[NB];
#pragma unroll
  for (uint32_t b = 0; b < NB; ++b) {
    const uint64_t dlane = lane0 + b * 64u + l;
    dact[b] = DM && dlane < G.B;
    dx[b] = dact[b] ? (G.S ? dlane / G.S : dlane) * G.xstride : 0;
    dbase[b] = DM == 1 ? dlane * G.nin : 0;
  }
  uint64_t dx0 = 0;
  bool dact0 = false;
  if (DM && PT && l < PTS) {
    const uint64_t ln = lane0 + l * LPP;
    dact0 = ln < G.B;
    dx0 = dact0 ? (G.S ? ln / G.S : ln) * G.xstride : 0;
  }
  DirBlock db[NB];  // direct 1: the blocks of these lanes' last signs

  // Stage registers.  Wave w stages terms w RPT .. w RPT + RPT - 1, lane l
  // of each, and its own members' weights for the stage's KC terms.
  // The descriptors travel one stage ahead of the rows: lane r < RPT of the
  // wave holds term w RPT + r's word of the stage after the one loading.
  float sb[RPT][NQ][NB], sc[RPT], sa[KC / 2];
  uint32_t dsc, din = 0u;
  auto load_desc = [&](uint32_t k0) {
    const uint32_t p = k0 + w * RPT + l;
    dsc = (l < RPT && p < T) ? PR[p] : (T_IMM << P_KIND_SHIFT);
    if (DM) din = (l < RPT && p < T) ? PI[p] : 0u;
  };
  auto load_stage = [&](uint32_t k0) {
    // direct 1: the stage's RPT signs of each staged lane first (bit r set:
    // term r draws -1)
    uint32_t sg[NB];
#pragma unroll
    for (uint32_t b = 0; b < NB; ++b) sg[b] = 0u;
    if (DM == 1) {
      // terms k0 + w RPT + r are inputs of the same index (below nseq): the
      // lane's four signs are bits idx0 .. idx0 + 3 of its stream, idx0 a
      // multiple of 4, so within one 32-bit word of one block
      const uint32_t p0 = k0 + w * RPT;
      const uint32_t nv = G.nseq > p0 ? min(G.nseq - p0, RPT) : 0u;
#pragma unroll
      for (uint32_t b = 0; b < NB; ++b) {
        if (!dact[b] || nv == 0u) continue;
        const uint64_t idx0 = dbase[b] + p0, blk = idx0 >> 7;
        if (blk != db[b].blk) direction_fill(RNG_RADEMACHER, G.seed, blk, db[b]);
        const uint32_t bi = (uint32_t)idx0 & 127u;
        const uint32_t lo = (bi & 32u) ? db[b].w1 : db[b].w0, hi2 = (bi & 32u) ? db[b].w3 : db[b].w2;
        sg[b] = ((((bi & 64u) ? hi2 : lo) >> (bi & 31u)) & 0xFu) & ((1u << nv) - 1u);
      }
    }
#pragma unroll
    for (uint32_t r = 0; r < RPT; ++r) {
#pragma unroll
      for (uint32_t q = 0; q < NQ; ++q)
#pragma unroll
        for (uint32_t b = 0; b < NB; ++b) sb[r][q][b] = 0.f;
      sc[r] = 0.f;
      const uint32_t kr = __builtin_amdgcn_readlane(dsc, r);
      if ((kr >> P_KIND_SHIFT) != T_IMM) {
        if (DM) {
          const uint32_t in = __builtin_amdgcn_readlane(din, r);
#pragma unroll
          for (uint32_t b = 0; b < NB; ++b) {
            float d = 0.f;
            if (dact[b]) {
              const uint64_t at = (uint64_t)in * G.B + lane0 + b * 64u + l;
              if (DM == 1) {
                d = ((sg[b] >> r) & 1u) ? -1.f : 1.f;
                if (G.dirs_out && by == 0) G.dirs_out[at] = d;
              } else {
                d = G.dcm[at];
              }
            }
            if (Q0 == 0) {
              sb[r][0][b] = dact[b] ? G.x[dx[b] + in] : 0.f;
              sb[r][NQ - 1][b] = d;
            } else {
              sb[r][0][b] = d;
            }
          }
          if (PT && l < PTS) sc[r] = dact0 ? G.x[dx0 + in] : 0.f;
        } else {
          const uint32_t reg = kr & P_REG_MASK;
#pragma unroll
          for (uint32_t q = 0; q < NQ; ++q)
#pragma unroll
            for (uint32_t b = 0; b < NB; ++b) sb[r][q][b] = row_of(G, reg, Q0 + q)[lane0 + b * 64u + l];
          if (PT && l < PTS) sc[r] = row_of(G, reg, 0)[lane0 + l * LPP];
        }
      }
    }
    // this lane's 8 packed weights of the stage: two 16-byte loads
    if (live) {
      const float4* wp = reinterpret_cast<const float4*>(G.pool + G.off_wp) +
                         (((size_t)(by * (MT / 32u) + w) * nst + k0 / KC) * 64u + l) * 2u;
      const float4 w0 = wp[0], w1 = wp[1];
      sa[0] = w0.x, sa[1] = w0.y, sa[2] = w0.z, sa[3] = w0.w;
      sa[4] = w1.x, sa[5] = w1.y, sa[6] = w1.z, sa[7] = w1.w;
    } else {
#pragma unroll
      for (uint32_t kk = 0; kk < KC / 2; ++kk) sa[kk] = 0.f;
    }
  };
  auto store_stage = [&]() {
#pragma unroll
    for (uint32_t r = 0; r < RPT; ++r) {
      const uint32_t i = w * RPT + r;
#pragma unroll
      for (uint32_t b = 0; b < NB; ++b) {
        bool fin = true;
#pragma unroll
        for (uint32_t q = 0; q < NQ; ++q) {
          Bs[q][i][b * 64u + l] = sb[r][q][b];
          fin = fin && isfinite(sb[r][q][b]);
        }
        if (!fin) atomicOr(&bad[2u * b + hi], 1u << n);
      }
      if (PT && l < PTS) {
        C0s[i][l] = sc[r];
        if (!isfinite(sc[r]))  // the whole group
          for (uint32_t t = l * LPP; t < (l + 1) * LPP; ++t) atomicOr(&bad[t >> 5], 1u << (t & 31u));
      }
    }
#pragma unroll
    for (uint32_t kk = 0; kk < KC / 2; ++kk) As[w][2u * kk + hi][n] = sa[kk];
  };

  load_desc(0);
  load_stage(0);
  load_desc(KC);
  for (uint32_t k0 = 0; k0 < T; k0 += KC) {
    __syncthreads();  // the previous stage's LDS reads are done
    store_stage();
    __syncthreads();
    if (k0 + KC < T) {  // in flight during the products below
      load_stage(k0 + KC);
      load_desc(k0 + 2 * KC);
    }
    if (live) {
#pragma unroll
      for (uint32_t kk = 0; kk < KC / 2; ++kk) {
        const uint32_t k = 2u * kk + hi;
        const float a = As[w][k][n];
#pragma unroll
        for (uint32_t q = 0; q < NQ; ++q)
#pragma unroll
          for (uint32_t t = 0; t < NTC; ++t)
            acc[q][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, Bs[q][k][t * 32u + n], acc[q][t], 0, 0, 0);
        if (PT) {
#pragma unroll
          for (uint32_t t = 0; t < PTS; ++t) c0acc[t] = fmaf(a, C0s[k][t], c0acc[t]);
        }
      }
    }
  }
  if (PT) {
    // the two term parities meet; lanes 0..31 hold member jw + n's c0 per group
#pragma unroll
    for (uint32_t t = 0; t < PTS; ++t) c0acc[t] += __shfl_xor(c0acc[t], 32, 64);
    if (live && hi == 0) {
      float imm = 0.f;
      if (j < m)
        for (uint32_t i = 0; i < G.nimm; ++i) imm += PW[(uint64_t)G.imm[i] * m + j];
      const uint32_t op = (PE && j < m) ? PE[j] : 0u;
#pragma unroll
      for (uint32_t t = 0; t < PTS; ++t) {
        const float c0 = c0acc[t] + imm;
        C0r[w * 32u + n][t] = c0;
        float p0 = 0.f, p1 = 0.f;
        if (op) epilogue_primal(op, c0, p0, p1);
        P0r[w * 32u + n][t] = p0;
        P1r[w * 32u + n][t] = p1;
      }
    }
  }
  __syncthreads();  // bad[] and C0r complete
  uint32_t badm[NTC], anybad = 0u;
#pragma unroll
  for (uint32_t t = 0; t < NTC; ++t) anybad |= (badm[t] = bad[t]);
  float ro[NTC][NC];  // read-out partial sums of this lane's columns
#pragma unroll
  for (uint32_t t = 0; t < NTC; ++t)
#pragma unroll
    for (uint32_t q = 0; q < NC; ++q) ro[t][q] = 0.f;
  if (live) {
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const uint32_t rl = 8u * (v >> 2) + 4u * hi + (v & 3), row = jw + rl;
      if (row >= m) continue;
      float c0[NTC];
      if (PT) {
#pragma unroll
        for (uint32_t t = 0; t < NTC; ++t) c0[t] = C0r[w * 32u + rl][(t * 32u + n) / LPP];
      } else {
#pragma unroll
        for (uint32_t t = 0; t < NTC; ++t) c0[t] = acc[0][t][v];
        for (uint32_t i = 0; i < G.nimm; ++i) {
          const float wi = PW[(uint64_t)G.imm[i] * m + row];
#pragma unroll
          for (uint32_t t = 0; t < NTC; ++t) c0[t] = c0[t] + wi;
        }
      }
      const uint32_t dst = PR[T + row];
      const uint32_t op = PE ? PE[row] : 0u;
#pragma unroll
      for (uint32_t t = 0; t < NTC; ++t) {
        if ((badm[t] >> n) & 1u) continue;  // recomputed below
        const uint64_t ln = lane0 + t * 32u + n;
        float c[NC];
        c[0] = c0[t];
#pragma unroll
        for (uint32_t q = 1; q < NC; ++q) c[q] = (q >= Q0 && q - Q0 < NQ) ? acc[q - Q0][t][v] : 0.f;
        if (op) {
          if (PT) {
            const uint32_t grp = (t * 32u + n) / LPP;
            epilogue_tail(op, P0r[w * 32u + rl][grp], P1r[w * 32u + rl][grp], c);
          } else {
            epilogue_op(op, c);
          }
        }
        if (PRO) {
          const float wr = PRO[row];
#pragma unroll
          for (uint32_t q = 0; q < NC; ++q) ro[t][q] = ro[t][q] + wr * c[q];
        } else {
#pragma unroll
          for (uint32_t q = 0; q < NC; ++q) row_of(G, dst, q)[ln] = c[q];
        }
      }
    }
  }
  if (PRO) {
    // lanes n and n + 32 hold the other 16 rows of the same columns; then
    // the four waves' 32-member sums meet in LDS, in wave order
#pragma unroll
    for (uint32_t t = 0; t < NTC; ++t)
#pragma unroll
      for (uint32_t q = 0; q < NC; ++q) ro[t][q] = ro[t][q] + __shfl_xor(ro[t][q], 32, 64);
    if (hi == 0)
#pragma unroll
      for (uint32_t t = 0; t < NTC; ++t)
#pragma unroll
        for (uint32_t q = 0; q < NC; ++q) RO[w][t * 32u + n][q] = ro[t][q];
    __syncthreads();
    for (uint32_t e = tid; e < NT * NC; e += blockDim.x) {
      const uint32_t col = e / NC, q = e % NC;
      if (!((bad[col >> 5] >> (col & 31u)) & 1u)) {  // recomputed lanes: below
        const float v = ((RO[0][col][q] + RO[1][col][q]) + RO[2][col][q]) + RO[3][col][q];
        row_of(G, PR[T + by * MT], q)[lane0 + col] = v;
      }
    }
  }
  if (anybad == 0u) return;
  // Lanes with a non-finite operand: every member of this workgroup term by
  // term with full Taylor products, in program order, as the interpreter's
  // exact path computes them.
  const uint32_t jend = min(m, (by + 1) * MT);
  for (uint32_t ln = 0; ln < NT; ++ln) {
    if (!((bad[ln >> 5] >> (ln & 31u)) & 1u)) continue;
    if (PRO) __syncthreads();  // the previous lane's ROB reads are done
    for (uint32_t jj = by * MT + tid; jj < jend; jj += blockDim.x) {
      float s[NC] = {0.f, 0.f, 0.f};
      for (uint32_t p = 0; p < T; ++p) {
        const uint32_t kr = PR[p];
        const uint32_t kind = kr >> P_KIND_SHIFT;
        const float c[NC] = {PW[(uint64_t)p * m + jj], 0.f, 0.f};
        float x[NC] = {0.f, 0.f, 0.f}, term[NC];
        if (kind != T_IMM) {
          if (DM) {
            const uint64_t lp = lane0 + ln;
            if (lp < G.B) {
              x[0] = G.x[(G.S ? lp / G.S : lp) * G.xstride + PI[p]];
              x[1] = DM == 1 ? (float)philox_direction(RNG_RADEMACHER, G.seed, lp * G.nin + PI[p])
                             : G.dcm[(uint64_t)PI[p] * G.B + lp];
            }
          } else {
#pragma unroll
            for (uint32_t q = 0; q < NZ; ++q) x[q] = row_of(G, kr & P_REG_MASK, q)[lane0 + ln];
          }
        }
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
      const uint32_t dst = PR[T + jj];
      if (PE && PE[jj]) epilogue_op(PE[jj], s);
      if (PRO) {
#pragma unroll
        for (uint32_t q = 0; q < NC; ++q) ROB[jj - by * MT][q] = s[q];
      } else {
#pragma unroll
        for (uint32_t q = 0; q < NC; ++q) row_of(G, dst, q)[lane0 + ln] = s[q];
      }
    }
    if (PRO) {
      __syncthreads();
      if (tid < NC) {  // the row's members in order: the full Taylor product with [w, 0, 0] each
        float acc = 0.f;
        for (uint32_t jj = by * MT; jj < jend; ++jj) {
          const uint32_t k = jj - by * MT;
          const float cw[NC] = {PRO[jj], 0.f, 0.f};
          float x[NC], term[NC];
#pragma unroll
          for (uint32_t q = 0; q < NC; ++q) x[q] = ROB[k][q];
          jt::mul<float, NC>(cw, x, term);
          acc = (jj == by * MT) ? term[tid] : acc + term[tid];
        }
        row_of(G, PR[T + by * MT], tid)[lane0 + ln] = acc;
      }
    }
  }
}

template <uint32_t NZ, uint32_t DM>
const void* pick(uint32_t pt, uint32_t nt) {
  if (NZ == 2 && nt == 128) {
    switch (pt) {
      case 1: return (const void*)k_dense_gemm<2, 1, DM, 128>;
      case 2: return (const void*)k_dense_gemm<2, 2, DM, 128>;
      case 4: return (const void*)k_dense_gemm<2, 4, DM, 128>;
      case 8: return (const void*)k_dense_gemm<2, 8, DM, 128>;
      case 16: return (const void*)k_dense_gemm<2, 16, DM, 128>;
      default: return nullptr;
    }
  }
  switch (pt) {
    case 1: return (const void*)k_dense_gemm<NZ, 1, DM, 64>;
    case 2: return (const void*)k_dense_gemm<NZ, 2, DM, 64>;
    case 4: return (const void*)k_dense_gemm<NZ, 4, DM, 64>;
    case 8: return (const void*)k_dense_gemm<NZ, 8, DM, 64>;
    case 16: return (const void*)k_dense_gemm<NZ, 16, DM, 64>;
    default: return (const void*)k_dense_gemm<NZ, 0, DM, 64>;
  }
}

// Lane p's directions over inputs [8 y, 8 y + 8): consecutive components
// share their Philox block.
__global__ __launch_bounds__(256) void k_draw_dirs(uint32_t kind, uint64_t seed, uint32_t S, uint64_t B, uint32_t n,
                                                   float* out) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= B) return;
  const uint32_t i0 = blockIdx.y * 8u, i1 = min(n, i0 + 8u);
  DirBlock db;
  for (uint32_t i = i0; i < i1; ++i) out[(uint64_t)i * B + p] = (float)lane_direction(kind, seed, S, p, n, i, db);
}

}  // namespace

hipError_t launch_draw_dirs(uint32_t kind, uint64_t seed, uint32_t S, uint64_t B, uint32_t nin, float* out,
                            hipStream_t s) {
  if (B == 0 || nin == 0) return hipSuccess;
  const uint64_t gx = (B + 255) / 256, gy = (nin + 7) / 8;
  if (gx > 0x7FFFFFFFull || gy > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_draw_dirs, dim3((uint32_t)gx, (uint32_t)gy), dim3(256), 0, s, kind, seed, S, B, nin, out);
  return hipGetLastError();
}

hipError_t launch_dense_gemm(const DenseGemm& G, hipStream_t s) {
  if (G.m == 0 || G.lanes == 0) return hipSuccess;
  if (G.lanes % 64 != 0) return hipErrorInvalidValue;
  // c0 groups per lane tile: groups that divide the tile, or tiles inside one group
  auto groups = [&](uint32_t nt) -> uint32_t {
    if (G.share >= nt && G.share % nt == 0) return 1;
    if (G.share >= 4 && nt % G.share == 0) return nt / G.share;
    return 0;
  };
  // 128-lane tiles when the matrix cores form c1 alone (inputs' c0 shared, c2 zero)
  const bool inputs = (G.flags & D_INPUT_OPERANDS) != 0;
  uint32_t nt = 64, pt = groups(64);
  if (inputs && G.lanes % 128 == 0) {
    const uint32_t p = groups(128);
    if (p >= 1 && p <= 16) nt = 128, pt = p;
  }
  const uint64_t tiles = G.lanes / nt, nrow = (G.m + MT - 1) / MT;
  const uint64_t wgs = (tiles + 7) / 8 * 8 * nrow;  // groups of eight lane tiles
  if (wgs > 0xFFFFFFFFull) return hipErrorInvalidValue;
  const dim3 grid((uint32_t)wgs);
  void* args[] = {const_cast<DenseGemm*>(&G)};
  if (G.direct > 2 || (G.direct && (!(G.flags & D_INPUT_OPERANDS) || G.off_in == ~0u || !G.x))) return hipErrorInvalidValue;
  if ((G.direct == 2 && !G.dcm) || (G.direct == 1 && (G.nseq == 0 || G.nin % 4 != 0))) return hipErrorInvalidValue;
  const void* fn = !inputs            ? pick<3, 0>(pt, nt)
                   : G.direct == 1   ? pick<2, 1>(pt, nt)
                   : G.direct == 2   ? pick<2, 2>(pt, nt)
                                     : pick<2, 0>(pt, nt);
  return hipLaunchKernel(fn, grid, dim3(256), args, 0, s);
}

}  // namespace ehip
