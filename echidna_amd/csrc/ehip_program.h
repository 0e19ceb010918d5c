// Register program compiled from a BytecodeTape at upload time.
//
// The reference interprets the tape entry by entry with one slot per entry
// and per point (src/gpu/kernels/tape_eval.cu:86-352).  Here the host turns
// the tape, once, into a straight-line program over a small register file:
//   * Const entries disappear: their values become instruction immediates;
//   * entries that cannot reach an output (forward) or the seeded output
//     (reverse) are dropped — they never influence a result;
//   * the reverse sweep is unrolled into explicit adjoint instructions in the
//     reference's descending order, and only operands the partials read are
//     kept alive;
//   * values and adjoints share registers by linear-scan allocation over
//     their live ranges, so the per-point state is the peak live set, not the
//     tape length;
//   * runs of `acc + x_k * c_k` (constant-weight dot products, the shape of
//     scalar-op MLP tapes) fuse into one DOT instruction with the same
//     per-term rounding and the same adjoint order.
// The interpreter kernels execute these words wave-uniformly.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace ehip {

// Instruction: 8 x u32.
//   w0 = opcode | flags << 8
//   forward tape op : w1 dst, w2/w3 a, w4/w5 b, w7 powi exponent
//   reverse tape op : w1 g (adjoint of the op), w2 a (w3 = adj of a when a
//                     is a register, else immediate high word), w4 b (w5 =
//                     adj of b likewise), w6 r (result value), w7 exponent
//   LOAD            : w1 dst, w2 input index
//   STORE_OUT       : w1 output index, w2/w3 value
//   SEED            : w1 dst (adjoint seeded to one)
//   STORE_GRAD      : w1 input index, w2/w3 adjoint (or immediate zero)
//   DOT / DOT_REV   : w1 dst (or g), w2 term count, w3 first term index
//                     into the side table below
//   LOADN           : w2 count, w3 first side-table index; reg[] holds the
//                     destination registers, imm[] the input indices
//   KINK            : w1 kink index, w2/w3 a, w4/w5 b (Max/Min), w6 opcode
//   DENSE           : w1 m members, w2 T terms, w3 pool offset of the term
//                     words then the m destination registers, w4 pool offset
//                     of the weights [T][m] (term-major); IW words per weight
//   DENSE_REV       : w1 m members, w2 T terms, w3 pool offset of the term
//                     words (adjoint registers) then the m member adjoints
//                     g_j in application order, w4 weights [m][T], w5 flags
// A DENSE replaces m DOT instructions that share one term structure (same
// kinds, same operand registers, all distinct; weights differ): member j
// computes exactly what its DOT computed, term by term in the same order, so
// results are bit-identical (a NaN's sign bit, which IEEE 754 leaves open,
// aside); the operand registers are read once per block of
// members instead of once per member.  DENSE_REV likewise replaces the m
// DOT_REVs, applying member contributions to each adjoint in their original
// order.  Pool term word: kind << 30 | first << 29 | const_left << 28 |
// register.
constexpr uint32_t NW = 8;

enum : uint32_t {
  I_LOAD = 64,
  I_STORE_OUT = 65,
  I_SEED = 66,
  I_STORE_GRAD = 67,
  I_DOT = 68,
  I_DOT_REV = 69,
  I_LOADN = 70,
  I_WEND = 71,   // end of a stream window: continue with the next one
  I_END = 72,    // end of the program
  I_DENSE = 73,
  I_DENSE_REV = 74,
  I_KINK = 75,
  I_REV_BASE = 128,
};

enum : uint32_t {
  F_A_IMM = 1u << 0,
  F_B_IMM = 1u << 1,
  F_FIRST_A = 1u << 2,
  F_FIRST_B = 1u << 3,
  F_ACC_A = 1u << 4,
  F_ACC_B = 1u << 5,
  F_HAS_B = 1u << 6,
  F_RD_A = 1u << 7,
  F_RD_B = 1u << 8,
  F_RD_R = 1u << 9,
};

// DOT side table, stored as three arrays of T entries each (T = total terms):
//   meta[T]  : kind (bits 0-1: 0 reg*imm product, 1 plain reg, 2 imm)
//              | first (bit 2: adjoint starts at zero, DOT_REV)
//              | const_left (bit 3: product written c * x on the tape)
//              | run (bits 4-11: how many consecutive terms from here can be
//                processed as one independent block, capped at 8)
//              | depth (bits 16-31: Add links between the DOT_REV root and
//                the link where the term entered, for the Dual replay)
//   reg[T]   : value register (DOT) or adjoint register (DOT_REV)
//   imm[T*IW]: immediate, IW = 1 word (f32) or 2 words (f64)
// A DOT instruction covers terms [w3, w3 + w2).  DOT_REV terms are stored in
// the order their adjoint contributions happen (descending tape index).
// F_DOT_CONT on a DOT continues an accumulator already in register w1 (a
// chain emitted in several pieces); otherwise term 0 initialises it.
enum : uint32_t { T_PROD = 0, T_REG = 1, T_IMM = 2 };
enum : uint32_t { M_FIRST = 1u << 2, M_CONST_LEFT = 1u << 3 };
constexpr uint32_t M_RUN_SHIFT = 4, M_DEPTH_SHIFT = 16;
constexpr uint32_t F_DOT_CONT = 1u << 10;

struct TapeHost {
  std::vector<uint32_t> opcodes, arg0, arg1;
  std::vector<double> values;  // constants (as uploaded, widened)
  uint32_t num_inputs = 0;
  std::vector<uint32_t> outputs;
  bool f64 = false;
};

constexpr uint32_t P_KIND_SHIFT = 30, P_FIRST = 1u << 29, P_CLEFT = 1u << 28, P_REG_MASK = (1u << 28) - 1;
// DENSE_REV term words of dual programs use bit 28 for P_DEEP instead: the
// term reaches its operand through at least one Add link of its chain (DOT_REV
// depth >= 1), so its adjoint arrives normalised once as 0 + 1 * g.  One
// normalisation is all that matters: applying it again changes no value.
constexpr uint32_t P_DEEP = 1u << 28;
// DENSE_REV w5: every term is a product and no weight is zero (in the
// program's dtype), so only the adjoints can trigger the reference guards.
constexpr uint32_t D_ALL_PROD_NZ = 1u;
// DENSE w5: every operand is a tape input and every weight is finite (in
// the program's dtype).  Jet interpreters seed inputs as (x, v, 0, ..), so
// the operands' coefficients of order >= 2 are +0 there.  w6 then holds the
// pool offset of the T operands' input indices.
constexpr uint32_t D_INPUT_OPERANDS = 2u;
// DENSE w5: every weight (plain register terms: 1) is finite in the
// program's dtype; the f32 jet interpreter may then run the block on FP32
// MFMA (ehip_set_jet_mma).
constexpr uint32_t D_FINITE_WEIGHTS = 4u;

struct Program {
  std::vector<uint32_t> words;  // NW words per instruction
  std::vector<uint32_t> terms;  // meta[T] | reg[T] | imm[T*IW]
  std::vector<uint32_t> pool;   // DENSE / DENSE_REV operand lists and weights
  uint32_t ninstr = 0;
  uint32_t nregs = 0;
  uint64_t dot_terms = 0;
  uint32_t nterms = 0;          // T
};

// PM_KINK: forward values plus one KINK record per nonsmooth entry of the
// whole tape (forward_nonsmooth, src/bytecode_tape/forward.rs:87-187).
enum ProgMode { PM_FORWARD = 0, PM_GRADIENT = 1, PM_KINK = 2 };

inline bool is_nonsmooth_op(uint32_t op) { return op == 10 || op == 11 || (op >= 36 && op <= 42); }

// Windowed instruction stream executed by the interpreters.  The program is
// cut into windows of WIN_WORDS u32 words that a workgroup stages in LDS (two
// buffers: one executing, the next arriving).  Every instruction lies inside
// one window; its first word is op | flags << 8 | length << 24.  Ordinary
// instructions keep their 8 words; DOT / DOT_REV / LOADN carry their terms
// inline after a 3-word header (w1 register, w2 term count):
// (meta, reg, imm[IW]) per term, and a chain longer than a window allows is
// split into continuation pieces.  I_WEND closes a window, I_END the program.
constexpr uint32_t WIN_WORDS = 512;
std::vector<uint32_t> encode_stream(const Program& p, bool f64);
// Instructions [i_begin, i_end) only, as a stream of its own.
std::vector<uint32_t> encode_stream(const Program& p, bool f64, uint32_t i_begin, uint32_t i_end);

// The f32 order-3 jet program cut at every DENSE the matrix cores can run
// (finite weights, at least JET_GEMM_MIN_MEMBERS members): those run as one
// batch GEMM over every lane of the launch (ehip_set_jet_mma), the
// instructions between them on the interpreter, with the whole register
// file in HBM so that the launches hand their registers on.
constexpr uint32_t JET_GEMM_MIN_MEMBERS = 32;
// Unary ops a DENSE's batch GEMM applies to its members in its epilogue
// when the program computes them in place right after the DENSE (the
// activation of an MLP layer): neg 12, exp 17, sin 24, cos 25, tanh 32.
inline bool jet_gemm_epilogue_op(uint32_t op) { return op == 12 || op == 17 || op == 24 || op == 25 || op == 32; }
struct JetSegment {
  bool gemm;
  uint32_t i0, i1;                         // interpreter: instructions [i0, i1)
  uint32_t m, T, off_regs, off_w, flags;   // gemm: the DENSE's words 1..5
};
std::vector<JetSegment> segment_jet_program(const Program& p);

// Build the program; `fuse` enables DOT fusion, `dense` the grouping of
// fused dot products into DENSE / DENSE_REV.  Never fails on a validated tape.
// `dual`: the program runs on the dual (HVP) interpreter, whose DENSE_REV
// also needs each term's P_DEEP flag.
// `dense_chunk`: members per forward DENSE (0: the default 16; a whole
// layer for the segmented jet program).
Program compile_program(const TapeHost& t, ProgMode mode, bool fuse, bool dense = true, bool dual = false,
                        size_t dense_chunk = 0);

// GpuTapeData::validate (src/gpu/mod.rs:397-564); returns "" when valid.
std::string validate_tape(const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1,
                          uint32_t num_ops, uint32_t num_inputs, uint32_t num_variables,
                          uint32_t num_constants, uint32_t output_index, const uint32_t* outs,
                          uint32_t num_outputs);

}  // namespace ehip
