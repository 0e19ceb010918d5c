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
//   DOT / DOT_REV   : w1 dst (or g), w2 term count, w3 first term word;
//                     terms live in the side table (see DotTerm)
constexpr uint32_t NW = 8;

enum : uint32_t {
  I_LOAD = 64,
  I_STORE_OUT = 65,
  I_SEED = 66,
  I_STORE_GRAD = 67,
  I_DOT = 68,
  I_DOT_REV = 69,
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

// DOT side-table entry, 4 x u32:
//   t0 = kind | flags << 8; kind 0 = reg*imm product, 1 = plain reg, 2 = imm
//   t1 = reg (value register for DOT, adjoint register for DOT_REV)
//   t2/t3 = immediate (f32 in t2, f64 in t2/t3)
// DOT_REV terms are stored in the order their adjoint contributions happen
// (descending tape index); flag bit 0 = first contribution (adjoint starts
// at zero).  The DOT_REV chain length (number of Add links between the
// result and each term) lives in t0 bits 16..31 for the Dual-number replay.
constexpr uint32_t TW = 4;
enum : uint32_t { T_PROD = 0, T_REG = 1, T_IMM = 2 };

struct TapeHost {
  std::vector<uint32_t> opcodes, arg0, arg1;
  std::vector<double> values;  // constants (as uploaded, widened)
  uint32_t num_inputs = 0;
  std::vector<uint32_t> outputs;
  bool f64 = false;
};

struct Program {
  std::vector<uint32_t> words;  // NW words per instruction
  std::vector<uint32_t> terms;  // TW words per DOT term
  uint32_t ninstr = 0;
  uint32_t nregs = 0;
  uint64_t dot_terms = 0;
};

enum ProgMode { PM_FORWARD = 0, PM_GRADIENT = 1 };

// Build the program; `fuse` enables DOT fusion.  Never fails on a validated
// tape.
Program compile_program(const TapeHost& t, ProgMode mode, bool fuse);

// GpuTapeData::validate (src/gpu/mod.rs:397-564); returns "" when valid.
std::string validate_tape(const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1,
                          uint32_t num_ops, uint32_t num_inputs, uint32_t num_variables,
                          uint32_t num_constants, uint32_t output_index, const uint32_t* outs,
                          uint32_t num_outputs);

}  // namespace ehip
