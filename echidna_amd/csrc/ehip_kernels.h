// Kernel launch interface shared by the host API and the HIP interpreters.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ehip_program.h"

namespace ehip {

struct Launch {
  const uint32_t* prog;   // windowed instruction stream (device, encode_stream)
  const uint32_t* pool;   // DENSE operand lists and weights (device)
  uint32_t nwin;          // windows of WIN_WORDS words in the stream
  uint32_t R_lds;         // registers held in LDS
  uint32_t nin, nout;
  uint64_t B;             // points (lanes) in this launch
  const void* x;          // primal inputs, row stride xstride (0 = shared row)
  uint64_t xstride;
  const void* dirs;       // tangent / direction seeds [B][nin]
  void* y;                // outputs
  void* g;                // gradients
  void* h;                // Hessian-vector products
  void* gregs;            // HBM register file for registers >= R_lds
  void* ksw;              // kink switching values [B][nkinks] (KINK program)
  int8_t* kbr;            // kink branches [B][nkinks]
  uint32_t nkinks;
  uint32_t S;             // lanes per x row (0: one row per lane)
  uint32_t rng;           // RNG_NONE: read dirs; else generate (philox.h)
  uint64_t seed;
  void* dirs_out;         // jets drawing directions also write them here, coordinate-major
                          // [nin][B] (lane p, coordinate i at i * B + p: a wave's writes are
                          // contiguous), or null
};

// One DENSE of the segmented f32 order-3 jet program as a batch GEMM over
// every lane (ehip_jetgemm.hip).  The register file is [reg][3][lanes]
// floats, every register in HBM.
struct DenseGemm {
  const uint32_t* pool;     // the segmented program's DENSE pool (device)
  uint32_t off_regs, off_w;  // term words then member registers; weights [T][m]
  uint32_t m, T, flags;
  uint32_t nimm;            // immediate terms (at most 4) and their term indices
  uint32_t imm[4];
  float* regs;
  uint64_t lanes;           // lanes per register row (a multiple of 64)
  uint32_t share;           // c0 rows repeat over groups of this many lanes (0: they need not)
  uint32_t off_epi;         // pool offset of m epilogue ops (jet_gemm_epilogue_op, 0: none), or ~0u
  // Read-out: the DOT after the DENSE (and its activations) sums w_j x_j over
  // every member, which nothing else reads.  The GEMM then stores, instead of
  // its members, one partial sum per workgroup row of members (members
  // [y JET_GEMM_MT, (y + 1) JET_GEMM_MT)) into the register of that row's
  // first member, and the program's DOT adds the partials.
  uint32_t off_ro;          // pool offset of the m read-out weights (f32 bits), or ~0u
  // Direct operands: every operand is a tape input (jet (x_i, v_i, 0)), and
  // the GEMM reads x and the lane's direction itself instead of the input
  // registers, which the program then need not write.
  uint32_t off_in;          // pool offset of the T terms' input indices, or ~0u (no direct form)
  uint32_t direct;          // per launch: 0 operand registers; 1 x, Rademacher signs drawn here
                            // (needs nseq); 2 x, directions from dcm
  const float* x;           // primal rows: lane p reads row (S ? p / S : p), stride xstride
  uint64_t xstride;
  uint32_t S, nin;
  uint64_t B;               // lanes in the batch (lanes at or past B read zeros)
  uint64_t seed;            // direct 1: the directions' Philox seed (philox.h)
  const float* dcm;         // direct 2: directions, coordinate-major [nin][B]
  float* dirs_out;          // direct 1: the drawn signs written coordinate-major [nin][B], or null
  // Weights packed for the matrix cores (pack_jet_gemm_weights): per block
  // of 32 members, per stage of JET_GEMM_KC terms, per lane l of a wave the
  // 8 weights it supplies (member 32 b + l % 32, terms 2 kk + l / 32), zero
  // past T and m.  Pool offset, a multiple of 4 words.
  uint32_t off_wp;
  // Terms [0, nseq) are the inputs 0 .. nseq - 1 in order and the rest
  // immediates, nin a multiple of 4 (0: not so): direct 1 takes a lane's
  // four signs of a stage as one 4-bit field of its Philox block.
  uint32_t nseq;
};

constexpr uint32_t JET_GEMM_MT = 128;  // members per workgroup of the jet batch GEMM
constexpr uint32_t JET_GEMM_KC = 16;   // terms per stage of the jet batch GEMM

// The packed weights (DenseGemm::off_wp) of a DENSE with T terms and m
// members whose weights are W[p * m + j] (f32 bits).
inline void pack_jet_gemm_weights(const uint32_t* W, uint32_t T, uint32_t m, uint32_t* out) {
  const uint32_t nst = (T + JET_GEMM_KC - 1) / JET_GEMM_KC, nmb = (m + 31) / 32;
  for (uint32_t jb = 0; jb < nmb; ++jb)
    for (uint32_t s = 0; s < nst; ++s)
      for (uint32_t l = 0; l < 64; ++l)
        for (uint32_t kk = 0; kk < JET_GEMM_KC / 2; ++kk) {
          const uint32_t p = s * JET_GEMM_KC + 2 * kk + l / 32, j = jb * 32 + l % 32;
          out[(((size_t)jb * nst + s) * 64 + l) * (JET_GEMM_KC / 2) + kk] = (p < T && j < m) ? W[(size_t)p * m + j] : 0u;
        }
}
inline size_t jet_gemm_packed_words(uint32_t T, uint32_t m) {
  return (size_t)((m + 31) / 32) * ((T + JET_GEMM_KC - 1) / JET_GEMM_KC) * 64 * (JET_GEMM_KC / 2);
}

hipError_t launch_dense_gemm(const DenseGemm& G, hipStream_t s);

// Directions of lanes [0, B) drawn on the device (philox.h lane_direction,
// rounded to f32), written coordinate-major: out[i * B + p].
hipError_t launch_draw_dirs(uint32_t kind, uint64_t seed, uint32_t S, uint64_t B, uint32_t nin, float* out,
                            hipStream_t s);

// Per-point STDE reduction over S samples (f64 Welford/Chan).
struct StdeReduce {
  const void* jets;       // [npoints*S][3] interleaved c0,c1,c2 (scalar output)
  uint64_t npoints;
  uint32_t S, nin;
  uint32_t rng;           // how directions were made (for control variates)
  uint64_t seed;
  const void* dirs;       // [npoints*S][nin] when rng == RNG_NONE (user directions)
  const void* dirs_cmajor;  // or the jets' own directions, coordinate-major [nin][npoints*S]
  const void* diag;       // diagonal control variate [.. ][nin] or null
  uint64_t diag_stride;   // 0: one diagonal shared by all points
  double* out;            // [npoints][5]: value, estimate, variance, std error, count
  const void* lcv;        // learned control variate L(x) [npoints][nin*rank] or null: bfloat16 with
                          // dirs_cmajor, else float
  uint32_t rank;          // r <= 32
  uint32_t npos;          // columns j < npos enter with +1, the rest with -1
  double* out_plain;      // plain-estimator statistics of the same samples, or null
  double sample_scale;    // samples are sample_scale * 2 c2 (0 means 1)
};

hipError_t launch_stde_reduce(bool f64, const StdeReduce& R, hipStream_t s);

// Persistent multi-tape queue: tile `it` evaluates points
// [first, first + workgroup size) of job `job`.
struct QueueTile {
  uint32_t job;
  uint32_t pad;
  uint64_t first;
};

struct Queue {
  const Launch* jobs;      // device array; each job's B masks its last tile
  const QueueTile* tiles;  // device array, in the order workgroups take them
  uint32_t ntiles;
  uint32_t* counter;       // next tile; cleared by launch_queue
};

// The interpreters are compiled twice, once per program-fetch mode
// (ehip_kernels.hip with EHIP_SFETCH 0 / 1): lds_fetch stages instruction
// windows in LDS per workgroup, scalar_fetch reads the stream through the
// scalar cache.  `lds` is the register file's LDS; the launchers add the
// window buffers where the mode needs them.
#define EHIP_FETCH_LAUNCHERS                                                                                \
  const void* queue_kernel(bool f64);                                                                      \
  size_t queue_lds_bytes(size_t reg_lds);                                                                  \
  hipError_t launch_queue(bool f64, const Queue& Q, uint32_t grid, uint32_t wg, size_t lds, hipStream_t s); \
  hipError_t launch_scalar(bool f64, const Launch& L, uint32_t wg, size_t lds, hipStream_t s);             \
  hipError_t launch_dual(bool f64, const Launch& L, uint32_t wg, size_t lds, hipStream_t s);               \
  hipError_t launch_jet(bool f64, int K, const Launch& L, uint32_t wg, size_t lds, hipStream_t s);
namespace lds_fetch {
EHIP_FETCH_LAUNCHERS
}
namespace scalar_fetch {
EHIP_FETCH_LAUNCHERS
}
#undef EHIP_FETCH_LAUNCHERS

// LDS the instruction windows take beside the register file.
inline size_t window_lds(bool sfetch) { return sfetch ? 0 : 2 * (size_t)WIN_WORDS * 4; }

inline const void* queue_kernel(bool f64, bool sfetch) {
  return sfetch ? scalar_fetch::queue_kernel(f64) : lds_fetch::queue_kernel(f64);
}
inline size_t queue_lds_bytes(bool sfetch, size_t reg_lds) {
  return sfetch ? scalar_fetch::queue_lds_bytes(reg_lds) : lds_fetch::queue_lds_bytes(reg_lds);
}
inline hipError_t launch_queue(bool f64, bool sfetch, const Queue& Q, uint32_t grid, uint32_t wg, size_t lds,
                               hipStream_t s) {
  return sfetch ? scalar_fetch::launch_queue(f64, Q, grid, wg, lds, s) : lds_fetch::launch_queue(f64, Q, grid, wg, lds, s);
}
inline hipError_t launch_scalar(bool f64, bool sfetch, const Launch& L, uint32_t wg, size_t lds, hipStream_t s) {
  return sfetch ? scalar_fetch::launch_scalar(f64, L, wg, lds, s) : lds_fetch::launch_scalar(f64, L, wg, lds, s);
}
inline hipError_t launch_dual(bool f64, bool sfetch, const Launch& L, uint32_t wg, size_t lds, hipStream_t s) {
  return sfetch ? scalar_fetch::launch_dual(f64, L, wg, lds, s) : lds_fetch::launch_dual(f64, L, wg, lds, s);
}
inline hipError_t launch_jet(bool f64, bool sfetch, int K, const Launch& L, uint32_t wg, size_t lds, hipStream_t s) {
  return sfetch ? scalar_fetch::launch_jet(f64, K, L, wg, lds, s) : lds_fetch::launch_jet(f64, K, L, wg, lds, s);
}

}  // namespace ehip
