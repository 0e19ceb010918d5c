// Host-side internals shared by the C ABI translation units: status and
// error text, device buffers, the context and tape handles, program upload
// and the interpreter launch.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/echidna_hip.h"
#include "ehip_kernels.h"
#include "ehip_program.h"
#include "philox.h"

using namespace ehip;

namespace ehip_detail {

inline thread_local std::string g_last_error;

inline int32_t fail(int32_t code, const char* fmt, ...) noexcept {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  // the message must not be able to throw: with no memory for its text the
  // status alone reports the failure
  try {
    g_last_error = buf;
  } catch (...) {
    g_last_error.clear();
  }
  return code;
}

// Every extern "C" entry point that can allocate is a function-try-block
// whose handler returns this: no C++ exception crosses the boundary.
// std::bad_alloc (host memory exhausted while compiling a program, building
// a kernel's source or parsing a wire payload) is GpuError::OutOfMemory;
// anything else is GpuError::Other with the exception's text.
inline int32_t exception_status() noexcept {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    return fail(EHIP_OUT_OF_MEMORY, "GPU out of memory (host allocation failed)");
  } catch (const std::exception& e) {
    return fail(EHIP_OTHER, "GPU error: %s", e.what());
  } catch (...) {
    return fail(EHIP_OTHER, "GPU error: unknown exception");
  }
}

inline int32_t ok() {
  g_last_error.clear();
  return EHIP_OK;
}

inline int32_t hip_fail(hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory) return fail(EHIP_OUT_OF_MEMORY, "GPU out of memory (%s)", what);
  return fail(EHIP_OTHER, "GPU error: %s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                     \
  do {                                          \
    hipError_t _e = (expr);                     \
    if (_e != hipSuccess) return hip_fail(_e, what); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  hipError_t reserve(size_t bytes) {
    if (bytes <= n) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) n = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace ehip_detail

using namespace ehip_detail;

struct ehip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;       // guards device state and knobs
  std::mutex host_mu;  // serializes host-pointer calls (staging buffers)
  DevBuf regs;                 // HBM register file
  DevBuf io[6];                // staging for host-pointer entry points
  DevBuf jets;                 // STDE per-sample jets
  DevBuf qjobs, qtiles, qcount;  // multi-tape queue tables
  DevBuf tup[4];               // control-variate tuple scratch
  DevBuf cvx, cvl;             // control-variate inputs and factors
  DevBuf cvdirs;               // control-variate pass: the samples' directions
  uint32_t ncu = 256;
  // untuned launches: 256-lane workgroups and 24 KB of register LDS, the
  // shape the autotuner picks for the MLP fwd+grad and order-3 jet programs
  uint32_t wg = 256;
  uint32_t lds_budget = 24576;
  uint32_t fetch = 0;  // program fetch of untuned launches and the queue (0 LDS, 1 scalar)
  uint32_t spec_max = 0, spec_mask = 0;  // specialise uploads of at most spec_max entries (ehip_set_specialize)
  bool fuse = true;
  bool dense = true;
  bool jet_mma = false;  // f32 order-3 jets: DENSE blocks on FP32 MFMA (ehip_set_jet_mma)
  size_t max_lds = 65536;
};

struct DevProgram {
  Program host;
  uint32_t* d_stream = nullptr;  // windowed stream (encode_stream)
  uint32_t* d_pool = nullptr;    // DENSE operand lists and weights
  uint32_t nwin = 0;
};

// Launch shape chosen by the autotuner for one evaluation kind of one tape.
struct Tuned {
  uint32_t wg = 0;          // 0: not tuned, use the context defaults
  uint32_t lds_budget = 0;  // bytes of LDS for the register file
  uint32_t fetch = 0;       // program fetch: 0 LDS windows, 1 scalar cache, 2 tape-specialised kernel
  uint32_t ppl = 0;         // points per lane of the tape-specialised kernel (0: one)
};

// A tape-specialised kernel (straight-line code compiled for one tape by
// hipRTC, ehip_spec.cpp): ehip_spec(X, xstride, Y, G, B) for values and
// gradients.
struct SpecKernel {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  uint32_t vgprs = 0, scratch = 0, code_bytes = 0;
  double compile_ms = 0;
  bool dual = false;  // HVP kernel: ehip_spec(X, xstride, V, G, H, B)
  uint32_t order = 0;  // Taylor kernel of this order: ehip_spec(X, xstride, V, Y, B, S, rng, seed)
  uint32_t waves = 0;  // occupancy target it was compiled for (0: the compiler's choice)
};

// Longest tape the specialiser accepts (compile time and code size grow
// with the tape).
constexpr uint32_t SPEC_MAX_ENTRIES = 1u << 16;

// ehip_tape_specialize `what` of the value-and-gradient kernel of an f32
// tape whose adjoints are accumulated in binary64 (it serves the gradient
// evaluations, slot 1).
constexpr int32_t SPEC_GRAD_ACC64 = 8;

namespace ehip_detail {
int32_t noalloc_setup(int device);  // ehip_noalloc.hip: the fixed scratch's address on `device` (current)
hipError_t spec_launch(const SpecKernel& s, const Launch& L, uint32_t wg, uint32_t ppl, hipStream_t st);
void spec_free(SpecKernel& s);
}  // namespace ehip_detail

// The f32 order-3 jet program cut for the matrix cores (ehip_set_jet_mma):
// interpreter streams and batch-GEMM DENSE blocks in program order.
struct JetSegDev {
  bool gemm = false;
  uint32_t* d_stream = nullptr;  // interpreter segment
  uint32_t nwin = 0;
  // The same segment in the direct plan (GEMMs read x and device-drawn
  // directions themselves, the LOADN loads only inputs others read): its
  // own stream when it holds that LOADN, or nothing left to run.
  uint32_t* d_stream_direct = nullptr;
  uint32_t nwin_direct = 0;
  bool skip_direct = false;
  DenseGemm g{};                 // gemm segment (regs and lanes set per launch)
};
struct JetSegProgram {
  Program host;                  // the forward program with whole-layer DENSE groups
  std::vector<JetSegDev> segs;   // empty: no DENSE the matrix cores can run
  uint32_t* d_pool = nullptr;
  uint32_t direct_gemms = 0;     // GEMMs with a direct form (0: no direct plan)
  uint32_t direct_dropped = 0;   // inputs the direct plan's LOADN leaves out
  bool direct_seq = false;       // every direct GEMM reads the inputs in order (DenseGemm::nseq)
};

struct ehip_tape {
  ehip_ctx* ctx = nullptr;
  bool f64 = false;
  uint32_t nin = 0, nout = 0;
  TapeHost host;
  DevProgram prog[4];  // forward, gradient, HVP (dual interpreter), kinks; all with DENSE groups
  uint32_t nkinks = 0;
  Tuned tuned[8];  // forward, gradient, hvp, taylor order 1..5
  SpecKernel spec[8];  // tape-specialised kernels, once compiled, indexed like `tuned`
  std::vector<SpecKernel> retired;  // kernels replaced by a later build, unloaded with the tape
  std::string spec_error[8];  // why an upload-time specialisation failed (the interpreters run instead)
  JetSegProgram jseg;         // f32 tapes: order-3 jets with DENSE blocks on the matrix cores
};

namespace ehip_detail {

inline int32_t upload_program(DevProgram& dp, bool f64) {
  const std::vector<uint32_t> st = encode_stream(dp.host, f64);
  dp.nwin = (uint32_t)(st.size() / WIN_WORDS);
  HIP_TRY(hipMalloc((void**)&dp.d_stream, st.size() * 4), "program upload");
  HIP_TRY(hipMemcpy(dp.d_stream, st.data(), st.size() * 4, hipMemcpyHostToDevice), "program upload");
  const std::vector<uint32_t>& pool = dp.host.pool;
  HIP_TRY(hipMalloc((void**)&dp.d_pool, std::max<size_t>(pool.size(), 4) * 4), "program upload");
  if (!pool.empty())
    HIP_TRY(hipMemcpy(dp.d_pool, pool.data(), pool.size() * 4, hipMemcpyHostToDevice), "program upload");
  return EHIP_OK;
}

inline void free_program(DevProgram& dp) {
  if (dp.d_stream) (void)hipFree(dp.d_stream);
  if (dp.d_pool) (void)hipFree(dp.d_pool);
  dp.d_stream = nullptr;
  dp.d_pool = nullptr;
}

inline void free_jet_segments(JetSegProgram& js) {
  for (auto& sg : js.segs) {
    if (sg.d_stream) (void)hipFree(sg.d_stream);
    if (sg.d_stream_direct) (void)hipFree(sg.d_stream_direct);
  }
  if (js.d_pool) (void)hipFree(js.d_pool);
  js.segs.clear();
  js.d_pool = nullptr;
  js.direct_gemms = js.direct_dropped = 0;
  js.direct_seq = false;
}

// The registers instruction i of a forward program reads (rd) and then
// writes (wr), in that order.  False for an instruction it does not know.
template <class RD, class WR>
inline bool visit_registers(const Program& p, uint32_t i, RD rd, WR wr) {
  const uint32_t T = p.nterms;
  const uint32_t* w = &p.words[(size_t)i * NW];
  const uint32_t op = w[0] & 0xFFu, fl = w[0] >> 8;
  if (op < 64u) {
    if (!(fl & F_A_IMM)) rd(w[2]);
    if ((fl & F_HAS_B) && !(fl & F_B_IMM)) rd(w[4]);
    wr(w[1]);
  } else if (op == I_DOT) {
    for (uint32_t k = 0; k < w[2]; ++k)
      if ((p.terms[w[3] + k] & 3u) != T_IMM) rd(p.terms[T + w[3] + k]);
    if (fl & F_DOT_CONT) rd(w[1]);
    wr(w[1]);
  } else if (op == I_DENSE) {
    for (uint32_t k = 0; k < w[2]; ++k) {
      const uint32_t tw = p.pool[w[3] + k];
      if ((tw >> P_KIND_SHIFT) != T_IMM) rd(tw & P_REG_MASK);
    }
    for (uint32_t j = 0; j < w[1]; ++j) wr(p.pool[w[3] + w[2] + j]);
  } else if (op == I_LOADN) {
    for (uint32_t k = 0; k < w[2]; ++k) wr(p.terms[T + w[3] + k]);
  } else if (op == I_STORE_OUT) {
    if (!(fl & F_A_IMM)) rd(w[2]);
  } else {
    return false;
  }
  return true;
}

// Read-out fusion check: true when no instruction from `from` on reads any
// of the registers `regs` before overwriting it.
inline bool registers_dead_from(const Program& p, uint32_t from, const std::vector<uint32_t>& regs) {
  std::vector<char> live(p.nregs + 1, 0);
  for (uint32_t r : regs)
    if (r < live.size()) live[r] = 1;
  for (uint32_t i = from; i < p.ninstr; ++i) {
    bool read = false;
    const bool known = visit_registers(
        p, i, [&](uint32_t r) { read = read || (r < live.size() && live[r]); },
        [&](uint32_t r) { if (r < live.size()) live[r] = 0; });
    if (!known || read) return false;
  }
  return true;
}

// Compile and upload the segmented order-3 jet program of an f32 tape: the
// forward program with each layer's members in one DENSE, cut at every
// DENSE the matrix cores can run (finite weights, at least 32 members, at
// most four immediate terms).  The in-place unary ops the compiler placed
// right after such a DENSE (its layer's activations) move into the GEMM's
// epilogue.  A tape without such a DENSE keeps js.segs empty.
inline int32_t upload_jet_segments(JetSegProgram& js, const TapeHost& h, bool fuse, bool dense) {
  js.segs.clear();
  if (h.f64 || !fuse || !dense) return EHIP_OK;
  js.host = compile_program(h, PM_FORWARD, fuse, dense, false, SIZE_MAX);
  std::vector<JetSegment> cut = segment_jet_program(js.host);
  std::vector<uint32_t>& pool = js.host.pool;
  auto immediates = [&](const JetSegment& c, uint32_t* idx) {
    uint32_t nimm = 0;
    for (uint32_t p = 0; p < c.T && nimm <= 4; ++p)
      if ((pool[c.off_regs + p] >> P_KIND_SHIFT) == T_IMM) {
        if (nimm < 4 && idx) idx[nimm] = p;
        ++nimm;
      }
    return nimm;
  };
  // per GEMM: the epilogue ops' pool offset and the first instruction after them
  std::vector<uint32_t> epi_off(cut.size(), ~0u), ro_off(cut.size(), ~0u), next(cut.size(), 0);
  bool any = false;
  for (size_t ci = 0; ci < cut.size(); ++ci) {
    const JetSegment& c = cut[ci];
    if (!c.gemm || immediates(c, nullptr) > 4) continue;
    any = true;
    std::vector<uint32_t> member_reg(c.m), ops(c.m, 0);
    for (uint32_t k = 0; k < c.m; ++k) member_reg[k] = pool[c.off_regs + c.T + k];
    std::vector<uint32_t> order(c.m);
    for (uint32_t k = 0; k < c.m; ++k) order[k] = k;
    std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return member_reg[x] < member_reg[y]; });
    uint32_t i = c.i1;
    bool taken = false;
    for (; i < js.host.ninstr; ++i) {
      const uint32_t* w = &js.host.words[(size_t)i * NW];
      const uint32_t op = w[0] & 0xFFu, fl = w[0] >> 8;
      if (op >= 64u || !jet_gemm_epilogue_op(op) || (fl & (F_A_IMM | F_HAS_B)) || w[1] != w[2]) break;
      auto it = std::lower_bound(order.begin(), order.end(), w[2],
                                 [&](uint32_t k, uint32_t r) { return member_reg[k] < r; });
      if (it == order.end() || member_reg[*it] != w[2] || ops[*it]) break;
      ops[*it] = op;
      taken = true;
    }
    next[ci] = i;
    if (taken) {
      epi_off[ci] = (uint32_t)pool.size();
      pool.insert(pool.end(), ops.begin(), ops.end());
    }
    // Read-out: the next instruction is a DOT whose products take every
    // member once (finite constant weights) and nothing later reads a
    // member.  The GEMM sums each workgroup row's members into the register
    // of the row's first member, and the DOT keeps its other terms (a
    // continued chain, immediates, other registers) after those partials.
    if (i < js.host.ninstr) {
      uint32_t* w = &js.host.words[(size_t)i * NW];
      const uint32_t Tn = js.host.nterms, n = w[2], first = w[3];
      if ((w[0] & 0xFFu) == I_DOT) {
        std::vector<uint32_t> wt(c.m, 0u), keep_meta, keep_reg, keep_bits;
        std::vector<char> seen(c.m, 0);
        uint32_t nprod = 0;
        bool ok = true;
        for (uint32_t k = 0; k < n && ok; ++k) {
          const uint32_t meta = js.host.terms[first + k], kind = meta & 3u;
          const uint32_t reg = js.host.terms[Tn + first + k], bits = js.host.terms[2 * Tn + first + k];
          auto it = std::lower_bound(order.begin(), order.end(), reg,
                                     [&](uint32_t x, uint32_t r) { return member_reg[x] < r; });
          const bool member = kind != T_IMM && it != order.end() && member_reg[*it] == reg;
          if (!member) {
            keep_meta.push_back(meta & ~(0xFFu << M_RUN_SHIFT));
            keep_reg.push_back(reg);
            keep_bits.push_back(bits);
            continue;
          }
          float wf;
          std::memcpy(&wf, &bits, 4);
          if (kind != T_PROD || seen[*it] || !std::isfinite(wf)) ok = false;
          else {
            seen[*it] = 1;
            wt[*it] = bits;
            ++nprod;
          }
        }
        const uint32_t tiles = (c.m + JET_GEMM_MT - 1) / JET_GEMM_MT;
        if (ok && nprod == c.m && tiles + keep_meta.size() <= n &&
            registers_dead_from(js.host, i + 1, member_reg)) {
          ro_off[ci] = (uint32_t)pool.size();
          pool.insert(pool.end(), wt.begin(), wt.end());
          uint32_t k = 0;
          for (uint32_t y = 0; y < tiles; ++y, ++k) {
            js.host.terms[first + k] = T_REG | (1u << M_RUN_SHIFT);
            js.host.terms[Tn + first + k] = member_reg[y * JET_GEMM_MT];
            js.host.terms[2 * Tn + first + k] = 0u;
          }
          for (size_t q = 0; q < keep_meta.size(); ++q, ++k) {
            js.host.terms[first + k] = keep_meta[q] | (1u << M_RUN_SHIFT);
            js.host.terms[Tn + first + k] = keep_reg[q];
            js.host.terms[2 * Tn + first + k] = keep_bits[q];
          }
          w[2] = k;
        }
      }
    }
  }
  if (!any) return EHIP_OK;
  // Direct plan.  A GEMM whose operands are all inputs the program's LOADN
  // loads (and nothing rewrites before it) can read x and device-drawn
  // directions itself; the LOADN of that plan then loads only the inputs
  // some other instruction reads.
  std::vector<uint32_t> in_off(cut.size(), ~0u);
  Program dprog;  // the direct plan's program: the same, with that LOADN
  uint32_t li = ~0u, dkept = 0;
  for (uint32_t i = 0; i < js.host.ninstr && li == ~0u; ++i)
    if ((js.host.words[(size_t)i * NW] & 0xFFu) == I_LOADN) li = i;
  js.direct_gemms = js.direct_dropped = 0;
  js.direct_seq = false;
  if (li != ~0u) {
    const uint32_t Tn = js.host.nterms, nl = js.host.words[(size_t)li * NW + 2], lo = js.host.words[(size_t)li * NW + 3];
    std::unordered_map<uint32_t, uint32_t> in_of;  // register -> input index
    for (uint32_t k = 0; k < nl; ++k) in_of[js.host.terms[Tn + lo + k]] = js.host.terms[2 * Tn + lo + k];
    std::vector<char> is_direct(js.host.ninstr, 0);
    for (size_t ci = 0; ci < cut.size(); ++ci) {
      const JetSegment& c = cut[ci];
      if (!c.gemm || !(c.flags & D_INPUT_OPERANDS) || immediates(c, nullptr) > 4 || c.i0 <= li) continue;
      std::vector<uint32_t> idx(c.T, 0u);
      std::unordered_map<uint32_t, char> ops;
      bool ok = true;
      for (uint32_t p = 0; p < c.T && ok; ++p) {
        const uint32_t tw = pool[c.off_regs + p];
        if ((tw >> P_KIND_SHIFT) == T_IMM) continue;
        auto it = in_of.find(tw & P_REG_MASK);
        if (it == in_of.end()) ok = false;
        else idx[p] = it->second, ops[tw & P_REG_MASK] = 1;
      }
      for (uint32_t i = li + 1; i < c.i0 && ok; ++i)
        ok = visit_registers(js.host, i, [](uint32_t) {}, [&](uint32_t r) { if (ops.count(r)) ok = false; });
      if (!ok) continue;
      in_off[ci] = (uint32_t)pool.size();
      pool.insert(pool.end(), idx.begin(), idx.end());
      is_direct[c.i0] = 1;
      ++js.direct_gemms;
    }
    if (js.direct_gemms) {
      // inputs read by an instruction other than a direct GEMM
      std::unordered_map<uint32_t, char> live, needed;
      for (auto& e : in_of) live[e.first] = 1;
      for (uint32_t i = li + 1; i < js.host.ninstr; ++i) {
        auto wr = [&](uint32_t r) { live.erase(r); };
        bool known;
        if (is_direct[i]) known = visit_registers(js.host, i, [](uint32_t) {}, wr);
        else known = visit_registers(js.host, i, [&](uint32_t r) { if (live.count(r)) needed[r] = 1; }, wr);
        if (!known) {
          for (auto& e : live) needed[e.first] = 1;
          break;
        }
      }
      dprog = js.host;
      const uint32_t IW = 1;  // f32 immediates: one word per term
      for (uint32_t k = 0; k < nl; ++k) {
        const uint32_t reg = js.host.terms[Tn + lo + k];
        if (!needed.count(reg)) continue;
        dprog.terms[lo + dkept] = js.host.terms[lo + k];
        dprog.terms[Tn + lo + dkept] = reg;
        dprog.terms[2 * Tn + (size_t)(lo + dkept) * IW] = js.host.terms[2 * Tn + (size_t)(lo + k) * IW];
        ++dkept;
      }
      for (uint32_t k = 0; k < dkept; ++k)  // runs of up to 8 terms, as the compiler sets them for a LOADN
        dprog.terms[lo + k] = (dprog.terms[lo + k] & ~(0xFFu << M_RUN_SHIFT)) | (std::min(8u, dkept - k) << M_RUN_SHIFT);
      dprog.words[(size_t)li * NW + 2] = dkept;
      js.direct_dropped = nl - dkept;
    }
  }
  // Packed weights of every GEMM (16-byte aligned), and the sequential
  // operand count of the direct ones (DenseGemm::nseq).
  std::vector<uint32_t> wp_off(cut.size(), 0u), nseq(cut.size(), 0u);
  for (size_t ci = 0; ci < cut.size(); ++ci) {
    const JetSegment& c = cut[ci];
    if (!c.gemm || immediates(c, nullptr) > 4) continue;
    pool.resize((pool.size() + 3) / 4 * 4, 0u);
    wp_off[ci] = (uint32_t)pool.size();
    pool.resize(pool.size() + jet_gemm_packed_words(c.T, c.m), 0u);
    pack_jet_gemm_weights(pool.data() + c.off_w, c.T, c.m, pool.data() + wp_off[ci]);
    if (in_off[ci] != ~0u && h.num_inputs % 4 == 0) {
      uint32_t q = 0;
      while (q < c.T && (pool[c.off_regs + q] >> P_KIND_SHIFT) != T_IMM && pool[in_off[ci] + q] == q) ++q;
      bool rest_imm = true;
      for (uint32_t p = q; p < c.T; ++p) rest_imm = rest_imm && (pool[c.off_regs + p] >> P_KIND_SHIFT) == T_IMM;
      if (rest_imm) nseq[ci] = q;
    }
  }
  js.direct_seq = js.direct_gemms > 0;
  for (size_t ci = 0; ci < cut.size(); ++ci)
    if (in_off[ci] != ~0u && nseq[ci] == 0) js.direct_seq = false;
  HIP_TRY(hipMalloc((void**)&js.d_pool, std::max<size_t>(pool.size(), 4) * 4), "program upload");
  HIP_TRY(hipMemcpy(js.d_pool, pool.data(), pool.size() * 4, hipMemcpyHostToDevice), "program upload");
  uint32_t i0 = 0;  // instructions not yet in a segment (a DENSE the cores cannot take stays interpreted)
  auto flush = [&](uint32_t i1) -> int32_t {
    if (i1 <= i0) return EHIP_OK;
    js.segs.push_back(JetSegDev{});
    JetSegDev& d = js.segs.back();
    const std::vector<uint32_t> st = encode_stream(js.host, false, i0, i1);
    d.nwin = (uint32_t)(st.size() / WIN_WORDS);
    HIP_TRY(hipMalloc((void**)&d.d_stream, st.size() * 4), "program upload");
    HIP_TRY(hipMemcpy(d.d_stream, st.data(), st.size() * 4, hipMemcpyHostToDevice), "program upload");
    if (js.direct_dropped && i0 <= li && li < i1) {
      if (dkept == 0 && i1 == i0 + 1) {
        d.skip_direct = true;  // the segment is that LOADN, with nothing left to load
      } else {
        const std::vector<uint32_t> sd = encode_stream(dprog, false, i0, i1);
        d.nwin_direct = (uint32_t)(sd.size() / WIN_WORDS);
        HIP_TRY(hipMalloc((void**)&d.d_stream_direct, sd.size() * 4), "program upload");
        HIP_TRY(hipMemcpy(d.d_stream_direct, sd.data(), sd.size() * 4, hipMemcpyHostToDevice), "program upload");
      }
    }
    i0 = i1;
    return EHIP_OK;
  };
  for (size_t ci = 0; ci < cut.size(); ++ci) {
    const JetSegment& c = cut[ci];
    if (!c.gemm) continue;
    DenseGemm g{};
    g.pool = js.d_pool;
    g.off_regs = c.off_regs;
    g.off_w = c.off_w;
    g.m = c.m;
    g.T = c.T;
    g.flags = c.flags;
    g.nimm = immediates(c, g.imm);
    if (g.nimm > 4) continue;
    g.off_epi = epi_off[ci];
    g.off_ro = ro_off[ci];
    g.off_in = in_off[ci];
    g.off_wp = wp_off[ci];
    g.nseq = nseq[ci];
    int32_t rc = flush(c.i0);
    if (rc) return rc;
    JetSegDev d;
    d.gemm = true;
    d.g = g;
    js.segs.push_back(d);
    i0 = next[ci];
  }
  return flush(js.host.ninstr);
}

// validate_batch_args (src/gpu/mod.rs:572-582)
inline int32_t check_batch(const ehip_tape* t, uint64_t batch) {
  if (batch == 0) return fail(EHIP_OTHER, "GPU error: batch_size must be non-zero");
  if (t->nin == 0)
    return fail(EHIP_OTHER, "GPU error: tape has no inputs (constant function); evaluate it on the CPU instead");
  return EHIP_OK;
}

// Slot of `tuned` for an evaluation: 0 forward, 1 gradient, 2 hvp, 2+K taylor.
// Program of an evaluation: kind 0 scalar forward / gradient, 1 dual (HVP),
// 2 jet, 3 scalar forward with kink records.
inline int prog_index(int kind, bool grad) {
  return kind == 0 ? (grad ? 1 : 0) : (kind == 1 ? 2 : (kind == 3 ? 3 : 0));
}

inline int tune_key(int kind, int K, bool grad) {
  return kind == 0 ? (grad ? 1 : 0) : (kind == 1 ? 2 : (kind == 3 ? 0 : 2 + K));
}

// HBM register-file bytes one lane of an evaluation needs, and the
// workgroup size it launches with (0 bytes: the tape-specialised kernel, or
// every register in LDS).  Mirrors the shape `run` picks.
inline size_t regfile_lane_bytes(const ehip_ctx* c, const ehip_tape* t, int kind, int K, bool grad,
                                 uint32_t* wg_out = nullptr) {
  const DevProgram& dp = t->prog[prog_index(kind, grad)];
  const uint32_t NC = (kind == 0 || kind == 3) ? 1u : (kind == 1 ? 2u : (uint32_t)K);
  const size_t S = t->f64 ? 8 : 4;
  const Tuned& tu = t->tuned[tune_key(kind, K, grad)];
  const uint32_t wg = tu.wg ? tu.wg : c->wg;
  if (wg_out) *wg_out = wg;
  const int slot = kind == 3 ? -1 : tune_key(kind, K, grad);
  const bool has_spec = slot >= 0 && t->spec[slot].fn != nullptr;
  const uint32_t fm = tu.wg ? tu.fetch : (has_spec ? 2u : c->fetch);
  if (fm == 2 && slot >= 0) return 0;
  const size_t per_reg = (size_t)NC * wg * S;
  const size_t win_lds = window_lds(fm == 1);
  const size_t budget =
      std::min<size_t>(tu.wg ? tu.lds_budget : c->lds_budget, c->max_lds > win_lds ? c->max_lds - win_lds : 0);
  const uint32_t R = dp.host.nregs;
  const uint32_t R_lds = (uint32_t)std::min<size_t>(R, budget / per_reg);
  return (size_t)(R - R_lds) * NC * S;
}

// Launch one interpreter over `L.B` lanes.  kind 0 scalar, 1 dual, 2 jet(K).
// `ws` (optional): a caller-owned HBM register file used instead of the
// context's grow-only one (graphs keep their own so a later reallocation
// cannot pull the memory from under a recorded launch).  `fixed`
// (optional): device memory of `fixed_cap` bytes that holds the register
// file instead; nothing is allocated and a launch that needs more fails.
inline int32_t run(ehip_ctx* c, const ehip_tape* t, int kind, int K, Launch L, hipStream_t s, const Tuned* force = nullptr,
                   DevBuf* ws = nullptr, void* fixed = nullptr, size_t fixed_cap = 0) {
  const DevProgram& dp = t->prog[prog_index(kind, L.g != nullptr)];
  L.prog = dp.d_stream;
  L.nwin = dp.nwin;
  L.pool = dp.d_pool;
  const uint32_t NC = (kind == 0 || kind == 3) ? 1u : (kind == 1 ? 2u : (uint32_t)K);
  const size_t S = t->f64 ? 8 : 4;
  const Tuned& tu = force ? *force : t->tuned[tune_key(kind, K, L.g != nullptr)];
  const uint32_t wg = tu.wg ? tu.wg : c->wg;
  // untuned evaluations run the tape-specialised kernel once it is built
  // (kinks have none; the other kinds share the tuning key's slot)
  const int slot = kind == 3 ? -1 : tune_key(kind, K, L.g != nullptr);
  const bool has_spec = slot >= 0 && t->spec[slot].fn != nullptr;
  const uint32_t fm = tu.wg ? tu.fetch : (has_spec ? 2u : c->fetch);
  if (fm == 2 && slot >= 0) {
    const SpecKernel& sk = t->spec[slot];
    if (!sk.fn) return fail(EHIP_OTHER, "GPU error: no specialised kernel for this evaluation (ehip_tape_specialize)");
    hipError_t e = spec_launch(sk, L, wg, tu.wg ? tu.ppl : 0u, s);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return EHIP_OK;
  }
  const bool sf = fm == 1;
  if (kind == 2 && K == 3 && !t->f64 && c->jet_mma && !t->jseg.segs.empty()) {
    const JetSegProgram& js = t->jseg;
    // every register in HBM: the launches hand their registers on
    const uint64_t grid = (L.B + wg - 1) / wg, lanes = grid * wg;
    const size_t bytes = (size_t)js.host.nregs * NC * lanes * S;
    // Direct plan for device-drawn directions: Rademacher signs of inputs
    // in order are drawn in the GEMMs; other directions once,
    // coordinate-major, into the caller's direction buffer or after the
    // register file.
    uint32_t direct = 0;
    if (js.direct_gemms && L.rng == RNG_RADEMACHER && js.direct_seq) direct = 1;
    else if (js.direct_dropped && (L.rng == RNG_RADEMACHER || L.rng == RNG_GAUSSIAN)) direct = 2;
    size_t dbytes = (direct == 2 && !L.dirs_out) ? ((size_t)L.nin * L.B * 4 + 255) / 256 * 256 : 0;
    if (fixed) {
      if (bytes + dbytes > fixed_cap && dbytes) direct = 0, dbytes = 0;  // no room: the plain plan
      if (bytes > fixed_cap)
        return fail(EHIP_OTHER, "GPU error: the register file needs %zu bytes, above the %zu of the fixed scratch", bytes,
                    fixed_cap);
      L.gregs = fixed;
    } else {
      DevBuf& rf = ws ? *ws : c->regs;
      HIP_TRY(rf.reserve(bytes + dbytes), "register file");
      L.gregs = rf.p;
    }
    float* dcm = nullptr;
    if (direct == 2) {
      dcm = L.dirs_out ? (float*)L.dirs_out : (float*)((char*)L.gregs + bytes);
      hipError_t e = launch_draw_dirs(L.rng, L.seed, L.S, L.B, L.nin, dcm, s);
      if (e != hipSuccess) return hip_fail(e, "kernel launch");
    }
    L.R_lds = 0;
    L.pool = js.d_pool;
    for (const JetSegDev& sg : js.segs) {
      hipError_t e;
      if (sg.gemm) {
        DenseGemm g = sg.g;
        g.regs = (float*)L.gregs;
        g.lanes = lanes;
        // c0 is the primal: equal on every lane that reads the same x row
        g.share = L.xstride == 0 ? 64u : (L.S >= 2 ? L.S : 0u);
        g.direct = (direct && g.off_in != ~0u) ? direct : 0u;
        g.x = (const float*)L.x;
        g.xstride = L.xstride;
        g.S = L.S;
        g.nin = L.nin;
        g.B = L.B;
        g.seed = L.seed;
        g.dcm = dcm;
        g.dirs_out = (float*)L.dirs_out;
        e = launch_dense_gemm(g, s);
      } else {
        if (direct && sg.skip_direct) continue;
        const bool alt = direct && sg.d_stream_direct;
        L.prog = alt ? sg.d_stream_direct : sg.d_stream;
        L.nwin = alt ? sg.nwin_direct : sg.nwin;
        e = launch_jet(false, sf, 3, L, wg, 0, s);
      }
      if (e != hipSuccess) return hip_fail(e, "kernel launch");
    }
    return EHIP_OK;
  }
  const size_t per_reg = (size_t)NC * wg * S;
  // LDS fetch: two instruction windows share the workgroup's LDS with the registers
  const size_t win_lds = window_lds(sf);
  size_t budget = std::min<size_t>(tu.wg ? tu.lds_budget : c->lds_budget, c->max_lds > win_lds ? c->max_lds - win_lds : 0);
  uint32_t R = dp.host.nregs;
  uint32_t R_lds = (uint32_t)std::min<size_t>(R, budget / per_reg);
  L.R_lds = R_lds;
  size_t lds = (size_t)R_lds * per_reg;
  uint64_t grid = (L.B + wg - 1) / wg;
  uint64_t lanes = grid * wg;
  if (R > R_lds) {
    size_t bytes = (size_t)(R - R_lds) * NC * lanes * S;
    if (fixed) {
      if (bytes > fixed_cap)
        return fail(EHIP_OTHER, "GPU error: the register file needs %zu bytes, above the %zu of the fixed scratch", bytes,
                    fixed_cap);
      L.gregs = fixed;
    } else {
      DevBuf& rf = ws ? *ws : c->regs;
      HIP_TRY(rf.reserve(bytes), "register file");
      L.gregs = rf.p;
    }
  } else {
    L.gregs = nullptr;
  }
  hipError_t e;
  if (kind == 0 || kind == 3) e = launch_scalar(t->f64, sf, L, wg, lds, s);
  else if (kind == 1) e = launch_dual(t->f64, sf, L, wg, lds, s);
  else e = launch_jet(t->f64, sf, K, L, wg, lds, s);
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return EHIP_OK;
}

inline hipStream_t pick(ehip_ctx* c, void* stream) { return stream ? (hipStream_t)stream : c->stream; }

}  // namespace ehip_detail

using namespace ehip_detail;
