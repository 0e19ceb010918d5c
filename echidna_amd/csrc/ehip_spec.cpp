// Tape-specialised kernels: the whole forward (and reverse) sweep of one
// tape written out as straight-line HIP and compiled at run time by hipRTC
// for the device's architecture.  Every tape entry becomes a named value
// that the compiler keeps in VGPRs (no register file in LDS or HBM), every
// constant becomes an instruction literal, and the program stream
// disappears.  Worth it for short tapes evaluated over many points, where the
// interpreter's register traffic and dispatch cost dominate; the compile
// costs a fraction of a second to tens of seconds, paid once per tape.
//
// The arithmetic is the interpreters' own: every entry calls the same
// s_eval / s_partials / s_powi / s_powi_partial device functions (the
// embedded text of ehip_math.h), add/sub/mul/neg are written as the plain
// operators those functions apply, and the sweeps follow the reference's
// order exactly:
//   forward   forward_into/forward_dispatch (src/bytecode_tape/forward.rs:11-43, 194-229)
//   reverse   reverse_sweep_core (src/bytecode_tape/reverse.rs:32-105): adjoints
//             start at zero, output 0 is seeded with 1, entries are visited
//             last to first, an entry whose adjoint is zero is skipped, and
//             a partial that is zero adds nothing.
// Two exact simplifications shape the generated reverse sweep:
//   * an entry whose adjoint nothing has written yet is zero at run time for
//     every point, so it is skipped when generating the code;
//   * an update g_a = g_a + c * g with c a finite non-zero constant (add,
//     sub, neg, multiplication by a recorded constant) is emitted without the
//     zero test: when g is +-0, c * g is +-0 and g_a + (+-0) == g_a bit for bit,
//     because an adjoint that starts at +0 and only ever has values added to
//     it is never -0 under round-to-nearest.
// Updates to entries that depend on no input are dropped: their adjoints
// flow only into other such entries, never into the gradient.
#include <hip/hiprtc.h>

#include <chrono>
#include <cinttypes>
#include <cmath>

#include "ehip_internal.h"

// ehip_math.h and philox.h, embedded when the library is built
extern "C" const char ehip_math_h_text[];
extern "C" const char philox_h_text[];

namespace ehip_detail {

namespace {

// opcodes (src/opcode.rs:21-130) the generator treats specially
constexpr uint32_t OP_INPUT = 0, OP_CONST = 1, OP_ADD = 2, OP_SUB = 3, OP_MUL = 4, OP_NEG = 12, OP_POWI = 16,
                   OP_TANH = 32;
constexpr uint32_t UNUSED = 0xFFFFFFFFu;

struct Gen {
  const TapeHost& h;
  const bool f64;
  std::string out;

  Gen(const TapeHost& th) : h(th), f64(th.f64) { out.reserve(h.opcodes.size() * 96); }

  // printf-style append, sized exactly
  void put(const char* fmt, ...) __attribute__((format(printf, 2, 3))) {
    va_list ap, ap2;
    va_start(ap, fmt);
    va_copy(ap2, ap);
    const int n = std::vsnprintf(nullptr, 0, fmt, ap);
    va_end(ap);
    if (n > 0) {
      const size_t at = out.size();
      out.resize(at + (size_t)n + 1);
      std::vsnprintf(&out[at], (size_t)n + 1, fmt, ap2);
      out.resize(at + (size_t)n);
    }
    va_end(ap2);
  }

  // A constant as a bit-exact literal of the tape's dtype.
  std::string lit(double v) const {
    char b[64];
    if (f64) {
      uint64_t u;
      std::memcpy(&u, &v, 8);
      std::snprintf(b, sizeof b, "__builtin_bit_cast(double, 0x%016" PRIx64 "ull)", u);
    } else {
      const float f = (float)v;
      uint32_t u;
      std::memcpy(&u, &f, 4);
      std::snprintf(b, sizeof b, "__builtin_bit_cast(float, 0x%08" PRIx32 "u)", u);
    }
    return b;
  }

  // Value of constant entry `i` in the tape's dtype.
  double cval(uint32_t i) const { return f64 ? h.values[i] : (double)(float)h.values[i]; }
};

// The order the forward sweep's entries are written in: tape order, except
// that an entry whose operands are all written and which ends the life of at
// least one of them is written at once, before the entries after it.  The
// values are the same in any order that respects the operands; this one
// keeps live ranges short where the tape records a reduction late, as a
// model's output sum over hidden units computed after every unit (each unit's
// value then dies as soon as it is made, rather than all of them staying live
// until the sum).  Constants are literals and hold no register: they are
// written when an entry that uses them is.
std::vector<uint32_t> emission_order(const TapeHost& h) {
  const uint32_t N = (uint32_t)h.opcodes.size();
  auto operands = [&](uint32_t i, uint32_t* o) {
    const uint32_t op = h.opcodes[i];
    uint32_t k = 0;
    if (op == OP_INPUT || op == OP_CONST) return k;
    if (h.arg0[i] != UNUSED) o[k++] = h.arg0[i];
    if (op != OP_POWI && h.arg1[i] != UNUSED) o[k++] = h.arg1[i];
    return k;
  };
  std::vector<uint32_t> remaining(N, 0), first(N + 1, 0), cons;
  uint32_t o[2];
  for (uint32_t i = 0; i < N; ++i) {
    const uint32_t k = operands(i, o);
    for (uint32_t t = 0; t < k; ++t) ++first[o[t] + 1];
  }
  for (uint32_t i = 0; i < N; ++i) first[i + 1] += first[i];
  cons.resize(first[N]);
  std::vector<uint32_t> fill(first.begin(), first.end() - 1);
  for (uint32_t i = 0; i < N; ++i) {
    const uint32_t k = operands(i, o);
    for (uint32_t t = 0; t < k; ++t) {
      cons[fill[o[t]]++] = i;
      ++remaining[o[t]];
    }
  }
  for (uint32_t out : h.outputs) remaining[out] += 1;  // outputs stay live to the end
  std::vector<uint8_t> done(N, 0);
  std::vector<uint32_t> order, stack;
  order.reserve(N);
  auto is_const = [&](uint32_t i) { return h.opcodes[i] == OP_CONST; };
  auto emit = [&](uint32_t i) {
    done[i] = 1;
    order.push_back(i);
    const uint32_t k = operands(i, o);
    for (uint32_t t = 0; t < k; ++t) --remaining[o[t]];
  };
  // written early: every operand written, and one of them used for the last time
  auto ready_and_kills = [&](uint32_t c) {
    uint32_t oc[2];
    const uint32_t k = operands(c, oc);
    bool kills = false;
    for (uint32_t t = 0; t < k; ++t) {
      if (!done[oc[t]] && !is_const(oc[t])) return false;
      const uint32_t uses = (k == 2 && oc[0] == oc[1]) ? 2u : 1u;
      if (!is_const(oc[t]) && remaining[oc[t]] == uses) kills = true;
    }
    return kills;
  };
  for (uint32_t i = 0; i < N; ++i) {
    if (done[i]) continue;
    emit(i);
    stack.push_back(i);
    while (!stack.empty()) {
      const uint32_t x = stack.back();
      stack.pop_back();
      for (uint32_t e = first[x]; e < first[x + 1]; ++e) {
        const uint32_t c = cons[e];
        if (!done[c] && ready_and_kills(c)) {
          uint32_t oc[2];
          const uint32_t k = operands(c, oc);
          for (uint32_t t = 0; t < k; ++t)
            if (!done[oc[t]]) emit(oc[t]);  // a constant recorded later
          emit(c);
          stack.push_back(c);
        }
      }
    }
  }
  return order;
}

}  // namespace

// Taylor jets of order K for every output (forward_tangent at T =
// Taylor<F, K>, src/bytecode_tape/tangent.rs:97-147, as taylor_forward_kth
// does per lane): lane p seeds input i with (x_i, v_i, 0, ..) where x is row
// p / S of X (S = 0: row p) and v is read from V or drawn on the device
// (philox.h lane_direction, through a block cache; also written to
// DO[i * B + p] when that is not null), the interpreters' launch convention.
std::string spec_source_jets(const TapeHost& h, int K, uint32_t waves) {
  Gen g(h);
  const uint32_t N = (uint32_t)h.opcodes.size(), n = h.num_inputs, m = (uint32_t)h.outputs.size();
  g.put("// generated: %u entries, %u inputs, %u outputs, %s, Taylor jets of order %d\n", N, n, m,
        h.f64 ? "f64" : "f32", K);
  g.put("#include \"ehip_math.h\"\n#include \"philox.h\"\n");
  if (waves) g.put("#define EHIP_OCC __attribute__((amdgpu_waves_per_eu(%u)))\n", waves);
  else g.put("#define EHIP_OCC\n");
  g.put("typedef %s F;\n", h.f64 ? "double" : "float");
  g.put("typedef ehip::Jet<F, %d> J;\n", K);
  g.put("extern \"C\" __global__ void __launch_bounds__(256) EHIP_OCC ehip_spec(const F* __restrict__ X, "
        "unsigned long long xs, const F* __restrict__ V, F* __restrict__ Y, unsigned long long B, unsigned S, "
        "unsigned rng, unsigned long long seed, F* __restrict__ DO) {\n");
  // points per lane: a grid-stride loop over the points (the launch sizes
  // the grid)
  g.put("for (unsigned long long p = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; p < B;\n"
        "     p += (unsigned long long)gridDim.x * blockDim.x) {\n");
  g.put("const F* __restrict__ x = X + (S ? p / S : p) * xs;\n");
  if (K > 1) g.put("ehip::DirBlock db;  // consecutive inputs share a direction block\n");
  g.put("auto seed_jet = [&](unsigned i) {\n  J r = ehip::jc<F, %d>(x[i]);\n", K);
  if (K > 1)
    g.put("  if (rng) {\n    r.c[1] = (F)ehip::lane_direction(rng, seed, S, p, %uu, i, db);\n"
          "    if (DO) DO[(unsigned long long)i * B + p] = r.c[1];\n  } else {\n    r.c[1] = V[p * %uull + i];\n"
          "  }\n", n, n);
  g.put("  return r;\n};\n");
  for (uint32_t i = 0; i < N; ++i) {
    const uint32_t op = h.opcodes[i], a = h.arg0[i], b = h.arg1[i];
    if (op == OP_INPUT) g.put("const J v%u = seed_jet(%u);\n", i, i);
    else if (op == OP_CONST)
      g.out += "const J v" + std::to_string(i) + " = ehip::jc<F, " + std::to_string(K) + ">(" + g.lit(h.values[i]) +
               ");\n";
    else if (op == OP_POWI) g.put("const J v%u = ehip::j_powi<F, %d>(v%u, %d);\n", i, K, a, (int32_t)b);
    else if (b != UNUSED) g.put("const J v%u = ehip::j_eval<F, %d>(%uu, v%u, v%u);\n", i, K, op, a, b);
    else g.put("const J v%u = ehip::j_eval<F, %d>(%uu, v%u, ehip::jc<F, %d>(F(0)));\n", i, K, op, a, K);
  }
  g.put("F* __restrict__ y = Y + p * %uull;\n", m * K);
  for (uint32_t j = 0; j < m; ++j)
    for (int q = 0; q < K; ++q) g.put("y[%u] = v%u.c[%d];\n", j * K + q, h.outputs[j], q);
  g.put("}\n}\n");
  return g.out;
}

// what: 0 values of every output; 1 values and the gradient of output 0;
// 2 gradient and Hessian-vector product of output 0 (tangent-over-reverse
// on Dual numbers: forward_tangent, src/bytecode_tape/tangent.rs:97-147, then
// reverse_tangent_inner, tangent.rs:191-262, as in hvp, tangent.rs:268-310);
// 3..7 Taylor jets of order what - 2.
std::string spec_source(const TapeHost& h, int what, uint32_t waves) {
  // what 8: value and gradient with binary64 adjoints on an f32 tape
  const bool acc64 = what == SPEC_GRAD_ACC64 && !h.f64;
  if (what == SPEC_GRAD_ACC64) what = 1;
  if (what >= 3) return spec_source_jets(h, what - 2, waves);
  Gen g(h);
  const bool dual = what == 2;
  const uint32_t N = (uint32_t)h.opcodes.size(), n = h.num_inputs, m = (uint32_t)h.outputs.size();
  static const char* kWhat[] = {"values", "value and gradient of output 0", "gradient and HVP of output 0"};
  g.put("// generated: %u entries, %u inputs, %u outputs, %s, %s%s\n", N, n, m, h.f64 ? "f64" : "f32", kWhat[what],
        acc64 ? ", binary64 adjoints" : "");
  g.put("#include \"ehip_math.h\"\n");
  // occupancy target: the compiler keeps at most 512 / waves registers per
  // lane, spilling the rest
  if (waves) g.put("#define EHIP_OCC __attribute__((amdgpu_waves_per_eu(%u)))\n", waves);
  else g.put("#define EHIP_OCC\n");
  g.put("typedef %s F;\n", h.f64 ? "double" : "float");
  g.put("typedef ehip::Dual<F> D;\n");
  if (dual)
    g.put("extern \"C\" __global__ void __launch_bounds__(256) EHIP_OCC ehip_spec(const F* __restrict__ X, "
          "unsigned long long xs, const F* __restrict__ V, F* __restrict__ G, F* __restrict__ H, "
          "unsigned long long B) {\n");
  else
    g.put("extern \"C\" __global__ void __launch_bounds__(256) EHIP_OCC ehip_spec(const F* __restrict__ X, "
          "unsigned long long xs, F* __restrict__ Y, F* __restrict__ G, unsigned long long B) {\n");
  // points per lane: a grid-stride loop over the points (the launch sizes
  // the grid)
  g.put("for (unsigned long long p = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; p < B;\n"
        "     p += (unsigned long long)gridDim.x * blockDim.x) {\n");
  g.put("const F* __restrict__ x = X + p * xs;\n");
  if (dual) g.put("const F* __restrict__ v = V + p * %uull;\n", n);
  const char* T = dual ? "D" : "F";
  for (const uint32_t i : emission_order(h)) {
    const uint32_t op = h.opcodes[i], a = h.arg0[i], b = h.arg1[i];
    if (op == OP_INPUT) {
      if (dual) g.put("const D v%u = {x[%u], v[%u]};\n", i, i, i);
      else g.put("const F v%u = x[%u];\n", i, i);
    } else if (op == OP_CONST) {
      g.out += std::string("const ") + T + " v" + std::to_string(i) + " = " + (dual ? "{" : "") +
               g.lit(h.values[i]) + (dual ? ", F(0)};\n" : ";\n");
    } else if (op == OP_POWI) {
      g.put("const %s v%u = ehip::%s<F>(v%u, %d);\n", T, i, dual ? "d_powi" : "s_powi", a, (int32_t)b);
    } else if (dual) {
      if (b != UNUSED) g.put("const D v%u = ehip::d_eval<F>(%uu, v%u, v%u);\n", i, op, a, b);
      else g.put("const D v%u = ehip::d_eval<F>(%uu, v%u, D{F(0), F(0)});\n", i, op, a);
    } else {
      switch (op) {
        case OP_TANH:
          // binary64: tanh and its derivative from one reduction, the
          // derivative kept for the reverse sweep (the values s_tanh and
          // s_dtanh give)
          if (g.f64 && what == 1) g.put("F v%u, t%u;\nehip::s_tanh_dtanh<F>(v%u, v%u, t%u);\n", i, i, a, i, i);
          else g.put("const F v%u = ehip::s_eval<F>(%uu, v%u, F(0));\n", i, op, a);
          break;
        case OP_ADD: g.put("const F v%u = v%u + v%u;\n", i, a, b); break;
        case OP_SUB: g.put("const F v%u = v%u - v%u;\n", i, a, b); break;
        case OP_MUL: g.put("const F v%u = v%u * v%u;\n", i, a, b); break;
        case OP_NEG: g.put("const F v%u = -v%u;\n", i, a); break;
        default:
          if (b != UNUSED) g.put("const F v%u = ehip::s_eval<F>(%uu, v%u, v%u);\n", i, op, a, b);
          else g.put("const F v%u = ehip::s_eval<F>(%uu, v%u, F(0));\n", i, op, a);
      }
    }
  }
  if (!dual) {
    g.put("if (Y) {\nF* __restrict__ y = Y + p * %uull;\n", m);
    for (uint32_t k = 0; k < m; ++k) g.put("y[%u] = v%u;\n", k, h.outputs[k]);
    g.put("}\n");
  }
  if (what >= 1) {
    // active[i]: entry i depends on an input
    std::vector<uint8_t> active(N, 0), touched(N, 0);
    for (uint32_t i = 0; i < N; ++i) {
      const uint32_t op = h.opcodes[i];
      if (op == OP_INPUT) active[i] = 1;
      else if (op != OP_CONST)
        active[i] = active[h.arg0[i]] || (op != OP_POWI && h.arg1[i] != UNUSED && active[h.arg1[i]]);
    }
    const uint32_t seed = h.outputs[0];
    // Adjoint type: F, or binary64 for an f32 tape (acc64) -- every partial
    // is still formed in FP32 from the FP32 forward values, and the sums of
    // an entry's contributions (its fan-in, one per use) are kept in
    // binary64, so contributions that cancel lose no FP32 bits.
    const char* A = acc64 ? "double" : "F";
    const std::string cast = acc64 ? "(double)" : "";
    if (dual) g.put("D g%u = {F(1), F(0)};\n", seed);
    else g.put("%s g%u = %s(1);\n", A, seed, A);
    touched[seed] = 1;
    auto declare = [&](uint32_t a) {
      if (!touched[a]) {
        if (dual) g.put("D g%u = {F(0), F(0)};\n", a);
        else g.put("%s g%u = %s(0);\n", A, a, A);
        touched[a] = 1;
      }
    };
    // g_a = g_a + P * g_i with a partial P fixed when generating
    auto static_update = [&](uint32_t a, uint32_t i, const std::string& P) {
      declare(a);
      const std::string ga = "g" + std::to_string(a), gi = "g" + std::to_string(i);
      if (dual) g.out += ga + " = ehip::d_add(" + ga + ", ehip::d_mul(" + P + ", " + gi + "));\n";
      else g.out += ga + " = " + ga + " + " + cast + P + " * " + gi + ";\n";
    };
    const std::string one = g.lit(1.0), mone = g.lit(-1.0), mzero = g.lit(-0.0);
    for (uint32_t ii = N; ii-- > 0;) {
      const uint32_t op = h.opcodes[ii];
      if (op == OP_INPUT || op == OP_CONST || !touched[ii]) continue;
      const uint32_t a = h.arg0[ii], b = h.arg1[ii];
      const bool ua = active[a];
      const bool ub = op != OP_POWI && b != UNUSED && active[b];
      if (!ua && !ub) continue;
      // Partials the op fixes, as reverse_partials computes them (ehip_math.h
      // s_partials / d_partials): finite and non-zero, or zero (no update).
      bool stat = false;
      std::string pa, pb;  // empty: zero partial
      auto scal = [&](double c) { return dual ? "D{" + g.lit(c) + ", F(0)}" : g.lit(c); };
      const std::string p_one = dual ? "D{" + one + ", F(0)}" : one;
      const std::string p_mone = dual ? "D{" + mone + ", " + mzero + "}" : mone;  // d_neg(one) = {-1, -0}
      if (op == OP_ADD) stat = true, pa = p_one, pb = p_one;
      else if (op == OP_SUB) stat = true, pa = p_one, pb = p_mone;
      else if (op == OP_NEG) stat = true, pa = p_mone;
      else if (op == OP_MUL && h.opcodes[b] == OP_CONST && std::isfinite(g.cval(b))) {
        stat = true;
        if (g.cval(b) != 0.0) pa = scal(g.cval(b));
      } else if (op == OP_MUL && h.opcodes[a] == OP_CONST && std::isfinite(g.cval(a))) {
        stat = true;
        if (g.cval(a) != 0.0) pb = scal(g.cval(a));
      }
      if (stat) {
        if (ua && !pa.empty()) static_update(a, ii, pa);
        if (ub && !pb.empty()) static_update(b, ii, pb);
        continue;
      }
      if (ua) declare(a);
      if (ub) declare(b);
      const std::string gi = "g" + std::to_string(ii);
      const std::string live = dual ? "!ehip::d_zero(" + gi + ")" : gi + " != " + A + "(0)";
      auto nz = [&](const char* d) { return dual ? std::string("!ehip::d_zero(") + d + ")" : std::string(d) + " != F(0)"; };
      auto upd = [&](uint32_t t, const char* d) {
        const std::string gt = "g" + std::to_string(t);
        return dual ? gt + " = ehip::d_add(" + gt + ", ehip::d_mul(" + d + ", " + gi + "));"
                    : gt + " = " + gt + " + " + cast + d + " * " + gi + ";";
      };
      if (op == OP_POWI) {
        g.put("{ const %s da = ehip::%s<F>(%d, v%u, v%u);\n", T, dual ? "d_powi_partial" : "s_powi_partial",
              (int32_t)b, a, ii);
        g.out += "if (" + live + " && " + nz("da") + ") " + upd(a, "da") + " }\n";
        continue;
      }
      if (op == OP_TANH && g.f64 && !dual) {
        g.put("{ const F da = t%u;\n", ii);
        g.out += "if (" + live + " && " + nz("da") + ") " + upd(a, "da") + " }\n";
        continue;
      }
      const char* zero = dual ? "D{F(0), F(0)}" : "F(0)";
      if (b != UNUSED)
        g.put("{ %s da, db; ehip::%s<F>(%uu, v%u, v%u, v%u, da, db);\n", T, dual ? "d_partials" : "s_partials", op,
              a, b, ii);
      else
        g.put("{ %s da, db; ehip::%s<F>(%uu, v%u, %s, v%u, da, db);\n", T, dual ? "d_partials" : "s_partials", op,
              a, zero, ii);
      g.out += "const bool lv = " + live + ";\n";
      if (ua) g.out += "if (lv && " + nz("da") + ") " + upd(a, "da") + "\n";
      if (ub) g.out += "if (lv && " + nz("db") + ") " + upd(b, "db") + "\n";
      g.put("}\n");
    }
    g.put("F* __restrict__ gp = G + p * %uull;\n", n);
    if (dual) g.put("F* __restrict__ hp = H + p * %uull;\n", n);
    for (uint32_t k = 0; k < n; ++k) {
      if (!touched[k]) {
        g.put("gp[%u] = F(0);\n", k);
        if (dual) g.put("hp[%u] = F(0);\n", k);
      } else if (dual) {
        g.put("gp[%u] = g%u.re;\nhp[%u] = g%u.eps;\n", k, k, k, k);
      } else {
        g.put("gp[%u] = (F)g%u;\n", k, k);
      }
    }
  }
  g.put("}\n}\n");
  return g.out;
}

int32_t spec_compile(const std::string& src, const char* arch, std::vector<char>& code, std::string& log,
                     double* ms) {
  const auto t0 = std::chrono::steady_clock::now();
  hiprtcProgram prog;
  const char* hdrs[] = {ehip_math_h_text, philox_h_text};
  const char* names[] = {"ehip_math.h", "philox.h"};
  if (hiprtcCreateProgram(&prog, src.c_str(), "ehip_spec.hip", 2, hdrs, names) != HIPRTC_SUCCESS)
    return fail(EHIP_SHADER_COMPILATION, "hipRTC: could not create the program");
  std::string archopt = std::string("--offload-arch=") + arch;
  const char* opts[] = {archopt.c_str(), "-O3", "-std=c++17",
                        // reference arithmetic: no a*b+c contraction, denormals kept
                        "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
                        // no packed FP32 instructions
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  log.assign(ls, '\0');
  if (ls) hiprtcGetProgramLog(prog, &log[0]);
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return fail(EHIP_SHADER_COMPILATION, "hipRTC compile failed: %.400s", log.c_str());
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code.resize(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return EHIP_OK;
}

void spec_free(SpecKernel& s) {
  if (s.mod) {
    (void)hipDeviceSynchronize();  // no launch of the old code may still be in flight
    (void)hipModuleUnload(s.mod);
  }
  s = SpecKernel{};
}

hipError_t spec_launch(const SpecKernel& s, const Launch& L, uint32_t wg, uint32_t ppl, hipStream_t st) {
  const void* X = L.x;
  unsigned long long xs = L.xstride;
  const void* V = L.dirs;
  void* Y = L.y;
  void* G = L.g;
  void* H = L.h;
  unsigned long long B = L.B;
  unsigned S = L.S, rng = L.rng;
  unsigned long long seed = L.seed;
  void* DO = L.dirs_out;
  void* args_fwd[] = {(void*)&X, (void*)&xs, (void*)&Y, (void*)&G, (void*)&B};
  void* args_hvp[] = {(void*)&X, (void*)&xs, (void*)&V, (void*)&G, (void*)&H, (void*)&B};
  void* args_jet[] = {(void*)&X, (void*)&xs, (void*)&V, (void*)&Y, (void*)&B, (void*)&S, (void*)&rng, (void*)&seed,
                      (void*)&DO};
  const uint64_t per = (uint64_t)wg * (ppl ? ppl : 1u);  // points per workgroup
  const uint64_t grid = (L.B + per - 1) / per;
  if (grid == 0 || grid > 0xFFFFFFFFull) return hipErrorInvalidConfiguration;
  return hipModuleLaunchKernel(s.fn, (uint32_t)grid, 1, 1, wg, 1, 1, 0, st,
                               s.order ? args_jet : (s.dual ? args_hvp : args_fwd), nullptr);
}

}  // namespace ehip_detail

using namespace ehip_detail;

namespace {

// Waves per SIMD a kernel of `regs` registers per lane (VGPRs and AGPRs,
// allocated in granules of 8 out of 512) can keep resident.
uint32_t occupancy(uint32_t regs) {
  const uint32_t g = (std::max<uint32_t>(regs, 1) + 7) / 8 * 8;
  return std::min<uint32_t>(8, 512 / g);
}

int32_t spec_build(ehip_ctx* c, const TapeHost& h, int what, uint32_t waves, const std::string& arch, SpecKernel& s) {
  const std::string src = spec_source(h, what, waves);
  std::vector<char> code;
  std::string log;
  double ms = 0;
  int32_t rc = spec_compile(src, arch.c_str(), code, log, &ms);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  HIP_TRY(hipModuleLoadData(&s.mod, code.data()), "load specialised kernel");
  hipError_t e = hipModuleGetFunction(&s.fn, s.mod, "ehip_spec");
  if (e != hipSuccess) {
    spec_free(s);
    return hip_fail(e, "load specialised kernel");
  }
  int v = 0;
  if (hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, s.fn) == hipSuccess) s.vgprs = (uint32_t)v;
  if (hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, s.fn) == hipSuccess) s.scratch = (uint32_t)v;
  s.code_bytes = (uint32_t)code.size();
  s.compile_ms = ms;
  s.dual = what == 2;
  s.order = what >= 3 && what <= 7 ? (uint32_t)(what - 2) : 0u;
  s.waves = waves;
  return EHIP_OK;
}

}  // namespace

int32_t ehip_tape_specialize(ehip_tape* t, int32_t what, uint32_t waves_per_eu, uint32_t* out5) try {
  if (!t || !out5) return fail(EHIP_OTHER, "GPU error: null argument");
  if (what < 0 || what > SPEC_GRAD_ACC64)
    return fail(EHIP_OTHER,
                "GPU error: specialised kernels exist for forward (0), gradient (1), hvp (2), taylor order 1..5 (3..7) "
                "and gradient with binary64 adjoints (8)");
  if (what == 2 && t->nout != 1) return fail(EHIP_OTHER, "GPU error: hvp is defined for scalar-output tapes only");
  if (what == SPEC_GRAD_ACC64 && t->f64)
    return fail(EHIP_OTHER, "GPU error: binary64 adjoint accumulation is for f32 tapes (f64 adjoints already are)");
  const int slot = what == SPEC_GRAD_ACC64 ? 1 : what;  // the evaluation it serves
  if (waves_per_eu > 8) return fail(EHIP_OTHER, "GPU error: waves_per_eu must be 0 (automatic) or 1..8");
  if (t->host.opcodes.size() > SPEC_MAX_ENTRIES)
    return fail(EHIP_OTHER, "GPU error: tape too long to specialise (%zu entries > %u)", t->host.opcodes.size(),
                SPEC_MAX_ENTRIES);
  ehip_ctx* c = t->ctx;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, c->device), "device properties");
  std::string arch = prop.gcnArchName;
  arch = arch.substr(0, arch.find(':'));
  // compile without the context lock (seconds for a few thousand entries);
  // load under it
  SpecKernel s;
  int32_t rc = spec_build(c, t->host, what, waves_per_eu, arch, s);
  if (rc) return rc;
  if (waves_per_eu == 0) {
    // Automatic: a build that allows one wave per SIMD is built again for
    // two and that build is kept.  On the n=32 H=64 MLP: f64 fwd+grad 748 ->
    // 1,096 M points/s with 28 VGPRs spilled, f64 HVP 567 -> 744.  From two
    // waves to three the spills cost as much as the wave gains or more
    // (f32 HVP +10% once and -3% later, f32 fwd+grad -26%), so two-wave
    // builds stay as the compiler made them.
    const uint32_t occ = occupancy(s.vgprs);
    if (occ < 2) {
      SpecKernel s2;
      rc = spec_build(c, t->host, what, occ + 1, arch, s2);
      if (rc) {
        std::lock_guard<std::mutex> lk(c->mu);
        spec_free(s);
        return rc;
      }
      s2.compile_ms += s.compile_ms;
      std::lock_guard<std::mutex> lk(c->mu);
      spec_free(s);
      s = s2;
    }
  }
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  // a graph captured earlier may still launch the kernel this one
  // replaces: its module stays loaded until the tape is freed
  if (t->spec[slot].mod) t->retired.push_back(t->spec[slot]);
  t->spec[slot] = s;
  out5[0] = s.vgprs;
  out5[1] = s.scratch;
  out5[2] = (uint32_t)std::min(s.compile_ms * 1000.0, 4.0e9);
  out5[3] = s.code_bytes;
  out5[4] = s.waves;
  return ok();
}
catch (...) {
  return exception_status();
}

int32_t ehip_spec_source(int32_t dtype, const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1,
                         const double* constants, uint32_t num_ops, uint32_t num_inputs, uint32_t output_index,
                         const uint32_t* output_indices, uint32_t num_outputs, int32_t what,
                         uint32_t waves_per_eu, const char* arch, char* src, uint64_t src_cap, uint64_t* out3) try {
  if (!out3 || what < 0 || what > SPEC_GRAD_ACC64 || waves_per_eu > 8 || (dtype != EHIP_F32 && dtype != EHIP_F64) ||
      (what == SPEC_GRAD_ACC64 && dtype != EHIP_F32) ||
      (num_ops > 0 && (!opcodes || !arg0 || !arg1 || !constants)))
    return fail(EHIP_OTHER, "GPU error: bad argument");
  for (uint32_t i = 0; i < num_ops; ++i)
    if (opcodes[i] == 43u) return fail(EHIP_CUSTOM_OPS_NOT_SUPPORTED, "tape contains custom ops which cannot run on GPU");
  std::vector<uint32_t> outs;
  if (num_outputs == 0) outs.push_back(output_index);
  else outs.assign(output_indices, output_indices + num_outputs);
  std::string msg = validate_tape(opcodes, arg0, arg1, num_ops, num_inputs, num_ops, num_ops, output_index,
                                  outs.data(), (uint32_t)outs.size());
  if (!msg.empty()) return fail(EHIP_OTHER, "%s", msg.c_str());
  if (num_ops == 0) return fail(EHIP_OTHER, "invalid GpuTapeData: empty tape");
  if (what == 2 && outs.size() != 1) return fail(EHIP_OTHER, "GPU error: hvp is defined for scalar-output tapes only");
  if (num_ops > SPEC_MAX_ENTRIES)
    return fail(EHIP_OTHER, "GPU error: tape too long to specialise (%u entries > %u)", num_ops, SPEC_MAX_ENTRIES);
  TapeHost th;
  th.opcodes.assign(opcodes, opcodes + num_ops);
  th.arg0.assign(arg0, arg0 + num_ops);
  th.arg1.assign(arg1, arg1 + num_ops);
  th.values.assign(constants, constants + num_ops);
  th.num_inputs = num_inputs;
  th.outputs = outs;
  th.f64 = dtype == EHIP_F64;
  const std::string s = spec_source(th, what, waves_per_eu);
  out3[0] = s.size();
  out3[1] = 0;
  out3[2] = 0;
  if (arch) {
    std::vector<char> code;
    std::string log;
    double ms = 0;
    int32_t rc = spec_compile(s, arch, code, log, &ms);
    if (rc) return rc;
    out3[1] = code.size();
    out3[2] = (uint64_t)(ms * 1000.0);
  }
  if (src) {
    if (src_cap < s.size() + 1) return fail(EHIP_OTHER, "GPU error: source buffer too small");
    std::memcpy(src, s.c_str(), s.size() + 1);
  }
  return ok();
} catch (...) {
  return exception_status();
}
