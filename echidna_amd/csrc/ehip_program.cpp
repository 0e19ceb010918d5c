// Tape -> register program compiler (see ehip_program.h).
#include "ehip_program.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <queue>
#include <tuple>

namespace ehip {

namespace {

constexpr uint32_t UNUSED = 0xFFFFFFFFu;
constexpr uint32_t OPC_INPUT = 0, OPC_CONST = 1, OPC_ADD = 2, OPC_SUB = 3, OPC_MUL = 4, OPC_DIV = 5,
                   OPC_REM = 6, OPC_POWF = 7, OPC_ATAN2 = 8, OPC_HYPOT = 9, OPC_MAX = 10, OPC_MIN = 11,
                   OPC_NEG = 12, OPC_RECIP = 13, OPC_SQRT = 14, OPC_CBRT = 15, OPC_POWI = 16,
                   OPC_EXP = 17, OPC_EXP2 = 18, OPC_EXPM1 = 19, OPC_FRACT = 42, OPC_CUSTOM = 43;

inline bool is_binary(uint32_t op) { return op >= OPC_ADD && op <= OPC_MIN; }

// Which operand values the reverse partial of `op` reads when the
// adjoints of a (accA) and/or b (accB) are consumed.  Mirrors the operands
// each arm of reverse_partials touches (src/opcode.rs:331-483).
void rev_reads(uint32_t op, bool accA, bool accB, int32_t exp, bool& ra, bool& rb, bool& rr) {
  ra = rb = rr = false;
  switch (op) {
    case OPC_ADD: case OPC_SUB: case OPC_NEG: case OPC_FRACT:
    case 37: case 38: case 39: case 40: case 41:  // signum floor ceil round trunc
      return;
    case OPC_MUL: rb = accA; ra = accB; return;
    case OPC_DIV: rb = true; rr = accB; return;
    case OPC_REM: ra = rb = accB; return;
    case OPC_POWF: ra = rb = rr = true; return;
    case OPC_ATAN2: ra = rb = true; return;
    case OPC_HYPOT: ra = rb = rr = true; return;
    case OPC_MAX: case OPC_MIN: ra = rb = true; return;
    case OPC_SQRT: case OPC_CBRT: case OPC_EXP: case OPC_EXP2: case OPC_EXPM1: rr = true; return;
    case OPC_POWI: ra = true; rr = (exp == INT32_MIN); return;
    default: ra = true; return;  // recip, logs, trig, hyperbolic, abs
  }
}

struct Opnd {
  bool imm = false;
  double v = 0.0;
  int64_t vid = -1;
};

struct ATerm {
  uint32_t kind = T_REG;
  int64_t vid = -1;      // value vid (DOT) / adjoint vid (DOT_REV)
  double imm = 0.0;
  bool const_left = false;  // product written c * x on the tape
  bool first = false;
  uint32_t depth = 0;
};

struct AIns {
  uint32_t op = 0, flags = 0, aux = 0;
  int64_t dst = -1, g = -1, r = -1, adjA = -1, adjB = -1;
  Opnd a, b;
  std::vector<ATerm> terms;
  bool full = false;             // DOT / DOT_REV covering its whole chain
  uint32_t kop = 0;              // KINK: opcode of the nonsmooth entry
  std::vector<int64_t> members;  // DENSE: destinations; DENSE_REV: adjoints g_j
  std::vector<double> W;         // DENSE / DENSE_REV weights [member][term]
  bool inplace = false;          // unary op that takes its operand's register
};

constexpr size_t DENSE_MIN = 4;     // members worth a DENSE instruction
constexpr size_t DENSE_CHUNK = 16;  // members per forward DENSE

// Replace groups of full DOTs (and their DOT_REVs) that share one term
// structure by DENSE (DENSE_REV) instructions.  A forward group moves to its
// first member's position: every member reads the same operands, defined
// before that point.  A reverse group moves to its last member's position
// when no instruction in between touches the operand adjoints, so each
// adjoint receives the same contributions in the same order.
void group_dense(std::vector<AIns>& prog, int64_t adj_offset, size_t chunk, bool dual) {
  struct Key {
    std::vector<int64_t> k;
    bool operator<(const Key& o) const { return k < o.k; }
  };
  auto key_of = [](const AIns& x, bool& ok) {
    Key key;
    ok = true;
    std::vector<int64_t> vids;
    for (const ATerm& tm : x.terms) {
      key.k.push_back((int64_t)tm.kind | (tm.const_left ? 4 : 0));
      key.k.push_back(tm.kind == T_IMM ? -1 : tm.vid);
      if (tm.kind != T_IMM) vids.push_back(tm.vid);
    }
    std::sort(vids.begin(), vids.end());
    if (std::adjacent_find(vids.begin(), vids.end()) != vids.end()) ok = false;
    if (vids.empty()) ok = false;
    return key;
  };
  std::map<Key, std::vector<size_t>> groups;
  for (size_t p = 0; p < prog.size(); ++p) {
    const AIns& x = prog[p];
    if (x.op != I_DOT || !x.full) continue;
    bool ok;
    Key k = key_of(x, ok);
    if (ok) groups[k].push_back(p);
  }
  std::vector<char> drop(prog.size(), 0);
  std::vector<std::pair<size_t, AIns>> place;  // position -> new instruction
  // DOT_REV lookup by root adjoint
  std::map<int64_t, size_t> rev_at;
  std::map<int64_t, int> rev_count;
  for (size_t p = 0; p < prog.size(); ++p)
    if (prog[p].op == I_DOT_REV) {
      rev_count[prog[p].g]++;
      if (prog[p].full) rev_at[prog[p].g] = p;
    }
  for (auto& kv : groups) {
    const std::vector<size_t>& mem = kv.second;
    if (mem.size() < DENSE_MIN) continue;
    const AIns& x0 = prog[mem[0]];
    const size_t T = x0.terms.size();
    // Forward chunks of `chunk` members, each at its first member's position:
    // the kernels block members by at most 16 anyway, and a chunk keeps at
    // most `chunk` new values live at once (forward programs; a gradient
    // program keeps the layer's values for the reverse sweep regardless).
    for (size_t c0 = 0; c0 < mem.size(); c0 += chunk) {
      const size_t c1 = std::min(mem.size(), c0 + chunk);
      AIns d;
      d.op = I_DENSE;
      d.terms = x0.terms;
      for (size_t j = c0; j < c1; ++j) {
        const AIns& x = prog[mem[j]];
        d.members.push_back(x.dst);
        for (size_t q = 0; q < T; ++q) d.W.push_back(x.terms[q].imm);
        drop[mem[j]] = 1;
      }
      if (c1 - c0 == 1) {  // a lone member stays a DOT
        drop[mem[c0]] = 0;
        continue;
      }
      place.push_back({mem[c0], std::move(d)});
    }
    // reverse: every member has exactly one DOT_REV, the whole chain
    std::vector<size_t> rp;
    bool rev_ok = true;
    for (size_t j : mem) {
      // the DOT_REV of member j reads the adjoint of its root
      const int64_t g = prog[j].dst + adj_offset;
      auto it = rev_at.find(g);
      if (it == rev_at.end() || rev_count[g] != 1) { rev_ok = false; break; }
      rp.push_back(it->second);
    }
    if (!rev_ok || rp.empty()) continue;
    std::vector<size_t> order(rp);
    std::sort(order.begin(), order.end());
    // the operand adjoints: the same for every member (same operands)
    const AIns& r0 = prog[order[0]];
    std::vector<int64_t> adj;
    for (const ATerm& tm : r0.terms) adj.push_back(tm.vid);
    std::vector<int64_t> sadj(adj);
    std::sort(sadj.begin(), sadj.end());
    bool same = true;
    for (size_t q : order) {
      const AIns& r = prog[q];
      if (r.terms.size() != adj.size()) { same = false; break; }
      std::vector<int64_t> a2;
      for (const ATerm& tm : r.terms) a2.push_back(tm.vid);
      std::sort(a2.begin(), a2.end());
      if (a2 != sadj) { same = false; break; }
      for (size_t k = 0; k < r.terms.size(); ++k)
        if (r.terms[k].first && q != order[0]) same = false;
    }
    // dual programs: every member must reach each operand through the same
    // kind of link (depth 0 or >= 1), the one property of the chain shape
    // the dual DENSE_REV keeps
    if (dual) {
      std::map<int64_t, bool> deep;
      for (const ATerm& tm : r0.terms) deep[tm.vid] = tm.depth >= 1;
      for (size_t q : order)
        for (const ATerm& tm : prog[q].terms)
          if (deep[tm.vid] != (tm.depth >= 1)) same = false;
    }
    if (!same) continue;
    auto touches = [&](const AIns& x) {
      auto hit = [&](int64_t v) { return v >= 0 && std::binary_search(sadj.begin(), sadj.end(), v); };
      if (hit(x.dst) || hit(x.g) || hit(x.r) || hit(x.adjA) || hit(x.adjB)) return true;
      if (!x.a.imm && hit(x.a.vid)) return true;
      if (!x.b.imm && hit(x.b.vid)) return true;
      for (const ATerm& tm : x.terms)
        if (hit(tm.vid)) return true;
      for (int64_t v : x.members)
        if (hit(v)) return true;
      return false;
    };
    bool clear = true;
    for (size_t q = order.front() + 1; q < order.back() && clear; ++q)
      if (!std::binary_search(order.begin(), order.end(), q) && touches(prog[q])) clear = false;
    if (!clear) continue;
    // terms in r0's order (operands distinct, so the order across operands
    // does not matter); members in application order
    AIns dr;
    dr.op = I_DENSE_REV;
    dr.terms = r0.terms;
    for (size_t q : order) {
      const AIns& r = prog[q];
      dr.members.push_back(r.g);
      std::map<int64_t, const ATerm*> by;
      for (const ATerm& tm : r.terms) by[tm.vid] = &tm;
      for (const ATerm& tm : r0.terms) dr.W.push_back(by[tm.vid]->imm);
      drop[q] = 1;
    }
    place.push_back({order.back(), std::move(dr)});
  }
  if (place.empty()) return;
  std::map<size_t, AIns*> at;
  for (auto& pr : place) at[pr.first] = &pr.second;
  std::vector<AIns> out;
  out.reserve(prog.size());
  for (size_t p = 0; p < prog.size(); ++p) {
    auto it = at.find(p);
    if (it != at.end()) out.push_back(std::move(*it->second));
    if (!drop[p]) out.push_back(std::move(prog[p]));
  }
  prog.swap(out);
}

// One chain of `acc + term` links rooted at an Add.
struct Term {
  uint32_t kind;
  uint32_t slot;       // value slot (PROD: the non-constant factor)
  double imm;
  bool const_left;
  uint32_t contrib;    // tape index of the op whose reverse feeds this term
  uint32_t sub;        // 0 = arg0 side, 1 = arg1 side (tie order)
  uint32_t add_pos;    // index of the Add link where the term entered
};
struct Chain {
  std::vector<Term> terms;
  uint32_t lo = 0, hi = 0, count = 0;  // contiguous ordinal range of absorbed ops
  uint32_t nadds = 0;
  bool has_prod = false;
};

}  // namespace

std::string validate_tape(const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1,
                          uint32_t num_ops, uint32_t num_inputs, uint32_t num_variables,
                          uint32_t num_constants, uint32_t output_index, const uint32_t* outs,
                          uint32_t num_outputs) {
  char buf[256];
  auto err = [&](const char* fmt, auto... xs) {
    std::snprintf(buf, sizeof buf, fmt, xs...);
    return std::string("invalid GpuTapeData: ") + buf;
  };
  if (num_ops != num_variables) return err("num_ops (%u) != num_variables (%u)", num_ops, num_variables);
  if (num_constants != num_variables)
    return err("constants.len() (%u) != num_variables (%u)", num_constants, num_variables);
  if (num_inputs > num_variables) return err("num_inputs (%u) > num_variables (%u)", num_inputs, num_variables);
  for (uint32_t i = 0; i < num_ops; ++i) {
    uint32_t op = opcodes[i];
    if (op > OPC_FRACT) return err("opcodes[%u] (%u) is not a GPU-executable opcode", i, op);
    if ((op == OPC_INPUT) != (i < num_inputs)) {
      if (i < num_inputs) return err("opcodes[%u] (%u) should be Input", i, op);
      return err("opcodes[%u] is Input, but only the first %u entries may be inputs", i, num_inputs);
    }
  }
  for (uint32_t i = num_inputs; i < num_ops; ++i) {
    uint32_t op = opcodes[i];
    if (op == OPC_CONST) continue;
    if (arg0[i] >= i) return err("arg0[%u] (%u) must reference an earlier slot", i, arg0[i]);
    if (op == OPC_POWI) continue;
    if (is_binary(op)) {
      if (arg1[i] >= i) return err("arg1[%u] (%u) must reference an earlier slot", i, arg1[i]);
    } else if (arg1[i] != UNUSED) {
      return err("arg1[%u] (%u) should be the UNUSED sentinel for unary opcode %u", i, arg1[i], op);
    }
  }
  if (output_index >= num_variables) return err("output_index (%u) >= num_variables (%u)", output_index, num_variables);
  for (uint32_t j = 0; j < num_outputs; ++j)
    if (outs[j] >= num_variables) return err("output_indices[%u] (%u) >= num_variables (%u)", j, outs[j], num_variables);
  return "";
}

// The segmented jet program (whole-layer DENSE groups, forward): a unary
// op whose operand is a DENSE member that nothing else reads moves to just
// after its DENSE and computes in place -- the member's register takes the
// result -- so the batch GEMM that runs the DENSE can apply it in its
// epilogue (an activation layer).  The moved op reads only that member,
// which the DENSE defines, and its result is a value of its own, so moving
// it earlier changes no value.  Read-out DOT pieces of one chain that the
// move leaves adjacent merge into one DOT: the same terms in the same order,
// so every rounding is unchanged.
void hoist_member_unaries(std::vector<AIns>& prog, size_t NV) {
  std::vector<uint32_t> reads(NV, 0);
  auto rd = [&](int64_t v) { if (v >= 0) ++reads[(size_t)v]; };
  for (const AIns& x : prog) {
    if (x.op == I_LOADN) continue;
    if (!x.a.imm) rd(x.a.vid);
    if (!x.b.imm && (x.flags & F_HAS_B)) rd(x.b.vid);
    if (x.op == I_DOT || x.op == I_DENSE)
      for (const ATerm& tm : x.terms) rd(tm.vid);
    if (x.op == I_DOT && (x.flags & F_DOT_CONT)) rd(x.dst);
  }
  std::vector<int64_t> dense_of(NV, -1);  // member vid -> position of its DENSE
  for (size_t p = 0; p < prog.size(); ++p)
    if (prog[p].op == I_DENSE)
      for (int64_t v : prog[p].members) dense_of[(size_t)v] = (int64_t)p;
  std::vector<std::vector<size_t>> moved(prog.size());
  std::vector<char> is_moved(prog.size(), 0);
  for (size_t q = 0; q < prog.size(); ++q) {
    const AIns& x = prog[q];
    if (x.op >= 64u || x.op == OPC_POWI || (x.flags & (F_HAS_B | F_A_IMM)) || x.a.vid < 0) continue;
    const int64_t d = dense_of[(size_t)x.a.vid];
    if (d < 0 || reads[(size_t)x.a.vid] != 1) continue;
    moved[(size_t)d].push_back(q);
    is_moved[q] = 1;
  }
  std::vector<AIns> out;
  out.reserve(prog.size());
  for (size_t p = 0; p < prog.size(); ++p) {
    if (is_moved[p]) continue;
    out.push_back(std::move(prog[p]));
    for (size_t q : moved[p]) {
      out.push_back(std::move(prog[q]));
      out.back().inplace = true;
    }
  }
  prog.clear();
  for (AIns& x : out) {
    if (!prog.empty() && x.op == I_DOT && (x.flags & F_DOT_CONT) && prog.back().op == I_DOT &&
        prog.back().dst == x.dst) {
      AIns& y = prog.back();
      y.terms.insert(y.terms.end(), x.terms.begin(), x.terms.end());
      y.full = false;
      continue;
    }
    prog.push_back(std::move(x));
  }
}

Program compile_program(const TapeHost& t, ProgMode mode, bool fuse, bool dense, bool dual, size_t dense_chunk) {
  const uint32_t N = (uint32_t)t.opcodes.size();
  const uint32_t ni = t.num_inputs;
  const bool grad = (mode == PM_GRADIENT);
  auto is_const = [&](uint32_t s) { return t.opcodes[s] == OPC_CONST; };
  auto binop = [&](uint32_t i) { return is_binary(t.opcodes[i]); };

  // ---- reverse reachability from the seed (output 0) ----
  std::vector<char> reach(N, 0);
  const uint32_t seed = t.outputs[0];
  if (grad && !is_const(seed)) reach[seed] = 1;
  if (grad) {
    for (int64_t ii = (int64_t)N - 1; ii >= (int64_t)ni; --ii) {
      uint32_t i = (uint32_t)ii;
      if (!reach[i] || is_const(i)) continue;
      uint32_t a = t.arg0[i];
      if (!is_const(a)) reach[a] = 1;
      if (binop(i) && !is_const(t.arg1[i])) reach[t.arg1[i]] = 1;
    }
  }

  // ---- forward values needed: outputs, reverse reads, transitive operands ----
  std::vector<char> need(N, 0);
  for (uint32_t o : t.outputs)
    if (!is_const(o)) need[o] = 1;
  if (mode == PM_KINK)
    for (uint32_t i = ni; i < N; ++i)
      if (is_nonsmooth_op(t.opcodes[i])) need[i] = 1;
  if (grad) {
    for (uint32_t i = ni; i < N; ++i) {
      if (!reach[i] || is_const(i)) continue;
      uint32_t op = t.opcodes[i];
      bool accA = !is_const(t.arg0[i]);
      bool accB = binop(i) && !is_const(t.arg1[i]);
      if (op == OPC_POWI) accB = false;
      bool ra, rb, rr;
      rev_reads(op, accA, accB, (int32_t)t.arg1[i], ra, rb, rr);
      if (ra && !is_const(t.arg0[i])) need[t.arg0[i]] = 1;
      if (rb && binop(i) && !is_const(t.arg1[i])) need[t.arg1[i]] = 1;
      if (rr) need[i] = 1;
      need[i] = 1;  // a reached op is evaluated so its operands exist
    }
  }
  for (int64_t ii = (int64_t)N - 1; ii >= (int64_t)ni; --ii) {
    uint32_t i = (uint32_t)ii;
    if (!need[i] || is_const(i)) continue;
    if (!is_const(t.arg0[i])) need[t.arg0[i]] = 1;
    if (binop(i) && !is_const(t.arg1[i])) need[t.arg1[i]] = 1;
  }

  // ---- use counts over needed ops and outputs ----
  std::vector<uint32_t> uses(N, 0);
  std::vector<char> is_out(N, 0);
  for (uint32_t o : t.outputs) is_out[o] = 1;
  std::vector<uint32_t> ord(N, 0);
  uint32_t nord = 0;
  for (uint32_t i = ni; i < N; ++i) {
    if (!need[i] || is_const(i)) continue;
    ord[i] = nord++;
    if (!is_const(t.arg0[i])) uses[t.arg0[i]]++;
    if (binop(i) && !is_const(t.arg1[i])) uses[t.arg1[i]]++;
  }

  // ---- DOT fusion ----
  std::vector<char> absorbed(N, 0);
  std::vector<int64_t> chain_of(N, -1);
  std::vector<Chain> chains;
  auto prod_candidate = [&](uint32_t s, uint32_t consumer) -> bool {
    if (s < ni || is_const(s) || t.opcodes[s] != OPC_MUL) return false;
    if (uses[s] != 1 || is_out[s]) return false;
    bool c0 = is_const(t.arg0[s]), c1 = is_const(t.arg1[s]);
    if (c0 == c1) return false;
    (void)consumer;
    return true;
  };
  if (fuse) {
    for (uint32_t i = ni; i < N; ++i) {
      if (!need[i] || is_const(i) || t.opcodes[i] != OPC_ADD) continue;
      uint32_t ops[2] = {t.arg0[i], t.arg1[i]};
      int take_chain = -1;
      for (int k = 0; k < 2; ++k) {
        uint32_t o = ops[k];
        if (take_chain < 0 && o >= ni && !is_const(o) && t.opcodes[o] == OPC_ADD && chain_of[o] >= 0 &&
            uses[o] == 1 && !is_out[o]) {
          const Chain& c = chains[chain_of[o]];
          // contiguity with op i (and a possible absorbed product on the other side)
          uint32_t lo = c.lo, hi = c.hi, cnt = c.count;
          uint32_t other = ops[1 - k];
          auto extend = [&](uint32_t s) { lo = std::min(lo, ord[s]); hi = std::max(hi, ord[s]); cnt++; };
          extend(i);
          if (!is_const(other) && prod_candidate(other, i)) extend(other);
          if (cnt == hi - lo + 1) take_chain = k;
        }
      }
      Chain ch;
      if (take_chain >= 0) {
        uint32_t o = ops[take_chain];
        ch = chains[chain_of[o]];
        absorbed[o] = 1;
      }
      uint32_t add_pos = ch.nadds;
      auto add_ord = [&](uint32_t s) {
        if (ch.count == 0) { ch.lo = ch.hi = ord[s]; ch.count = 1; return; }
        ch.lo = std::min(ch.lo, ord[s]);
        ch.hi = std::max(ch.hi, ord[s]);
        ch.count++;
      };
      add_ord(i);
      // Products absorbed here must keep the absorbed ordinals contiguous.
      // When both operands are products (a chain's first link), they are
      // judged together: x0*c0, x1*c1, Add is contiguous as a pair though
      // neither product is alone.
      bool take_prod[2] = {false, false};
      {
        std::vector<uint32_t> cand;
        for (int k = 0; k < 2; ++k)
          if (k != take_chain && !is_const(ops[k]) && prod_candidate(ops[k], i)) cand.push_back((uint32_t)k);
        auto fits = [&](const std::vector<uint32_t>& ks) {
          uint32_t lo = ch.lo, hi = ch.hi, cnt = ch.count;
          for (uint32_t k : ks) { lo = std::min(lo, ord[ops[k]]); hi = std::max(hi, ord[ops[k]]); ++cnt; }
          return cnt == hi - lo + 1;
        };
        if (cand.size() == 2 && fits(cand)) {
          take_prod[0] = take_prod[1] = true;
        } else {
          // otherwise the later product first, then the other if it still fits
          std::sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return ord[ops[a]] > ord[ops[b]]; });
          std::vector<uint32_t> taken;
          for (uint32_t k : cand) {
            taken.push_back(k);
            if (fits(taken)) take_prod[k] = true;
            else taken.pop_back();
          }
        }
      }
      for (int k = 0; k < 2; ++k) {
        if (k == take_chain) continue;
        uint32_t o = ops[k];
        Term tm{};
        tm.sub = (uint32_t)k;
        tm.add_pos = add_pos;
        if (is_const(o)) {
          tm.kind = T_IMM; tm.slot = o; tm.imm = t.values[o]; tm.contrib = i;
        } else if (prod_candidate(o, i)) {
          if (take_prod[k]) {
            bool cl = is_const(t.arg0[o]);
            tm.kind = T_PROD;
            tm.slot = cl ? t.arg1[o] : t.arg0[o];
            tm.imm = t.values[cl ? t.arg0[o] : t.arg1[o]];
            tm.const_left = cl;
            tm.contrib = o;
            absorbed[o] = 1;
            add_ord(o);
            ch.has_prod = true;
          } else {
            tm.kind = T_REG; tm.slot = o; tm.contrib = i;
          }
        } else {
          tm.kind = T_REG; tm.slot = o; tm.contrib = i;
        }
        ch.terms.push_back(tm);
      }
      ch.nadds++;
      chain_of[i] = (int64_t)chains.size();
      chains.push_back(std::move(ch));
    }
  }
  auto dot_root = [&](uint32_t i) -> const Chain* {
    if (!fuse || chain_of[i] < 0 || absorbed[i]) return nullptr;
    const Chain& c = chains[chain_of[i]];
    if (c.terms.size() >= 3 || c.has_prod) return &c;
    return nullptr;
  };

  // ---- abstract instruction stream ----
  std::vector<AIns> prog;
  auto VAL = [&](uint32_t s) -> int64_t { return (int64_t)s; };
  auto ADJ = [&](uint32_t s) -> int64_t { return (int64_t)N + s; };
  auto opnd = [&](uint32_t s) {
    Opnd o;
    if (is_const(s)) { o.imm = true; o.v = t.values[s]; }
    else o.vid = VAL(s);
    return o;
  };

  {
    // One batched load of every needed input: the loads issue together.
    AIns x;
    x.op = I_LOADN;
    for (uint32_t i = 0; i < ni; ++i) {
      if (!need[i]) continue;
      ATerm at;
      at.kind = T_REG;
      at.vid = VAL(i);
      at.imm = (double)i;  // input index
      x.terms.push_back(at);
    }
    if (!x.terms.empty()) prog.push_back(std::move(x));
  }

  // ---- DOT piece schedule ----
  // A fused chain is emitted as one or more DOT pieces.  A term whose operand
  // is a single-use computed value may be accumulated right after that value
  // is produced (its register is then freed at once), which matters for
  // read-out sums over a whole layer; other terms wait for the chain's own
  // position.  Term order, and so every rounding, is unchanged.
  std::vector<std::vector<std::pair<uint32_t, std::pair<uint32_t, uint32_t>>>> pieces_at(N + 1);  // pos+1 -> (root, [lo,hi))
  std::vector<char> split_root(N, 0);
  if (fuse) {
    for (uint32_t r = ni; r < N; ++r) {
      if (!need[r] || is_const(r) || absorbed[r]) continue;
      const Chain* c = dot_root(r);
      if (!c) continue;
      const size_t T = c->terms.size();
      std::vector<int64_t> q(T, (int64_t)r);
      std::vector<char> is_imm(T, 0);
      for (size_t k = 0; k < T; ++k) {
        const Term& tm = c->terms[k];
        if (tm.kind == T_IMM) { is_imm[k] = 1; continue; }
        const uint32_t s_ = tm.slot;
        bool hoist = s_ >= ni && uses[s_] == 1 && !is_out[s_] && !absorbed[s_];
        q[k] = hoist ? (int64_t)s_ : (int64_t)r;
      }
      // immediates follow the next operand term (or the previous one)
      for (size_t k = 0; k < T; ++k) {
        if (!is_imm[k]) continue;
        int64_t v = -2;
        for (size_t j = k + 1; j < T; ++j)
          if (!is_imm[j]) { v = q[j]; break; }
        if (v == -2) v = (k > 0) ? q[k - 1] : (int64_t)r;
        q[k] = v;
      }
      for (size_t k = 1; k < T; ++k) q[k] = std::max(q[k], q[k - 1]);
      size_t lo = 0;
      for (size_t k = 1; k <= T; ++k) {
        if (k == T || q[k] != q[lo]) {
          pieces_at[(size_t)(q[lo] + 1)].push_back({r, {(uint32_t)lo, (uint32_t)k}});
          lo = k;
        }
      }
      split_root[r] = 1;
    }
  }
  uint64_t dot_terms = 0;
  auto emit_piece = [&](uint32_t r, uint32_t lo, uint32_t hi) {
    const Chain& c = chains[chain_of[r]];
    AIns x;
    x.op = I_DOT;
    x.dst = VAL(r);
    x.full = lo == 0 && hi == (uint32_t)c.terms.size();
    if (lo > 0) x.flags |= F_DOT_CONT;
    for (uint32_t k = lo; k < hi; ++k) {
      const Term& tm = c.terms[k];
      ATerm at;
      at.kind = tm.kind;
      at.imm = tm.imm;
      at.const_left = tm.const_left;
      if (tm.kind != T_IMM) at.vid = VAL(tm.slot);
      x.terms.push_back(at);
    }
    dot_terms += hi - lo;
    prog.push_back(std::move(x));
  };
  auto flush_pieces = [&](int64_t pos) {
    for (auto& pc : pieces_at[(size_t)(pos + 1)]) emit_piece(pc.first, pc.second.first, pc.second.second);
  };
  flush_pieces(-1);
  uint32_t nkink = 0;
  for (uint32_t i = ni; i < N; ++i) {
    if (!need[i] || is_const(i) || absorbed[i]) continue;
    if (split_root[i]) {
      flush_pieces(i);
      continue;
    }
    AIns x;
    uint32_t op = t.opcodes[i];
    x.op = op;
    x.dst = VAL(i);
    x.a = opnd(t.arg0[i]);
    if (x.a.imm) x.flags |= F_A_IMM;
    if (op == OPC_POWI) {
      x.aux = t.arg1[i];
    } else if (binop(i)) {
      x.flags |= F_HAS_B;
      x.b = opnd(t.arg1[i]);
      if (x.b.imm) x.flags |= F_B_IMM;
    }
    prog.push_back(std::move(x));
    if (mode == PM_KINK && is_nonsmooth_op(op)) {
      AIns k;
      k.op = I_KINK;
      k.aux = nkink++;
      k.kop = op;
      k.a = opnd(t.arg0[i]);
      if (k.a.imm) k.flags |= F_A_IMM;
      if (binop(i)) {
        k.flags |= F_HAS_B;
        k.b = opnd(t.arg1[i]);
        if (k.b.imm) k.flags |= F_B_IMM;
      }
      prog.push_back(std::move(k));
    }
    flush_pieces(i);
  }
  for (uint32_t j = 0; j < (uint32_t)t.outputs.size(); ++j) {
    AIns x;
    x.op = I_STORE_OUT;
    x.aux = j;
    x.a = opnd(t.outputs[j]);
    if (x.a.imm) x.flags |= F_A_IMM;
    prog.push_back(x);
  }

  if (grad) {
    std::vector<char> adj_def(N, 0);
    if (!is_const(seed)) {
      AIns x;
      x.op = I_SEED;
      x.dst = ADJ(seed);
      adj_def[seed] = 1;
      prog.push_back(x);
    }
    // DOT_REV pieces: a term whose operand is a single-use computed value
    // receives no other adjoint contribution, so its contribution may be
    // applied just before that operand's own reverse instruction instead of at
    // the chain root; its adjoint register then lives only briefly.  Terms
    // keep their relative order.
    std::vector<std::vector<std::pair<uint32_t, std::vector<const Term*>>>> rpieces(N + 1);
    std::vector<std::vector<const Term*>> rev_order(N);
    for (int64_t ii = (int64_t)N - 1; ii >= (int64_t)ni; --ii) {
      uint32_t r = (uint32_t)ii;
      if (!reach[r] || is_const(r) || absorbed[r]) continue;
      const Chain* c = dot_root(r);
      if (!c) continue;
      std::vector<const Term*> order;
      for (const Term& tm : c->terms)
        if (tm.kind != T_IMM) order.push_back(&tm);
      std::stable_sort(order.begin(), order.end(), [](const Term* p, const Term* q) {
        if (p->contrib != q->contrib) return p->contrib > q->contrib;
        return p->sub < q->sub;
      });
      // Only contributions to the same adjoint must keep their order, and a
      // sunk term's adjoint has no other contribution; every other term stays
      // at the root.  Terms keep their stored order within each position.
      std::vector<int64_t> qs;
      for (const Term* tm : order) {
        const uint32_t s_ = tm->slot;
        bool sink = s_ >= ni && uses[s_] == 1 && !is_out[s_] && !absorbed[s_] && reach[s_];
        qs.push_back(sink ? (int64_t)s_ : (int64_t)r);
      }
      std::vector<int64_t> positions(qs);
      std::sort(positions.begin(), positions.end());
      positions.erase(std::unique(positions.begin(), positions.end()), positions.end());
      std::vector<const Term*> group;
      for (int64_t pos : positions) {
        group.clear();
        for (size_t k = 0; k < order.size(); ++k)
          if (qs[k] == pos) group.push_back(order[k]);
        rpieces[(size_t)pos].push_back({r, group});
      }
    }
    auto emit_rpiece = [&](uint32_t r, const std::vector<const Term*>& terms_) {
      const Chain& c = chains[chain_of[r]];
      AIns x;
      x.op = I_DOT_REV;
      x.g = ADJ(r);
      size_t nonimm = 0;
      for (const Term& tm : c.terms) nonimm += tm.kind != T_IMM;
      x.full = terms_.size() == nonimm;
      for (const Term* tm : terms_) {
        ATerm at;
        at.kind = tm->kind;
        at.imm = tm->imm;
        at.const_left = tm->const_left;
        at.vid = ADJ(tm->slot);
        at.depth = (c.nadds - 1) - tm->add_pos;
        at.first = !adj_def[tm->slot];
        adj_def[tm->slot] = 1;
        x.terms.push_back(at);
      }
      prog.push_back(std::move(x));
    };
    for (int64_t ii = (int64_t)N - 1; ii >= (int64_t)ni; --ii) {
      uint32_t i = (uint32_t)ii;
      if (!reach[i] || is_const(i) || absorbed[i]) continue;
      // Pieces of other chains sunk onto this entry add to adj(i), so they go
      // first; then the entry's own reverse (for a chain root: its piece at
      // the root position).
      for (auto& pc : rpieces[i])
        if (pc.first != i) emit_rpiece(pc.first, pc.second);
      if (dot_root(i)) {
        for (auto& pc : rpieces[i])
          if (pc.first == i) emit_rpiece(pc.first, pc.second);
        continue;
      }
      uint32_t op = t.opcodes[i];
      uint32_t a = t.arg0[i], b = t.arg1[i];
      bool accA = !is_const(a);
      bool accB = binop(i) && !is_const(b);
      bool ra, rb, rr;
      rev_reads(op, accA, accB, (int32_t)b, ra, rb, rr);
      AIns x;
      x.op = I_REV_BASE + op;
      x.g = ADJ(i);
      x.a = opnd(a);
      if (x.a.imm) x.flags |= F_A_IMM;
      else if (ra) x.flags |= F_RD_A;
      if (op == OPC_POWI) {
        x.aux = b;
      } else if (binop(i)) {
        x.flags |= F_HAS_B;
        x.b = opnd(b);
        if (x.b.imm) x.flags |= F_B_IMM;
        else if (rb) x.flags |= F_RD_B;
      }
      if (rr) { x.flags |= F_RD_R; x.r = VAL(i); }
      if (accA) {
        x.flags |= F_ACC_A;
        x.adjA = ADJ(a);
        if (!adj_def[a]) { x.flags |= F_FIRST_A; adj_def[a] = 1; }
      }
      if (accB) {
        x.flags |= F_ACC_B;
        x.adjB = ADJ(b);
        if (!adj_def[b]) { x.flags |= F_FIRST_B; adj_def[b] = 1; }
      }
      prog.push_back(std::move(x));
    }
    for (uint32_t i = 0; i < ni; ++i) {
      AIns x;
      x.op = I_STORE_GRAD;
      x.aux = i;
      if (adj_def[i]) x.a.vid = ADJ(i);
      else { x.a.imm = true; x.a.v = 0.0; x.flags |= F_A_IMM; }
      prog.push_back(x);
    }
  }

  if (fuse && dense) group_dense(prog, (int64_t)N, grad ? SIZE_MAX : (dense_chunk ? dense_chunk : DENSE_CHUNK), dual);
  const bool jet_segments = fuse && dense && mode == PM_FORWARD && !dual && dense_chunk == SIZE_MAX;
  if (jet_segments) hoist_member_unaries(prog, 2 * (size_t)N);

  // ---- live ranges ----
  const size_t NV = 2 * (size_t)N;
  std::vector<int64_t> def(NV, -1), last(NV, -1);
  auto rd = [&](int64_t v, int64_t p) { if (v >= 0) last[v] = std::max(last[v], p); };
  auto wr = [&](int64_t v, int64_t p) { if (v >= 0 && def[v] < 0) def[v] = p; };
  for (int64_t p = 0; p < (int64_t)prog.size(); ++p) {
    AIns& x = prog[p];
    if (!x.a.imm) {
      bool read_a = true;
      if (x.op >= I_REV_BASE) read_a = (x.flags & F_RD_A) != 0;
      if (x.op == I_LOAD || x.op == I_SEED) read_a = false;
      if (read_a) rd(x.a.vid, p);
    }
    if (!x.b.imm && (x.flags & F_HAS_B)) {
      bool read_b = true;
      if (x.op >= I_REV_BASE) read_b = (x.flags & F_RD_B) != 0;
      if (read_b) rd(x.b.vid, p);
    }
    rd(x.g, p);
    rd(x.r, p);
    if (x.op == I_LOADN) {
      for (auto& tm : x.terms) wr(tm.vid, p);
    } else if (x.op == I_DOT) {
      for (auto& tm : x.terms) rd(tm.vid, p);
      if (x.flags & F_DOT_CONT) rd(x.dst, p);
      wr(x.dst, p);
    } else if (x.op == I_DOT_REV) {
      for (auto& tm : x.terms) {
        if (tm.first) wr(tm.vid, p);
        else rd(tm.vid, p);
        rd(tm.vid, p);
      }
    } else if (x.op == I_DENSE) {
      for (auto& tm : x.terms) rd(tm.vid, p);
      for (int64_t v : x.members) wr(v, p);
    } else if (x.op == I_DENSE_REV) {
      for (int64_t v : x.members) rd(v, p);
      for (auto& tm : x.terms) {
        if (tm.first) wr(tm.vid, p);
        else rd(tm.vid, p);
        rd(tm.vid, p);
      }
    } else {
      wr(x.dst, p);
      if (x.adjA >= 0) { if (x.flags & F_FIRST_A) wr(x.adjA, p); else rd(x.adjA, p); rd(x.adjA, p); }
      if (x.adjB >= 0) { if (x.flags & F_FIRST_B) wr(x.adjB, p); else rd(x.adjB, p); rd(x.adjB, p); }
    }
  }
  // A value read before any write would be a compiler bug; treat as defined
  // at its first read to stay memory-safe.
  struct Iv { int64_t s, e; size_t v; };
  std::vector<Iv> ivs;
  for (size_t v = 0; v < NV; ++v) {
    if (def[v] < 0 && last[v] < 0) continue;
    int64_t s = def[v] >= 0 ? def[v] : last[v];
    int64_t e = std::max(last[v], s);
    ivs.push_back({s, e, v});
  }
  std::sort(ivs.begin(), ivs.end(), [](const Iv& a, const Iv& b) { return a.s < b.s || (a.s == b.s && a.v < b.v); });
  std::vector<uint32_t> reg(NV, 0);
  std::priority_queue<uint32_t, std::vector<uint32_t>, std::greater<uint32_t>> freeq;
  // (end, register, owner): an in-place op hands its operand's register to
  // its result, so a register is freed only when its current owner ends
  typedef std::tuple<int64_t, uint32_t, size_t> EndReg;
  std::priority_queue<EndReg, std::vector<EndReg>, std::greater<EndReg>> active;
  std::vector<size_t> owner;
  uint32_t nregs = 0;
  for (const Iv& iv : ivs) {
    while (!active.empty() && std::get<0>(active.top()) < iv.s) {
      const uint32_t r0 = std::get<1>(active.top());
      if (owner[r0] == std::get<2>(active.top())) freeq.push(r0);
      active.pop();
    }
    uint32_t r;
    const AIns* dx = def[iv.v] >= 0 ? &prog[(size_t)def[iv.v]] : nullptr;
    if (dx && dx->inplace && dx->dst == (int64_t)iv.v && last[(size_t)dx->a.vid] == iv.s) {
      r = reg[(size_t)dx->a.vid];  // the operand ends here: its register holds the result
    } else if (!freeq.empty()) {
      r = freeq.top();
      freeq.pop();
    } else {
      r = nregs++;
      owner.push_back(SIZE_MAX);
    }
    reg[iv.v] = r;
    owner[r] = iv.v;
    active.push(EndReg{iv.e, r, iv.v});
  }

  // ---- encode ----
  Program out;
  out.nregs = std::max<uint32_t>(nregs, 1);
  out.dot_terms = dot_terms;
  auto put_imm = [&](double v, uint32_t& lo, uint32_t& hi) {
    if (t.f64) {
      uint64_t bits;
      std::memcpy(&bits, &v, 8);
      lo = (uint32_t)bits;
      hi = (uint32_t)(bits >> 32);
    } else {
      float f = (float)v;
      std::memcpy(&lo, &f, 4);
      hi = 0;
    }
  };
  auto R = [&](int64_t v) -> uint32_t { return v >= 0 ? reg[v] : 0u; };
  std::vector<uint32_t> t_meta, t_reg, t_imm;
  for (AIns& x : prog) {
    uint32_t w[NW] = {0, 0, 0, 0, 0, 0, 0, 0};
    w[0] = x.op | (x.flags << 8);
    switch (x.op) {
      case I_LOAD: w[1] = R(x.dst); w[2] = x.aux; break;
      case I_SEED: w[1] = R(x.dst); break;
      case I_STORE_OUT:
      case I_STORE_GRAD:
        w[1] = x.aux;
        if (x.a.imm) put_imm(x.a.v, w[2], w[3]);
        else w[2] = R(x.a.vid);
        break;
      case I_LOADN:
      case I_DOT:
      case I_DOT_REV: {
        w[1] = (x.op == I_DOT) ? R(x.dst) : (x.op == I_DOT_REV ? R(x.g) : 0u);
        w[2] = (uint32_t)x.terms.size();
        w[3] = (uint32_t)t_meta.size();
        const size_t base = t_meta.size();
        for (auto& tm : x.terms) {
          uint32_t m = tm.kind | (tm.first ? M_FIRST : 0u) | (tm.const_left ? M_CONST_LEFT : 0u) |
                       (std::min<uint32_t>(tm.depth, 0xFFFFu) << M_DEPTH_SHIFT);
          t_meta.push_back(m);
          t_reg.push_back((tm.vid >= 0) ? R(tm.vid) : 0u);
          uint32_t lo = 0, hi = 0;
          if (x.op == I_LOADN) lo = (uint32_t)tm.imm;
          else put_imm(tm.imm, lo, hi);
          t_imm.push_back(lo);
          if (t.f64) t_imm.push_back(hi);
        }
        // run lengths: independent blocks of up to 8 product terms (reverse
        // blocks additionally need distinct adjoint registers)
        const size_t T = x.terms.size();
        for (size_t k = 0; k < T; ++k) {
          uint32_t run = 0;
          for (size_t j = k; j < T && run < 8; ++j) {
            bool okk = (x.op == I_LOADN) || x.terms[j].kind == T_PROD;
            if (okk && x.op == I_DOT_REV) {
              for (size_t q2 = k; q2 < j; ++q2)
                if (t_reg[base + q2] == t_reg[base + j]) okk = false;
            }
            if (!okk) break;
            ++run;
          }
          t_meta[base + k] |= run << M_RUN_SHIFT;
        }
        break;
      }
      case I_KINK:
        w[1] = x.aux;
        if (x.a.imm) put_imm(x.a.v, w[2], w[3]);
        else w[2] = R(x.a.vid);
        if (x.flags & F_HAS_B) {
          if (x.b.imm) put_imm(x.b.v, w[4], w[5]);
          else w[4] = R(x.b.vid);
        }
        w[6] = x.kop;
        break;
      case I_DENSE:
      case I_DENSE_REV: {
        const size_t T = x.terms.size(), m = x.members.size();
        w[1] = (uint32_t)m;
        w[2] = (uint32_t)T;
        w[3] = (uint32_t)out.pool.size();
        for (auto& tm : x.terms)
          out.pool.push_back((tm.kind << P_KIND_SHIFT) | (tm.first ? P_FIRST : 0u) |
                             ((dual && x.op == I_DENSE_REV) ? (tm.depth >= 1 ? P_DEEP : 0u)
                                                            : (tm.const_left ? P_CLEFT : 0u)) |
                             (tm.vid >= 0 ? R(tm.vid) : 0u));
        for (int64_t v : x.members) out.pool.push_back(R(v));
        if (t.f64 && (out.pool.size() & 1)) out.pool.push_back(0u);  // 8-byte aligned weights
        w[4] = (uint32_t)out.pool.size();
        // DENSE: [T][m] (a term's weights for all members together);
        // DENSE_REV: [m][T] (a member's weights for all adjoints together)
        bool prod_nz = true;
        bool in_ops = x.op == I_DENSE;
        bool finite_w = x.op == I_DENSE;
        for (auto& tm : x.terms)
          if (tm.kind != T_IMM && !(tm.vid >= 0 && tm.vid < (int64_t)t.num_inputs)) in_ops = false;
        for (size_t a = 0; a < (x.op == I_DENSE ? T : m); ++a)
          for (size_t b = 0; b < (x.op == I_DENSE ? m : T); ++b) {
            const size_t j = x.op == I_DENSE ? b : a, q = x.op == I_DENSE ? a : b;
            double v = x.W[j * T + q];
            // a plain register term adds its operand (forward) or the whole
            // adjoint (reverse): weight 1, so the interpreters run c * x or
            // c * g for every kind
            if (x.terms[q].kind == T_REG || (x.op == I_DENSE_REV && x.terms[q].kind != T_PROD)) v = 1.0;
            if (x.terms[q].kind != T_PROD) prod_nz = false;
            uint32_t lo = 0, hi = 0;
            put_imm(v, lo, hi);
            if ((t.f64 ? v : (double)(float)v) == 0.0) prod_nz = false;
            if (!std::isfinite(t.f64 ? v : (double)(float)v)) in_ops = finite_w = false;
            out.pool.push_back(lo);
            if (t.f64) out.pool.push_back(hi);
          }
        w[5] = (prod_nz ? D_ALL_PROD_NZ : 0u) | (in_ops ? D_INPUT_OPERANDS : 0u) | (finite_w ? D_FINITE_WEIGHTS : 0u);
        if (in_ops) {
          // the operands' input indices, for interpreters that read an
          // operand's value straight from the input row
          w[6] = (uint32_t)out.pool.size();
          for (auto& tm : x.terms) out.pool.push_back(tm.kind == T_IMM ? 0u : (uint32_t)tm.vid);
        }
        break;
      }
      default:
        if (x.op < I_REV_BASE) {
          w[1] = R(x.dst);
        } else {
          w[1] = R(x.g);
          w[6] = R(x.r);
        }
        if (x.a.imm) put_imm(x.a.v, w[2], w[3]);
        else { w[2] = R(x.a.vid); w[3] = R(x.adjA); }
        if (x.flags & F_HAS_B) {
          if (x.b.imm) put_imm(x.b.v, w[4], w[5]);
          else { w[4] = R(x.b.vid); w[5] = R(x.adjB); }
        }
        w[7] = x.aux;
        break;
    }
    out.words.insert(out.words.end(), w, w + NW);
  }
  out.ninstr = (uint32_t)prog.size();
  out.nterms = (uint32_t)t_meta.size();
  out.terms = t_meta;
  out.terms.insert(out.terms.end(), t_reg.begin(), t_reg.end());
  out.terms.insert(out.terms.end(), t_imm.begin(), t_imm.end());
  if (out.terms.empty()) out.terms.assign(4, 0);
  return out;
}


std::vector<uint32_t> encode_stream(const Program& p, bool f64) { return encode_stream(p, f64, 0, p.ninstr); }

std::vector<JetSegment> segment_jet_program(const Program& p) {
  std::vector<JetSegment> out;
  uint32_t i0 = 0;
  for (uint32_t i = 0; i < p.ninstr; ++i) {
    const uint32_t* w = &p.words[(size_t)i * NW];
    if ((w[0] & 0xFFu) != I_DENSE || !(w[5] & D_FINITE_WEIGHTS) || w[1] < JET_GEMM_MIN_MEMBERS) continue;
    if (i > i0) out.push_back(JetSegment{false, i0, i, 0, 0, 0, 0, 0});
    out.push_back(JetSegment{true, i, i + 1, w[1], w[2], w[3], w[4], w[5]});
    i0 = i + 1;
  }
  if (i0 < p.ninstr || out.empty()) out.push_back(JetSegment{false, i0, p.ninstr, 0, 0, 0, 0, 0});
  return out;
}

std::vector<uint32_t> encode_stream(const Program& p, bool f64, uint32_t i_begin, uint32_t i_end) {
  const uint32_t IW = f64 ? 2u : 1u;
  const uint32_t TWW = 2 + IW;                     // words per inline term
  const uint32_t MAXLEN = std::min<uint32_t>(255u, WIN_WORDS - 1);
  const uint32_t max_terms = (MAXLEN - 3) / TWW;
  const uint32_t T = p.nterms;
  std::vector<uint32_t> out;
  out.reserve(p.words.size() + (size_t)T * TWW + WIN_WORDS);
  uint32_t used = 0;  // words used in the current window
  auto need = [&](uint32_t len) {
    // leave room for the closing I_WEND
    if (used + len + 1 > WIN_WORDS) {
      out.push_back(I_WEND);
      ++used;
      while (used < WIN_WORDS) { out.push_back(0u); ++used; }
      used = 0;
    }
  };
  auto emit = [&](const uint32_t* w, uint32_t len) {
    need(len);
    out.insert(out.end(), w, w + len);
    used += len;
  };
  for (uint32_t i = i_begin; i < i_end; ++i) {
    const uint32_t* w = &p.words[(size_t)i * NW];
    const uint32_t op = w[0] & 0xFFu;
    const uint32_t fl = (w[0] >> 8) & 0xFFFFu;
    if (op == I_DOT || op == I_DOT_REV || op == I_LOADN) {
      const uint32_t n = w[2], off = w[3];
      uint32_t done = 0;
      while (done < n || (n == 0 && done == 0)) {
        const uint32_t k = std::min(max_terms, n - done);
        uint32_t pfl = fl;
        if (op == I_DOT && done > 0) pfl |= F_DOT_CONT;
        std::vector<uint32_t> ins;
        ins.push_back(op | (pfl << 8) | ((3 + k * TWW) << 24));
        ins.push_back(w[1]);
        ins.push_back(k);
        for (uint32_t j = 0; j < k; ++j) {
          const uint32_t t = off + done + j;
          uint32_t meta = p.terms[t];
          // runs must not cross a piece boundary
          uint32_t run = (meta >> M_RUN_SHIFT) & 0xFFu;
          run = std::min(run, k - j);
          meta = (meta & ~(0xFFu << M_RUN_SHIFT)) | (run << M_RUN_SHIFT);
          ins.push_back(meta);
          ins.push_back(p.terms[T + t]);
          for (uint32_t q = 0; q < IW; ++q) ins.push_back(p.terms[2 * T + (size_t)t * IW + q]);
        }
        emit(ins.data(), (uint32_t)ins.size());
        done += k;
        if (n == 0) break;
      }
    } else {
      uint32_t ins[NW];
      std::memcpy(ins, w, sizeof ins);
      ins[0] = (ins[0] & 0x00FFFFFFu) | (NW << 24);
      emit(ins, NW);
    }
  }
  need(1);
  out.push_back(I_END);
  ++used;
  while (used < WIN_WORDS) { out.push_back(0u); ++used; }
  return out;
}

}  // namespace ehip
