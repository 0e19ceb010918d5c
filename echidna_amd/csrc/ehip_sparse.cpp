// Structure for the sparse derivative paths: Hessian and Jacobian sparsity
// patterns from input-dependency bitsets, and the greedy colorings that
// compress them, bit-exact with the reference (src/bytecode_tape/sparse.rs:
// 463-776, src/sparse.rs:57-413); then the GPU sparse Jacobian (colored
// tangent sweeps on the Taylor order-2 interpreter) and sparse Hessian
// (colored HVPs), src/gpu/cuda_backend.rs:306-441.  Pattern detection runs
// once per tape on the host, as in the reference; the sweeps run on the GPU.
#include "ehip_internal.h"

namespace {

// OpCode discriminants (src/opcode.rs:21-130)
enum : uint32_t {
  OP_INPUT = 0, OP_CONST = 1, OP_ADD = 2, OP_SUB = 3, OP_MUL = 4, OP_DIV = 5, OP_REM = 6, OP_POWF = 7,
  OP_ATAN2 = 8, OP_HYPOT = 9, OP_MAX = 10, OP_MIN = 11, OP_NEG = 12, OP_ABS = 36, OP_SIGNUM = 37, OP_FLOOR = 38,
  OP_CEIL = 39, OP_ROUND = 40, OP_TRUNC = 41, OP_FRACT = 42
};

enum class OpClass { Linear, UnaryNonlinear, BinaryNonlinear, ZeroDerivative };

// classify_op (src/bytecode_tape/sparse.rs:643-700)
OpClass classify(uint32_t op) {
  switch (op) {
    case OP_ADD: case OP_SUB: case OP_NEG: case OP_FRACT: return OpClass::Linear;
    case OP_MUL: case OP_DIV: case OP_POWF: case OP_ATAN2: case OP_HYPOT: return OpClass::BinaryNonlinear;
    case OP_ABS: case OP_SIGNUM: case OP_FLOOR: case OP_CEIL: case OP_ROUND: case OP_TRUNC:
    case OP_MAX: case OP_MIN: case OP_REM: return OpClass::ZeroDerivative;
    default: return OpClass::UnaryNonlinear;  // recip, sqrt, cbrt, powi, exp*, logs, trig, hyperbolic
  }
}

// is_binary_op (src/bytecode_tape/sparse.rs:702-718): powi's second
// argument is its exponent.
bool binary(uint32_t op) { return op >= OP_ADD && op <= OP_MIN; }

struct Bits {
  size_t words;
  std::vector<uint64_t> b;  // [node][words]
  uint64_t* row(size_t i) { return b.data() + i * words; }
  void unite(size_t dst, size_t src) {
    if (dst == src) return;
    uint64_t* d = row(dst);
    const uint64_t* s = row(src);
    for (size_t w = 0; w < words; ++w) d[w] |= s[w];
  }
  std::vector<uint32_t> list(size_t i, size_t n) {
    std::vector<uint32_t> out;
    const uint64_t* r = row(i);
    for (size_t w = 0; w < words; ++w) {
      uint64_t x = r[w];
      while (x) {
        const uint32_t pos = (uint32_t)(w * 64 + __builtin_ctzll(x));
        if (pos < n) out.push_back(pos);
        x &= x - 1;
      }
    }
    return out;
  }
};

Bits dependencies(const TapeHost& t, bool hessian, std::vector<std::pair<uint32_t, uint32_t>>* pairs) {
  const size_t N = t.opcodes.size(), n = t.num_inputs;
  Bits d{(n + 63) / 64, {}};
  d.b.assign(N * d.words, 0);
  auto upper = [&](const std::vector<uint32_t>& bits) {
    for (size_t a = 0; a < bits.size(); ++a)
      for (size_t c = 0; c <= a; ++c)
        pairs->push_back({std::max(bits[a], bits[c]), std::min(bits[a], bits[c])});
  };
  uint32_t input = 0;
  for (size_t i = 0; i < N; ++i) {
    const uint32_t op = t.opcodes[i];
    if (op == OP_INPUT) {
      d.row(i)[input / 64] |= 1ull << (input % 64);
      ++input;
      continue;
    }
    if (op == OP_CONST) continue;
    const size_t a = t.arg0[i], b = t.arg1[i];
    const bool has_b = binary(op) && t.arg1[i] != 0xFFFFFFFFu;
    if (!hessian) {
      d.unite(i, a);
      if (has_b) d.unite(i, b);
      continue;
    }
    switch (classify(op)) {
      case OpClass::Linear:
      case OpClass::ZeroDerivative:
        d.unite(i, a);
        if (has_b) d.unite(i, b);
        break;
      case OpClass::UnaryNonlinear:
        d.unite(i, a);
        upper(d.list(i, n));
        break;
      case OpClass::BinaryNonlinear: {
        const std::vector<uint32_t> ba = d.list(a, n), bb = d.list(b, n);
        d.unite(i, a);
        d.unite(i, b);
        for (uint32_t x : ba)
          for (uint32_t y : bb) pairs->push_back({std::max(x, y), std::min(x, y)});
        if (op != OP_MUL) {
          upper(ba);
          upper(bb);
        }
        break;
      }
    }
  }
  return d;
}

// greedy_distance1_coloring (src/sparse.rs:357-413): vertices by decreasing
// degree (stable), each takes the lowest color unused by its neighbours.
uint32_t distance1(const std::vector<std::vector<uint32_t>>& adj, uint32_t* colors) {
  const size_t n = adj.size();
  std::vector<uint32_t> order(n);
  for (size_t v = 0; v < n; ++v) order[v] = (uint32_t)v;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return adj[a].size() > adj[b].size(); });
  for (size_t v = 0; v < n; ++v) colors[v] = 0xFFFFFFFFu;
  uint32_t nc = 0;
  std::vector<uint32_t> used;
  for (uint32_t v : order) {
    used.clear();
    for (uint32_t u : adj[v])
      if (colors[u] != 0xFFFFFFFFu) used.push_back(colors[u]);
    std::sort(used.begin(), used.end());
    used.erase(std::unique(used.begin(), used.end()), used.end());
    uint32_t c = 0;
    for (uint32_t u : used) {
      if (u != c) break;
      ++c;
    }
    colors[v] = c;
    nc = std::max(nc, c + 1);
  }
  return nc;
}

void sort_dedup(std::vector<uint32_t>& v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
}

// intersection_graph_coloring (src/sparse.rs:318-354)
uint32_t intersection_coloring(const uint32_t* group, const uint32_t* color, uint64_t nnz, uint32_t group_dim,
                               uint32_t color_dim, uint32_t* colors) {
  if (color_dim == 0) return 0;
  std::vector<std::vector<uint32_t>> groups(group_dim), adj(color_dim);
  for (uint64_t k = 0; k < nnz; ++k) groups[group[k]].push_back(color[k]);
  for (const auto& m : groups)
    for (size_t i = 0; i < m.size(); ++i)
      for (size_t j = i + 1; j < m.size(); ++j) {
        adj[m[i]].push_back(m[j]);
        adj[m[j]].push_back(m[i]);
      }
  for (auto& l : adj) sort_dedup(l);
  return distance1(adj, colors);
}

int32_t check_pattern(const uint32_t* rows, const uint32_t* cols, uint64_t nnz, uint32_t nr, uint32_t nc) {
  if (nnz && (!rows || !cols)) return fail(EHIP_OTHER, "GPU error: null pattern arrays");
  for (uint64_t k = 0; k < nnz; ++k)
    if (rows[k] >= nr || cols[k] >= nc)
      return fail(EHIP_OTHER, "GPU error: pattern entry %llu (%u, %u) out of range", (unsigned long long)k, rows[k],
                  cols[k]);
  return EHIP_OK;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

int32_t raw_tape(const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1, uint32_t num_ops,
                 uint32_t num_inputs, const uint32_t* outs, uint32_t nout, TapeHost& th) {
  if (num_ops && (!opcodes || !arg0 || !arg1)) return fail(EHIP_OTHER, "GPU error: null argument");
  if (nout && !outs) return fail(EHIP_OTHER, "GPU error: null argument");
  std::string msg = validate_tape(opcodes, arg0, arg1, num_ops, num_inputs, num_ops, num_ops, nout ? outs[0] : 0,
                                  outs, nout);
  if (!msg.empty()) return fail(EHIP_OTHER, "%s", msg.c_str());
  th.opcodes.assign(opcodes, opcodes + num_ops);
  th.arg0.assign(arg0, arg0 + num_ops);
  th.arg1.assign(arg1, arg1 + num_ops);
  th.num_inputs = num_inputs;
  th.outputs.assign(outs, outs + nout);
  return EHIP_OK;
}

int32_t hessian_pattern(const TapeHost& th, uint32_t* rows, uint32_t* cols, uint64_t cap, uint64_t* nnz) {
  std::vector<std::pair<uint32_t, uint32_t>> pairs;
  (void)dependencies(th, true, &pairs);
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  *nnz = pairs.size();
  if (pairs.size() > cap || (!pairs.empty() && (!rows || !cols)))
    return fail(EHIP_OTHER, "GPU error: pattern buffers too small (need %llu)", (unsigned long long)pairs.size());
  for (size_t k = 0; k < pairs.size(); ++k) {
    rows[k] = pairs[k].first;
    cols[k] = pairs[k].second;
  }
  return ok();
}

int32_t jacobian_pattern(const TapeHost& th, uint32_t* rows, uint32_t* cols, uint64_t cap, uint64_t* nnz) {
  Bits d = dependencies(th, false, nullptr);
  std::vector<uint32_t> r, c;
  for (size_t o = 0; o < th.outputs.size(); ++o)
    for (uint32_t in : d.list(th.outputs[o], th.num_inputs)) {
      r.push_back((uint32_t)o);
      c.push_back(in);
    }
  *nnz = r.size();
  if (r.size() > cap || (!r.empty() && (!rows || !cols)))
    return fail(EHIP_OTHER, "GPU error: pattern buffers too small (need %llu)", (unsigned long long)r.size());
  std::copy(r.begin(), r.end(), rows);
  std::copy(c.begin(), c.end(), cols);
  return ok();
}

}  // namespace

extern "C" {

int32_t ehip_hessian_sparsity(const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1, uint32_t num_ops,
                              uint32_t num_inputs, uint32_t* rows, uint32_t* cols, uint64_t cap, uint64_t* nnz) try {
  if (!nnz) return fail(EHIP_OTHER, "GPU error: null argument");
  TapeHost th;
  int32_t rc = raw_tape(opcodes, arg0, arg1, num_ops, num_inputs, nullptr, 0, th);
  return rc ? rc : hessian_pattern(th, rows, cols, cap, nnz);
} catch (...) {
  return exception_status();
}

int32_t ehip_jacobian_sparsity(const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1, uint32_t num_ops,
                               uint32_t num_inputs, const uint32_t* output_indices, uint32_t num_outputs,
                               uint32_t* rows, uint32_t* cols, uint64_t cap, uint64_t* nnz) try {
  if (!nnz) return fail(EHIP_OTHER, "GPU error: null argument");
  TapeHost th;
  int32_t rc = raw_tape(opcodes, arg0, arg1, num_ops, num_inputs, output_indices, num_outputs, th);
  return rc ? rc : jacobian_pattern(th, rows, cols, cap, nnz);
} catch (...) {
  return exception_status();
}

// greedy_coloring (src/sparse.rs:57-98): distance-2 coloring of the
// interaction graph.
int32_t ehip_greedy_coloring(uint32_t dim, const uint32_t* rows, const uint32_t* cols, uint64_t nnz,
                             uint32_t* colors, uint32_t* num_colors) try {
  if (!num_colors || (dim && !colors)) return fail(EHIP_OTHER, "GPU error: null argument");
  int32_t rc = check_pattern(rows, cols, nnz, dim, dim);
  if (rc) return rc;
  *num_colors = 0;
  if (dim == 0) return ok();
  std::vector<std::vector<uint32_t>> adj(dim), adj2(dim);
  for (uint64_t k = 0; k < nnz; ++k)
    if (rows[k] != cols[k]) {
      adj[rows[k]].push_back(cols[k]);
      adj[cols[k]].push_back(rows[k]);
    }
  for (uint32_t v = 0; v < dim; ++v) {
    for (uint32_t u : adj[v]) adj2[v].push_back(u);
    for (uint32_t u : adj[v])
      for (uint32_t w : adj[u])
        if (w != v) adj2[v].push_back(w);
    sort_dedup(adj2[v]);
  }
  *num_colors = distance1(adj2, colors);
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_column_coloring(uint32_t num_outputs, uint32_t num_inputs, const uint32_t* rows, const uint32_t* cols,
                             uint64_t nnz, uint32_t* colors, uint32_t* num_colors) try {
  if (!num_colors || (num_inputs && !colors)) return fail(EHIP_OTHER, "GPU error: null argument");
  int32_t rc = check_pattern(rows, cols, nnz, num_outputs, num_inputs);
  if (rc) return rc;
  *num_colors = intersection_coloring(rows, cols, nnz, num_outputs, num_inputs, colors);
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_row_coloring(uint32_t num_outputs, uint32_t num_inputs, const uint32_t* rows, const uint32_t* cols,
                          uint64_t nnz, uint32_t* colors, uint32_t* num_colors) try {
  if (!num_colors || (num_outputs && !colors)) return fail(EHIP_OTHER, "GPU error: null argument");
  int32_t rc = check_pattern(rows, cols, nnz, num_outputs, num_inputs);
  if (rc) return rc;
  *num_colors = intersection_coloring(cols, rows, nnz, num_inputs, num_outputs, colors);
  return ok();
} catch (...) {
  return exception_status();
}

}  // extern "C"

namespace {

template <class F>
int32_t sparse_hessian_t(ehip_ctx* c, const ehip_tape* t, const F* x, F* value, F* grad, uint32_t* rows,
                         uint32_t* cols, F* vals, uint64_t cap, uint64_t* nnz) {
  const uint32_t n = t->nin;
  uint64_t m = 0;
  hessian_pattern(t->host, nullptr, nullptr, 0, &m);
  *nnz = m;
  if (m > cap || (m && (!rows || !cols || !vals)))
    return fail(EHIP_OTHER, "GPU error: pattern buffers too small (need %llu)", (unsigned long long)m);
  int32_t rc = hessian_pattern(t->host, rows, cols, cap, nnz);
  if (rc) return rc;
  std::vector<uint32_t> colors(n);
  uint32_t nc = 0;
  rc = ehip_greedy_coloring(n, rows, cols, m, colors.data(), &nc);
  if (rc) return rc;
  if (nc == 0) {
    // no interactions: value and gradient from one fwd+grad point
    return ehip_gradient_batch(c, t, x, 1, value, grad);
  }
  std::vector<F> dirs((size_t)nc * n, F(0)), g((size_t)nc * n), hv((size_t)nc * n);
  for (uint32_t col = 0; col < nc; ++col)
    for (uint32_t i = 0; i < n; ++i)
      if (colors[i] == col) dirs[(size_t)col * n + i] = F(1);
  rc = ehip_hvp_batch(c, t, x, dirs.data(), nc, g.data(), hv.data());
  if (rc) return rc;
  std::copy(g.begin(), g.begin() + n, grad);
  for (uint64_t k = 0; k < m; ++k) vals[k] = hv[(size_t)colors[cols[k]] * n + rows[k]];
  F y;
  std::vector<F> gg(n);
  rc = ehip_gradient_batch(c, t, x, 1, &y, gg.data());
  if (rc) return rc;
  *value = y;
  return ok();
}

template <class F>
int32_t sparse_jacobian_t(ehip_ctx* c, const ehip_tape* t, const F* x, F* outs, uint32_t* rows, uint32_t* cols,
                          F* vals, uint64_t cap, uint64_t* nnz) {
  const uint32_t n = t->nin, no = t->nout;
  uint64_t m = 0;
  jacobian_pattern(t->host, nullptr, nullptr, 0, &m);
  *nnz = m;
  if (m > cap || (m && (!rows || !cols || !vals)))
    return fail(EHIP_OTHER, "GPU error: pattern buffers too small (need %llu)", (unsigned long long)m);
  int32_t rc = jacobian_pattern(t->host, rows, cols, cap, nnz);
  if (rc) return rc;
  std::vector<uint32_t> colors(n);
  uint32_t nc = 0;
  rc = ehip_column_coloring(no, n, rows, cols, m, colors.data(), &nc);
  if (rc) return rc;
  if (nc == 0) return ehip_forward_batch(c, t, x, 1, outs);
  // one order-2 Taylor sweep (value, directional derivative) per color
  std::vector<F> xs((size_t)nc * n), seeds((size_t)nc * n, F(0)), coef(2 * (size_t)nc * no);
  for (uint32_t col = 0; col < nc; ++col)
    for (uint32_t i = 0; i < n; ++i) {
      xs[(size_t)col * n + i] = x[i];
      if (colors[i] == col) seeds[(size_t)col * n + i] = F(1);
    }
  rc = ehip_taylor_kth_batch(c, t, xs.data(), seeds.data(), nc, 2, coef.data());
  if (rc) return rc;
  const F* c1 = coef.data() + (size_t)nc * no;
  for (uint64_t k = 0; k < m; ++k) vals[k] = c1[(size_t)colors[cols[k]] * no + rows[k]];
  for (uint32_t j = 0; j < no; ++j) outs[j] = coef[j];
  return ok();
}

}  // namespace

extern "C" {

// GpuBackend::sparse_hessian (src/gpu/cuda_backend.rs:397-441): pattern and
// distance-2 coloring on the host, one HVP per color on the GPU; entries
// recovered as hvp[color(col)][row].
int32_t ehip_sparse_hessian(ehip_ctx* c, const ehip_tape* t, const void* x, void* value, void* grad, uint32_t* rows,
                            uint32_t* cols, void* vals, uint64_t cap, uint64_t* nnz) try {
  if (!c || !t || !x || !value || !grad || !nnz) return fail(EHIP_OTHER, "GPU error: null argument");
  if (t->nout != 1)
    return fail(EHIP_OTHER, "GPU error: sparse_hessian requires a scalar-output tape; got %u outputs", t->nout);
  if (t->nin == 0) return fail(EHIP_OTHER, "GPU error: tape has no inputs (constant function); evaluate it on the CPU instead");
  if (t->f64)
    return sparse_hessian_t<double>(c, t, (const double*)x, (double*)value, (double*)grad, rows, cols,
                                    (double*)vals, cap, nnz);
  return sparse_hessian_t<float>(c, t, (const float*)x, (float*)value, (float*)grad, rows, cols, (float*)vals, cap,
                                 nnz);
} catch (...) {
  return exception_status();
}

// GpuBackend::sparse_jacobian (src/gpu/cuda_backend.rs:306-395): pattern and
// column coloring on the host, one tangent sweep per color on the GPU;
// entries recovered as tangent_out[color(col)][row].
int32_t ehip_sparse_jacobian(ehip_ctx* c, const ehip_tape* t, const void* x, void* outputs, uint32_t* rows,
                             uint32_t* cols, void* vals, uint64_t cap, uint64_t* nnz) try {
  if (!c || !t || !x || !outputs || !nnz) return fail(EHIP_OTHER, "GPU error: null argument");
  if (t->nin == 0) return fail(EHIP_OTHER, "GPU error: tape has no inputs (constant function); evaluate it on the CPU instead");
  if (t->f64)
    return sparse_jacobian_t<double>(c, t, (const double*)x, (double*)outputs, rows, cols, (double*)vals, cap, nnz);
  return sparse_jacobian_t<float>(c, t, (const float*)x, (float*)outputs, rows, cols, (float*)vals, cap, nnz);
} catch (...) {
  return exception_status();
}

}  // extern "C"
