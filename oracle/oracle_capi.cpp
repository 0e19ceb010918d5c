// CPU ORACLE — test infrastructure only (see echidna_oracle.hpp).
//
// C entry points over the templated restatement, called by the tests through
// ctypes and by bench.py's cpu_baseline leg.  Batch variants spread points
// over `nthreads` std::threads the way the reference spreads them over rayon
// workers (src/bytecode_tape/parallel.rs:240-253).
#include <algorithm>
#include <thread>
#include <vector>

#include "echidna_oracle.hpp"

using namespace orc;

namespace {

TapeView view(const uint32_t* opc, const uint32_t* a0, const uint32_t* a1, const double* vals,
              uint32_t nops, uint32_t nin, const uint32_t* outs, uint32_t nout) {
  TapeView t;
  t.opcodes = opc;
  t.arg0 = a0;
  t.arg1 = a1;
  t.values = vals;
  t.num_ops = nops;
  t.num_inputs = nin;
  t.output_indices = outs;
  t.num_outputs = nout;
  return t;
}

template <class Fn> void par_for(uint64_t B, int nthreads, Fn fn) {
  if (nthreads <= 1 || B < 2) {
    fn(0, B);
    return;
  }
  std::vector<std::thread> th;
  uint64_t chunk = (B + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    uint64_t lo = t * chunk, hi = std::min<uint64_t>(B, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([=] { fn(lo, hi); });
  }
  for (auto& x : th) x.join();
}

template <class F> void forward_batch(const TapeView& t, const F* x, uint64_t B, F* y, int nt) {
  par_for(B, nt, [&](uint64_t lo, uint64_t hi) {
    std::vector<F> v;
    for (uint64_t p = lo; p < hi; ++p) {
      forward_into<F>(t, x + p * t.num_inputs, v);
      for (uint32_t j = 0; j < t.num_outputs; ++j) y[p * t.num_outputs + j] = v[t.output_indices[j]];
    }
  });
}

template <class F> void gradient_batch(const TapeView& t, const F* x, uint64_t B, F* y, F* g, int nt) {
  par_for(B, nt, [&](uint64_t lo, uint64_t hi) {
    std::vector<F> v, adj;
    for (uint64_t p = lo; p < hi; ++p) {
      forward_into<F>(t, x + p * t.num_inputs, v);
      for (uint32_t j = 0; j < t.num_outputs; ++j) y[p * t.num_outputs + j] = v[t.output_indices[j]];
      reverse_from<F>(t, v, t.output_indices[0], adj);
      for (uint32_t i = 0; i < t.num_inputs; ++i) g[p * t.num_inputs + i] = adj[i];
    }
  });
}

template <class F> void hvp_batch(const TapeView& t, const F* x, uint64_t xstride, const F* v, uint64_t B,
                                  F* g, F* hv, int nt) {
  par_for(B, nt, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t p = lo; p < hi; ++p) {
      uint64_t n = t.num_inputs;
      hvp<F>(t, x + p * xstride, v + p * n, g + p * n, hv + p * n);
    }
  });
}

template <class F, int K> void taylor_batch(const TapeView& t, const F* x, uint64_t xstride, const F* v,
                                            uint64_t B, F* out, int nt) {
  par_for(B, nt, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t p = lo; p < hi; ++p) {
      taylor_jets<F, K>(t, x + p * xstride, v + p * t.num_inputs, out + p * t.num_outputs * K);
    }
  });
}

template <class F> int taylor_dispatch(const TapeView& t, const F* x, uint64_t xs, const F* v, uint64_t B,
                                       F* out, int K, int nt) {
  switch (K) {
    case 1: taylor_batch<F, 1>(t, x, xs, v, B, out, nt); return 0;
    case 2: taylor_batch<F, 2>(t, x, xs, v, B, out, nt); return 0;
    case 3: taylor_batch<F, 3>(t, x, xs, v, B, out, nt); return 0;
    case 4: taylor_batch<F, 4>(t, x, xs, v, B, out, nt); return 0;
    case 5: taylor_batch<F, 5>(t, x, xs, v, B, out, nt); return 0;
    default: return 5;
  }
}

}  // namespace

#define TAPE_ARGS                                                                               \
  const uint32_t *opc, const uint32_t *a0, const uint32_t *a1, const double *vals, uint32_t nops, \
      uint32_t nin, const uint32_t *outs, uint32_t nout
#define TAPE view(opc, a0, a1, vals, nops, nin, outs, nout)

extern "C" {

int orc_forward_f64(TAPE_ARGS, const double* x, uint64_t B, double* y, int nt) {
  forward_batch<double>(TAPE, x, B, y, nt);
  return 0;
}
int orc_forward_f32(TAPE_ARGS, const float* x, uint64_t B, float* y, int nt) {
  forward_batch<float>(TAPE, x, B, y, nt);
  return 0;
}
int orc_gradient_f64(TAPE_ARGS, const double* x, uint64_t B, double* y, double* g, int nt) {
  gradient_batch<double>(TAPE, x, B, y, g, nt);
  return 0;
}
int orc_gradient_f32(TAPE_ARGS, const float* x, uint64_t B, float* y, float* g, int nt) {
  gradient_batch<float>(TAPE, x, B, y, g, nt);
  return 0;
}
// Per point: magnitudes of the terms behind every output value, each
// gradient and H v component, and the jet coefficients c1, c2 along v
// (orc::magnitudes).  out: val [B][m], grad [B][n], hv [B][n], c12 [B][2].
int orc_magnitudes_f64(TAPE_ARGS, const double* x, uint64_t xstride, const double* v, uint64_t B, double* val,
                       double* grad, double* hv, double* c12, int nt) {
  TapeView t = TAPE;
  par_for(B, nt, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t p = lo; p < hi; ++p) {
      const uint64_t n = t.num_inputs;
      magnitudes<double>(t, x + p * xstride, v + p * n, val + p * t.num_outputs, grad + p * n, hv + p * n,
                         c12 + 2 * p, c12 + 2 * p + 1);
    }
  });
  return 0;
}

// Per-point x when xstride == nin; a single shared x when xstride == 0.
int orc_hvp_f64(TAPE_ARGS, const double* x, uint64_t xstride, const double* v, uint64_t B, double* g,
                double* hv, int nt) {
  hvp_batch<double>(TAPE, x, xstride, v, B, g, hv, nt);
  return 0;
}
int orc_hvp_f32(TAPE_ARGS, const float* x, uint64_t xstride, const float* v, uint64_t B, float* g,
                float* hv, int nt) {
  hvp_batch<float>(TAPE, x, xstride, v, B, g, hv, nt);
  return 0;
}
int orc_taylor_f64(TAPE_ARGS, const double* x, uint64_t xstride, const double* v, uint64_t B, int K,
                   double* out, int nt) {
  return taylor_dispatch<double>(TAPE, x, xstride, v, B, out, K, nt);
}
int orc_taylor_f32(TAPE_ARGS, const float* x, uint64_t xstride, const float* v, uint64_t B, int K,
                   float* out, int nt) {
  return taylor_dispatch<float>(TAPE, x, xstride, v, B, out, K, nt);
}

// Welford over `n` samples: out = {estimate, sample_variance, standard_error, count}.
void orc_welford_f64(const double* s, uint64_t n, double* out) {
  Welford<double> w;
  for (uint64_t i = 0; i < n; ++i) w.update(s[i]);
  w.finalize(&out[0], &out[1], &out[2]);
  out[3] = (double)w.count;
}
void orc_welford_f32(const float* s, uint64_t n, float* out) {
  Welford<float> w;
  for (uint64_t i = 0; i < n; ++i) w.update(s[i]);
  w.finalize(&out[0], &out[1], &out[2]);
  out[3] = (float)w.count;
}

// laplacian_with_stats (src/stde/laplacian.rs:56-62 via pipeline.rs:19-38):
// res = {value, estimate, sample_variance, standard_error, num_samples}.
int orc_laplacian_with_stats_f64(TAPE_ARGS, const double* x, const double* dirs, uint64_t S, double* res) {
  TapeView t = TAPE;
  Welford<double> w;
  double value = 0.0, c[3];
  for (uint64_t s = 0; s < S; ++s) {
    taylor_jets<double, 3>(t, x, dirs + s * nin, c);
    value = c[0];
    w.update(2.0 * c[2]);
  }
  res[0] = value;
  w.finalize(&res[1], &res[2], &res[3]);
  res[4] = (double)w.count;
  return 0;
}

// laplacian_with_control (src/stde/laplacian.rs:90-129)
int orc_laplacian_with_control_f64(TAPE_ARGS, const double* x, const double* dirs, uint64_t S,
                                   const double* diag, double* res) {
  TapeView t = TAPE;
  double tr = 0.0;
  for (uint32_t i = 0; i < nin; ++i) tr = tr + diag[i];
  Welford<double> w;
  double value = 0.0, c[3];
  for (uint64_t s = 0; s < S; ++s) {
    const double* v = dirs + s * nin;
    taylor_jets<double, 3>(t, x, v, c);
    value = c[0];
    double raw = 2.0 * c[2];
    double cv = 0.0;
    for (uint32_t i = 0; i < nin; ++i) cv = cv + diag[i] * v[i] * v[i];
    w.update(raw - cv + tr);
  }
  res[0] = value;
  w.finalize(&res[1], &res[2], &res[3]);
  res[4] = (double)w.count;
  return 0;
}

// hessian_diagonal (src/stde/laplacian.rs:137-171)
int orc_hessian_diagonal_f64(TAPE_ARGS, const double* x, double* value, double* diag) {
  TapeView t = TAPE;
  std::vector<double> e(nin, 0.0);
  double c[3];
  *value = 0.0;
  for (uint32_t j = 0; j < nin; ++j) {
    e[j] = 1.0;
    taylor_jets<double, 3>(t, x, e.data(), c);
    *value = c[0];
    diag[j] = 2.0 * c[2];
    e[j] = 0.0;
  }
  return 0;
}

// Dense Hessian by n HVPs (src/bytecode_tape/tangent.rs:334-360), row-major [n][n].
int orc_hessian_f64(TAPE_ARGS, const double* x, double* H) {
  TapeView t = TAPE;
  std::vector<double> e(nin, 0.0), g(nin), hv(nin);
  for (uint32_t j = 0; j < nin; ++j) {
    e[j] = 1.0;
    hvp<double>(t, x, e.data(), g.data(), hv.data());
    for (uint32_t i = 0; i < nin; ++i) H[j * nin + i] = hv[i];
    e[j] = 0.0;
  }
  return 0;
}


// Structure.  Sizes first: call with cap 0 to get *nnz.
int orc_hessian_pattern(TAPE_ARGS, uint32_t* rows, uint32_t* cols, uint64_t cap, uint64_t* nnz) {
  std::vector<uint32_t> r, c;
  hessian_pattern(TAPE, r, c);
  *nnz = r.size();
  if (r.size() > cap) return 1;
  std::copy(r.begin(), r.end(), rows);
  std::copy(c.begin(), c.end(), cols);
  return 0;
}

int orc_jacobian_pattern(TAPE_ARGS, uint32_t* rows, uint32_t* cols, uint64_t cap, uint64_t* nnz) {
  std::vector<uint32_t> r, c;
  jacobian_pattern(TAPE, r, c);
  *nnz = r.size();
  if (r.size() > cap) return 1;
  std::copy(r.begin(), r.end(), rows);
  std::copy(c.begin(), c.end(), cols);
  return 0;
}

uint32_t orc_greedy_coloring(uint32_t n, const uint32_t* rows, const uint32_t* cols, uint64_t nnz, uint32_t* colors) {
  std::vector<uint32_t> r(rows, rows + nnz), c(cols, cols + nnz), out;
  uint32_t k = greedy_coloring(n, r, c, out);
  std::copy(out.begin(), out.end(), colors);
  return k;
}

uint32_t orc_intersection_coloring(const uint32_t* gk, const uint32_t* ck, uint64_t nnz, uint32_t gdim, uint32_t cdim,
                                   uint32_t* colors) {
  std::vector<uint32_t> g(gk, gk + nnz), c(ck, ck + nnz), out;
  uint32_t k = intersection_coloring(g, c, gdim, cdim, out);
  std::copy(out.begin(), out.end(), colors);
  return k;
}

// Kinks of B points: values [B][nout], switching [B][K], branch [B][K].
int orc_nonsmooth_f64(TAPE_ARGS, const double* x, uint64_t B, uint32_t K, double* y, double* sw, int8_t* br) {
  TapeView t = TAPE;
  std::vector<double> v, s;
  std::vector<int8_t> b;
  for (uint64_t p = 0; p < B; ++p) {
    forward_nonsmooth<double>(t, x + p * t.num_inputs, v, s, b);
    if (s.size() != K) return 1;
    for (uint32_t j = 0; j < t.num_outputs; ++j) y[p * t.num_outputs + j] = v[t.output_indices[j]];
    std::copy(s.begin(), s.end(), sw + p * K);
    std::copy(b.begin(), b.end(), br + p * K);
  }
  return 0;
}

int orc_nonsmooth_f32(TAPE_ARGS, const float* x, uint64_t B, uint32_t K, float* y, float* sw, int8_t* br) {
  TapeView t = TAPE;
  std::vector<float> v, s;
  std::vector<int8_t> b;
  for (uint64_t p = 0; p < B; ++p) {
    forward_nonsmooth<float>(t, x + p * t.num_inputs, v, s, b);
    if (s.size() != K) return 1;
    for (uint32_t j = 0; j < t.num_outputs; ++j) y[p * t.num_outputs + j] = v[t.output_indices[j]];
    std::copy(s.begin(), s.end(), sw + p * K);
    std::copy(b.begin(), b.end(), br + p * K);
  }
  return 0;
}

}  // extern "C"
