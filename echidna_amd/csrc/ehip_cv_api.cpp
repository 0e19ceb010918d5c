// Host side of the learned control variate and of batched exact traces
// (include/echidna_hip.h).  No counterpart in the reference; the nearest are
// laplacian_with_control (src/stde/laplacian.rs:90-129) and its GPU twin
// (src/gpu/stde_gpu.rs:135-200), whose diagonal control variate this
// generalises to a learned low-rank one.
//
// Network: x (n) -> tanh(W1 x + b1) (H) -> tanh(W2 h1 + b2) (H) -> W3 h2 + b3
// (n*r) read as L(x) in R^{n x r}, row-major (L[i][j] at i*r + j).  Training
// minimises the batch mean of ||Hv - L L^T v||^2 / n over (x, v, Hv) tuples,
// an unbiased estimate of ||H - L L^T||_F^2 / n for E[v v^T] = I, i.e. of
// the variance the control variate leaves in v^T H v for Gaussian v.
// Parameters live in one flat FP32 vector [W1 | b1 | W2 | b2 | W3 | b3]
// (weights row-major [out][in]); Adam keeps FP32 moments.
#include "ehip_cv.h"
#include "ehip_internal.h"
#include "philox.h"

#include <cmath>

struct ehip_cv {
  ehip_ctx* ctx = nullptr;
  uint32_t n = 0, r = 0, H = 0;
  uint64_t nparams = 0;
  uint64_t off[7] = {};  // W1 b1 W2 b2 W3 b3 end
  DevBuf params, grads, mom1, mom2;
  DevBuf act;       // H1, H2, Lf, dLf, dZ2, dZ1, loss for `cap` tuples
  DevBuf ws;        // split-K partial sums of the dZ2 GEMM (16 slices)
  DevBuf inf;       // inference: H1, H2 for `inf_cap` points
  uint32_t cap = 0, inf_cap = 0;
  uint64_t step = 0;
  float gain = 1.0f;  // L = gain * (W3 h2 + b3)
  uint32_t npos = 0;  // columns with sign +1 (the rest -1); r = L L^T
};

namespace {

constexpr uint32_t CV_CHUNK = 2048;       // tuples per forward/backward pass
constexpr uint32_t CV_INF_CHUNK = 16384;  // points per inference pass (hidden activations only)

struct Act {
  float *H1, *H2, *Lf, *dLf, *dZ2, *dZ1, *loss;
};

int32_t cv_reserve(ehip_cv* cv, uint32_t T) {
  if (T <= cv->cap) return EHIP_OK;
  const size_t per = 4ull * cv->H + 2ull * cv->n * cv->r + 1;
  HIP_TRY(cv->act.reserve(per * T * sizeof(float)), "control variate activations");
  HIP_TRY(cv->ws.reserve(16ull * T * cv->H * sizeof(float)), "control variate workspace");
  cv->cap = T;
  return EHIP_OK;
}

// Hidden activations of an inference pass over up to T points (L goes
// straight to the caller's buffer): larger chunks than training keeps fill
// the chip in the two narrow hidden-layer products.
int32_t cv_reserve_inf(ehip_cv* cv, uint32_t T, Act& a) {
  if (T > cv->inf_cap) {
    HIP_TRY(cv->inf.reserve(2ull * cv->H * T * sizeof(float)), "control variate activations");
    cv->inf_cap = T;
  }
  a = Act{};
  a.H1 = (float*)cv->inf.p;
  a.H2 = a.H1 + (size_t)cv->H * T;
  return EHIP_OK;
}

Act cv_act(ehip_cv* cv, uint32_t T) {
  float* p = (float*)cv->act.p;
  Act a;
  const size_t H = cv->H, nr = (size_t)cv->n * cv->r;
  a.H1 = p;
  a.H2 = a.H1 + H * T;
  a.Lf = a.H2 + H * T;
  a.dLf = a.Lf + nr * T;
  a.dZ2 = a.dLf + nr * T;
  a.dZ1 = a.dZ2 + H * T;
  a.loss = a.dZ1 + H * T;
  return a;
}

float* P(ehip_cv* cv, int k) { return (float*)cv->params.p + cv->off[k]; }
float* G(ehip_cv* cv, int k) { return (float*)cv->grads.p + cv->off[k]; }

Gemm gemm(const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk, int64_t sbn, float* C, int64_t ldc,
          int M, int N, int K, int epi = EPI_NONE, const float* bias = nullptr, const float* aux = nullptr,
          int64_t ldaux = 0) {
  Gemm g{};
  g.A = A;
  g.sam = sam;
  g.sak = sak;
  g.B = B;
  g.sbk = sbk;
  g.sbn = sbn;
  g.C = C;
  g.ldc = ldc;
  g.M = M;
  g.N = N;
  g.K = K;
  g.epi = epi;
  g.bias = bias;
  g.aux = aux;
  g.ldaux = ldaux;
  g.alpha = 1.0f;
  return g;
}

// Forward pass of T <= cap tuples: H1, H2 and L (into `Lout`, row pitch n*r,
// FP32); for_pass: L as bfloat16 written past the caches (a whole batch's
// factors, read once by the reduction, which forms its samples and |L|^2
// from the same rounded values).
int32_t cv_forward(ehip_cv* cv, const float* X, uint32_t T, void* Lout, const Act& a, hipStream_t s,
                   bool for_pass = false) {
  const int n = (int)cv->n, H = (int)cv->H, nr = (int)(cv->n * cv->r);
  // H1[t][j] = tanh(sum_i X[t][i] W1[j][i] + b1[j])
  HIP_TRY(launch_gemm(gemm(X, n, 1, P(cv, 0), 1, n, a.H1, H, (int)T, H, n, EPI_BIAS_TANH, P(cv, 1)), s), "cv forward");
  HIP_TRY(launch_gemm(gemm(a.H1, H, 1, P(cv, 2), 1, H, a.H2, H, (int)T, H, H, EPI_BIAS_TANH, P(cv, 3)), s),
          "cv forward");
  Gemm g3 = gemm(a.H2, H, 1, P(cv, 4), 1, H, (float*)Lout, nr, (int)T, nr, H, EPI_BIAS, P(cv, 5));
  g3.alpha = cv->gain;
  g3.stream_out = for_pass ? 1 : 0;
  g3.out_bf16 = for_pass ? 1 : 0;
  HIP_TRY(launch_gemm(g3, s), "cv forward");
  return EHIP_OK;
}

// Loss and parameter gradients of one chunk; gradients accumulate into
// cv->grads when `accumulate`, with the chunk's loss sum added to loss_dev.
int32_t cv_backward_chunk(ehip_cv* cv, const float* X, const float* V, const float* HV, uint32_t T, uint64_t Ttotal,
                          int accumulate, const Act& a, hipStream_t s) {
  const int n = (int)cv->n, H = (int)cv->H, nr = (int)(cv->n * cv->r);
  int32_t rc = cv_forward(cv, X, T, a.Lf, a, s);
  if (rc) return rc;
  CvLoss l{};
  l.L = a.Lf;
  l.v = V;
  l.hv = HV;
  l.dL = a.dLf;
  l.loss = a.loss;
  l.n = cv->n;
  l.r = cv->r;
  l.npos = cv->npos;
  l.T = T;
  // gradient with respect to the pre-gain output W3 h2 + b3
  l.gscale = -2.0f * cv->gain / ((float)n * (float)Ttotal);
  HIP_TRY(launch_cv_loss(l, s), "cv loss");
  // dW3[j][k] = sum_t dLf[t][j] H2[t][k]; db3 = colsum(dLf)
  Gemm g = gemm(a.dLf, 1, nr, a.H2, H, 1, G(cv, 4), H, nr, H, (int)T);
  g.accumulate = accumulate;
  HIP_TRY(launch_gemm(g, s), "cv backward");
  HIP_TRY(launch_colsum(a.dLf, nr, T, nr, 1.0f, G(cv, 5), accumulate, s), "cv backward");
  // dZ2 = (dLf W3) * (1 - H2^2): K = n r is long and M x N small, so the
  // launcher may split K over the workspace
  g = gemm(a.dLf, nr, 1, P(cv, 4), H, 1, a.dZ2, H, (int)T, H, nr, EPI_DTANH, nullptr, a.H2, H);
  g.part = (float*)cv->ws.p;
  g.part_cap = (int64_t)(cv->ws.n / sizeof(float));
  HIP_TRY(launch_gemm(g, s), "cv backward");
  g = gemm(a.dZ2, 1, H, a.H1, H, 1, G(cv, 2), H, H, H, (int)T);
  g.accumulate = accumulate;
  HIP_TRY(launch_gemm(g, s), "cv backward");
  HIP_TRY(launch_colsum(a.dZ2, H, T, H, 1.0f, G(cv, 3), accumulate, s), "cv backward");
  // dZ1 = (dZ2 W2) * (1 - H1^2)
  HIP_TRY(launch_gemm(gemm(a.dZ2, H, 1, P(cv, 2), H, 1, a.dZ1, H, (int)T, H, H, EPI_DTANH, nullptr, a.H1, H), s),
          "cv backward");
  g = gemm(a.dZ1, 1, H, X, n, 1, G(cv, 0), n, H, n, (int)T);
  g.accumulate = accumulate;
  HIP_TRY(launch_gemm(g, s), "cv backward");
  HIP_TRY(launch_colsum(a.dZ1, H, T, H, 1.0f, G(cv, 1), accumulate, s), "cv backward");
  return EHIP_OK;
}

// Loss and gradients over `T` tuples (chunked); the batch-mean loss lands in
// *loss_dev (device float) when non-null.
int32_t cv_loss_grad(ehip_cv* cv, const float* X, const float* V, const float* HV, uint64_t T, float* loss_dev,
                     hipStream_t s) {
  const uint32_t chunk = (uint32_t)std::min<uint64_t>(T, CV_CHUNK);
  int32_t rc = cv_reserve(cv, chunk);
  if (rc) return rc;
  const Act a = cv_act(cv, chunk);
  const size_t n = cv->n;
  for (uint64_t t0 = 0; t0 < T; t0 += chunk) {
    const uint32_t c = (uint32_t)std::min<uint64_t>(chunk, T - t0);
    rc = cv_backward_chunk(cv, X + t0 * n, V + t0 * n, HV + t0 * n, c, T, t0 > 0, a, s);
    if (rc) return rc;
    if (loss_dev) HIP_TRY(launch_colsum(a.loss, 1, c, 1, 1.0f / (float)T, loss_dev, t0 > 0, s), "cv loss");
  }
  return EHIP_OK;
}

int32_t cv_adam(ehip_cv* cv, float lr, float b1, float b2, float eps, hipStream_t s) {
  cv->step += 1;
  const double bc1 = 1.0 - std::pow((double)b1, (double)cv->step);
  const double bc2 = 1.0 - std::pow((double)b2, (double)cv->step);
  HIP_TRY(launch_adam((float*)cv->params.p, (const float*)cv->grads.p, (float*)cv->mom1.p, (float*)cv->mom2.p,
                      cv->nparams, (float)(lr / bc1), b1, b2, (float)std::sqrt(bc2), eps, s),
          "adam");
  return EHIP_OK;
}

}  // namespace

extern "C" {

int32_t ehip_cv_create(ehip_ctx* c, uint32_t n, uint32_t rank, uint32_t hidden, uint64_t seed, float out_scale,
                       ehip_cv** out) try {
  if (!c || !out) return fail(EHIP_OTHER, "GPU error: null argument");
  *out = nullptr;
  if (n == 0 || hidden == 0) return fail(EHIP_OTHER, "GPU error: control variate needs n > 0 and hidden > 0");
  if (rank == 0 || rank > 32) return fail(EHIP_OTHER, "GPU error: control variate rank must be 1..=32, got %u", rank);
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  ehip_cv* cv = new ehip_cv();
  cv->ctx = c;
  cv->n = n;
  cv->r = rank;
  cv->H = hidden;
  cv->npos = rank;
  const uint64_t H = hidden, nr = (uint64_t)n * rank;
  const uint64_t sizes[6] = {H * n, H, H * H, H, nr * H, nr};
  cv->off[0] = 0;
  for (int k = 0; k < 6; ++k) cv->off[k + 1] = cv->off[k] + sizes[k];
  cv->nparams = cv->off[6];
  // W ~ N(0, 1/fan_in) from the (seed, STREAM_CV) Philox stream at each
  // parameter's flat index; the output layer scaled by out_scale; biases 0.
  std::vector<float> host(cv->nparams, 0.0f);
  const double sd[3] = {1.0 / std::sqrt((double)n), 1.0 / std::sqrt((double)H), out_scale / std::sqrt((double)H)};
  for (int k = 0; k < 3; ++k)
    for (uint64_t i = cv->off[2 * k]; i < cv->off[2 * k + 1]; ++i)
      host[i] = (float)(sd[k] * philox_draw(RNG_GAUSSIAN, seed, STREAM_CV, i));
  const size_t bytes = cv->nparams * sizeof(float);
  hipError_t e = cv->params.reserve(bytes);
  if (e == hipSuccess) e = cv->grads.reserve(bytes);
  if (e == hipSuccess) e = cv->mom1.reserve(bytes);
  if (e == hipSuccess) e = cv->mom2.reserve(bytes);
  if (e == hipSuccess) e = hipMemcpy(cv->params.p, host.data(), bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(cv->grads.p, 0, bytes);
  if (e == hipSuccess) e = hipMemset(cv->mom1.p, 0, bytes);
  if (e == hipSuccess) e = hipMemset(cv->mom2.p, 0, bytes);
  if (e != hipSuccess) {
    cv->params.release();
    cv->grads.release();
    cv->mom1.release();
    cv->mom2.release();
    delete cv;
    return hip_fail(e, "control variate create");
  }
  *out = cv;
  return ok();
} catch (...) {
  return exception_status();
}

void ehip_cv_destroy(ehip_cv* cv) try {
  if (!cv) return;
  std::lock_guard<std::mutex> lk(cv->ctx->mu);
  (void)hipSetDevice(cv->ctx->device);
  (void)hipStreamSynchronize(cv->ctx->stream);
  cv->params.release();
  cv->grads.release();
  cv->mom1.release();
  cv->mom2.release();
  cv->act.release();
  cv->ws.release();
  cv->inf.release();
  delete cv;
} catch (...) {
  (void)exception_status();
}

int32_t ehip_cv_info(const ehip_cv* cv, uint64_t* out5) try {
  if (!cv || !out5) return fail(EHIP_OTHER, "GPU error: null argument");
  out5[0] = cv->n;
  out5[1] = cv->r;
  out5[2] = cv->H;
  out5[3] = cv->nparams;
  out5[4] = cv->step;
  return ok();
} catch (...) {
  return exception_status();
}

// which: 0 parameters, 1 gradients, 2 first moments, 3 second moments.
int32_t ehip_cv_get(ehip_cv* cv, int32_t which, float* host_out) try {
  if (!cv || !host_out || which < 0 || which > 3) return fail(EHIP_OTHER, "GPU error: bad argument");
  std::lock_guard<std::mutex> lk(cv->ctx->mu);
  (void)hipSetDevice(cv->ctx->device);
  DevBuf* b[4] = {&cv->params, &cv->grads, &cv->mom1, &cv->mom2};
  HIP_TRY(hipStreamSynchronize(cv->ctx->stream), "cv get");
  HIP_TRY(hipMemcpy(host_out, b[which]->p, cv->nparams * sizeof(float), hipMemcpyDeviceToHost), "cv get");
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_cv_set_gain(ehip_cv* cv, float gain) try {
  if (!cv) return fail(EHIP_OTHER, "GPU error: null argument");
  if (!(gain > 0.0f) || !std::isfinite(gain)) return fail(EHIP_OTHER, "GPU error: gain must be positive and finite");
  std::lock_guard<std::mutex> lk(cv->ctx->mu);
  cv->gain = gain;
  return ok();
} catch (...) {
  return exception_status();
}

float ehip_cv_gain(const ehip_cv* cv) { return cv ? cv->gain : 0.0f; }

int32_t ehip_cv_set_signs(ehip_cv* cv, uint32_t npos) try {
  if (!cv) return fail(EHIP_OTHER, "GPU error: null argument");
  if (npos > cv->r) return fail(EHIP_OTHER, "GPU error: npos must be <= rank (%u), got %u", cv->r, npos);
  std::lock_guard<std::mutex> lk(cv->ctx->mu);
  cv->npos = npos;
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_cv_set_params(ehip_cv* cv, const float* host_in) try {
  if (!cv || !host_in) return fail(EHIP_OTHER, "GPU error: null argument");
  std::lock_guard<std::mutex> lk(cv->ctx->mu);
  (void)hipSetDevice(cv->ctx->device);
  HIP_TRY(hipStreamSynchronize(cv->ctx->stream), "cv set");
  HIP_TRY(hipMemcpy(cv->params.p, host_in, cv->nparams * sizeof(float), hipMemcpyHostToDevice), "cv set");
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_cv_forward_dev(ehip_cv* cv, const float* d_x, uint64_t batch, float* d_L, void* stream) try {
  if (!cv || !d_x || !d_L) return fail(EHIP_OTHER, "GPU error: null argument");
  if (batch == 0) return ok();
  std::lock_guard<std::mutex> lk(cv->ctx->mu);
  (void)hipSetDevice(cv->ctx->device);
  hipStream_t s = pick(cv->ctx, stream);
  const uint32_t chunk = (uint32_t)std::min<uint64_t>(batch, CV_INF_CHUNK);
  Act a;
  int32_t rc = cv_reserve_inf(cv, chunk, a);
  if (rc) return rc;
  const size_t n = cv->n, nr = (size_t)cv->n * cv->r;
  for (uint64_t t0 = 0; t0 < batch; t0 += chunk) {
    const uint32_t c = (uint32_t)std::min<uint64_t>(chunk, batch - t0);
    rc = cv_forward(cv, d_x + t0 * n, c, d_L + t0 * nr, a, s);
    if (rc) return rc;
  }
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_cv_loss_grad_dev(ehip_cv* cv, const float* d_x, const float* d_v, const float* d_hv, uint64_t batch,
                              float* d_loss, void* stream) try {
  if (!cv || !d_x || !d_v || !d_hv) return fail(EHIP_OTHER, "GPU error: null argument");
  if (batch == 0) return fail(EHIP_OTHER, "GPU error: batch_size must be non-zero");
  std::lock_guard<std::mutex> lk(cv->ctx->mu);
  (void)hipSetDevice(cv->ctx->device);
  int32_t rc = cv_loss_grad(cv, d_x, d_v, d_hv, batch, d_loss, pick(cv->ctx, stream));
  return rc ? rc : ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_cv_adam_step(ehip_cv* cv, float lr, float beta1, float beta2, float eps, void* stream) try {
  if (!cv) return fail(EHIP_OTHER, "GPU error: null argument");
  std::lock_guard<std::mutex> lk(cv->ctx->mu);
  (void)hipSetDevice(cv->ctx->device);
  int32_t rc = cv_adam(cv, lr, beta1, beta2, eps, pick(cv->ctx, stream));
  return rc ? rc : ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_cv_train_dev(ehip_cv* cv, const float* d_x, const float* d_v, const float* d_hv, uint64_t ntuples,
                          uint32_t batch, uint32_t epochs, float lr, float* d_loss_hist, void* stream) try {
  if (!cv || !d_x || !d_v || !d_hv) return fail(EHIP_OTHER, "GPU error: null argument");
  if (ntuples == 0 || batch == 0) return fail(EHIP_OTHER, "GPU error: ntuples and batch must be non-zero");
  std::lock_guard<std::mutex> lk(cv->ctx->mu);
  (void)hipSetDevice(cv->ctx->device);
  hipStream_t s = pick(cv->ctx, stream);
  const size_t n = cv->n;
  const uint64_t steps = (ntuples + batch - 1) / batch;
  for (uint32_t e = 0; e < epochs; ++e) {
    for (uint64_t k = 0; k < steps; ++k) {
      const uint64_t t0 = k * batch;
      const uint64_t T = std::min<uint64_t>(batch, ntuples - t0);
      float* lh = d_loss_hist ? d_loss_hist + (uint64_t)e * steps + k : nullptr;
      int32_t rc = cv_loss_grad(cv, d_x + t0 * n, d_v + t0 * n, d_hv + t0 * n, T, lh, s);
      if (rc) return rc;
      rc = cv_adam(cv, lr, 0.9f, 0.999f, 1e-8f, s);
      if (rc) return rc;
    }
  }
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_cv_make_tuples_dev(ehip_ctx* c, const ehip_tape* t, uint64_t ntuples, uint64_t seed, float* d_x,
                                float* d_v, float* d_hv, void* stream) try {
  if (!c || !t || !d_x || !d_v || !d_hv) return fail(EHIP_OTHER, "GPU error: null argument");
  int32_t rc = check_batch(t, ntuples);
  if (rc) return rc;
  if (t->nout != 1)
    return fail(EHIP_OTHER, "GPU error: control variate tuples need a scalar-output tape; got %u outputs", t->nout);
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = pick(c, stream);
  const uint64_t cnt = ntuples * t->nin;
  const size_t Sb = t->f64 ? 8 : 4;
  // x ~ N(0, I) from (seed, STREAM_POINTS), v ~ N(0, I) from (seed, STREAM_DIRS)
  void *x = d_x, *v = d_v, *g = nullptr, *hv = d_hv;
  if (t->f64) {
    HIP_TRY(c->tup[0].reserve(cnt * Sb), "tuple scratch");
    HIP_TRY(c->tup[1].reserve(cnt * Sb), "tuple scratch");
    HIP_TRY(c->tup[3].reserve(cnt * Sb), "tuple scratch");
    x = c->tup[0].p;
    v = c->tup[1].p;
    hv = c->tup[3].p;
  }
  HIP_TRY(c->tup[2].reserve(cnt * Sb), "tuple scratch");
  g = c->tup[2].p;
  HIP_TRY(launch_philox_fill(x, t->f64, cnt, RNG_GAUSSIAN, seed, STREAM_POINTS, 0, s), "tuple draw");
  HIP_TRY(launch_philox_fill(v, t->f64, cnt, RNG_GAUSSIAN, seed, STREAM_DIRS, 0, s), "tuple draw");
  Launch L{};
  L.nin = t->nin;
  L.nout = 1;
  L.B = ntuples;
  L.x = x;
  L.xstride = t->nin;
  L.dirs = v;
  L.g = g;
  L.h = hv;
  rc = run(c, t, 1, 2, L, s);
  if (rc) return rc;
  if (t->f64) {
    HIP_TRY(launch_to_f32(x, true, cnt, d_x, s), "tuple convert");
    HIP_TRY(launch_to_f32(v, true, cnt, d_v, s), "tuple convert");
    HIP_TRY(launch_to_f32(hv, true, cnt, d_hv, s), "tuple convert");
    HIP_TRY(hipStreamSynchronize(s), "tuple convert");  // scratch is reused by later calls
  }
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_laplacian_cv_batch_dev(ehip_ctx* c, const ehip_tape* t, ehip_cv* cv, const void* d_x, uint64_t npoints,
                                    uint32_t S, uint32_t dir_kind, uint64_t seed, double* d_results,
                                    double* d_plain_results, void* stream) try {
  if (!c || !t || !d_x || !d_results) return fail(EHIP_OTHER, "GPU error: null argument");
  if (t->nout != 1)
    return fail(EHIP_OTHER, "GPU error: laplacian_gpu requires a scalar-output tape; got %u outputs", t->nout);
  if (t->nin == 0) return fail(EHIP_OTHER, "GPU error: laplacian_gpu requires a non-empty x");
  if (S == 0) return fail(EHIP_OTHER, "GPU error: no directions provided");
  if (dir_kind != RNG_RADEMACHER && dir_kind != RNG_GAUSSIAN)
    return fail(EHIP_OTHER, "GPU error: the learned control variate needs device-drawn directions (1 Rademacher, "
                            "2 Gaussian)");
  if (cv && cv->n != t->nin)
    return fail(EHIP_OTHER, "GPU error: control variate built for n=%u, tape has %u inputs", cv->n, t->nin);
  if (npoints == 0) return ok();
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = pick(c, stream);
  const size_t Sb = t->f64 ? 8 : 4;
  const uint64_t lanes = npoints * (uint64_t)S;
  HIP_TRY(c->jets.reserve(lanes * 3 * Sb), "jet scratch");
  // With a control variate the reduction needs every sample's direction
  // again (u = L^T v): the jets write the directions they draw to a buffer
  // that the reduction reads.
  void* dirs = nullptr;
  if (cv) {
    HIP_TRY(c->cvdirs.reserve(lanes * t->nin * Sb), "directions");
    dirs = c->cvdirs.p;
  }
  Launch L{};
  L.nin = t->nin;
  L.nout = 1;
  L.B = lanes;
  L.x = d_x;
  L.xstride = t->nin;
  L.S = S;
  L.rng = dir_kind;
  L.dirs_out = dirs;
  L.seed = seed;
  L.y = c->jets.p;
  int32_t rc = run(c, t, 2, 3, L, s);
  if (rc) return rc;
  StdeReduce R{};
  R.jets = c->jets.p;
  R.npoints = npoints;
  R.S = S;
  R.nin = t->nin;
  R.rng = dirs ? RNG_NONE : dir_kind;
  R.dirs_cmajor = dirs;
  R.seed = seed;
  R.out = d_results;
  R.out_plain = d_plain_results;
  if (cv) {
    // L(x) for every point: FP32 inputs for the network
    const uint64_t cnt = npoints * t->nin;
    const float* xf = (const float*)d_x;
    if (t->f64) {
      HIP_TRY(c->cvx.reserve(cnt * sizeof(float)), "cv inputs");
      HIP_TRY(launch_to_f32(d_x, true, cnt, (float*)c->cvx.p, s), "cv inputs");
      xf = (const float*)c->cvx.p;
    }
    HIP_TRY(c->cvl.reserve(npoints * (uint64_t)cv->n * cv->r * 2), "cv factors");
    uint16_t* Lp = (uint16_t*)c->cvl.p;  // bfloat16
    const uint32_t chunk = (uint32_t)std::min<uint64_t>(npoints, CV_INF_CHUNK);
    Act a;
    rc = cv_reserve_inf(cv, chunk, a);
    if (rc) return rc;
    for (uint64_t t0 = 0; t0 < npoints; t0 += chunk) {
      const uint32_t cc = (uint32_t)std::min<uint64_t>(chunk, npoints - t0);
      rc = cv_forward(cv, xf + t0 * t->nin, cc, Lp + t0 * (uint64_t)cv->n * cv->r, a, s, true);
      if (rc) return rc;
    }
    R.lcv = Lp;
    R.rank = cv->r;
    R.npos = cv->npos;
  }
  HIP_TRY(launch_stde_reduce(t->f64, R, s), "stde reduce");
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_hessian_trace_batch_dev(ehip_ctx* c, const ehip_tape* t, const void* d_x, uint64_t npoints,
                                     double* d_results, void* stream) try {
  if (!c || !t || !d_x || !d_results) return fail(EHIP_OTHER, "GPU error: null argument");
  if (t->nout != 1)
    return fail(EHIP_OTHER, "GPU error: hessian trace requires a scalar-output tape; got %u outputs", t->nout);
  if (t->nin == 0) return fail(EHIP_OTHER, "GPU error: hessian trace requires a non-empty x");
  if (npoints == 0) return ok();
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = pick(c, stream);
  const size_t Sb = t->f64 ? 8 : 4;
  const uint32_t S = t->nin;
  const uint64_t lanes = npoints * (uint64_t)S;
  HIP_TRY(c->jets.reserve(lanes * 3 * Sb), "jet scratch");
  Launch L{};
  L.nin = t->nin;
  L.nout = 1;
  L.B = lanes;
  L.x = d_x;
  L.xstride = t->nin;
  L.S = S;
  L.rng = RNG_BASIS;
  L.y = c->jets.p;
  int32_t rc = run(c, t, 2, 3, L, s);
  if (rc) return rc;
  StdeReduce R{};
  R.jets = c->jets.p;
  R.npoints = npoints;
  R.S = S;
  R.nin = t->nin;
  R.rng = RNG_BASIS;
  R.sample_scale = (double)S;  // mean of n * H_ii over the n basis samples = tr(H)
  R.out = d_results;
  HIP_TRY(launch_stde_reduce(t->f64, R, s), "stde reduce");
  return ok();
} catch (...) {
  return exception_status();
}

}  // extern "C"
