// Host side of the C ABI in include/echidna_hip.h.
//
// Mirrors the reference's CUDA backend host (src/gpu/cuda_backend.rs):
// context acquisition, tape upload with GpuTapeData::validate
// (src/gpu/mod.rs:397-564), validate_batch_args (src/gpu/mod.rs:572-582),
// and the batched entry points.  Unlike the reference, a tape is compiled
// once at upload into register programs (ehip_program.h) and the per-call
// scratch lives in a grow-only context workspace instead of per-call
// allocations.
#include "ehip_internal.h"


extern "C" {

const char* ehip_last_error(void) { return g_last_error.c_str(); }
const char* ehip_version(void) { return "echidna-hip 0.1 (gfx950)"; }

int32_t ehip_ctx_create(int32_t device, ehip_ctx** out) try {
  if (!out) return fail(EHIP_OTHER, "GPU error: null output pointer");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(EHIP_NO_DEVICE, "no suitable GPU device found");
  if (device < 0 || device >= n) return fail(EHIP_NO_DEVICE, "no suitable GPU device found (device %d of %d)", device, n);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(EHIP_NO_DEVICE, "no suitable GPU device found");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(EHIP_NO_DEVICE, "no suitable GPU device found (device %d is %s; the library targets gfx950)", device,
                prop.gcnArchName);
  HIP_TRY(hipSetDevice(device), "hipSetDevice");
  // the allocation-free entry point's scratch on this device; a failure is
  // reported by that entry point, the rest of the library does not need it
  (void)noalloc_setup(device);
  ehip_ctx* c = new ehip_ctx();
  c->device = device;
  c->max_lds = prop.sharedMemPerBlock ? prop.sharedMemPerBlock : 65536;
  c->ncu = prop.multiProcessorCount > 0 ? (uint32_t)prop.multiProcessorCount : 256;
  // A blocking stream: work enqueued on it waits for what the caller put on
  // the legacy default stream before (torch's default stream, for one), so
  // a device input produced there is complete when a kernel reads it.
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamDefault);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "stream create");
  }
  *out = c;
  return ok();
} catch (...) {
  return exception_status();
}

void ehip_ctx_destroy(ehip_ctx* c) try {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
  }
  c->regs.release();
  c->jets.release();
  c->qjobs.release();
  c->qtiles.release();
  c->qcount.release();
  for (auto& b : c->tup) b.release();
  c->cvx.release();
  c->cvl.release();
  c->cvdirs.release();
  for (auto& b : c->io) b.release();
  delete c;
} catch (...) {
  (void)exception_status();
}

int32_t ehip_set_launch(ehip_ctx* c, uint32_t wg, uint32_t lds_budget) try {
  if (!c) return fail(EHIP_OTHER, "GPU error: null context");
  if (wg == 0) wg = 256;
  if (wg != 64 && wg != 128 && wg != 256) return fail(EHIP_OTHER, "GPU error: workgroup size must be 64, 128 or 256");
  std::lock_guard<std::mutex> lk(c->mu);
  c->wg = wg;
  c->lds_budget = lds_budget ? lds_budget : 24576;
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_set_fetch(ehip_ctx* c, int32_t mode) try {
  if (!c) return fail(EHIP_OTHER, "GPU error: null context");
  if (mode != 0 && mode != 1) return fail(EHIP_OTHER, "GPU error: program fetch must be 0 (LDS) or 1 (scalar)");
  std::lock_guard<std::mutex> lk(c->mu);
  c->fetch = (uint32_t)mode;
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_set_dense(ehip_ctx* c, int32_t enable) try {
  if (!c) return fail(EHIP_OTHER, "GPU error: null context");
  std::lock_guard<std::mutex> lk(c->mu);
  c->dense = enable != 0;
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_set_jet_mma(ehip_ctx* c, int32_t enable) try {
  if (!c) return fail(EHIP_OTHER, "GPU error: null context");
  std::lock_guard<std::mutex> lk(c->mu);
  c->jet_mma = enable != 0;
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_set_fusion(ehip_ctx* c, int32_t enable) try {
  if (!c) return fail(EHIP_OTHER, "GPU error: null context");
  std::lock_guard<std::mutex> lk(c->mu);
  c->fuse = enable != 0;
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_tape_upload(ehip_ctx* c, int32_t dtype, const uint32_t* opcodes, const uint32_t* arg0,
                         const uint32_t* arg1, const void* constants, uint32_t num_ops, uint32_t num_inputs,
                         uint32_t num_variables, uint32_t output_index, const uint32_t* output_indices,
                         uint32_t num_outputs, ehip_tape** out) try {
  if (!c || !out) return fail(EHIP_OTHER, "GPU error: null argument");
  *out = nullptr;
  if (dtype != EHIP_F32 && dtype != EHIP_F64) return fail(EHIP_OTHER, "GPU error: dtype must be EHIP_F32 or EHIP_F64");
  if (num_ops > 0 && (!opcodes || !arg0 || !arg1 || !constants))
    return fail(EHIP_OTHER, "GPU error: null tape array");
  for (uint32_t i = 0; i < num_ops; ++i)
    if (opcodes[i] == 43u) return fail(EHIP_CUSTOM_OPS_NOT_SUPPORTED, "tape contains custom ops which cannot run on GPU");
  std::vector<uint32_t> outs;
  if (num_outputs == 0) outs.push_back(output_index);
  else outs.assign(output_indices, output_indices + num_outputs);
  std::string msg = validate_tape(opcodes, arg0, arg1, num_ops, num_inputs, num_variables, num_variables,
                                  output_index, outs.data(), (uint32_t)outs.size());
  if (!msg.empty()) return fail(EHIP_OTHER, "%s", msg.c_str());
  if (num_ops == 0) return fail(EHIP_OTHER, "invalid GpuTapeData: output_index (%u) >= num_variables (0)", output_index);
  ehip_tape* t = new ehip_tape();
  t->ctx = c;
  t->f64 = dtype == EHIP_F64;
  t->nin = num_inputs;
  t->nout = (uint32_t)outs.size();
  t->host.opcodes.assign(opcodes, opcodes + num_ops);
  t->host.arg0.assign(arg0, arg0 + num_ops);
  t->host.arg1.assign(arg1, arg1 + num_ops);
  t->host.values.resize(num_ops);
  if (t->f64) {
    const double* v = (const double*)constants;
    for (uint32_t i = 0; i < num_ops; ++i) t->host.values[i] = v[i];
  } else {
    const float* v = (const float*)constants;
    for (uint32_t i = 0; i < num_ops; ++i) t->host.values[i] = (double)v[i];
  }
  t->host.num_inputs = num_inputs;
  t->host.outputs = outs;
  t->host.f64 = t->f64;
  bool fuse, dense;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    fuse = c->fuse;
    dense = c->dense;
  }
  t->prog[0].host = compile_program(t->host, PM_FORWARD, fuse, dense);
  t->prog[1].host = compile_program(t->host, PM_GRADIENT, fuse, dense);
  t->prog[2].host = compile_program(t->host, PM_GRADIENT, fuse, dense, true);
  t->prog[3].host = compile_program(t->host, PM_KINK, fuse, dense);
  for (uint32_t i = t->host.num_inputs; i < num_ops; ++i)
    if (is_nonsmooth_op(t->host.opcodes[i])) t->nkinks++;
  (void)hipSetDevice(c->device);
  for (int m = 0; m < 4; ++m) {
    int32_t rc = upload_program(t->prog[m], t->f64);
    if (rc != EHIP_OK) {
      for (auto& p : t->prog) free_program(p);
      delete t;
      return rc;
    }
  }
  {
    int32_t rc = upload_jet_segments(t->jseg, t->host, fuse, dense);
    if (rc != EHIP_OK) {
      for (auto& p : t->prog) free_program(p);
      free_jet_segments(t->jseg);
      delete t;
      return rc;
    }
  }
  uint32_t smax, smask;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    smax = c->spec_max;
    smask = c->spec_mask;
  }
  if (num_ops <= smax)
    for (int32_t what = 0; what < 8; ++what) {
      if (!(smask & (1u << what)) || (what == 2 && t->nout != 1)) continue;
      // a failed build leaves this evaluation on the interpreters, which
      // run every tape; the reason is kept on the tape
      uint32_t info[5];
      if (ehip_tape_specialize(t, what, 0, info) != EHIP_OK) t->spec_error[what] = g_last_error;
    }
  *out = t;
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_set_specialize(ehip_ctx* c, uint32_t max_entries, uint32_t what_mask) try {
  if (!c) return fail(EHIP_OTHER, "GPU error: null argument");
  if (what_mask > 255)
    return fail(EHIP_OTHER, "GPU error: what_mask holds bits 0 forward, 1 gradient, 2 hvp, 3..7 taylor order 1..5");
  if (max_entries > SPEC_MAX_ENTRIES)
    return fail(EHIP_OTHER, "GPU error: max_entries above the specialiser's limit (%u)", SPEC_MAX_ENTRIES);
  std::lock_guard<std::mutex> lk(c->mu);
  c->spec_max = what_mask ? max_entries : 0;
  c->spec_mask = what_mask;
  return ok();
} catch (...) {
  return exception_status();
}

void ehip_tape_free(ehip_tape* t) try {
  if (!t) return;
  if (t->ctx) (void)hipSetDevice(t->ctx->device);
  for (auto& p : t->prog) free_program(p);
  free_jet_segments(t->jseg);
  for (auto& k : t->spec) spec_free(k);
  for (auto& k : t->retired) spec_free(k);
  delete t;
} catch (...) {
  (void)exception_status();
}

const char* ehip_spec_error(const ehip_tape* t, int32_t what) {
  return (t && what >= 0 && what < 8) ? t->spec_error[what].c_str() : "";
}

uint32_t ehip_num_outputs(const ehip_tape* t) { return t ? t->nout : 0; }
uint32_t ehip_num_inputs(const ehip_tape* t) { return t ? t->nin : 0; }
int32_t ehip_tape_dtype(const ehip_tape* t) { return t ? (t->f64 ? EHIP_F64 : EHIP_F32) : -1; }

int32_t ehip_program_info(const ehip_tape* t, int32_t mode, uint64_t* out4) try {
  if (!t || !out4 || mode < 0 || mode > 4) return fail(EHIP_OTHER, "GPU error: bad argument");
  if (mode == 4) {
    const JetSegProgram& js = t->jseg;
    out4[0] = js.direct_gemms;
    out4[1] = js.direct_dropped;
    out4[2] = out4[3] = 0;
    for (const JetSegDev& sg : js.segs) out4[2] += sg.skip_direct ? 1u : 0u;
    return ok();
  }
  if (mode == 3) {
    const JetSegProgram& js = t->jseg;
    out4[0] = js.segs.size();
    out4[1] = out4[2] = out4[3] = 0;
    for (const JetSegDev& sg : js.segs) {
      out4[1] += sg.gemm ? 1u : 0u;
      out4[2] += (sg.gemm && sg.g.off_epi != ~0u) ? 1u : 0u;
      out4[3] += (sg.gemm && sg.g.off_ro != ~0u) ? 1u : 0u;
    }
    return ok();
  }
  const Program& p = t->prog[mode].host;
  out4[0] = p.ninstr;
  out4[1] = p.nregs;
  out4[2] = p.words.size();
  out4[3] = p.dot_terms;
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_compile_program(int32_t dtype, const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1,
                             const double* constants, uint32_t num_ops, uint32_t num_inputs,
                             uint32_t output_index, const uint32_t* output_indices, uint32_t num_outputs,
                             int32_t mode, int32_t fuse, uint32_t* words, uint64_t words_cap, uint32_t* terms,
                             uint64_t terms_cap, uint32_t* pool, uint64_t pool_cap, uint64_t* out5) try {
  uint64_t* out4 = out5;
  if (!out4 || mode < 0 || mode > 15 || (mode & 5) == 4 || (mode & 9) == 9 || fuse < 0 || fuse > 2 ||
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
  TapeHost th;
  th.opcodes.assign(opcodes, opcodes + num_ops);
  th.arg0.assign(arg0, arg0 + num_ops);
  th.arg1.assign(arg1, arg1 + num_ops);
  th.values.assign(constants, constants + num_ops);
  if (dtype == EHIP_F32)
    for (auto& v : th.values) v = (double)(float)v;
  th.num_inputs = num_inputs;
  th.outputs = outs;
  th.f64 = dtype == EHIP_F64;
  Program p = compile_program(th, (mode & 1) == 0 ? PM_FORWARD : PM_GRADIENT, fuse != 0, fuse == 1, (mode & 4) != 0,
                              (mode & 8) ? SIZE_MAX : 0);
  out4[0] = p.ninstr;
  out4[1] = p.nregs;
  out4[3] = p.nterms;
  out5[4] = p.pool.size();
  std::vector<uint32_t> st;
  if (mode & 2) st = encode_stream(p, th.f64);  // the windowed stream the kernels execute
  const std::vector<uint32_t>& wv = (mode & 2) ? st : p.words;
  out4[2] = wv.size();
  const bool small = wv.size() > words_cap || !words || (!(mode & 2) && (p.terms.size() > terms_cap || !terms)) ||
                     (!p.pool.empty() && (p.pool.size() > pool_cap || !pool));
  if (small) return fail(EHIP_OTHER, "GPU error: program buffers too small");
  std::memcpy(words, wv.data(), wv.size() * 4);
  if (!(mode & 2)) std::memcpy(terms, p.terms.data(), p.terms.size() * 4);
  if (!p.pool.empty()) std::memcpy(pool, p.pool.data(), p.pool.size() * 4);
  return ok();
} catch (...) {
  return exception_status();
}

// ---------------------------- device-pointer API ----------------------------

int32_t ehip_forward_batch_dev(ehip_ctx* c, const ehip_tape* t, const void* d_in, uint64_t batch, void* d_out,
                               void* stream) try {
  if (!c || !t) return fail(EHIP_OTHER, "GPU error: null argument");
  int32_t rc = check_batch(t, batch);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  Launch L{};
  L.nin = t->nin;
  L.nout = t->nout;
  L.B = batch;
  L.x = d_in;
  L.xstride = t->nin;
  L.y = d_out;
  rc = run(c, t, 0, 1, L, pick(c, stream));
  return rc ? rc : ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_gradient_batch_dev(ehip_ctx* c, const ehip_tape* t, const void* d_in, uint64_t batch, void* d_out,
                                void* d_grads, void* stream) try {
  if (!c || !t || !d_grads) return fail(EHIP_OTHER, "GPU error: null argument");
  int32_t rc = check_batch(t, batch);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  Launch L{};
  L.nin = t->nin;
  L.nout = t->nout;
  L.B = batch;
  L.x = d_in;
  L.xstride = t->nin;
  L.y = d_out;
  L.g = d_grads;
  rc = run(c, t, 0, 1, L, pick(c, stream));
  return rc ? rc : ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_hvp_batch_dev(ehip_ctx* c, const ehip_tape* t, const void* d_x, uint64_t x_stride, const void* d_dirs,
                           uint64_t batch, void* d_grads, void* d_hvps, void* stream) try {
  if (!c || !t || !d_x || !d_dirs || !d_grads || !d_hvps) return fail(EHIP_OTHER, "GPU error: null argument");
  int32_t rc = check_batch(t, batch);
  if (rc) return rc;
  if (t->nout != 1)
    return fail(EHIP_OTHER, "GPU error: hvp is defined for scalar-output tapes only; this tape has %u outputs", t->nout);
  if (x_stride != 0 && x_stride != t->nin) return fail(EHIP_OTHER, "GPU error: x_stride must be 0 or num_inputs");
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  Launch L{};
  L.nin = t->nin;
  L.nout = t->nout;
  L.B = batch;
  L.x = d_x;
  L.xstride = x_stride;
  L.dirs = d_dirs;
  L.g = d_grads;
  L.h = d_hvps;
  rc = run(c, t, 1, 2, L, pick(c, stream));
  return rc ? rc : ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_taylor_kth_batch_dev(ehip_ctx* c, const ehip_tape* t, const void* d_x, uint64_t x_stride,
                                  const void* d_dirs, uint64_t batch, uint32_t order, void* d_coeffs,
                                  void* stream) try {
  if (!c || !t || !d_x || !d_dirs || !d_coeffs) return fail(EHIP_OTHER, "GPU error: null argument");
  if (order < 1 || order > 5)
    return fail(EHIP_OTHER, "GPU error: taylor_forward_kth_batch supports order 1..=5, got %u", order);
  int32_t rc = check_batch(t, batch);
  if (rc) return rc;
  if (x_stride != 0 && x_stride != t->nin) return fail(EHIP_OTHER, "GPU error: x_stride must be 0 or num_inputs");
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  Launch L{};
  L.nin = t->nin;
  L.nout = t->nout;
  L.B = batch;
  L.x = d_x;
  L.xstride = x_stride;
  L.dirs = d_dirs;
  L.y = d_coeffs;
  rc = run(c, t, 2, (int)order, L, pick(c, stream));
  return rc ? rc : ok();
} catch (...) {
  return exception_status();
}

// --------------------------- nonsmooth kink records ---------------------------

uint32_t ehip_kink_count(const ehip_tape* t) { return t ? t->nkinks : 0; }

int32_t ehip_kink_info(const ehip_tape* t, uint32_t* tape_index, uint32_t* opcode, uint32_t cap) try {
  if (!t || (t->nkinks && (!tape_index || !opcode))) return fail(EHIP_OTHER, "GPU error: null argument");
  if (cap < t->nkinks) return fail(EHIP_OTHER, "GPU error: kink buffers too small (need %u)", t->nkinks);
  uint32_t k = 0;
  for (uint32_t i = t->host.num_inputs; i < (uint32_t)t->host.opcodes.size(); ++i)
    if (is_nonsmooth_op(t->host.opcodes[i])) {
      tape_index[k] = i;
      opcode[k] = t->host.opcodes[i];
      ++k;
    }
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_forward_nonsmooth_dev(ehip_ctx* c, const ehip_tape* t, const void* d_x, uint64_t batch, void* d_out,
                                   void* d_switch, int8_t* d_branch, void* stream) try {
  if (!c || !t || !d_x || (t->nkinks && (!d_switch || !d_branch))) return fail(EHIP_OTHER, "GPU error: null argument");
  int32_t rc = check_batch(t, batch);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  Launch L{};
  L.nin = t->nin;
  L.nout = t->nout;
  L.B = batch;
  L.x = d_x;
  L.xstride = t->nin;
  L.y = d_out;
  L.ksw = d_switch;
  L.kbr = d_branch;
  L.nkinks = t->nkinks;
  rc = run(c, t, 3, 1, L, pick(c, stream));
  return rc ? rc : ok();
} catch (...) {
  return exception_status();
}

// ------------------------------ multi-tape queue ------------------------------

int32_t ehip_multi_batch_dev(ehip_ctx* c, const ehip_job* jobs, uint32_t njobs, uint32_t wg, uint32_t lds_budget,
                             void* stream) try {
  if (!c || (!jobs && njobs)) return fail(EHIP_OTHER, "GPU error: null argument");
  if (njobs == 0) return ok();
  if (wg == 0) wg = 256;
  if (wg != 64 && wg != 128 && wg != 256) return fail(EHIP_OTHER, "GPU error: workgroup size must be 64, 128 or 256");
  if (lds_budget == 0) lds_budget = 24576;
  const bool f64 = jobs[0].tape && jobs[0].tape->f64;
  for (uint32_t j = 0; j < njobs; ++j) {
    const ehip_job& J = jobs[j];
    if (!J.tape || !J.d_x) return fail(EHIP_OTHER, "GPU error: job %u: null argument", j);
    if (J.tape->f64 != f64) return fail(EHIP_OTHER, "GPU error: job %u: all tapes in one queue share a dtype", j);
    if (J.what != 0 && J.what != 1) return fail(EHIP_OTHER, "GPU error: job %u: what must be 0 (forward) or 1 (gradient)", j);
    if (J.what == 1 && !J.d_g) return fail(EHIP_OTHER, "GPU error: job %u: gradient job without a gradient buffer", j);
    if (J.what == 0 && !J.d_y) return fail(EHIP_OTHER, "GPU error: job %u: forward job without an output buffer", j);
    int32_t rc = check_batch(J.tape, J.batch);
    if (rc) return fail(rc, "job %u: %s", j, g_last_error.c_str());
    if (J.batch > (1ull << 40)) return fail(EHIP_OTHER, "GPU error: job %u: batch too large", j);
  }
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = pick(c, stream);
  const size_t S = f64 ? 8 : 4;
  const size_t per_reg = (size_t)wg * S;
  const bool sf = c->fetch != 0;
  const size_t win_lds = window_lds(sf);
  const size_t cap = c->max_lds > win_lds ? c->max_lds - win_lds : 0;
  const size_t budget = std::min<size_t>(lds_budget, cap);
  // jobs, longest program first
  std::vector<uint32_t> order(njobs);
  for (uint32_t j = 0; j < njobs; ++j) order[j] = j;
  auto prog_of = [&](uint32_t j) -> const DevProgram& {
    return jobs[j].tape->prog[prog_index(0, jobs[j].what == 1)];
  };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return prog_of(a).nwin > prog_of(b).nwin; });
  std::vector<Launch> L(njobs);
  std::vector<QueueTile> tiles;
  uint32_t max_hbm_regs = 0, max_rlds = 0;
  for (uint32_t j = 0; j < njobs; ++j) {
    const ehip_job& J = jobs[j];
    const DevProgram& dp = prog_of(j);
    Launch& l = L[j];
    l = Launch{};
    l.prog = dp.d_stream;
    l.pool = dp.d_pool;
    l.nwin = dp.nwin;
    l.nin = J.tape->nin;
    l.nout = J.tape->nout;
    l.B = J.batch;
    l.x = J.d_x;
    l.xstride = J.tape->nin;
    l.y = J.d_y;
    l.g = J.what == 1 ? J.d_g : nullptr;
    const uint32_t R = dp.host.nregs;
    l.R_lds = (uint32_t)std::min<size_t>(R, budget / per_reg);
    max_rlds = std::max(max_rlds, l.R_lds);
    max_hbm_regs = std::max(max_hbm_regs, R - l.R_lds);
  }
  for (uint32_t j : order)
    for (uint64_t f = 0; f < jobs[j].batch; f += wg) tiles.push_back(QueueTile{j, 0u, f});
  if (tiles.size() > 0xFFFFFFF0ull) return fail(EHIP_OTHER, "GPU error: too many tiles in one queue");
  const size_t lds = (size_t)max_rlds * per_reg;
  int per_cu = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, queue_kernel(f64, sf), (int)wg, queue_lds_bytes(sf, lds)),
          "queue occupancy");
  uint64_t grid = (uint64_t)std::max(per_cu, 1) * c->ncu;
  grid = std::min<uint64_t>(grid, tiles.size());
  const uint64_t lanes = grid * wg;
  void* gregs = nullptr;
  if (max_hbm_regs) {
    HIP_TRY(c->regs.reserve((size_t)max_hbm_regs * lanes * S), "register file");
    gregs = c->regs.p;
  }
  for (auto& l : L) l.gregs = gregs;
  HIP_TRY(c->qjobs.reserve(L.size() * sizeof(Launch)), "queue tables");
  HIP_TRY(c->qtiles.reserve(tiles.size() * sizeof(QueueTile)), "queue tables");
  HIP_TRY(c->qcount.reserve(256), "queue tables");
  HIP_TRY(hipMemcpyAsync(c->qjobs.p, L.data(), L.size() * sizeof(Launch), hipMemcpyHostToDevice, s), "queue tables");
  HIP_TRY(hipMemcpyAsync(c->qtiles.p, tiles.data(), tiles.size() * sizeof(QueueTile), hipMemcpyHostToDevice, s),
          "queue tables");
  Queue Q;
  Q.jobs = (const Launch*)c->qjobs.p;
  Q.tiles = (const QueueTile*)c->qtiles.p;
  Q.ntiles = (uint32_t)tiles.size();
  Q.counter = (uint32_t*)c->qcount.p;
  hipError_t e = launch_queue(f64, sf, Q, (uint32_t)grid, wg, lds, s);
  if (e != hipSuccess) return hip_fail(e, "queue launch");
  // the host tables above are released on return
  HIP_TRY(hipStreamSynchronize(s), "queue");
  return ok();
} catch (...) {
  return exception_status();
}

// --------------------------------- autotuner ---------------------------------

namespace {

// Launch fields for an evaluation `what` (0 forward, 1 gradient, 2 hvp,
// 3 taylor of `order`) over caller buffers.
int32_t tune_launch(const ehip_tape* t, int32_t what, uint32_t order, const void* d_x, uint64_t x_stride,
                    const void* d_dirs, uint64_t batch, void* d_y, void* d_g, void* d_h, Launch& L, int& kind,
                    int& K) {
  L = Launch{};
  L.nin = t->nin;
  L.nout = t->nout;
  L.B = batch;
  L.x = d_x;
  L.xstride = x_stride;
  switch (what) {
    case 0: kind = 0; K = 1; L.y = d_y; break;
    case 1: kind = 0; K = 1; L.y = d_y; L.g = d_g; if (!d_g) return fail(EHIP_OTHER, "GPU error: null argument"); break;
    case 2:
      kind = 1; K = 2; L.dirs = d_dirs; L.g = d_g; L.h = d_h;
      if (!d_dirs || !d_g || !d_h) return fail(EHIP_OTHER, "GPU error: null argument");
      if (t->nout != 1) return fail(EHIP_OTHER, "GPU error: hvp is defined for scalar-output tapes only");
      break;
    case 3:
      if (order < 1 || order > 5) return fail(EHIP_OTHER, "GPU error: taylor order must be 1..=5");
      kind = 2; K = (int)order; L.dirs = d_dirs; L.y = d_y;
      if (!d_dirs || !d_y) return fail(EHIP_OTHER, "GPU error: null argument");
      break;
    default: return fail(EHIP_OTHER, "GPU error: unknown evaluation kind %d", what);
  }
  if (x_stride != 0 && x_stride != t->nin) return fail(EHIP_OTHER, "GPU error: x_stride must be 0 or num_inputs");
  return EHIP_OK;
}

}  // namespace

int32_t ehip_autotune_dev(ehip_ctx* c, ehip_tape* t, int32_t what, uint32_t order, const void* d_x,
                          uint64_t x_stride, const void* d_dirs, uint64_t batch, void* d_y, void* d_g, void* d_h,
                          uint32_t* out5) try {
  if (!c || !t || !d_x || !out5) return fail(EHIP_OTHER, "GPU error: null argument");
  int32_t rc = check_batch(t, batch);
  if (rc) return rc;
  Launch L;
  int kind = 0, K = 1;
  rc = tune_launch(t, what, order, d_x, x_stride, d_dirs, batch, d_y, d_g, d_h, L, kind, K);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  const uint32_t NC = (kind == 0) ? 1u : (kind == 1 ? 2u : (uint32_t)K);
  const size_t S = t->f64 ? 8 : 4;
  const uint32_t R = t->prog[prog_index(kind, what == 1)].host.nregs;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0), "autotune");
  HIP_TRY(hipEventCreate(&e1), "autotune");
  Tuned best;
  float best_ms = 1e30f;
  static const uint32_t budgets[] = {4096, 8192, 16384, 24576, 32768, 49152, 65536, 98304, 131072};
  for (uint32_t fetch : {0u, 1u})
  for (uint32_t wg : {64u, 128u, 256u}) {
    std::vector<uint32_t> seen;
    for (uint32_t b : budgets) {
      const size_t per_reg = (size_t)NC * wg * S;
      const size_t win_lds = window_lds(fetch != 0);
      const size_t cap = c->max_lds > win_lds ? c->max_lds - win_lds : 0;
      if (b > cap) continue;
      const uint32_t rl = (uint32_t)std::min<size_t>(R, b / per_reg);
      if (std::find(seen.begin(), seen.end(), rl) != seen.end()) continue;
      seen.push_back(rl);
      Tuned cand;
      cand.wg = wg;
      cand.lds_budget = b;
      cand.fetch = fetch;
      float ms = 1e30f;
      for (int rep = 0; rep < 3 && rc == EHIP_OK; ++rep) {
        HIP_TRY(hipEventRecord(e0, c->stream), "autotune");
        rc = run(c, t, kind, K, L, c->stream, &cand);
        HIP_TRY(hipEventRecord(e1, c->stream), "autotune");
        HIP_TRY(hipEventSynchronize(e1), "autotune");
        float m = 0.f;
        HIP_TRY(hipEventElapsedTime(&m, e0, e1), "autotune");
        if (rep > 0) ms = std::min(ms, m);  // first run warms caches
      }
      if (rc) break;
      if (ms < best_ms) { best_ms = ms; best = cand; }
    }
    if (rc) break;
  }
  // the tape-specialised kernel, where one has been compiled (ehip_tape_specialize)
  if (rc == EHIP_OK && t->spec[tune_key(kind, K, what == 1)].fn)
    for (uint32_t wg : {64u, 128u, 256u})
    for (uint32_t ppl : {1u, 2u, 4u}) {
      Tuned cand;
      cand.wg = wg;
      cand.fetch = 2;
      cand.ppl = ppl;
      float ms = 1e30f;
      for (int rep = 0; rep < 3 && rc == EHIP_OK; ++rep) {
        HIP_TRY(hipEventRecord(e0, c->stream), "autotune");
        rc = run(c, t, kind, K, L, c->stream, &cand);
        HIP_TRY(hipEventRecord(e1, c->stream), "autotune");
        HIP_TRY(hipEventSynchronize(e1), "autotune");
        float m = 0.f;
        HIP_TRY(hipEventElapsedTime(&m, e0, e1), "autotune");
        if (rep > 0) ms = std::min(ms, m);
      }
      if (rc) break;
      if (ms < best_ms) { best_ms = ms; best = cand; }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  t->tuned[tune_key(kind, K, what == 1)] = best;
  out5[0] = best.wg;
  out5[1] = best.lds_budget;
  out5[2] = (uint32_t)std::min(best_ms * 1000.f, 4.0e9f);
  out5[3] = best.fetch;
  out5[4] = best.ppl ? best.ppl : 1u;
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_set_tuned(ehip_tape* t, int32_t what, uint32_t order, uint32_t wg, uint32_t lds_budget,
                       uint32_t fetch, uint32_t points_per_lane) try {
  if (!t) return fail(EHIP_OTHER, "GPU error: null argument");
  if (wg != 0 && wg != 64 && wg != 128 && wg != 256)
    return fail(EHIP_OTHER, "GPU error: workgroup size must be 64, 128 or 256");
  if (fetch > 2)
    return fail(EHIP_OTHER, "GPU error: program fetch must be 0 (LDS), 1 (scalar) or 2 (specialised kernel)");
  int key;
  if (what >= 0 && what <= 2) key = what;
  else if (what == 3 && order >= 1 && order <= 5) key = 2 + (int)order;
  else return fail(EHIP_OTHER, "GPU error: unknown evaluation kind %d", what);
  if (fetch == 2 && wg && !t->spec[key].fn)
    return fail(EHIP_OTHER, "GPU error: no specialised kernel for this evaluation (ehip_tape_specialize)");
  if (points_per_lane > 64 || (points_per_lane > 1 && fetch != 2))
    return fail(EHIP_OTHER, "GPU error: points_per_lane must be 0 or 1, or up to 64 with the specialised kernel");
  t->tuned[key].ppl = (wg && fetch == 2) ? points_per_lane : 0u;
  t->tuned[key].wg = wg;
  t->tuned[key].lds_budget = wg ? lds_budget : 0;
  t->tuned[key].fetch = wg ? fetch : 0;
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_get_tuned(const ehip_tape* t, int32_t what, uint32_t order, uint32_t* out4) try {
  if (!t || !out4) return fail(EHIP_OTHER, "GPU error: null argument");
  int key;
  if (what >= 0 && what <= 2) key = what;
  else if (what == 3 && order >= 1 && order <= 5) key = 2 + (int)order;
  else return fail(EHIP_OTHER, "GPU error: unknown evaluation kind %d", what);
  out4[0] = t->tuned[key].wg;
  out4[1] = t->tuned[key].lds_budget;
  out4[2] = t->tuned[key].fetch;
  out4[3] = t->tuned[key].ppl ? t->tuned[key].ppl : 1u;
  return ok();
} catch (...) {
  return exception_status();
}

// ------------------------------- hipGraph capture -------------------------------

struct ehip_graph {
  ehip_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  DevBuf regs;  // the graph's own HBM register file
};

int32_t ehip_graph_capture(ehip_ctx* c, ehip_tape* t, int32_t what, uint32_t order, const void* d_x,
                           uint64_t x_stride, const void* d_dirs, uint64_t batch, void* d_y, void* d_g, void* d_h,
                           uint32_t repeats, ehip_graph** out) try {
  if (!c || !t || !d_x || !out) return fail(EHIP_OTHER, "GPU error: null argument");
  *out = nullptr;
  if (repeats == 0) return fail(EHIP_OTHER, "GPU error: repeats must be non-zero");
  int32_t rc = check_batch(t, batch);
  if (rc) return rc;
  Launch L;
  int kind = 0, K = 1;
  rc = tune_launch(t, what, order, d_x, x_stride, d_dirs, batch, d_y, d_g, d_h, L, kind, K);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  ehip_graph* gr = new ehip_graph();
  gr->ctx = c;
  // one eager launch sizes the graph's register file (no allocation may
  // happen while the stream is being captured)
  rc = run(c, t, kind, K, L, c->stream, nullptr, &gr->regs);
  hipError_t e = hipSuccess;
  if (!rc) e = hipStreamSynchronize(c->stream);
  if (!rc && e == hipSuccess) e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (!rc && e == hipSuccess) {
    for (uint32_t k = 0; k < repeats && !rc; ++k) rc = run(c, t, kind, K, L, c->stream, nullptr, &gr->regs);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    gr->graph = g;
    if (e2 != hipSuccess && !rc) e = e2;
  }
  if (!rc && e == hipSuccess) e = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
  if (rc || e != hipSuccess) {
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    gr->regs.release();
    delete gr;
    return rc ? rc : hip_fail(e, "graph capture");
  }
  *out = gr;
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_graph_launch(ehip_graph* gr, void* stream) try {
  if (!gr) return fail(EHIP_OTHER, "GPU error: null argument");
  std::lock_guard<std::mutex> lk(gr->ctx->mu);
  (void)hipSetDevice(gr->ctx->device);
  HIP_TRY(hipGraphLaunch(gr->exec, pick(gr->ctx, stream)), "graph launch");
  return ok();
} catch (...) {
  return exception_status();
}

void ehip_graph_destroy(ehip_graph* gr) try {
  if (!gr) return;
  {
    std::lock_guard<std::mutex> lk(gr->ctx->mu);
    (void)hipSetDevice(gr->ctx->device);
    (void)hipDeviceSynchronize();
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    gr->regs.release();
  }
  delete gr;
} catch (...) {
  (void)exception_status();
}

// ----------------------------- host-pointer API -----------------------------

namespace {
struct HostCall {
  ehip_ctx* c;
  size_t S;
  int32_t stage(int slot, const void* src, size_t bytes, void** dptr) {
    hipError_t e = c->io[slot].reserve(bytes);
    if (e != hipSuccess) return hip_fail(e, "staging buffer");
    *dptr = c->io[slot].p;
    if (src) {
      e = hipMemcpyAsync(*dptr, src, bytes, hipMemcpyHostToDevice, c->stream);
      if (e != hipSuccess) return hip_fail(e, "host to device copy");
    }
    return EHIP_OK;
  }
  int32_t fetch(void* dst, const void* dsrc, size_t bytes) {
    hipError_t e = hipMemcpyAsync(dst, dsrc, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return hip_fail(e, "device to host copy");
    return EHIP_OK;
  }
  int32_t sync() {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "synchronize");
    return EHIP_OK;
  }
};
}  // namespace

#define TRY(x)              \
  do {                      \
    int32_t _r = (x);       \
    if (_r) return _r;      \
  } while (0)

int32_t ehip_forward_batch(ehip_ctx* c, const ehip_tape* t, const void* inputs, uint32_t batch, void* outputs) try {
  if (!c || !t || !inputs || !outputs) return fail(EHIP_OTHER, "GPU error: null argument");
  TRY(check_batch(t, batch));
  std::lock_guard<std::mutex> hlk(c->host_mu);
  HostCall h{c, t->f64 ? 8u : 4u};
  void *dx, *dy;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    TRY(h.stage(0, inputs, (size_t)batch * t->nin * h.S, &dx));
    TRY(h.stage(1, nullptr, (size_t)batch * t->nout * h.S, &dy));
  }
  TRY(ehip_forward_batch_dev(c, t, dx, batch, dy, nullptr));
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(h.fetch(outputs, dy, (size_t)batch * t->nout * h.S));
  TRY(h.sync());
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_forward_nonsmooth(ehip_ctx* c, const ehip_tape* t, const void* inputs, uint32_t batch, void* outputs,
                               void* switching, int8_t* branches) try {
  if (!c || !t || !inputs || !outputs || (t->nkinks && (!switching || !branches)))
    return fail(EHIP_OTHER, "GPU error: null argument");
  TRY(check_batch(t, batch));
  std::lock_guard<std::mutex> hlk(c->host_mu);
  HostCall h{c, t->f64 ? 8u : 4u};
  void *dx, *dy, *ds, *db;
  const size_t nk = (size_t)batch * t->nkinks;
  TRY(h.stage(0, inputs, (size_t)batch * t->nin * h.S, &dx));
  TRY(h.stage(1, nullptr, (size_t)batch * t->nout * h.S, &dy));
  TRY(h.stage(2, nullptr, std::max<size_t>(nk * h.S, 8), &ds));
  TRY(h.stage(3, nullptr, std::max<size_t>(nk, 8), &db));
  TRY(ehip_forward_nonsmooth_dev(c, t, dx, batch, dy, ds, (int8_t*)db, nullptr));
  TRY(h.fetch(outputs, dy, (size_t)batch * t->nout * h.S));
  if (nk) {
    TRY(h.fetch(switching, ds, nk * h.S));
    TRY(h.fetch(branches, db, nk));
  }
  TRY(h.sync());
  return ok();
} catch (...) {
  return exception_status();
}

int32_t ehip_gradient_batch(ehip_ctx* c, const ehip_tape* t, const void* inputs, uint32_t batch, void* outputs,
                            void* grads) try {
  if (!c || !t || !inputs || !outputs || !grads) return fail(EHIP_OTHER, "GPU error: null argument");
  TRY(check_batch(t, batch));
  std::lock_guard<std::mutex> hlk(c->host_mu);
  HostCall h{c, t->f64 ? 8u : 4u};
  void *dx, *dy, *dg;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    TRY(h.stage(0, inputs, (size_t)batch * t->nin * h.S, &dx));
    TRY(h.stage(1, nullptr, (size_t)batch * t->nout * h.S, &dy));
    TRY(h.stage(2, nullptr, (size_t)batch * t->nin * h.S, &dg));
  }
  TRY(ehip_gradient_batch_dev(c, t, dx, batch, dy, dg, nullptr));
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(h.fetch(outputs, dy, (size_t)batch * t->nout * h.S));
  TRY(h.fetch(grads, dg, (size_t)batch * t->nin * h.S));
  TRY(h.sync());
  return ok();
} catch (...) {
  return exception_status();
}

static int32_t hvp_host(ehip_ctx* c, const ehip_tape* t, const void* x, uint64_t xrows, const void* dirs,
                        uint32_t batch, void* grads, void* hvps) {
  if (!c || !t || !x || !dirs || !grads || !hvps) return fail(EHIP_OTHER, "GPU error: null argument");
  TRY(check_batch(t, batch));
  std::lock_guard<std::mutex> hlk(c->host_mu);
  HostCall h{c, t->f64 ? 8u : 4u};
  size_t row = (size_t)t->nin * h.S;
  void *dx, *dv, *dg, *dh;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    TRY(h.stage(0, x, xrows * row, &dx));
    TRY(h.stage(1, dirs, (size_t)batch * row, &dv));
    TRY(h.stage(2, nullptr, (size_t)batch * row, &dg));
    TRY(h.stage(3, nullptr, (size_t)batch * row, &dh));
  }
  TRY(ehip_hvp_batch_dev(c, t, dx, xrows == 1 ? 0 : t->nin, dv, batch, dg, dh, nullptr));
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(h.fetch(grads, dg, (size_t)batch * row));
  TRY(h.fetch(hvps, dh, (size_t)batch * row));
  TRY(h.sync());
  return ok();
}

int32_t ehip_hvp_batch(ehip_ctx* c, const ehip_tape* t, const void* x, const void* dirs, uint32_t batch,
                       void* grads, void* hvps) try {
  return hvp_host(c, t, x, 1, dirs, batch, grads, hvps);
} catch (...) {
  return exception_status();
}

int32_t ehip_hvp_batch_points(ehip_ctx* c, const ehip_tape* t, const void* xs, const void* dirs, uint32_t batch,
                              void* grads, void* hvps) try {
  return hvp_host(c, t, xs, batch, dirs, batch, grads, hvps);
} catch (...) {
  return exception_status();
}

int32_t ehip_taylor_kth_batch(ehip_ctx* c, const ehip_tape* t, const void* primals, const void* seeds,
                              uint32_t batch, uint32_t order, void* coeffs) try {
  if (!c || !t || !primals || !seeds || !coeffs) return fail(EHIP_OTHER, "GPU error: null argument");
  if (order < 1 || order > 5)
    return fail(EHIP_OTHER, "GPU error: taylor_forward_kth_batch supports order 1..=5, got %u", order);
  TRY(check_batch(t, batch));
  std::lock_guard<std::mutex> hlk(c->host_mu);
  HostCall h{c, t->f64 ? 8u : 4u};
  size_t row = (size_t)t->nin * h.S;
  size_t total_out = (size_t)batch * t->nout;
  void *dx, *dv, *dc;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    TRY(h.stage(0, primals, (size_t)batch * row, &dx));
    TRY(h.stage(1, seeds, (size_t)batch * row, &dv));
    TRY(h.stage(2, nullptr, total_out * order * h.S, &dc));
  }
  TRY(ehip_taylor_kth_batch_dev(c, t, dx, t->nin, dv, batch, order, dc, nullptr));
  std::vector<unsigned char> raw(total_out * order * h.S);
  {
    std::lock_guard<std::mutex> lk(c->mu);
    TRY(h.fetch(raw.data(), dc, raw.size()));
    TRY(h.sync());
  }
  // deinterleave_coefficients (src/gpu/mod.rs:247-259)
  unsigned char* dst = (unsigned char*)coeffs;
  for (size_t i = 0; i < total_out; ++i)
    for (uint32_t k = 0; k < order; ++k)
      std::memcpy(dst + ((size_t)k * total_out + i) * h.S, raw.data() + (i * order + k) * h.S, h.S);
  return ok();
} catch (...) {
  return exception_status();
}

// ------------------------------- GPU STDE ----------------------------------

int32_t ehip_laplacian_batch_dev(ehip_ctx* c, const ehip_tape* t, const void* d_x, uint64_t npoints, uint32_t S,
                                 uint32_t dir_kind, uint64_t seed, const void* d_dirs, const void* d_diag,
                                 uint64_t diag_stride, double* d_results, void* stream) try {
  if (!c || !t || !d_x || !d_results) return fail(EHIP_OTHER, "GPU error: null argument");
  if (t->nout != 1)
    return fail(EHIP_OTHER, "GPU error: laplacian_gpu requires a scalar-output tape; got %u outputs", t->nout);
  if (t->nin == 0)
    return fail(EHIP_OTHER, "GPU error: laplacian_gpu requires a non-empty x (a zero-input tape has no Laplacian); "
                            "zero-sized GPU buffers would panic");
  if (S == 0) return fail(EHIP_OTHER, "GPU error: no directions provided");
  if (dir_kind > 2) return fail(EHIP_OTHER, "GPU error: dir_kind must be 0 (supplied), 1 (Rademacher) or 2 (Gaussian)");
  if (dir_kind == 0 && !d_dirs) return fail(EHIP_OTHER, "GPU error: null directions");
  if (npoints == 0) return ok();
  if (d_diag && diag_stride != 0 && diag_stride != t->nin)
    return fail(EHIP_OTHER, "GPU error: diag_stride must be 0 or num_inputs");
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  hipStream_t s = pick(c, stream);
  const size_t Sb = t->f64 ? 8 : 4;
  const uint64_t lanes = npoints * (uint64_t)S;
  HIP_TRY(c->jets.reserve(lanes * 3 * Sb), "jet scratch");
  Launch L{};
  L.nin = t->nin;
  L.nout = 1;
  L.B = lanes;
  L.x = d_x;
  L.xstride = t->nin;
  L.S = S;
  L.rng = dir_kind;
  L.seed = seed;
  L.dirs = d_dirs;
  L.y = c->jets.p;
  int32_t rc = run(c, t, 2, 3, L, s);
  if (rc) return rc;
  StdeReduce R{};
  R.jets = c->jets.p;
  R.npoints = npoints;
  R.S = S;
  R.nin = t->nin;
  R.rng = dir_kind;
  R.seed = seed;
  R.dirs = d_dirs;
  R.diag = d_diag;
  R.diag_stride = diag_stride;
  R.out = d_results;
  HIP_TRY(launch_stde_reduce(t->f64, R, s), "stde reduce");
  return ok();
} catch (...) {
  return exception_status();
}

static int32_t laplacian_host(ehip_ctx* c, const ehip_tape* t, const void* x, uint64_t npoints, const void* dirs,
                              uint32_t S, uint32_t dir_kind, uint64_t seed, const void* diag, uint64_t diag_rows,
                              double* results) {
  if (!c || !t || !x || !results) return fail(EHIP_OTHER, "GPU error: null argument");
  std::lock_guard<std::mutex> hlk(c->host_mu);
  const size_t Sb = t->f64 ? 8 : 4;
  void *dx, *dv = nullptr, *dd = nullptr, *dr;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    HostCall h{c, Sb};
    TRY(h.stage(0, x, npoints * t->nin * Sb, &dx));
    if (dir_kind == 0 && dirs) TRY(h.stage(1, dirs, npoints * (uint64_t)S * t->nin * Sb, &dv));
    if (diag) TRY(h.stage(2, diag, diag_rows * t->nin * Sb, &dd));
    TRY(h.stage(3, nullptr, npoints * 5 * sizeof(double), &dr));
  }
  TRY(ehip_laplacian_batch_dev(c, t, dx, npoints, S, dir_kind, seed, dv, dd, diag_rows > 1 ? t->nin : 0,
                               (double*)dr, nullptr));
  std::lock_guard<std::mutex> lk(c->mu);
  HostCall h{c, Sb};
  TRY(h.fetch(results, dr, npoints * 5 * sizeof(double)));
  TRY(h.sync());
  return ok();
}

int32_t ehip_laplacian(ehip_ctx* c, const ehip_tape* t, const void* x, const void* dirs, uint32_t S,
                       double* result5) try {
  if (S > 0 && !dirs) return fail(EHIP_OTHER, "GPU error: null directions");
  return laplacian_host(c, t, x, 1, dirs, S, 0, 0, nullptr, 0, result5);
} catch (...) {
  return exception_status();
}

int32_t ehip_laplacian_with_control(ehip_ctx* c, const ehip_tape* t, const void* x, const void* dirs, uint32_t S,
                                    const void* control_diagonal, double* result5) try {
  if (!control_diagonal) return fail(EHIP_OTHER, "GPU error: null control diagonal");
  if (S > 0 && !dirs) return fail(EHIP_OTHER, "GPU error: null directions");
  return laplacian_host(c, t, x, 1, dirs, S, 0, 0, control_diagonal, 1, result5);
} catch (...) {
  return exception_status();
}

int32_t ehip_laplacian_batch(ehip_ctx* c, const ehip_tape* t, const void* xs, uint64_t npoints, uint32_t S,
                             uint32_t dir_kind, uint64_t seed, const void* dirs, double* results) try {
  return laplacian_host(c, t, xs, npoints, dirs, S, dir_kind, seed, nullptr, 0, results);
} catch (...) {
  return exception_status();
}

int32_t ehip_hessian_diagonal(ehip_ctx* c, const ehip_tape* t, const void* x, void* value, void* diag) try {
  if (!c || !t || !x || !value || !diag) return fail(EHIP_OTHER, "GPU error: null argument");
  if (t->nout != 1)
    return fail(EHIP_OTHER, "GPU error: hessian_diagonal_gpu requires a scalar-output tape; got %u outputs", t->nout);
  if (t->nin == 0)
    return fail(EHIP_OTHER, "GPU error: hessian_diagonal_gpu requires a non-empty x (a zero-input tape has no "
                            "Hessian diagonal); zero-sized GPU buffers would panic");
  const uint32_t n = t->nin;
  const size_t Sb = t->f64 ? 8 : 4;
  std::vector<unsigned char> xs((size_t)n * n * Sb), seeds((size_t)n * n * Sb, 0), out((size_t)3 * n * Sb);
  for (uint32_t j = 0; j < n; ++j) {
    std::memcpy(xs.data() + (size_t)j * n * Sb, x, n * Sb);
    if (t->f64) ((double*)seeds.data())[(size_t)j * n + j] = 1.0;
    else ((float*)seeds.data())[(size_t)j * n + j] = 1.0f;
  }
  TRY(ehip_taylor_kth_batch(c, t, xs.data(), seeds.data(), n, 3, out.data()));
  // diag(H)[j] = 2 * c2[j]  (src/gpu/stde_gpu.rs:103-129)
  if (t->f64) {
    const double* o = (const double*)out.data();
    *(double*)value = o[0];
    for (uint32_t j = 0; j < n; ++j) ((double*)diag)[j] = 2.0 * o[2 * n + j];
  } else {
    const float* o = (const float*)out.data();
    *(float*)value = o[0];
    for (uint32_t j = 0; j < n; ++j) ((float*)diag)[j] = 2.0f * o[2 * n + j];
  }
  return ok();
} catch (...) {
  return exception_status();
}

}  // extern "C"
