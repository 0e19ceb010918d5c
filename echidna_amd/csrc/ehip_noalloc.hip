// The allocation-free variant of ehip_gradient_batch_dev: its interpreter
// register file lives in a fixed device array instead of the context's
// grow-only workspace, so the call allocates nothing (callers under graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// `scratch` is one file-scope device array per GPU, held until the process
// ends.  Host code never passes the array's name to a kernel (in host code
// it is a stand-in, not a device address): ehip_ctx_create calls
// noalloc_setup once per device with that device current, which looks up
// the array's address on that device with hipGetSymbolAddress and keeps it;
// every later use goes through that address.
#include "ehip_internal.h"

namespace {

constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_mu;
void* g_scratch[MAX_DEVICES];        // the array's address on each device
hipError_t g_setup_error[MAX_DEVICES];

}  // namespace

namespace ehip_detail {

// Called by ehip_ctx_create with `device` current.
int32_t noalloc_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) return fail(EHIP_OTHER, "GPU error: device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return EHIP_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_setup_error[device] = e;
  if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch symbol address");
  g_scratch[device] = p;
  return EHIP_OK;
}

}  // namespace ehip_detail

extern "C" {

uint64_t ehip_noalloc_scratch_bytes(void) { return SCRATCH_BYTES; }

int32_t ehip_gradient_batch_dev_noalloc(ehip_ctx* c, const ehip_tape* t, const void* d_in, uint64_t batch,
                                        void* d_out, void* d_grads, void* stream) try {
  if (!c || !t || !d_in || !d_out || !d_grads) return fail(EHIP_OTHER, "GPU error: null argument");
  int32_t rc = check_batch(t, batch);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  void* base;
  {
    std::lock_guard<std::mutex> g(g_mu);
    base = (c->device >= 0 && c->device < MAX_DEVICES) ? g_scratch[c->device] : nullptr;
    if (!base)
      return fail(EHIP_OTHER, "GPU error: the fixed scratch is unavailable on device %d (%s)", c->device,
                  hipGetErrorString(g_setup_error[c->device]));
  }
  uint32_t wg = 256;
  const size_t lane = regfile_lane_bytes(c, t, 0, 1, true, &wg);
  // points per launch: whole workgroups whose register files fit the scratch
  uint64_t chunk = batch;
  if (lane > 0) {
    const uint64_t groups = SCRATCH_BYTES / (lane * wg);
    if (groups == 0)
      return fail(EHIP_OTHER,
                  "GPU error: one workgroup of this tape needs %zu bytes of register file, above the %zu-byte fixed "
                  "scratch; use ehip_gradient_batch_dev",
                  lane * wg, SCRATCH_BYTES);
    chunk = std::min<uint64_t>(batch, groups * wg);
  }
  const size_t S = t->f64 ? 8 : 4;
  const hipStream_t s = pick(c, stream);
  for (uint64_t off = 0; off < batch; off += chunk) {
    Launch L{};
    L.nin = t->nin;
    L.nout = t->nout;
    L.B = std::min<uint64_t>(chunk, batch - off);
    L.x = (const char*)d_in + off * t->nin * S;
    L.xstride = t->nin;
    L.y = (char*)d_out + off * t->nout * S;
    L.g = (char*)d_grads + off * t->nin * S;
    rc = run(c, t, 0, 1, L, s, nullptr, nullptr, base, SCRATCH_BYTES);
    if (rc) return rc;
  }
  return ok();
} catch (...) {
  return exception_status();
}

}  // extern "C"
