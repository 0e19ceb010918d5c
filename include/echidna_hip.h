/*
 * echidna_hip.h — C ABI of the MI355X (gfx950) batched BytecodeTape backend.
 *
 * Drop-in boundary for the reference's GPU backend trait `GpuBackend`
 * (src/gpu/mod.rs:92-225) and the CUDA backend's f64 inherent methods
 * (src/gpu/cuda_backend.rs:566-953).  Plain pointers and sizes only; every
 * entry point returns an `int32_t` status (EHIP_OK or a `GpuError` variant,
 * src/gpu/mod.rs:276-287) and leaves a thread-local message readable with
 * ehip_last_error().  No call panics or throws across the boundary.
 *
 * Scalar type: a tape is uploaded as f32 or f64 (EHIP_F32 / EHIP_F64); every
 * `void*` data argument of a call on that tape points to that scalar type.
 * f32 tapes serve the trait's f32 methods, f64 tapes the CUDA backend's
 * `*_f64` inherent methods.
 *
 * Layouts at the boundary follow the reference: point-major rows
 * (`inputs[b * num_inputs + i]`), outputs `[b * num_outputs + j]`, Taylor
 * coefficients de-interleaved into `coeffs[k * (batch * num_outputs) + ...]`
 * (TaylorKthBatchResult, src/gpu/mod.rs:266-272).
 *
 * Host-pointer entry points (`ehip_*_batch`) copy in, run and copy out
 * synchronously, like the reference's per-call clone_htod / launch /
 * synchronize / clone_dtoh.  Device-pointer entry points (`*_dev`) take
 * buffers already resident in HBM and enqueue on the caller's HIP stream
 * (NULL = the context's stream) without synchronizing.  The context's stream
 * is a blocking stream: its work is ordered after work enqueued before it on
 * the legacy default (null) stream, and the null stream's later work waits
 * for it.  Inputs produced on any other stream must be complete (or the
 * caller's stream passed) before the call.
 *
 * Threading: calls on one context are serialized by the context; the
 * context's scratch registers are reused across calls, so `*_dev` work on
 * different streams through ONE context must not overlap — use one context
 * per stream for concurrent streams.
 */
#ifndef ECHIDNA_HIP_H
#define ECHIDNA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: 0 ok, then the GpuError variants (src/gpu/mod.rs:276-287). */
enum {
  EHIP_OK = 0,
  EHIP_NO_DEVICE = 1,
  EHIP_SHADER_COMPILATION = 2,
  EHIP_OUT_OF_MEMORY = 3,
  EHIP_CUSTOM_OPS_NOT_SUPPORTED = 4,
  EHIP_OTHER = 5
};

enum { EHIP_F32 = 0, EHIP_F64 = 1 };

typedef struct ehip_ctx ehip_ctx;
typedef struct ehip_tape ehip_tape;

/* Message of the last failing call on this thread ("" when none). */
const char* ehip_last_error(void);

/* Library version string. */
const char* ehip_version(void);

/* CudaContext::new (src/gpu/cuda_backend.rs:503-552): acquire device
 * `device`.  Returns EHIP_NO_DEVICE when no gfx950 device is present. */
int32_t ehip_ctx_create(int32_t device, ehip_ctx** out);
void ehip_ctx_destroy(ehip_ctx* ctx);

/* GpuBackend::upload_tape (src/gpu/mod.rs:107) / CudaContext::upload_tape_f64
 * (src/gpu/cuda_backend.rs:566-611).  Runs the GpuTapeData::validate checks
 * (src/gpu/mod.rs:397-564) before touching the device; invalid data returns
 * EHIP_OTHER with the reference's message (the Rust shim turns this into the
 * trait's panic).  An opcode 43 (Custom) returns
 * EHIP_CUSTOM_OPS_NOT_SUPPORTED.  `constants` has `num_variables` entries of
 * the tape's scalar type.  `num_outputs == 0` synthesizes
 * `[output_index]` (src/gpu/cuda_backend.rs:638-647). */
int32_t ehip_tape_upload(ehip_ctx* ctx, int32_t dtype, const uint32_t* opcodes, const uint32_t* arg0,
                         const uint32_t* arg1, const void* constants, uint32_t num_ops,
                         uint32_t num_inputs, uint32_t num_variables, uint32_t output_index,
                         const uint32_t* output_indices, uint32_t num_outputs, ehip_tape** out);
void ehip_tape_free(ehip_tape* tape);

/* GpuBackend::num_outputs (src/gpu/mod.rs:115). */
uint32_t ehip_num_outputs(const ehip_tape* tape);
uint32_t ehip_num_inputs(const ehip_tape* tape);
int32_t ehip_tape_dtype(const ehip_tape* tape);

/* GpuBackend::forward_batch (src/gpu/mod.rs:118-124), forward_batch_f64
 * (src/gpu/cuda_backend.rs:878).  inputs [batch*num_inputs],
 * outputs [batch*num_outputs]. */
int32_t ehip_forward_batch(ehip_ctx* ctx, const ehip_tape* tape, const void* inputs, uint32_t batch,
                           void* outputs);

/* GpuBackend::gradient_batch (src/gpu/mod.rs:126-137), gradient_batch_f64
 * (src/gpu/cuda_backend.rs:896).  Seeds output_indices[0]. */
int32_t ehip_gradient_batch(ehip_ctx* ctx, const ehip_tape* tape, const void* inputs, uint32_t batch,
                            void* outputs, void* grads);

/* GpuBackend::hvp_batch (src/gpu/mod.rs:149-160), hvp_batch_f64
 * (src/gpu/cuda_backend.rs:935): ONE point x [num_inputs], `batch`
 * directions [batch*num_inputs] -> grads, hvps [batch*num_inputs]. */
int32_t ehip_hvp_batch(ehip_ctx* ctx, const ehip_tape* tape, const void* x, const void* tangent_dirs,
                       uint32_t batch, void* grads, void* hvps);

/* Extension of hvp_batch with one point per direction: xs [batch*num_inputs]
 * (no reference counterpart; the reference replicates one x on the host,
 * src/gpu/cuda_backend.rs:244-247). */
int32_t ehip_hvp_batch_points(ehip_ctx* ctx, const ehip_tape* tape, const void* xs,
                              const void* tangent_dirs, uint32_t batch, void* grads, void* hvps);

/* GpuBackend::taylor_forward_kth_batch (src/gpu/mod.rs:206-224),
 * taylor_forward_kth_batch_f64 (src/gpu/cuda_backend.rs:854-873).
 * order in 1..=5; inputs seeded (x_i, v_i, 0, ...); coeffs receives
 * order * batch * num_outputs values, coefficient-major. */
int32_t ehip_taylor_kth_batch(ehip_ctx* ctx, const ehip_tape* tape, const void* primal_inputs,
                              const void* direction_seeds, uint32_t batch, uint32_t order, void* coeffs);

/* ---- device-pointer variants: buffers already in HBM, enqueued on `stream`
 * (a hipStream_t; NULL = the context's stream), no synchronization. ---- */
int32_t ehip_forward_batch_dev(ehip_ctx* ctx, const ehip_tape* tape, const void* d_inputs, uint64_t batch,
                               void* d_outputs, void* stream);
int32_t ehip_gradient_batch_dev(ehip_ctx* ctx, const ehip_tape* tape, const void* d_inputs, uint64_t batch,
                                void* d_outputs, void* d_grads, void* stream);
/* x_stride 0: one shared point; x_stride num_inputs: one point per row. */
int32_t ehip_hvp_batch_dev(ehip_ctx* ctx, const ehip_tape* tape, const void* d_x, uint64_t x_stride,
                           const void* d_dirs, uint64_t batch, void* d_grads, void* d_hvps, void* stream);
/* coeffs interleaved [batch*num_outputs][order] on the device (the kernel's
 * native layout; the host entry point de-interleaves). */
int32_t ehip_taylor_kth_batch_dev(ehip_ctx* ctx, const ehip_tape* tape, const void* d_x, uint64_t x_stride,
                                  const void* d_dirs, uint64_t batch, uint32_t order, void* d_coeffs,
                                  void* stream);

/* ---- GPU STDE (src/gpu/stde_gpu.rs) ----
 * Results are binary64, reduced on the device: result5 = {value, estimate,
 * sample_variance, standard_error, num_samples} (EstimatorResult,
 * src/stde/types.rs:9-22).  All require a scalar-output tape. */

/* laplacian_gpu (src/gpu/stde_gpu.rs:51-81): one point x [n], S directions
 * [S*n]; samples 2*c2 through the Welford accumulator. */
int32_t ehip_laplacian(ehip_ctx* ctx, const ehip_tape* tape, const void* x, const void* dirs, uint32_t S,
                       double* result5);

/* laplacian_with_control_gpu (src/gpu/stde_gpu.rs:135-200): residual
 * 2*c2 - sum_j d_j v_j^2, estimate tr(D) + mean. */
int32_t ehip_laplacian_with_control(ehip_ctx* ctx, const ehip_tape* tape, const void* x, const void* dirs,
                                    uint32_t S, const void* control_diagonal, double* result5);

/* hessian_diagonal_gpu (src/gpu/stde_gpu.rs:103-129): n basis jets,
 * value [1] and diag [n] in the tape's scalar type. */
int32_t ehip_hessian_diagonal(ehip_ctx* ctx, const ehip_tape* tape, const void* x, void* value, void* diag);

/* Batched STDE over many points (no reference counterpart; the reference
 * handles one x per call): xs [npoints*n], S directions per point either
 * supplied (dir_kind 0, dirs [npoints*S*n]) or drawn on the device from the
 * Philox stream `seed` (1 Rademacher, 2 Gaussian; component i of direction s
 * of point p is direction component (p*S+s)*n+i: Rademacher component k is
 * bit k mod 128 of Philox block k / 128, Gaussian components 4k .. 4k+3
 * are the cosine and sine branches of block k's two binary32 Box-Muller
 * pairs, radius and angle from words (0, 1) and (2, 3), each uniform
 * (top 23 bits + 1/2) 2^-23).  results [npoints*5]. */
int32_t ehip_laplacian_batch(ehip_ctx* ctx, const ehip_tape* tape, const void* xs, uint64_t npoints, uint32_t S,
                             uint32_t dir_kind, uint64_t seed, const void* dirs, double* results);
/* Device-pointer form; d_diag (optional) is a diagonal control variate with
 * row stride diag_stride (0 = shared). */
int32_t ehip_laplacian_batch_dev(ehip_ctx* ctx, const ehip_tape* tape, const void* d_xs, uint64_t npoints,
                                 uint32_t S, uint32_t dir_kind, uint64_t seed, const void* d_dirs,
                                 const void* d_diag, uint64_t diag_stride, double* d_results, void* stream);

/* Introspection of the compiled register program (for tests and the
 * traffic model): mode 0 forward, 1 gradient, 2 HVP. out[0]=instructions,
 * out[1]=registers, out[2]=program words, out[3]=fused dot terms.  Mode 3,
 * the f32 order-3 jet program cut for the matrix cores (ehip_set_jet_mma):
 * out[0]=segments, out[1]=batch-GEMM segments, out[2]=GEMMs that apply
 * their layer's activation in the epilogue, out[3]=GEMMs that also sum the
 * read-out dot product over their members (which are then never stored).
 * Mode 4, that program's direct plan (used when the directions are drawn on
 * the device): out[0]=GEMMs that read x and the directions themselves,
 * out[1]=inputs its LOADN no longer loads, out[2]=interpreter segments it
 * skips, out[3]=0. */
int32_t ehip_program_info(const ehip_tape* tape, int32_t mode, uint64_t* out4);

/* Compile a tape into its register program without a device (inspection
 * and tests).  Same validation as ehip_tape_upload.  fuse: 0 none, 1 dot
 * fusion and DENSE grouping (what uploads use), 2 dot fusion only.  Mode 0
 * forward, 1 gradient: writes up to `words_cap` program words (8 per
 * instruction) and `terms_cap` side-table words; the side table needs
 * max(4, T*(2+IW)) words (IW = 1 f32, 2 f64).  Mode 2 forward, 3 gradient:
 * writes the windowed instruction stream the kernels execute (512-word
 * windows, dot terms inline) into `words`; `terms` is unused.  Every mode
 * writes the DENSE pool (operand lists and weights) into `pool`.  Adding 4
 * to a gradient mode compiles the HVP program of the dual interpreter
 * (DENSE_REV terms carry P_DEEP).  Adding 8 to a forward mode groups
 * each layer's members into one DENSE, the program the f32 order-3 jets
 * run cut into interpreter and batch-GEMM segments (ehip_set_jet_mma).
 * out5 =
 * {instructions, registers, program or stream words, side-table entries T,
 * pool words}.  Returns EHIP_OTHER if a buffer is too small (out5 still
 * holds the sizes). */
int32_t ehip_compile_program(int32_t dtype, const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1,
                             const double* constants, uint32_t num_ops, uint32_t num_inputs,
                             uint32_t output_index, const uint32_t* output_indices, uint32_t num_outputs,
                             int32_t mode, int32_t fuse, uint32_t* words, uint64_t words_cap, uint32_t* terms,
                             uint64_t terms_cap, uint32_t* pool, uint64_t pool_cap, uint64_t* out5);

/* Launch shape of untuned launches: workgroup size (64/128/256) and the
 * per-workgroup LDS budget for the register file in bytes.  0 = default
 * (256 lanes, 24576 bytes). */
int32_t ehip_set_launch(ehip_ctx* ctx, uint32_t wg_size, uint32_t lds_budget_bytes);

/* Program fetch for untuned launches and the multi-tape queue: 0 stages the
 * instruction stream in LDS windows per workgroup, 1 reads it through the
 * scalar cache per wave (no window LDS).  Results are identical. */
int32_t ehip_set_fetch(ehip_ctx* ctx, int32_t mode);

/* One job of a multi-tape queue: `batch` points of `tape`, inputs [batch][nin]
 * on the device; what = 0 forward (d_y [batch][nout]) or 1 gradient (d_g
 * [batch][nin], d_y optional). */
typedef struct ehip_job {
  const ehip_tape* tape;
  int32_t what;
  const void* d_x;
  uint64_t batch;
  void* d_y;
  void* d_g;
} ehip_job;

/* Many tapes of uneven length and point count in one persistent launch (no
 * counterpart in the reference, which launches one tape per call:
 * src/gpu/cuda_backend.rs:861-913).  Workgroups take (tape, point-tile) items
 * from a device work queue, longest program first; tile width = workgroup
 * size (64/128/256, 0 = 256) and `lds_budget` bytes of LDS hold each tile's
 * first registers (0 = 24576).  All tapes share one dtype.  Results equal
 * the per-tape *_dev calls bit for bit.  Synchronizes `stream` before return. */
int32_t ehip_multi_batch_dev(ehip_ctx* ctx, const ehip_job* jobs, uint32_t njobs, uint32_t wg, uint32_t lds_budget,
                             void* stream);

/* Autotuner (no counterpart in the reference, whose launch shape is fixed at
 * 256 threads per block in src/gpu/cuda_backend.rs).  Times every launch
 * shape -- workgroup size 64/128/256 x LDS budget for the register file --
 * on the caller's device buffers for evaluation `what` (0 forward, 1
 * gradient, 2 hvp, 3 taylor of `order`), keeps the fastest on the tape for
 * later calls of that kind, and returns out5 = {workgroup size, LDS budget
 * bytes, microseconds per launch, program fetch (0 LDS, 1 scalar, 2 the
 * tape-specialised kernel), points per lane}.  The shapes are workgroup size
 * 64/128/256 x LDS budget x program fetch, plus the specialised kernel at
 * each workgroup size with 1, 2 or 4 points per lane (a grid-stride loop)
 * when ehip_tape_specialize has built one; the interpreters run one point
 * per lane.  Buffers are as for the matching *_dev
 * entry point (d_y may be null for gradient/hvp); their contents are
 * overwritten.  Runs on the context stream and synchronizes. */
int32_t ehip_autotune_dev(ehip_ctx* ctx, ehip_tape* tape, int32_t what, uint32_t order, const void* d_x,
                          uint64_t x_stride, const void* d_dirs, uint64_t batch, void* d_y, void* d_g, void* d_h,
                          uint32_t* out5);

/* Allocation-free ehip_gradient_batch_dev.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  Same arguments,
 * results and launch shape as ehip_gradient_batch_dev, but the interpreters'
 * HBM register file lives in a fixed device array of
 * ehip_noalloc_scratch_bytes() bytes (64 MiB) per GPU, held until the
 * process ends, and the call allocates nothing: batches whose register file
 * exceeds it run as several launches of whole workgroups on `stream`.  A
 * tape whose single workgroup needs more than the array is refused
 * (EHIP_OTHER).  Every call on one GPU shares that GPU's array (across
 * contexts too): calls that overlap on one GPU race; serialize them. */
int32_t ehip_gradient_batch_dev_noalloc(ehip_ctx* ctx, const ehip_tape* tape, const void* d_inputs, uint64_t batch,
                                        void* d_outputs, void* d_grads, void* stream);
uint64_t ehip_noalloc_scratch_bytes(void);

/* hipGraph capture of repeated evaluations (no counterpart in
 * the reference).  Records `repeats` launches of evaluation `what` (as for
 * ehip_autotune_dev) over fixed device buffers into one graph with its own
 * register workspace; ehip_graph_launch replays it on `stream` (null: the
 * context stream).  The recorded launches contain no memset nodes.  A graph
 * must be destroyed before its tape and context. */
typedef struct ehip_graph ehip_graph;
int32_t ehip_graph_capture(ehip_ctx* ctx, ehip_tape* tape, int32_t what, uint32_t order, const void* d_x,
                           uint64_t x_stride, const void* d_dirs, uint64_t batch, void* d_y, void* d_g, void* d_h,
                           uint32_t repeats, ehip_graph** out);
int32_t ehip_graph_launch(ehip_graph* graph, void* stream);
void ehip_graph_destroy(ehip_graph* graph);

/* Restore (wg != 0) or clear (wg == 0) a tuned launch shape, e.g. from a
 * persisted autotuner cache; ehip_get_tuned reads it back
 * (out4 = {wg, lds, fetch, points per lane}).  fetch 2 selects the
 * tape-specialised kernel (every evaluation but kinks, once
 * ehip_tape_specialize has built it), which alone takes points_per_lane
 * above 1 (up to 64; 0 means 1). */
int32_t ehip_set_tuned(ehip_tape* tape, int32_t what, uint32_t order, uint32_t wg, uint32_t lds_budget,
                       uint32_t fetch, uint32_t points_per_lane);
int32_t ehip_get_tuned(const ehip_tape* tape, int32_t what, uint32_t order, uint32_t* out4);

/* Tape-specialised kernels (short tapes keep their slots in registers; no
 * counterpart in the reference, whose CUDA backend generates
 * source at run time per Taylor order, src/gpu/taylor_codegen.rs, but never
 * per tape).  Writes the
 * tape's forward (what 0), value-and-gradient (what 1), gradient-and-HVP
 * (what 2, tangent-over-reverse on dual numbers) or Taylor-jet sweep of
 * order 1..5 (what 3..7, also serving the STDE entry points) as straight-line HIP,
 * every entry a named value kept in VGPRs and every constant an instruction
 * literal, and compiles it with hipRTC for the tape's device.  Arithmetic
 * and order are the interpreters', so results are the same.  Once built it
 * runs every untuned evaluation of that kind; ehip_set_tuned (fetch 2, or
 * 0/1 for the interpreters) and ehip_autotune_dev choose explicitly.
 * waves_per_eu (1..8) asks the compiler for that occupancy, spilling what
 * does not fit; 0 = automatic: when the compiler's own choice allows one
 * wave per SIMD, the kernel is built again for two and that build is kept.  out5 = {VGPRs and AGPRs, scratch bytes per
 * lane, compile microseconds (all builds), code object bytes, the
 * occupancy target used (0: the compiler's)}.  Tapes longer than 65536
 * entries are refused.  The compile runs outside the context lock.  Building
 * a kernel again makes the new one serve later calls; the one it replaces
 * stays loaded until the tape is freed, so graphs captured with it replay
 * unchanged.
 * what 8 (f32 tapes only) builds the value-and-gradient kernel with binary64
 * adjoints: the forward sweep and every partial stay FP32, and each entry's
 * adjoint -- the sum of one contribution per use of the entry -- is kept in
 * binary64 and rounded to FP32 once, at the gradient store, so cancelling
 * contributions lose no FP32 bits.  It serves the gradient evaluations in
 * place of the what 1 kernel (rebuild what 1 to go back).  The reference
 * accumulates f32 adjoints in f32 (src/gpu/kernels/tape_eval.cu:340-344), so
 * results differ from it by that rounding; no counterpart in its API. */
int32_t ehip_tape_specialize(ehip_tape* tape, int32_t what, uint32_t waves_per_eu, uint32_t* out5);

/* Specialise every tape uploaded after this call whose entry count is at
 * most max_entries (<= 65536; 0 = off, the default), for the evaluations in
 * what_mask (bit 0 forward, bit 1 gradient, bit 2 hvp, bits 3..7 taylor
 * order 1..5; hvp only for scalar-output tapes).  A compile failure does
 * not fail the upload: that evaluation runs on the interpreters, and
 * ehip_spec_error gives the reason. */
int32_t ehip_set_specialize(ehip_ctx* ctx, uint32_t max_entries, uint32_t what_mask);

/* Why the upload-time specialisation of evaluation `what` (as for
 * ehip_tape_specialize) failed on this tape; "" when it did not fail or was
 * not asked for.  The text lives as long as the tape. */
const char* ehip_spec_error(const ehip_tape* tape, int32_t what);

/* The source ehip_tape_specialize compiles, from the flattened tape arrays
 * with no device (tape arguments as ehip_compile_program).  When `arch` is
 * non-null (e.g. "gfx950") the source is also compiled with hipRTC for it.
 * out3 = {source bytes (without the terminating zero), code object bytes,
 * compile microseconds}; `src` (may be null) receives the text. */
int32_t ehip_spec_source(int32_t dtype, const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1,
                         const double* constants, uint32_t num_ops, uint32_t num_inputs, uint32_t output_index,
                         const uint32_t* output_indices, uint32_t num_outputs, int32_t what,
                         uint32_t waves_per_eu, const char* arch, char* src, uint64_t src_cap, uint64_t* out3);

/* DENSE grouping for tapes uploaded after this call (default on): fused dot
 * products sharing one operand list run as one register-blocked instruction
 * (ehip_program.h); results are identical with it off. */
int32_t ehip_set_dense(ehip_ctx* ctx, int32_t enable);

/* Order-3 Taylor jets of f32 tapes (the STDE samples): run each DENSE
 * block whose weights are all finite on the FP32 matrix cores
 * (v_mfma_f32_32x32x2_f32, a wave's 64 jets as the N dimension) instead of
 * the exact per-term path (enable != 0; default off).  The sums are the
 * same products in another order, fused, so results agree with the exact
 * path to f32 rounding, not bit for bit; a wave whose operands hold an Inf
 * or NaN takes the exact path.  No counterpart in the reference. */
int32_t ehip_set_jet_mma(ehip_ctx* ctx, int32_t enable);

/* Dot-product fusion for tapes uploaded after this call (default on).  Off
 * runs every tape entry as its own instruction; results are identical. */
int32_t ehip_set_fusion(ehip_ctx* ctx, int32_t enable);

/* ------------------------------------------------------------------------
 * Structure for the sparse derivative paths (bit-exact with the reference)
 * and the GPU sparse Jacobian / Hessian.  Pattern detection works on the
 * flattened tape arrays (no device needed).  Two-call pattern for sizes:
 * call with cap 0 to learn *nnz (the call then returns EHIP_OTHER), then
 * again.
 * ------------------------------------------------------------------------ */
/* BytecodeTape::detect_sparsity (src/bytecode_tape/sparse.rs:42-52,
 * 463-600): lower-triangle COO (rows >= cols), sorted and deduplicated. */
int32_t ehip_hessian_sparsity(const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1, uint32_t num_ops,
                              uint32_t num_inputs, uint32_t* rows, uint32_t* cols, uint64_t cap, uint64_t* nnz);
/* BytecodeTape::detect_jacobian_sparsity (src/bytecode_tape/sparse.rs:54-67,
 * 733-776): (output, input) pairs, outputs in order, inputs ascending. */
int32_t ehip_jacobian_sparsity(const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1, uint32_t num_ops,
                               uint32_t num_inputs, const uint32_t* output_indices, uint32_t num_outputs,
                               uint32_t* rows, uint32_t* cols, uint64_t cap, uint64_t* nnz);
/* greedy_coloring (src/sparse.rs:57-98): distance-2 coloring of a symmetric
 * pattern; column_coloring / row_coloring (src/sparse.rs:268-300) of a
 * Jacobian pattern.  Decreasing-degree stable order, lowest free color. */
int32_t ehip_greedy_coloring(uint32_t dim, const uint32_t* rows, const uint32_t* cols, uint64_t nnz,
                             uint32_t* colors, uint32_t* num_colors);
int32_t ehip_column_coloring(uint32_t num_outputs, uint32_t num_inputs, const uint32_t* rows, const uint32_t* cols,
                             uint64_t nnz, uint32_t* colors, uint32_t* num_colors);
int32_t ehip_row_coloring(uint32_t num_outputs, uint32_t num_inputs, const uint32_t* rows, const uint32_t* cols,
                          uint64_t nnz, uint32_t* colors, uint32_t* num_colors);
/* GpuBackend::sparse_hessian (src/gpu/mod.rs:158-172, cuda_backend.rs:
 * 397-441): value, gradient [n], pattern and values at x (host, n). */
int32_t ehip_sparse_hessian(ehip_ctx* ctx, const ehip_tape* tape, const void* x, void* value, void* gradient,
                            uint32_t* rows, uint32_t* cols, void* values, uint64_t cap, uint64_t* nnz);
/* GpuBackend::sparse_jacobian (src/gpu/mod.rs:138-148, cuda_backend.rs:
 * 306-395): outputs [m], pattern and values at x (host, n). */
int32_t ehip_sparse_jacobian(ehip_ctx* ctx, const ehip_tape* tape, const void* x, void* outputs, uint32_t* rows,
                             uint32_t* cols, void* values, uint64_t cap, uint64_t* nnz);

/* ------------------------------------------------------------------------
 * Nonsmooth kink records per point (forward_nonsmooth,
 * src/bytecode_tape/forward.rs:87-187): one record per Abs / Min / Max /
 * Signum / Floor / Ceil / Round / Trunc / Fract entry of the whole tape, in
 * tape order.  The entries are structural (ehip_kink_info); the switching
 * values and branches (+1 / -1) are per point.
 * ------------------------------------------------------------------------ */
uint32_t ehip_kink_count(const ehip_tape* tape);
int32_t ehip_kink_info(const ehip_tape* tape, uint32_t* tape_index, uint32_t* opcode, uint32_t cap);
/* Forward values [batch][m] plus switching [batch][K] and branch [batch][K]
 * (K = ehip_kink_count); device buffers / host buffers. */
int32_t ehip_forward_nonsmooth_dev(ehip_ctx* ctx, const ehip_tape* tape, const void* d_x, uint64_t batch,
                                   void* d_outputs, void* d_switching, int8_t* d_branches, void* stream);
int32_t ehip_forward_nonsmooth(ehip_ctx* ctx, const ehip_tape* tape, const void* inputs, uint32_t batch,
                               void* outputs, void* switching, int8_t* branches);

/* ------------------------------------------------------------------------
 * Tape wire format: the serde JSON form of BytecodeTape
 * (src/bytecode_tape/serde_support.rs:11-100), host only (no device).
 * ------------------------------------------------------------------------ */
/* Serialize (serde_support.rs:11-31): serde_json's compact layout of
 * opcodes (variant names), arg_indices, values (non-finite -> null),
 * num_inputs, num_variables (= num_entries), output_index, output_indices,
 * custom_second_args ({}).  A Custom opcode returns
 * EHIP_CUSTOM_OPS_NOT_SUPPORTED with the reference's message.  `values` has
 * num_entries elements of `dtype`.  buf == NULL: *len = text length only;
 * otherwise cap must be >= *len + 1 (a NUL is written). */
int32_t ehip_tape_to_json(int32_t dtype, const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1,
                          const void* values, uint64_t num_entries, uint32_t num_inputs, uint32_t output_index,
                          const uint32_t* output_indices, uint64_t num_output_indices, char* buf, uint64_t cap,
                          uint64_t* len);
/* Deserialize (serde_support.rs:33-100): parse, refuse Custom opcodes,
 * non-empty custom_second_args and a num_variables stamp != opcodes.len(),
 * then BytecodeTape::validate (src/bytecode_tape/mod.rs:575-703); every
 * rejection is EHIP_OTHER with the reference's message.  Sizing call: all
 * array pointers NULL fills the four counts only.  output_indices is the
 * payload's list as written (possibly empty). */
int32_t ehip_tape_from_json(const char* json, uint64_t len, int32_t dtype, uint32_t* opcodes, uint32_t* arg0,
                            uint32_t* arg1, void* values, uint64_t cap_entries, uint32_t* output_indices,
                            uint64_t cap_outputs, uint64_t* num_entries, uint32_t* num_inputs,
                            uint32_t* output_index, uint64_t* num_output_indices);
/* serde_json::from_str::<BytecodeTape<F>> followed by upload_tape /
 * upload_tape_f64 (f32 values are the f64 parse rounded to f32). */
int32_t ehip_tape_upload_json(ehip_ctx* ctx, int32_t dtype, const char* json, uint64_t len, ehip_tape** out);
/* The same model as CBOR (RFC 8949), the binary form the reference
 * round-trips through ciborium (tests/serde.rs:156-170): a definite map with
 * text keys in field order, variant names as text, shortest unsigned heads,
 * floats in the shortest of binary16/32/64 that holds them exactly (NaN as
 * binary16 0x7e00; Inf and NaN round-trip).  Same two-call sizing, checks
 * and messages as the JSON pair; no terminator is written. */
int32_t ehip_tape_to_cbor(int32_t dtype, const uint32_t* opcodes, const uint32_t* arg0, const uint32_t* arg1,
                          const void* values, uint64_t num_entries, uint32_t num_inputs, uint32_t output_index,
                          const uint32_t* output_indices, uint64_t num_output_indices, uint8_t* buf, uint64_t cap,
                          uint64_t* len);
int32_t ehip_tape_from_cbor(const uint8_t* cbor, uint64_t len, int32_t dtype, uint32_t* opcodes, uint32_t* arg0,
                            uint32_t* arg1, void* values, uint64_t cap_entries, uint32_t* output_indices,
                            uint64_t cap_outputs, uint64_t* num_entries, uint32_t* num_inputs,
                            uint32_t* output_index, uint64_t* num_output_indices);

/* ------------------------------------------------------------------------
 * Learned low-rank control variate for STDE.  No counterpart in the
 * reference; it generalises laplacian_with_control (src/stde/laplacian.rs:
 * 90-129, src/gpu/stde_gpu.rs:135-200) from a fixed diagonal to
 * v^T L(x) L(x)^T v - tr(L L^T) with L(x) in R^{n x r} from an MLP
 * n -> hidden -> hidden -> n*r (tanh hidden layers), trained on GPU with
 * BF16-MFMA forward/backward GEMMs (FP32 accumulate) and Adam.
 * ------------------------------------------------------------------------ */
typedef struct ehip_cv ehip_cv;

/* New network with W ~ N(0, 1/fan_in) drawn from Philox (seed, stream 9) at
 * each parameter's flat index, the output layer scaled by out_scale, zero
 * biases.  rank 1..=32. */
int32_t ehip_cv_create(ehip_ctx* ctx, uint32_t n, uint32_t rank, uint32_t hidden, uint64_t seed, float out_scale,
                       ehip_cv** out);
void ehip_cv_destroy(ehip_cv* cv);
/* out5 = {n, rank, hidden, parameter count, Adam steps taken}. */
int32_t ehip_cv_info(const ehip_cv* cv, uint64_t* out5);
/* Copy the flat FP32 vector [W1 | b1 | W2 | b2 | W3 | b3] (which = 0), the
 * last gradients (1) or the Adam moments (2, 3) to the host. */
int32_t ehip_cv_get(ehip_cv* cv, int32_t which, float* host_out);
int32_t ehip_cv_set_params(ehip_cv* cv, const float* host_in);
/* Output gain g: L(x) = g * (W3 h2 + b3), so L L^T scales with g^2 (the
 * training loop then works on a target of unit scale).  Default 1. */
int32_t ehip_cv_set_gain(ehip_cv* cv, float gain);
float ehip_cv_gain(const ehip_cv* cv);
/* Signed factor: the control variate becomes v^T L S L^T v - tr(L S L^T)
 * with S = diag(+1 for the first npos columns, -1 for the rest), still mean
 * zero, able to follow indefinite Hessians.  Default npos = rank (L L^T). */
int32_t ehip_cv_set_signs(ehip_cv* cv, uint32_t npos);
/* L(x) for `batch` FP32 points: d_L [batch][n*r], L[i][j] at i*r + j. */
int32_t ehip_cv_forward_dev(ehip_cv* cv, const float* d_x, uint64_t batch, float* d_L, void* stream);
/* Gradients of mean_t ||hv_t - L(x_t) L(x_t)^T v_t||^2 / n into the
 * network's gradient buffer; the mean loss to *d_loss (device) if non-null. */
int32_t ehip_cv_loss_grad_dev(ehip_cv* cv, const float* d_x, const float* d_v, const float* d_hv, uint64_t batch,
                              float* d_loss, void* stream);
/* One Adam step (PyTorch form) on the last gradients. */
int32_t ehip_cv_adam_step(ehip_cv* cv, float lr, float beta1, float beta2, float eps, void* stream);
/* `epochs` passes over `ntuples` tuples in minibatches of `batch` (Adam,
 * betas 0.9/0.999, eps 1e-8); per-step mean losses to d_loss_hist
 * [epochs * ceil(ntuples / batch)] when non-null. */
int32_t ehip_cv_train_dev(ehip_cv* cv, const float* d_x, const float* d_v, const float* d_hv, uint64_t ntuples,
                          uint32_t batch, uint32_t epochs, float lr, float* d_loss_hist, void* stream);
/* Training tuples from the tape: x_t ~ N(0, I) (Philox seed, stream 1),
 * v_t ~ N(0, I) (seed, stream 7) and H(x_t) v_t from the HVP kernel, all
 * written as FP32 [ntuples][n]. */
int32_t ehip_cv_make_tuples_dev(ehip_ctx* ctx, const ehip_tape* tape, uint64_t ntuples, uint64_t seed, float* d_x,
                                float* d_v, float* d_hv, void* stream);
/* Batched STDE with the learned control variate (cv may be null: plain):
 * S device-drawn directions per point (1 Rademacher, 2 Gaussian), results
 * [npoints][5] as ehip_laplacian_batch_dev; d_plain_results (optional)
 * receives the plain estimator's statistics of the same samples.  The pass
 * holds L(x) in bfloat16 (the network's FP32 output rounded to nearest
 * even) and forms both v^T L L^T v and tr(L L^T) from those values. */
int32_t ehip_laplacian_cv_batch_dev(ehip_ctx* ctx, const ehip_tape* tape, ehip_cv* cv, const void* d_x,
                                    uint64_t npoints, uint32_t S, uint32_t dir_kind, uint64_t seed,
                                    double* d_results, double* d_plain_results, void* stream);
/* Exact tr(H) per point from n basis-direction Taylor jets (results[p][1];
 * layout as ehip_laplacian_batch_dev, with zero variance up to rounding). */
int32_t ehip_hessian_trace_batch_dev(ehip_ctx* ctx, const ehip_tape* tape, const void* d_x, uint64_t npoints,
                                     double* d_results, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* ECHIDNA_HIP_H */
