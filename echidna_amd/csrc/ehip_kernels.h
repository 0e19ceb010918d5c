// Kernel launch interface shared by the host API and the HIP interpreters.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ehip {

struct Launch {
  const uint32_t* prog;   // NW words per instruction (device)
  const uint32_t* terms;  // DOT side table (device)
  uint32_t ninstr;
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
};

hipError_t launch_scalar(bool f64, const Launch& L, uint32_t wg, size_t lds, hipStream_t s);
hipError_t launch_dual(bool f64, const Launch& L, uint32_t wg, size_t lds, hipStream_t s);
hipError_t launch_jet(bool f64, int K, const Launch& L, uint32_t wg, size_t lds, hipStream_t s);

}  // namespace ehip
