// pond_ctx.h -- the opaque context behind the C ABI (include/pond.h).
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "../../include/pond.h"
#include "lattice.h"
#include "pond_device.h"

namespace pond {

// device buffers of one state (load or store) of the plain FP64 path
struct F64Buf {
  double* f = nullptr;  // Qn planes
  double* m = nullptr;  // rho, ux, uy, uz, T planes
};

// device buffers of one state of the split path: FP32 populations and (rho, ux, uy, uz),
// FP64 temperature
struct SplitBuf {
  float* f = nullptr;   // Qn planes
  float* m = nullptr;   // rho, ux, uy, uz planes
  double* T = nullptr;  // T plane
};

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
};

hipError_t launch_step_f64(int lattice, int interp, const double* fin, const double* min, double* fout,
                           double* mout, int* niter, const StepArgs& a, hipStream_t st);

}  // namespace pond

struct pond_ctx {
  pond_params p{};
  pond::LatticeTables L{};
  pond::Grid g{};
  pond::StepArgs args{};
  int Qn = 0;
  hipStream_t stream = nullptr;
  int cur = 0;              // index of the buffer holding the current state
  bool has_state = false;
  pond::F64Buf f64[2];
  pond::SplitBuf sp[2];
  int* niter = nullptr;     // per-cell gauge iterations of the last step
  double* red = nullptr;    // reduction partials
  int64_t red_len = 0;
  int64_t step = 0;
  int64_t bytes = 0;
  // timing of the step kernel launches (HIP events on `stream`)
  std::vector<pond::EventPair> ev_pool;
  std::vector<pond::EventPair> ev_pending;
  double last_ms = 0, total_ms = 0, last_mlp_ms = 0;
  int64_t timed_steps = 0;
  std::string err;
};
