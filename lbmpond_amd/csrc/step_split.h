// step_split.h -- arguments of the FP32/FP64-split step kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "pond_device.h"

namespace pond {

struct SplitArgs {
  Grid g;
  double dtau;
  double tol_abs, tol_rel;  // FP64 gauge residual tolerances (thermal-speed / T scaled)
  int max_iter;
  double sqrt_tbase;
  double tlat;
  double inv_dim;
  int pass1_iters;  // iterations of the first pass of the D3Q27 step (the rest go to the second)
  int weno5c_variant;   // D3Q27 WENO5C kernel: 0 grouped targets, 1 grouped at 3 waves/SIMD, 2 per target,
                        // 3 grouped with one x slice per load
  float learned_trust;  // largest predicted change (thermal units) taken as a start
  int refill_chunk;     // D3Q27 WENO5C single pass: 64-cell tiles per wave with lane refill (0: one tile per wave)
  float w[kQmax];
};

struct SplitPtrs {
  float* f;   // Qn planes
  float* m;   // rho, ux, uy, uz planes
  double* T;  // temperature plane
};

using SplitPending = Pending;

// node gauge records (ux, uy, uz, sqrt(|T|)*sign(T)), 16 bytes per cell, of a split state
hipError_t launch_node_gauge(const float* m, const double* T, float* g4, int64_t n, int64_t plane, hipStream_t st);
hipError_t launch_step_split(int lattice, int interp, const SplitPtrs& in, const SplitPtrs& out, const float* g4,
                             int* niter, const float* guess, const SplitPending& pend, const SplitArgs& a,
                             hipStream_t st);

}  // namespace pond
