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
  float learned_trust;  // largest predicted change (thermal units) taken as a start
  int refill_chunk;     // D3Q27 WENO5C single pass: 64-cell tiles per wave with lane refill (0: one tile per wave)
  float w[kQmax];
};

// The predicted (learned) start is taken only in cells whose previous step took
// kLearnedPrevMin .. kLearnedPrevMax gauge iterations (step_split.hip gives the measurements).
constexpr int kLearnedPrevMin = 3, kLearnedPrevMax = 4;

// D3Q27 split populations are stored in seven quad planes: population k of cell c at float index
// ((k / 4) * plane + c) * 4 + k % 4 (16 bytes per cell and quad plane, the 28th slot unused), so
// a stencil node's 27 populations are seven 16-byte loads.  D3Q125 stores each cell's 125
// populations as one 512-byte record (index 128 c + k, three floats unused): the step's block for
// a cell reads each of its 27 neighbours' populations as one contiguous record (one population
// per plane put every read in its own cache line and page: at 256^3 the step was bound by that
// gather).
constexpr int kSplit27Floats = 28;   // floats per cell of the D3Q27 split population storage
constexpr int kSplit125Floats = 128;  // floats per cell of the D3Q125 split population storage
__host__ __device__ inline int64_t split27_fidx(int k, int64_t c, int64_t plane) {
  return (((int64_t)(k >> 2) * plane + c) << 2) + (k & 3);
}
__host__ __device__ inline int64_t split_fidx(int lattice, int k, int64_t c, int64_t plane) {
  return lattice == kD3Q27 ? split27_fidx(k, c, plane) : c * kSplit125Floats + k;
}
// floats of one split population buffer
inline int64_t split_fcount(int lattice, int qn, int64_t plane) {
  return (lattice == kD3Q27 ? kSplit27Floats : kSplit125Floats) * plane;
}

struct SplitPtrs {
  float* f;   // populations (layout above)
  float* m;   // rho, ux, uy, uz planes
  double* T;  // temperature plane
};

using SplitPending = Pending;

// node gauge records (ux, uy, uz, sqrt(|T|)*sign(T)), 16 bytes per cell, of a D3Q27 split state;
// also sets its non-finite node data to the values that add only finite transfer terms
hipError_t launch_node_gauge(float* f, const float* m, const double* T, float* g4, int64_t n, int64_t plane,
                             hipStream_t st);
hipError_t launch_step_split(int lattice, int interp, const SplitPtrs& in, const SplitPtrs& out, const float* g4,
                             int* niter, const float* guess, const SplitPending& pend, const SplitArgs& a,
                             hipStream_t st);

}  // namespace pond
