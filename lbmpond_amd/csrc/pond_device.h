// pond_device.h -- device-side helpers and kernel argument structs shared by the step kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "lattice.h"

namespace pond {

// Cell-grid geometry.  Cells are stored x fastest (structs.cuh:153-166); every field is a
// plane of `plane` elements (plane >= n, padded so each plane starts 256-byte aligned).
struct Grid {
  int nx, ny, nz;
  int64_t n;
  int64_t plane;
};

// Parameters of one step, passed by value to the kernels.
struct StepArgs {
  Grid g;
  double dtau;
  double tol_abs, tol_rel;
  int max_iter;
  double sqrt_tbase;  // sqrt(TLat), evaluated once on the host (structs.cuh:530)
  double tlat;
  double inv_dim;     // 1./Dim (structs.cuh:330)
  double w[kQmax];    // lattice weights for the collide and set_base
};

__device__ __forceinline__ int isign_d(double v) { return (0.0 < v) - (v < 0.0); }

// params.h:110-112
__device__ __forceinline__ double sqrts_d(double v) { return sqrt(fabs(v)) * (double)isign_d(v); }

__device__ __forceinline__ bool finite_d(double v) { return !__builtin_isnan(v) && !__builtin_isinf(v); }

// periodic wrap, exactly (v % n + n) % n for any v (steps.cu:1015-1017)
__device__ __forceinline__ int wrapi(int v, int n) {
  if (v >= 0 && v < n) return v;
  int r = v % n;
  return r < 0 ? r + n : r;
}

// floor(v) as a stencil coordinate.  A double->int conversion of NaN or of a value outside the
// int range is undefined (LLVM: poison), and a poisoned coordinate turns into an arbitrary
// address after the periodic wrap.  NaN maps to 0, as CUDA's conversion does, and values are
// clamped to +-2^30 cells so that small offsets added to the result cannot overflow.  For every
// finite departure point closer than 2^30 cells the result is exactly (int)floor(v).
__device__ __forceinline__ int floor_int(double v) {
  const double f = floor(v);
  if (!(f == f)) return 0;
  if (f > 1073741824.0) return 1073741824;
  if (f < -1073741824.0) return -1073741824;
  return (int)f;
}

__device__ __forceinline__ int64_t cell_index(const Grid& g, int x, int y, int z) {
  return (int64_t)x + (int64_t)g.nx * ((int64_t)y + (int64_t)g.ny * (int64_t)z);
}

}  // namespace pond
