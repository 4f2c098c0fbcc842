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
  // 64-cell wave tiles of the one-lane-per-cell kernels: tile b holds 64 cells of one x row
  // (rp > 0) or the cells 64b .. 64b+63 of the linear order, across row ends (rp = 0)
  int64_t ntiles;
  // rows (y, z) taken in rp x rp patches (y fastest within a patch, patches y fastest then z)
  // instead of y fastest over the whole plane; 0 = plain row order.  Needs nx % 64 == 0,
  // ny % rp == 0 and nz % rp == 0.
  int rp;
  // mirror-boundary axes, bit d for axis d (0: periodic in every axis)
  int mirror;
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
  int pass1_iters;    // iterations of the first pass of the one-lane-per-cell step
  int global_loads;   // 1: global loads even where the buffer-resource loads apply
  int refill_chunk;   // D3Q27 first pass: 64-cell tiles per wave with lane refill (0: one tile per wave)
  const int* cell_set;  // Cinfo::set per cell (0: not stepped), or null: every cell is stepped
  double w[kQmax];    // lattice weights for the collide and set_base
};

// Cells left unconverged by the first pass of a two-pass step and their iteration state,
// structure of arrays with one entry per cell of the grid.
struct Pending {
  int* count;     // number of entries (device)
  int* cell;      // cell index
  double* g;      // gauge, 4 planes of cap entries
  double* res;    // one double per entry (split: residual of the last iteration; FP64: rho)
  int* meta;      // iteration count (| learned << 16 in the split step)
  int64_t cap;    // capacity
};

// Append the pending lanes of a wave to the list: one atomic per wave, entries in lane order.
// Returns the lane's entry index (valid where pending).
__device__ __forceinline__ int64_t pending_slot(const Pending& p, bool pending, bool& any) {
  const unsigned long long mask = __ballot(pending);
  any = mask != 0;
  if (!any) return 0;
  const int lane = threadIdx.x & 63;
  const int rank = __popcll(mask & ((1ull << lane) - 1ull));
  const int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(p.count, __popcll(mask));
  base = __shfl(base, leader, 64);
  return (int64_t)base + rank;
}

__global__ void k_zero_count(int* c);

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

// Mirror boundaries (propagate<N>, steps.cu:1004-1017): on a mirrored axis a coordinate below 0
// or at or above n is reflected about that face (-1-v, or n-1-v+n) and `rev` is set; then the
// periodic wrap, as in the reference (which wraps after reflecting).
__device__ __forceinline__ int mirror_wrap(int v, int n, bool mirrored, bool& rev) {
  rev = false;
  if (mirrored) {
    if (v < 0) { v = -1 - v; rev = true; }
    if (v >= n) { v = n - 1 - v + n; rev = true; }
  }
  return wrapi(v, n);
}

// mirror image of a D1Q5 / D3Q125 per-axis velocity class {0, +e1, -e1, +e2, -e2}
constexpr int kMir5[5] = {0, 2, 1, 4, 3};

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

// Workgroups are handed to the 8 XCDs of the chip round-robin (workgroup b runs on XCD b % 8),
// and each XCD has its own L2.  xcd_block maps the hardware order to a logical block order in
// which XCD k works through one contiguous slab of blocks, so the stencil neighbourhoods of the
// workgroups resident on one XCD overlap in its L2.  A bijection on [0, nblocks) for any count.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nblocks) {
  constexpr int kXcd = 8;
  const int64_t q = nblocks / kXcd, r = nblocks % kXcd;
  const int64_t x = b % kXcd, local = b / kXcd;
  return x * q + (x < r ? x : r) + local;
}

// cell (x, y, z) of lane `lane` in tile `b`; false outside the grid.  Without row patches the
// tiles are the linear order: tile b holds cells 64b .. 64b+63, across row ends.
__device__ __forceinline__ bool tile_cell(const Grid& g, int64_t b, int lane, int& x, int& y, int& z) {
  if (g.rp > 0) {
    // the workgroups resident on one XCD at a time then cover a compact (y, z) patch of rows,
    // whose stencil halo is smaller than that of a run of rows along y
    const int wpr = g.nx / 64;
    const int64_t row = b / wpr;
    const int64_t pp = (int64_t)g.rp * g.rp, t = row / pp;
    const int w = (int)(row % pp), npy = g.ny / g.rp;
    x = (int)(b % wpr) * 64 + lane;
    y = (int)(t % npy) * g.rp + w % g.rp;
    z = (int)(t / npy) * g.rp + w / g.rp;
    return row < (int64_t)g.ny * g.nz;
  }
  const int64_t c = b * 64 + lane;
  x = (int)(c % g.nx);
  y = (int)((c / g.nx) % g.ny);
  z = (int)(c / ((int64_t)g.nx * g.ny));
  return c < g.n;
}

// FP64 population storage.  D3Q27 keeps the populations in fourteen pair planes: population k
// of cell c at double index ((k / 2) * plane + c) * 2 + k % 2 (16 bytes per cell and pair
// plane, the 28th slot unused), so a stencil node's 27 populations are fourteen 16-byte loads.
// D1Q5 keeps one plane per population (index k * plane + c).  D3Q125 keeps each cell's 125
// populations as one 1 KB record (index 128 c + k, three doubles unused): its step's block for a
// cell reads the whole records of the cell's 27 neighbours.
constexpr int kF64Q125Doubles = 128;
__host__ __device__ inline int64_t f64_fidx(int lattice, int k, int64_t c, int64_t plane) {
  return lattice == kD3Q27    ? ((((int64_t)(k >> 1) * plane + c) << 1) + (k & 1))
         : lattice == kD3Q125 ? c * kF64Q125Doubles + k
                              : (int64_t)k * plane + c;
}
// doubles of one FP64 population buffer
inline int64_t f64_fcount(int lattice, int qn, int64_t plane) {
  return (lattice == kD3Q27 ? 28 : lattice == kD3Q125 ? kF64Q125Doubles : qn) * plane;
}

__device__ __forceinline__ int64_t cell_index(const Grid& g, int x, int y, int z) {
  return (int64_t)x + (int64_t)g.nx * ((int64_t)y + (int64_t)g.ny * (int64_t)z);
}

}  // namespace pond
