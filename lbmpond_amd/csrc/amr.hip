// amr.hip -- the 1D cell activation passes of amr.cu, run after a D1Q5 FP64 step when
// pond_params.amr is set (as calcStep would run them with the launches of update.cu:102-105
// enabled).
//
// Records: FP64 D1Q5 planes (population k at k*plane + cell, moments rho, ux, uy, uz, T at
// k*plane + cell).  `st` is the step's output (the reference's storeGroups), `ld` its input.
#include "pond_ctx.h"

namespace pond {

namespace {

// Longest line the removal pass stages in LDS (one byte per cell).
constexpr int kAmrMaxLine = 32768;

// AMR_add (amr.cu:1-27), one lane per cell: an active cell finds the next active cell along x
// (wrapping); where their new states differ (|du_x|^2 / min T > 0.01 or |1 - rho_m / rho_p| >
// 0.01) and the gap is an even number of cells, the cell half way gets the average of the two
// records in both buffers (Cell::save_12) and is switched on from the next step (setnew).  A lane
// reads `set` and the store records of active cells only, and writes only a cell strictly inside
// its own gap, so the lanes are independent.
__global__ __launch_bounds__(256) void k_amr_add(double* stf, double* stm, double* ldf, double* ldm,
                                                 const int* __restrict__ set, int* __restrict__ setnew, Grid g,
                                                 int qn) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.n || !set[c]) return;
  const int x = (int)(c % g.nx);
  const int64_t row = c - x;
  const int64_t P = g.plane;
  int xp = x, dist = 0;
  do {  // ends at the latest at the cell itself, which is active
    xp = xp + 1 == g.nx ? 0 : xp + 1;
    dist++;
  } while (!set[row + xp]);
  const int64_t cp = row + xp;
  const double vdif = stm[P + cp] - stm[P + c];
  const double Tm = stm[4 * P + c], Tp = stm[4 * P + cp];
  if (!(vdif * vdif / fmin(Tm, Tp) > 0.01 || fabs(1 - stm[c] / stm[cp]) > 0.01)) return;
  if (dist % 2 != 0) return;
  const int64_t mid = row + (x + dist / 2) % g.nx;
  for (int k = 0; k < 5; k++) {
    const double v = 0.5 * (stm[k * P + c] + stm[k * P + cp]);
    stm[k * P + mid] = v;
    ldm[k * P + mid] = v;
  }
  for (int i = 0; i < qn; i++) {
    const double v = 0.5 * (stf[i * P + c] + stf[i * P + cp]);
    stf[i * P + mid] = v;
    ldf[i * P + mid] = v;
  }
  setnew[mid] = 1;
}

// AMR_remove (amr.cu:28-55) on the row y = z = 0: one pass in increasing x, each decision seeing
// the earlier ones.  An active cell is switched off when its previous and next active cells lie
// within 3 cells and both match it (|du_x|^2 / T < 0.01 and |rho ratio - 1| < 0.01, with its own
// T and rho).  Whether the cell k places back (or ahead) matches does not depend on the pass, so
// the whole block first works out, for every cell, those six comparisons (bits 1-3: k = 1..3
// back, bits 4-6: ahead) next to its flag (bit 0) in LDS; one lane then makes the ordered pass
// over the bytes.  A search that finds no active cell within 3 decides "keep" as the
// reference's unbounded search does (it would find one 4 or more cells away), and on lines
// shorter than 4 cells it may come back to the cell itself, as the reference's does.
__global__ __launch_bounds__(1024) void k_amr_remove(const double* __restrict__ stm, int* __restrict__ set, Grid g) {
  __shared__ unsigned char L[kAmrMaxLine];
  const int nx = g.nx;
  const int64_t P = g.plane;
  for (int x = threadIdx.x; x < nx; x += blockDim.x) {
    const double r0 = stm[x], u0 = stm[P + x], T0 = stm[4 * P + x];
    unsigned b = set[x] ? 1u : 0u;
    for (int k = 1; k <= 3; k++) {
      const int xm = ((x - k) % nx + nx) % nx, xp = (x + k) % nx;
      const double vm = stm[P + xm] - u0, vp = stm[P + xp] - u0;
      if (vm * vm / T0 < 0.01 && fabs(stm[xm] / r0 - 1) < 0.01) b |= 1u << k;
      if (vp * vp / T0 < 0.01 && fabs(stm[xp] / r0 - 1) < 0.01) b |= 1u << (3 + k);
    }
    L[x] = (unsigned char)b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int x = 0; x < nx; x++) {
      const unsigned b = L[x];
      if (!(b & 1u)) continue;
      int dm = 0, dp = 0;
      for (int k = 1; k <= 3 && !dm; k++)
        if (L[((x - k) % nx + nx) % nx] & 1u) dm = k;
      for (int k = 1; k <= 3 && !dp; k++)
        if (L[(x + k) % nx] & 1u) dp = k;
      if (dm && dp && (b >> dm & 1u) && (b >> (3 + dp) & 1u)) L[x] = (unsigned char)(b & ~1u);
    }
  }
  __syncthreads();
  for (int x = threadIdx.x; x < nx; x += blockDim.x) set[x] = L[x] & 1u;
}

__global__ void k_copy_i32(const int* __restrict__ a, int* __restrict__ b, int64_t n) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) b[c] = a[c];
}

}  // namespace

int amr_max_line() { return kAmrMaxLine; }

hipError_t launch_amr(const F64Buf& st, const F64Buf& ld, int* set, int* setnew, const Grid& g, int qn,
                      hipStream_t stream) {
  if (g.ny != 1 || g.nz != 1 || g.nx > kAmrMaxLine) return hipErrorInvalidValue;
  const unsigned blocks = (unsigned)((g.n + 255) / 256);
  hipLaunchKernelGGL(k_copy_i32, dim3(blocks), dim3(256), 0, stream, set, setnew, g.n);  // cinfo_setnew_reset
  hipLaunchKernelGGL(k_amr_add, dim3(blocks), dim3(256), 0, stream, st.f, st.m, ld.f, ld.m, set, setnew, g, qn);
  hipLaunchKernelGGL(k_copy_i32, dim3(blocks), dim3(256), 0, stream, setnew, set, g.n);  // cinfo_set_new
  hipLaunchKernelGGL(k_amr_remove, dim3(1), dim3(1024), 0, stream, st.m, set, g);
  return hipGetLastError();
}

}  // namespace pond
