// step_f64.hip -- plain FP64 PonD step for gfx950, reference operation order.
//
// Replaces PonDstep -> LBMpull<> (update.cu:11-14, steps.cu:356-588) and PonDstepD3Q125
// (steps.cu:151-250).  Every rounding step of the reference is reproduced: loops run in the
// reference order, expression trees are kept, the build disables FMA contraction.  The only
// changes are ones that cannot alter a bit:
//   * pure sub-expressions are hoisted (the node's sqrt(|T0|)*sign(T0), the per-target
//     constants, the shared left prefixes W*gx, (W*gx)*gy of the left-associated products);
//   * a node's transfer to a given target is computed once and reused where the reference
//     recomputes it (the three overlapping WENO sub-stencils; 1D grids, where all y/z stencil
//     nodes are the same cell);
//   * populations of the iterate live in the output buffer instead of a register record, and
//     the moment sums (structs.cuh:323-327) are accumulated while the populations are produced,
//     in the same population order.
//
// D1Q5/D3Q27 mapping: one lane per cell, 64 consecutive cells (x fastest) per wave, so node
// loads of neighbouring lanes are neighbouring addresses of each population plane.
// D3Q125 mapping: one 128-lane block per cell, lane = target population (as the reference's
// 125-thread block), the fixed 3x3x3 neighbourhood staged in LDS.
#include <type_traits>
#include "pond_device.h"

namespace pond {

// ---------------------------------------------------------------------------------------------
// Gauge transfer, D3Q27 (structs.cuh:525-548)

struct Tgt27 {
  double u1x, u1y, u1z, T1;
  double W;           // (wk*wl)*wm
  double Cx, Cy, Cz;  // (e_inew*sqrt(|T1|))*sign(T1)
  bool kx, ky, kz;    // e_inew component != 0
  bool wok;           // |W| < 2^200
};

// magnitude bound of the fast path: five factors below 2^200 multiply to below 2^1000
constexpr double kSafe = 1.6069380442589903e60;  // 2^200

__device__ __forceinline__ Tgt27 make_tgt27(int inew, double u1x, double u1y, double u1z, double T1) {
  Tgt27 t;
  t.u1x = u1x; t.u1y = u1y; t.u1z = u1z; t.T1 = T1;
  const int ex = kE27[inew][0], ey = kE27[inew][1], ez = kE27[inew][2];
  double wk = 1. / (2 * T1), wl = wk, wm = wk;
  if (ex == 0) wk = -1. / T1;
  if (ey == 0) wl = -1. / T1;
  if (ez == 0) wm = -1. / T1;
  t.W = wk * wl * wm;
  const double sq1 = sqrt(fabs(T1));
  const double sg1 = (double)isign_d(T1);
  t.Cx = (double)ex * sq1 * sg1;
  t.Cy = (double)ey * sq1 * sg1;
  t.Cz = (double)ez * sq1 * sg1;
  t.kx = ex != 0; t.ky = ey != 0; t.kz = ez != 0;
  t.wok = fabs(t.W) < kSafe;
  return t;
}

// Node data of the load state: plane bases plus a per-node 32-bit byte offset.
struct NodeSrc {
  const double* __restrict__ f;          // population planes, stride g.plane
  const double* __restrict__ m;          // rho, ux, uy, uz, T planes
  const double* __restrict__ P;          // sqrt(|T|)*sign(T) per cell (params.h:110-112)
  const unsigned char* __restrict__ ok;  // populations finite and below 2^200
  Grid g;
};

__device__ __forceinline__ double ldp(const double* base, int64_t k, int64_t plane, uint32_t boff) {
  return *(const double*)((const char*)(base + k * plane) + boff);
}

__device__ __forceinline__ uint32_t node_off(const Grid& g, int x, int y, int z) {
  return (uint32_t)cell_index(g, x, y, z) * 8u;
}

// The same node data as raw buffer resources: population planes 9j .. 9j+8 in resource j, the
// plane selected by the scalar offset, the per-lane address a single 32-bit offset (a global
// load needs a 64-bit per-lane address: two VGPRs and a 64-bit add per load).  Resource sizes
// are exact, so the hardware range check returns 0 rather than reading past an array.  Word 3
// (0x00020000) is the raw-buffer configuration of the gfx9 family.  Usable while nine FP64
// planes fit in 4 GiB (plane <= 59,652,323 cells).
constexpr int kPlanesPerRsrc64 = 9;
constexpr int kRawBufferWord3 = 0x00020000;

struct NodeBuf {
  // populations: D1Q5 planes 9j .. 9j+8 in f[j]; D3Q27 pair planes 4j .. 4j+3 in f[j]
  __amdgpu_buffer_rsrc_t f[4], m, P, ok;
  uint32_t pbytes;                      // bytes per FP64 plane
  uint32_t qbytes;                      // bytes per D3Q27 pair plane
  Grid g;
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc64(const void* base, uint64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)(uint32_t)bytes, kRawBufferWord3);
}

template <int LAT>
__device__ __forceinline__ NodeBuf make_nodebuf(const NodeSrc& s) {
  NodeBuf b;
  const uint64_t pb = (uint64_t)s.g.plane * 8u;
  b.pbytes = (uint32_t)pb;
  b.qbytes = (uint32_t)(2 * pb);
  if (LAT == kD3Q27) {
    for (int j = 0; j < 4; j++) {
      const int cnt = 14 - 4 * j < 4 ? 14 - 4 * j : 4;
      b.f[j] = rsrc64(s.f + (int64_t)8 * j * s.g.plane, (uint64_t)cnt * 2 * pb);
    }
  } else {
    for (int j = 0; j < 3; j++) b.f[j] = rsrc64(s.f + (int64_t)j * kPlanesPerRsrc64 * s.g.plane, kPlanesPerRsrc64 * pb);
    b.f[3] = b.f[2];
  }
  b.m = rsrc64(s.m, 5 * pb);
  b.P = rsrc64(s.P, LAT == kD3Q27 ? 4 * pb : pb);
  b.ok = rsrc64(s.ok, (uint64_t)s.g.plane);
  b.g = s.g;
  return b;
}

// node accessors: population k, moment plane k, P, finite flag of the node at byte offset boff
__device__ __forceinline__ double node_f(const NodeSrc& s, int k, uint32_t boff) { return ldp(s.f, k, s.g.plane, boff); }
__device__ __forceinline__ double node_m(const NodeSrc& s, int k, uint32_t boff) { return ldp(s.m, k, s.g.plane, boff); }
__device__ __forceinline__ double node_P(const NodeSrc& s, uint32_t boff) { return ldp(s.P, 0, s.g.plane, boff); }
__device__ __forceinline__ bool node_ok(const NodeSrc& s, uint32_t boff) { return s.ok[boff >> 3] != 0; }

__device__ __forceinline__ double ldb64(__amdgpu_buffer_rsrc_t r, uint32_t boff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)boff, (int)soff, 0));
}
__device__ __forceinline__ double node_f(const NodeBuf& s, int k, uint32_t boff) {
  return ldb64(s.f[k / kPlanesPerRsrc64], boff, (uint32_t)(k % kPlanesPerRsrc64) * s.pbytes);
}
__device__ __forceinline__ double node_m(const NodeBuf& s, int k, uint32_t boff) {
  return ldb64(s.m, boff, (uint32_t)k * s.pbytes);
}
__device__ __forceinline__ double node_P(const NodeBuf& s, uint32_t boff) { return ldb64(s.P, boff, 0); }
__device__ __forceinline__ bool node_ok(const NodeBuf& s, uint32_t boff) {
  return __builtin_amdgcn_raw_buffer_load_b8(s.ok, (int)(boff >> 3), 0, 0) != 0;
}

// D3Q27: the node's gauge record (ux, uy, uz, sqrt(|T|)*sign(T)), 32 bytes per cell, written by
// k_node_prep into the P buffer, at byte offset boff of an FP64 plane (boff = 8 * cell)
// (64-bit byte offsets: 4 * boff leaves 32 bits from 2^27 cells on, the grids this path serves)
__device__ __forceinline__ void node_g4(const NodeSrc& s, uint32_t boff, double& ux, double& uy, double& uz,
                                       double& P) {
  const double2 a = *(const double2*)((const char*)s.P + 4ull * boff);
  const double2 b = *(const double2*)((const char*)s.P + 4ull * boff + 16ull);
  ux = a.x; uy = a.y; uz = b.x; P = b.y;
}
__device__ __forceinline__ double dbl_of(uint32_t lo, uint32_t hi) {
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ void node_g4(const NodeBuf& s, uint32_t boff, double& ux, double& uy, double& uz,
                                       double& P) {
  const auto a = __builtin_amdgcn_raw_buffer_load_b128(s.P, (int)(4u * boff), 0, 0);
  const auto b = __builtin_amdgcn_raw_buffer_load_b128(s.P, (int)(4u * boff), 16, 0);
  ux = dbl_of((uint32_t)a[0], (uint32_t)a[1]);
  uy = dbl_of((uint32_t)a[2], (uint32_t)a[3]);
  uz = dbl_of((uint32_t)b[0], (uint32_t)b[1]);
  P = dbl_of((uint32_t)b[2], (uint32_t)b[3]);
}

// the 27 populations of a D3Q27 node (pair-plane layout) at byte offset boff of an FP64 plane
__device__ __forceinline__ void node_f27(const NodeSrc& s, uint32_t boff, double f[27]) {
#pragma unroll
  for (int j = 0; j < 14; j++) {
    const double2 v = *(const double2*)((const char*)(s.f + (int64_t)2 * j * s.g.plane) + 2ull * boff);
    f[2 * j] = v.x;
    if (2 * j + 1 < 27) f[2 * j + 1] = v.y;
  }
}
__device__ __forceinline__ void node_f27(const NodeBuf& s, uint32_t boff, double f[27]) {
#pragma unroll
  for (int j = 0; j < 14; j++) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(s.f[j >> 2], (int)(2u * boff), (int)((uint32_t)(j & 3) * s.qbytes), 0);
    f[2 * j] = __builtin_bit_cast(double, ((uint64_t)(uint32_t)v[1] << 32) | (uint32_t)v[0]);
    if (2 * j + 1 < 27) f[2 * j + 1] = __builtin_bit_cast(double, ((uint64_t)(uint32_t)v[3] << 32) | (uint32_t)v[2]);
  }
}

// basis values of the node at byte offset boff for target t (structs.cuh:530-538).
// e_iold*sqrt(|T0|)*sign(T0) for e_iold = -1, 0, +1 is exactly {-P, 0*P, P}.
__device__ __forceinline__ bool basis27u(double u0x, double u0y, double u0z, double Pn, bool nok, const Tgt27& t,
                                         double sqTb, double gx[3], double gy[3], double gz[3]) {
  const double Dx = (t.u1x - u0x) * sqTb;
  const double Dy = (t.u1y - u0y) * sqTb;
  const double Dz = (t.u1z - u0z) * sqTb;
  const double E[3] = {-Pn, 0.0 * Pn, Pn};
  bool safe = nok && t.wok;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const double ax = Dx - E[j], ay = Dy - E[j], az = Dz - E[j];
    gx[j] = ax * ax - (t.kx ? t.Cx * ax : t.T1);
    gy[j] = ay * ay - (t.ky ? t.Cy * ay : t.T1);
    gz[j] = az * az - (t.kz ? t.Cz * az : t.T1);
    safe = safe && fabs(gx[j]) < kSafe && fabs(gy[j]) < kSafe && fabs(gz[j]) < kSafe;
  }
  return safe;
}

template <class SRC>
__device__ __forceinline__ bool basis27(const SRC& s, uint32_t boff, const Tgt27& t, double sqTb, double gx[3],
                                        double gy[3], double gz[3]) {
  double u0x, u0y, u0z, Pn;
  node_g4(s, boff, u0x, u0y, u0z, Pn);
  return basis27u(u0x, u0y, u0z, Pn, node_ok(s, boff), t, sqTb, gx, gy, gz);
}

// gauge transfer of the node at byte offset boff to target t (structs.cuh:525-548)
template <class SRC>
__device__ __forceinline__ double transfer27(const SRC& s, uint32_t boff, const Tgt27& t, double sqTb) {
  double gx[3], gy[3], gz[3];
  const bool safe = basis27(s, boff, t, sqTb, gx, gy, gz);
  double p2[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) p2[a][b] = t.W * gx[a] * gy[b];
  double f[27];
  node_f27(s, boff, f);
  double fnew = 0;
#pragma unroll
  for (int iold = 0; iold < 27; iold++) {
    const int jx = kE27[iold][0] + 1, jy = kE27[iold][1] + 1, jz = kE27[iold][2] + 1;
    fnew += p2[jx][jy] * gz[jz] * f[iold];
  }
  if (!safe) {
    // the reference's term loop with NaN/Inf terms dropped (structs.cuh:539-542), for the
    // nodes that fail the magnitude test; in line, so the basis values stay in registers
    fnew = 0;
#pragma unroll
    for (int iold = 0; iold < 27; iold++) {
      const int jx = kE27[iold][0] + 1, jy = kE27[iold][1] + 1, jz = kE27[iold][2] + 1;
      const double fadd = p2[jx][jy] * gz[jz] * f[iold];
      if (finite_d(fadd)) fnew += fadd;
    }
  }
  return fnew;
}

// transfer27 of the node's mirror image on the axes set in mx, my, mz (Cell::mirrX/Y/Z,
// structs.cuh:552-578): the axis velocity negated and every population exchanged with its mirror
// image (kIdx27 with that component negated), then the same term loop
template <class SRC>
__device__ double transfer27_mirrored(const SRC& s, uint32_t boff, const Tgt27& t, double sqTb, bool mx, bool my,
                                      bool mz) {
  double u0x, u0y, u0z, Pn;
  node_g4(s, boff, u0x, u0y, u0z, Pn);
  if (mx) u0x = -u0x;
  if (my) u0y = -u0y;
  if (mz) u0z = -u0z;
  double gx[3], gy[3], gz[3];
  const bool safe = basis27u(u0x, u0y, u0z, Pn, node_ok(s, boff), t, sqTb, gx, gy, gz);
  double p2[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) p2[a][b] = t.W * gx[a] * gy[b];
  double f0[27], f[27];
  node_f27(s, boff, f0);
#pragma unroll
  for (int i = 0; i < 27; i++) {
    const int jx = kE27[i][0] + 1, jy = kE27[i][1] + 1, jz = kE27[i][2] + 1;
    // f[i] = f0 of the mirrored velocity, as selects over compile-time indices
    const double a0 = mx ? f0[kIdx27.v[2 - jx][jy][jz]] : f0[i];
    const double a1 = mx ? f0[kIdx27.v[2 - jx][2 - jy][jz]] : f0[kIdx27.v[jx][2 - jy][jz]];
    const double b0 = my ? a1 : a0;
    const double a2 = mx ? f0[kIdx27.v[2 - jx][jy][2 - jz]] : f0[kIdx27.v[jx][jy][2 - jz]];
    const double a3 = mx ? f0[kIdx27.v[2 - jx][2 - jy][2 - jz]] : f0[kIdx27.v[jx][2 - jy][2 - jz]];
    const double b1 = my ? a3 : a2;
    f[i] = mz ? b1 : b0;
  }
  double fnew = 0;
#pragma unroll
  for (int iold = 0; iold < 27; iold++) {
    const int jx = kE27[iold][0] + 1, jy = kE27[iold][1] + 1, jz = kE27[iold][2] + 1;
    const double fadd = p2[jx][jy] * gz[jz] * f[iold];
    if (safe || finite_d(fadd)) fnew += fadd;
  }
  return fnew;
}

// ---------------------------------------------------------------------------------------------
// Gauge transfer, D1Q5 (structs.cuh:352-395)

struct Tgt5 {
  double u1x, et, wk, fnew0;
  double e1sq, e2sq, E, etet;
  bool k0;  // inew == 0
};

__device__ __forceinline__ Tgt5 make_tgt5(int inew, double u1x, double T1) {
  Tgt5 t;
  t.u1x = u1x;
  const double s1 = sqrts_d(T1 / 1.0);
  const double e1_to = kA5[1] * s1, e2_to = kA5[3] * s1;
  t.et = kA5[inew] * s1;
  const double ef = kA5[inew];
  double wk;
  if (ef == 0) wk = 1. / (e1_to * e1_to * e2_to * e2_to);
  else if (fabs(ef) == kEc1) wk = 1. / (2 * e1_to * e1_to * (e1_to * e1_to - e2_to * e2_to));
  else wk = 1. / (2 * e2_to * e2_to * (e2_to * e2_to - e1_to * e1_to));
  t.wk = wk;
  t.fnew0 = wk * 0.0;
  t.e1sq = e1_to * e1_to;
  t.e2sq = e2_to * e2_to;
  t.E = t.e1sq + t.e2sq;
  t.etet = t.et * t.et;
  t.k0 = inew == 0;
  return t;
}

// MIRX: the node's mirror image in x (Cell::mirrX, structs.cuh:552-559): u0x negated and the
// populations +e/-e exchanged
template <bool MIRX = false, class SRC>
__device__ __forceinline__ double transfer5(const SRC& s, uint32_t boff, const Tgt5& t, bool mx = false) {
  double u0x = node_m(s, 1, boff);
  double fm[5];
  if (MIRX) {
    if (mx) u0x = -u0x;
    double f0[5];
#pragma unroll
    for (int i = 0; i < 5; i++) f0[i] = node_f(s, i, boff);
#pragma unroll
    for (int i = 0; i < 5; i++) fm[i] = mx ? f0[kMir5[i]] : f0[i];
  }
  const double s0 = node_P(s, boff);  // sqrts(T0/1.0) = sqrt(|T0|)*sign(T0)
  const double et = t.et;
  const double difu = t.u1x - u0x;
  double fnew = t.fnew0;
#pragma unroll
  for (int iold = 0; iold < 5; iold++) {
    const double e = kA5[iold] * s0;
    double K0, K1, K2, K3;
    const double K4 = 1;
    K3 = -4 * e - et;
    if (t.k0) {
      K2 = 6 * e * e - t.E;
      K1 = 2 * e * (t.E - 2 * e * e);
      K0 = (t.e1sq - e * e) * (t.e2sq - e * e);
    } else {
      K2 = 6 * e * e + 3 * e * et - (t.E - t.etet);
      K1 = (2 * e + et) * (t.E - t.etet) - e * e * (4 * e + 3 * et);
      K0 = e * (e * e - t.E + t.etet) * (e + et);
    }
    double dm4 = u0x + e;
    dm4 = dm4 * dm4 * dm4 * dm4;
    dm4 *= 0;
    const double gx = K4 * difu * difu * difu * difu + K3 * difu * difu * difu + K2 * difu * difu + K1 * difu + K0 - dm4;
    const double fadd = t.wk * gx * (MIRX ? fm[iold] : node_f(s, iold, boff));
    if (finite_d(fadd)) fnew += fadd;
  }
  return fnew;
}

// ---------------------------------------------------------------------------------------------
// Lattice traits for the one-lane-per-cell kernel

template <int LAT>
struct Lat;

template <>
struct Lat<kD3Q27> {
  static constexpr int Q = 27;
  using Tgt = Tgt27;
  __device__ static double efc(int i, int d) { return (double)kE27[i][d]; }
  __device__ static Tgt tgt(int i, const double* g) { return make_tgt27(i, g[0], g[1], g[2], g[3]); }
  template <class SRC>
  __device__ static double xfer(const SRC& s, uint32_t boff, const Tgt& t, double sqTb) {
    return transfer27(s, boff, t, sqTb);
  }
  template <class SRC>
  __device__ static double xferm(const SRC& s, uint32_t boff, const Tgt& t, double sqTb, bool mx, bool my, bool mz) {
    return transfer27_mirrored(s, boff, t, sqTb, mx, my, mz);
  }
};

template <>
struct Lat<kD1Q5> {
  static constexpr int Q = 5;
  using Tgt = Tgt5;
  __device__ static double efc(int i, int d) { return d == 0 ? kA5[i] : 0.0; }
  __device__ static Tgt tgt(int i, const double* g) { return make_tgt5(i, g[0], g[3]); }
  template <class SRC>
  __device__ static double xfer(const SRC& s, uint32_t boff, const Tgt& t, double) {
    return transfer5(s, boff, t);
  }
  // D1Q5 velocities have no y or z component: a mirror in y or z only negates the unused uy, uz
  template <class SRC>
  __device__ static double xferm(const SRC& s, uint32_t boff, const Tgt& t, double, bool mx, bool, bool) {
    return transfer5<true>(s, boff, t, mx);
  }
};

// ---------------------------------------------------------------------------------------------
// Interpolators

// the two factors (s-p)/(i-p), p != i, p ascending, of the 3-node Lagrange basis (steps.cu:597)
__device__ __forceinline__ void lag3_factors(double s, double F[3][2]) {
  F[0][0] = (s - 1) / (0 - 1); F[0][1] = (s - 2) / (0 - 2);
  F[1][0] = (s - 0) / (1 - 0); F[1][1] = (s - 2) / (1 - 2);
  F[2][0] = (s - 0) / (2 - 0); F[2][1] = (s - 1) / (2 - 1);
}

__device__ __forceinline__ void lag4_factors(double s, double F[4][3]) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int c = 0;
#pragma unroll
    for (int p = 0; p < 4; p++)
      if (p != i) F[i][c++] = (s - p) / (double)(i - p);
  }
}

// the WENO5 combination of the three sub-stencil sums (S0: left window, S2: right) with the
// smoothness samples ub at the five window positions, in the operation order of form VAR
template <int VAR>
__device__ __forceinline__ double weno5_combine_f64(double cdist, const double (&ub)[5], double S0, double S1,
                                                    double S2) {
  const double u1 = ub[0], u2 = ub[1], u3 = ub[2], u4 = ub[3], u5 = ub[4];
  if (VAR == 2) {
    // S0, S1, S2 of steps.cu:617-619 are this function's right, middle and left sums
    const double d0 = (cdist * cdist + 3 * cdist + 2) / 12.;
    const double d2 = (cdist * cdist - 3 * cdist + 2) / 12.;
    const double d1 = 1. - d0 - d2;
    const double beta0 = (10 * u3 * u3 - 31 * u3 * u4 + 25 * u4 * u4 + 11 * u3 * u5 - 19 * u4 * u5 + 4 * u5 * u5) / 3.0;
    const double beta1 = (4 * u2 * u2 - 13 * u2 * u3 + 13 * u3 * u3 + 5 * u2 * u4 - 13 * u3 * u4 + 4 * u4 * u4) / 3.0;
    const double beta2 = (4 * u1 * u1 - 19 * u1 * u2 + 25 * u2 * u2 + 11 * u1 * u3 - 31 * u2 * u3 + 10 * u3 * u3) / 3.0;
    double w0 = d0 / ((1e-6 + beta0) * (1e-6 + beta0));
    double w1 = d1 / ((1e-6 + beta1) * (1e-6 + beta1));
    double w2 = d2 / ((1e-6 + beta2) * (1e-6 + beta2));
    const double wsum = w0 + w1 + w2;
    w0 /= wsum;
    w1 /= wsum;
    w2 /= wsum;
    return w0 * S2 + w1 * S1 + w2 * S0;
  }
  const double d0 = (cdist * cdist - 3 * cdist + 2) / 12.;
  const double d2 = (cdist * cdist + 3 * cdist + 2) / 12.;
  const double d1 = 1. - d0 - d2;
  const double beta0 = (10 * u3 * u3 + (11 * u1 - 31 * u2) * u3 + 25 * u2 * u2 - 19 * u1 * u2 + 4 * u1 * u1) / 3.0;
  const double beta1 = (4 * u4 * u4 + (5 * u2 - 13 * u3) * u4 + 13 * u3 * u3 - 13 * u2 * u3 + 4 * u2 * u2) / 3.0;
  const double beta2 = (4 * u5 * u5 + (11 * u3 - 19 * u4) * u5 + 25 * u4 * u4 - 31 * u3 * u4 + 10 * u3 * u3) / 3.0;
  double w0 = d0 * 1, w1 = d1 * 1, w2 = d2 * 1;
  const double sc = 1e-6;
  w0 /= (sc + beta0) * (sc + beta0);
  w1 /= (sc + beta1) * (sc + beta1);
  w2 /= (sc + beta2) * (sc + beta2);
  const double wsum = w0 + w1 + w2;
  w0 /= wsum;
  w1 /= wsum;
  w2 /= wsum;
  return w0 * S0 + w1 * S1 + w2 * S2;
}

// WENO5 interpolation, three forms (VAR):
// 0: propagateWENO<3,5> (steps.cu:774-825) with LagrEval<3> (steps.cu:924-937), as written:
//    y/z sub-stencil nodes floor(y+0.5)+0..2 and smoothness samples on the row
//    (r mod Ny, r mod Nz), r = floor(x+0.5) (steps.cu:776-783).  This extrapolates downwind in
//    y and z and is unstable in 3D (grid-scale modes grow every step).
// 1: the corrected variant (not in the reference): y/z nodes floor(y+0.5)-1..+1 and smoothness
//    samples on the departure point's own row, which is the middle node of every slice, so
//    they reuse that slice's transfers.
// 2: the generic propagateWENO<5,M> (steps.cu:610-654) with propagate3 (steps.cu:909-923): the
//    nodes, samples and interpolant of form 0, with the sub-stencils numbered from the right
//    (lshift = 1), the indicators written expanded and each weight formed in one expression;
//    only the rounding differs from form 0.
template <int LAT, int VAR, class SRC>
__device__ double interp_weno5(const SRC& src, double cx, double cy, double cz, const typename Lat<LAT>::Tgt& t,
                               double sqTb) {
  constexpr bool CENTRED = VAR == 1;
  const Grid& g = src.g;
  const int r = floor_int(cx + 0.5);
  const int ry = floor_int(cy + 0.5) - (CENTRED ? 1 : 0), rz = floor_int(cz + 0.5) - (CENTRED ? 1 : 0);
  const double sy = cy - ry, sz = cz - rz;
  double Fy[3][2], Fz[3][2];
  lag3_factors(sy, Fy);
  lag3_factors(sz, Fz);
  // x parts of LagrPol<3> for sub-stencils k = 0,1,2 with left node r-2+k
  double px[3][3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double Fx[3][2];
    lag3_factors(cx - (r - 2 + k), Fx);
#pragma unroll
    for (int ix = 0; ix < 3; ix++) px[k][ix] = 1 * Fx[ix][0] * Fx[ix][1];
  }
  int yy[3], zz[3];
#pragma unroll
  for (int a = 0; a < 3; a++) { yy[a] = wrapi(ry + a, g.ny); zz[a] = wrapi(rz + a, g.nz); }
  const bool line = (g.ny == 1) && (g.nz == 1);
  const int by = CENTRED ? yy[1] : wrapi(r, g.ny), bz = CENTRED ? zz[1] : wrapi(r, g.nz);
  double S0 = 0, S1 = 0, S2 = 0;
  double ub[5];
  for (int xs = 0; xs < 5; xs++) {
    const int x = wrapi(r - 2 + xs, g.nx);
    double T[9];
    if (line) {
      const double v = Lat<LAT>::xfer(src, (uint32_t)x * 8u, t, sqTb);
#pragma unroll
      for (int q = 0; q < 9; q++) T[q] = v;
      ub[xs] = v;  // smoothness row (r mod 1, r mod 1) = (0,0): same node
    } else {
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
          T[3 * a + b] = Lat<LAT>::xfer(src, node_off(g, x, yy[a], zz[b]), t, sqTb);
        }
      if (CENTRED) {
        ub[xs] = T[4];
      } else {
        ub[xs] = Lat<LAT>::xfer(src, node_off(g, x, by, bz), t, sqTb);
      }
    }
    // accumulate into every sub-stencil that contains this x slice, in (iy, iz) order
    if (xs <= 2) {
      const double p = xs == 0 ? px[0][0] : xs == 1 ? px[0][1] : px[0][2];
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) S0 += p * Fy[a][0] * Fy[a][1] * Fz[b][0] * Fz[b][1] * T[3 * a + b];
    }
    if (xs >= 1 && xs <= 3) {
      const double p = xs == 1 ? px[1][0] : xs == 2 ? px[1][1] : px[1][2];
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) S1 += p * Fy[a][0] * Fy[a][1] * Fz[b][0] * Fz[b][1] * T[3 * a + b];
    }
    if (xs >= 2) {
      const double p = xs == 2 ? px[2][0] : xs == 3 ? px[2][1] : px[2][2];
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) S2 += p * Fy[a][0] * Fy[a][1] * Fz[b][0] * Fz[b][1] * T[3 * a + b];
    }
  }
  return weno5_combine_f64<VAR>(cx - r, ub, S0, S1, S2);
}

// ---------------------------------------------------------------------------------------------
// D3Q27 targets in (ey, ez) groups: the three targets with ex = -1, 0, +1 share their departure
// point's y and z (v_y and v_z depend on ey, ez only), so their stencils share the y/z rows and
// differ in x only.  A node of the union of their windows is loaded once and transferred to each
// target whose window holds it, with exactly transfer27's operations for that target; the y and
// z basis values, which the three share, are evaluated once.  Each target's interpolation sums
// receive their terms in the order of the one-target code (slices in increasing x, rows (a, b)
// in order inside a slice), so every value is bitwise what interp_weno5 / interp_lagr4 give.

// The three targets of a group: the fields of Tgt27 they share (u1, T1 and the y, z target
// components) once, the x ones per target.
struct Tgt27g3 {
  double u1x, u1y, u1z, T1, Cy, Cz;
  bool ky, kz;
  double W[3], Cx[3];
  bool kx[3], wok[3];
};

// transfer27 of one node to the targets k with use[k] (T[k]; the others untouched)
template <class SRC>
__device__ __forceinline__ void transfer27_g3(const SRC& s, uint32_t boff, const Tgt27g3& t, const bool (&use)[3],
                                              double sqTb, double (&T)[3]) {
  double u0x, u0y, u0z, Pn;
  node_g4(s, boff, u0x, u0y, u0z, Pn);
  const bool nok = node_ok(s, boff);
  // basis27u, split into the x part (per target) and the shared y and z parts
  const double Dx = (t.u1x - u0x) * sqTb;
  const double Dy = (t.u1y - u0y) * sqTb;
  const double Dz = (t.u1z - u0z) * sqTb;
  const double E[3] = {-Pn, 0.0 * Pn, Pn};
  double ax[3], gy[3], gz[3];
  bool syz = true;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    ax[j] = Dx - E[j];
    const double ay = Dy - E[j], az = Dz - E[j];
    gy[j] = ay * ay - (t.ky ? t.Cy * ay : t.T1);
    gz[j] = az * az - (t.kz ? t.Cz * az : t.T1);
    syz = syz && fabs(gy[j]) < kSafe && fabs(gz[j]) < kSafe;
  }
  double f[27];
  node_f27(s, boff, f);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (!use[k]) continue;
    double gx[3];
    bool safe = nok && t.wok[k] && syz;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      gx[j] = ax[j] * ax[j] - (t.kx[k] ? t.Cx[k] * ax[j] : t.T1);
      safe = safe && fabs(gx[j]) < kSafe;
    }
    double p2[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) p2[a][b] = t.W[k] * gx[a] * gy[b];
    double fnew = 0;
#pragma unroll
    for (int iold = 0; iold < 27; iold++) {
      const int jx = kE27[iold][0] + 1, jy = kE27[iold][1] + 1, jz = kE27[iold][2] + 1;
      fnew += p2[jx][jy] * gz[jz] * f[iold];
    }
    if (!safe) {
      fnew = 0;
#pragma unroll
      for (int iold = 0; iold < 27; iold++) {
        const int jx = kE27[iold][0] + 1, jy = kE27[iold][1] + 1, jz = kE27[iold][2] + 1;
        const double fadd = p2[jx][jy] * gz[jz] * f[iold];
        if (finite_d(fadd)) fnew += fadd;
      }
    }
    T[k] = fnew;
  }
}

// x part of LagrPol<3> of window node ix of sub-stencil q for a departure point at cx whose
// window starts at r - 2 (the value interp_weno5 keeps in px[q][ix])
__device__ __forceinline__ double lag3x(double cx, int r, int q, int ix) {
  double Fx[3][2];
  lag3_factors(cx - (r - 2 + q), Fx);
  return ix == 0 ? 1 * Fx[0][0] * Fx[0][1] : ix == 1 ? 1 * Fx[1][0] * Fx[1][1] : 1 * Fx[2][0] * Fx[2][1];
}

__device__ __forceinline__ Tgt27 tgt27_of(const Tgt27g3& t, int k) {
  Tgt27 tk;
  tk.u1x = t.u1x; tk.u1y = t.u1y; tk.u1z = t.u1z; tk.T1 = t.T1;
  tk.W = t.W[k]; tk.Cx = t.Cx[k]; tk.Cy = t.Cy; tk.Cz = t.Cz;
  tk.kx = t.kx[k]; tk.ky = t.ky; tk.kz = t.kz; tk.wok = t.wok[k];
  return tk;
}

// interp_weno5<kD3Q27, VAR> for the three targets of a group (cx[k] of target ex = k - 1; cy, cz
// shared), on a grid that is not a line, without mirrors.  Window positions are taken relative
// to the cell's own x, so lanes of a wave whose flow is alike visit the same positions.  Taken
// only for s = sqrt(T1/TLat) <= 1.5 (f64_once): the windows of ex = -1 and +1 are 2s apart, so
// the union spans at most 9 slices.
template <int VAR, class SRC>
__device__ void interp_weno5_g3(const SRC& src, int x, const double (&cx)[3], double cy, double cz,
                                const Tgt27g3& t, double sqTb, double (&fn)[3]) {
  constexpr bool CENTRED = VAR == 1;
  const Grid& g = src.g;
  const int ry = floor_int(cy + 0.5) - (CENTRED ? 1 : 0), rz = floor_int(cz + 0.5) - (CENTRED ? 1 : 0);
  const double sy = cy - ry, sz = cz - rz;
  double Fy[3][2], Fz[3][2];
  lag3_factors(sy, Fy);
  lag3_factors(sz, Fz);
  int r[3], w0[3];  // window start relative to x
#pragma unroll
  for (int k = 0; k < 3; k++) {
    r[k] = floor_int(cx[k] + 0.5);
    w0[k] = (r[k] - 2) - x;
  }
  const int dlo = min(w0[0], min(w0[1], w0[2]));
  // (at most dlo + 8 for s <= 1.5; the bound keeps the loop finite whatever the caller passes)
  const int dhi = min(max(w0[0], max(w0[1], w0[2])) + 4, dlo + 8);
  // the smoothness samples of the three targets wait in LDS (one wave per block), which keeps
  // the sub-stencil sums and a node's populations in registers
  __shared__ double s_ub[3][5][64];
  const int lane = threadIdx.x & 63;
  double S[3][3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
#pragma unroll
    for (int q = 0; q < 3; q++) S[k][q] = 0;
#pragma unroll
    for (int j = 0; j < 5; j++) s_ub[k][j][lane] = 0;
  }
  for (int d = dlo; d <= dhi; d++) {
    const int xw = wrapi(x + d, g.nx);
    int xs[3];
    bool use[3];
    double pw[3][3];  // per target, the x weight of this slice in sub-stencils 0, 1, 2
#pragma unroll
    for (int k = 0; k < 3; k++) {
      xs[k] = d - w0[k];
      use[k] = xs[k] >= 0 && xs[k] <= 4;
      const int j = xs[k];
#pragma unroll
      for (int q = 0; q < 3; q++) pw[k][q] = (j - q >= 0 && j - q <= 2) ? lag3x(cx[k], r[k], q, j - q) : 0.0;
    }
#pragma unroll 1
    for (int a = 0; a < 3; a++) {
      const int ya = wrapi(ry + a, g.ny);
#pragma unroll 1
      for (int b = 0; b < 3; b++) {
        double T[3] = {0, 0, 0};
        transfer27_g3(src, node_off(g, xw, ya, wrapi(rz + b, g.nz)), t, use, sqTb, T);
        const double fy0 = a == 0 ? Fy[0][0] : a == 1 ? Fy[1][0] : Fy[2][0];
        const double fy1 = a == 0 ? Fy[0][1] : a == 1 ? Fy[1][1] : Fy[2][1];
        const double fz0 = b == 0 ? Fz[0][0] : b == 1 ? Fz[1][0] : Fz[2][0];
        const double fz1 = b == 0 ? Fz[0][1] : b == 1 ? Fz[1][1] : Fz[2][1];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          if (!use[k]) continue;
          const int j = xs[k];
#pragma unroll
          for (int q = 0; q < 3; q++)
            if (j - q >= 0 && j - q <= 2) S[k][q] += pw[k][q] * fy0 * fy1 * fz0 * fz1 * T[k];
          if (CENTRED && a == 1 && b == 1) s_ub[k][j][lane] = T[k];
        }
      }
    }
    if (!CENTRED) {
      // smoothness samples on the row (r mod Ny, r mod Nz) of each target
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (!use[k]) continue;
        s_ub[k][xs[k]][lane] =
            transfer27(src, node_off(g, xw, wrapi(r[k], g.ny), wrapi(r[k], g.nz)), tgt27_of(t, k), sqTb);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double ub[5];
#pragma unroll
    for (int j = 0; j < 5; j++) ub[j] = s_ub[k][j][lane];
    fn[k] = weno5_combine_f64<VAR>(cx[k] - r[k], ub, S[k][0], S[k][1], S[k][2]);
  }
}

// x part of LagrPol<4> of node ix for a departure point at cx whose window starts at lx (the
// value interp_lagr4 forms as p)
__device__ __forceinline__ double lag4x(double cx, int lx, int ix) {
  double Fx[4][3];
  lag4_factors(cx - lx, Fx);
  return ix == 0   ? 1 * Fx[0][0] * Fx[0][1] * Fx[0][2]
         : ix == 1 ? 1 * Fx[1][0] * Fx[1][1] * Fx[1][2]
         : ix == 2 ? 1 * Fx[2][0] * Fx[2][1] * Fx[2][2]
                   : 1 * Fx[3][0] * Fx[3][1] * Fx[3][2];
}

// interp_lagr4<kD3Q27, false> for the three targets of a group, as interp_weno5_g3 does for
// WENO5: each node of the union of their windows (at most 8 slices for s <= 1.5) loaded once,
// each target's sum in its own node order.
template <class SRC>
__device__ void interp_lagr4_g3(const SRC& src, int x, const double (&cx)[3], double cy, double cz,
                                const Tgt27g3& t, double sqTb, double (&fn)[3]) {
  const Grid& g = src.g;
  const int ly = floor_int(cy) - 4 / 2 + 1;
  const int lz = floor_int(cz) - 4 / 2 + 1;
  double Fy[4][3], Fz[4][3];
  lag4_factors(cy - ly, Fy);
  lag4_factors(cz - lz, Fz);
  int lx[3], w0[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lx[k] = floor_int(cx[k]) - 4 / 2 + 1;
    w0[k] = lx[k] - x;
    fn[k] = 0;
  }
  const int dlo = min(w0[0], min(w0[1], w0[2]));
  // (at most dlo + 7 for s <= 1.5; the bound keeps the loop finite whatever the caller passes)
  const int dhi = min(max(w0[0], max(w0[1], w0[2])) + 3, dlo + 7);
  for (int d = dlo; d <= dhi; d++) {
    const int xw = wrapi(x + d, g.nx);
    bool use[3];
    double p[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int ix = d - w0[k];
      use[k] = ix >= 0 && ix <= 3;
      p[k] = use[k] ? lag4x(cx[k], lx[k], ix) : 0.0;
    }
#pragma unroll 1
    for (int a = 0; a < 4; a++) {
      const int ya = wrapi(ly + a, g.ny);
      const double fy0 = a == 0 ? Fy[0][0] : a == 1 ? Fy[1][0] : a == 2 ? Fy[2][0] : Fy[3][0];
      const double fy1 = a == 0 ? Fy[0][1] : a == 1 ? Fy[1][1] : a == 2 ? Fy[2][1] : Fy[3][1];
      const double fy2 = a == 0 ? Fy[0][2] : a == 1 ? Fy[1][2] : a == 2 ? Fy[2][2] : Fy[3][2];
#pragma unroll 1
      for (int b = 0; b < 4; b++) {
        double T[3] = {0, 0, 0};
        transfer27_g3(src, node_off(g, xw, ya, wrapi(lz + b, g.nz)), t, use, sqTb, T);
        const double fz0 = b == 0 ? Fz[0][0] : b == 1 ? Fz[1][0] : b == 2 ? Fz[2][0] : Fz[3][0];
        const double fz1 = b == 0 ? Fz[0][1] : b == 1 ? Fz[1][1] : b == 2 ? Fz[2][1] : Fz[3][1];
        const double fz2 = b == 0 ? Fz[0][2] : b == 1 ? Fz[1][2] : b == 2 ? Fz[2][2] : Fz[3][2];
#pragma unroll
        for (int k = 0; k < 3; k++)
          if (use[k]) fn[k] += p[k] * fy0 * fy1 * fy2 * fz0 * fz1 * fz2 * T[k];
      }
    }
  }
}

// propagate<4> (steps.cu:986-1036): tricubic Lagrange on floor(x)-1 .. floor(x)+2.
// MIR: mirror boundaries on the axes of g.mirror (steps.cu:1004-1022): nodes beyond a mirrored
// face are the mirror images of the cells inside.
template <int LAT, bool MIR, class SRC>
__device__ double interp_lagr4(const SRC& src, double cx, double cy, double cz, const typename Lat<LAT>::Tgt& t,
                               double sqTb) {
  const Grid& g = src.g;
  const int lx = floor_int(cx) - 4 / 2 + 1;
  const int ly = floor_int(cy) - 4 / 2 + 1;
  const int lz = floor_int(cz) - 4 / 2 + 1;
  double Fx[4][3], Fy[4][3], Fz[4][3];
  lag4_factors(cx - lx, Fx);
  lag4_factors(cy - ly, Fy);
  lag4_factors(cz - lz, Fz);
  int yy[4], zz[4];
  bool ry[4] = {false, false, false, false}, rz[4] = {false, false, false, false};
#pragma unroll
  for (int a = 0; a < 4; a++) {
    if (MIR) {
      yy[a] = mirror_wrap(ly + a, g.ny, (g.mirror & 2) != 0, ry[a]);
      zz[a] = mirror_wrap(lz + a, g.nz, (g.mirror & 4) != 0, rz[a]);
    } else {
      yy[a] = wrapi(ly + a, g.ny);
      zz[a] = wrapi(lz + a, g.nz);
    }
  }
  const bool line = (g.ny == 1) && (g.nz == 1);
  double fpr = 0;
  for (int ix = 0; ix < 4; ix++) {
    bool rx = false;
    const int x = MIR ? mirror_wrap(lx + ix, g.nx, (g.mirror & 1) != 0, rx) : wrapi(lx + ix, g.nx);
    const double p = 1 * Fx[ix][0] * Fx[ix][1] * Fx[ix][2];
    double v1 = 0;
    if (line && !MIR) v1 = Lat<LAT>::xfer(src, (uint32_t)x * 8u, t, sqTb);
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        double T;
        if (MIR) {
          const uint32_t off = node_off(g, x, yy[a], zz[b]);
          T = (rx || ry[a] || rz[b]) ? Lat<LAT>::xferm(src, off, t, sqTb, rx, ry[a], rz[b])
                                     : Lat<LAT>::xfer(src, off, t, sqTb);
        } else if (line) {
          T = v1;
        } else {
          T = Lat<LAT>::xfer(src, node_off(g, x, yy[a], zz[b]), t, sqTb);
        }
        fpr += p * Fy[a][0] * Fy[a][1] * Fy[a][2] * Fz[b][0] * Fz[b][1] * Fz[b][2] * T;
      }
  }
  return fpr;
}

// ---------------------------------------------------------------------------------------------
// One lane per cell: LBMpull (steps.cu:356-365, 421-512, 557-564, 582)

// Iteration state of one cell: gauge c0, density rho0, count niter.  The iterate's
// populations live in fout.
struct F64Iter {
  double c0[4];
  double rho0;
  int niter;
};

// f64_once for D3Q27 WENO5 (INTERP 0, 2, 4) and Lagrange-4 (1) off a line, without mirrors: the
// targets in (ey, ez) groups of three (interp_weno5_g3, interp_lagr4_g3), each value stored as it
// is found; then the moments summed over the stored values in population order, as calcMoments
// does.
template <int INTERP, class SRC>
__device__ __forceinline__ bool f64_once_g3(const SRC& src, const NodeSrc& gsrc, double* __restrict__ fout,
                                            int64_t cell, int x, int y, int z, bool zero, F64Iter& st,
                                            const StepArgs& a) {
  constexpr int VAR = INTERP == 2 ? 1 : INTERP == 4 ? 2 : 0;
  const int64_t P = a.g.plane;
  double* c0 = st.c0;
  st.niter++;
  const int niter = st.niter;
  const double uT[4] = {c0[0], c0[1], c0[2], c0[3]};
  const double s = sqrt(uT[3] / a.tlat);
  bool conv = true;
#pragma unroll 1
  for (int gi = 0; gi < 9; gi++) {
    const int jy = gi / 3, jz = gi % 3;
    int idx[3];
    Tgt27g3 t;
    double cx[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      idx[k] = kIdx27.v[k][jy][jz];
      const double vx = s * Lat<kD3Q27>::efc(idx[k], 0) + uT[0];
      const Tgt27 tk = Lat<kD3Q27>::tgt(idx[k], uT);
      t.W[k] = tk.W; t.Cx[k] = tk.Cx; t.kx[k] = tk.kx; t.wok[k] = tk.wok;
      if (k == 0) {
        t.u1x = tk.u1x; t.u1y = tk.u1y; t.u1z = tk.u1z; t.T1 = tk.T1;
        t.Cy = tk.Cy; t.Cz = tk.Cz; t.ky = tk.ky; t.kz = tk.kz;
      }
      cx[k] = (double)x - vx;
    }
    const double vy = s * Lat<kD3Q27>::efc(idx[0], 1) + uT[1];
    const double vz = s * Lat<kD3Q27>::efc(idx[0], 2) + uT[2];
    const double cy = (double)y - vy, cz = (double)z - vz;
    double fn[3];
    if constexpr (INTERP == 1) interp_lagr4_g3(src, x, cx, cy, cz, t, a.sqrt_tbase, fn);
    else interp_weno5_g3<VAR>(src, x, cx, cy, cz, t, a.sqrt_tbase, fn);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int64_t o = f64_fidx(kD3Q27, idx[k], cell, P);
      const double fo = niter > 1 ? fout[o] : (zero ? a.w[idx[k]] * 1.0 : gsrc.f[o]);
      if (fabs(fn[k] - fo) >= a.tol_abs + a.tol_rel * fabs(fn[k])) conv = false;
      fout[o] = fn[k];
    }
  }
  double rho = 0, m0 = 0, m1 = 0, m2 = 0, m3 = 0;
  for (int i = 0; i < 27; i++) {
    const double vx = s * Lat<kD3Q27>::efc(i, 0) + uT[0];
    const double vy = s * Lat<kD3Q27>::efc(i, 1) + uT[1];
    const double vz = s * Lat<kD3Q27>::efc(i, 2) + uT[2];
    const double fn = fout[f64_fidx(kD3Q27, i, cell, P)];
    rho += fn;
    m0 += fn * vx;
    m1 += fn * vy;
    m2 += fn * vz;
    m3 += fn * (vx * vx + vy * vy + vz * vz);
  }
  if (rho != 0) { m0 /= rho; m1 /= rho; m2 /= rho; m3 /= rho; }
  m3 = (m3 - m0 * m0 - m1 * m1 - m2 * m2) * a.inv_dim;
  if (m3 < 0) m3 = c0[3];
  const double mn[4] = {m0, m1, m2, m3};
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (fabs(mn[k] - c0[k]) >= a.tol_abs + a.tol_rel * fabs(mn[k])) conv = false;
  st.rho0 = rho;
  c0[0] = m0; c0[1] = m1; c0[2] = m2; c0[3] = m3;
  return conv;
}

// One fixed-point iteration (LBMpull, steps.cu:421-512) of one cell: true when it has converged.
template <int LAT, int INTERP, class SRC>
__device__ __forceinline__ bool f64_once(const SRC& src, const NodeSrc& gsrc, double* __restrict__ fout, int64_t cell,
                                         int x, int y, int z, bool zero, F64Iter& st, const StepArgs& a) {
  if constexpr (LAT == kD3Q27 && (INTERP == 0 || INTERP == 1 || INTERP == 2 || INTERP == 4)) {
    // the grouped path off a line, for gauges whose targets' windows lie close (s <= 1.5, not
    // NaN); hotter or non-finite ones, one target at a time below
    if ((a.g.ny > 1 || a.g.nz > 1) && sqrt(st.c0[3] / a.tlat) <= 1.5)
      return f64_once_g3<INTERP>(src, gsrc, fout, cell, x, y, z, zero, st, a);
  }
  constexpr int Q = Lat<LAT>::Q;
  const int64_t P = a.g.plane;
  double* c0 = st.c0;
  {
    st.niter++;
    const int niter = st.niter;
    const double uT[4] = {c0[0], c0[1], c0[2], c0[3]};
    const double s = sqrt(uT[3] / a.tlat);
    double rho = 0, m0 = 0, m1 = 0, m2 = 0, m3 = 0;
    bool conv = true;
    for (int i = 0; i < Q; i++) {
      const double vx = s * Lat<LAT>::efc(i, 0) + uT[0];
      const double vy = s * Lat<LAT>::efc(i, 1) + uT[1];
      const double vz = s * Lat<LAT>::efc(i, 2) + uT[2];
      const typename Lat<LAT>::Tgt t = Lat<LAT>::tgt(i, uT);
      const double cx = (double)x - vx, cy = (double)y - vy, cz = (double)z - vz;
      // INTERP 3: Lagrange-4 with mirror boundaries; 4: the generic WENO5
      const double fn = INTERP == 1   ? interp_lagr4<LAT, false>(src, cx, cy, cz, t, a.sqrt_tbase)
                        : INTERP == 3 ? interp_lagr4<LAT, true>(src, cx, cy, cz, t, a.sqrt_tbase)
                        : INTERP == 2 ? interp_weno5<LAT, 1>(src, cx, cy, cz, t, a.sqrt_tbase)
                        : INTERP == 4 ? interp_weno5<LAT, 2>(src, cx, cy, cz, t, a.sqrt_tbase)
                                      : interp_weno5<LAT, 0>(src, cx, cy, cz, t, a.sqrt_tbase);
      rho += fn;
      m0 += fn * vx;
      m1 += fn * vy;
      m2 += fn * vz;
      m3 += fn * (vx * vx + vy * vy + vz * vz);
      const int64_t o = f64_fidx(LAT, i, cell, P);
      const double fo = niter > 1 ? fout[o] : (zero ? a.w[i] * 1.0 : gsrc.f[o]);
      if (fabs(fn - fo) >= a.tol_abs + a.tol_rel * fabs(fn)) conv = false;
      fout[o] = fn;
    }
    // Cell::calcMoments (structs.cuh:329-330)
    if (rho != 0) { m0 /= rho; m1 /= rho; m2 /= rho; m3 /= rho; }
    m3 = (m3 - m0 * m0 - m1 * m1 - m2 * m2) * a.inv_dim;
    if (m3 < 0) m3 = c0[3];  // steps.cu:475-478
    const double mn[4] = {m0, m1, m2, m3};
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (fabs(mn[k] - c0[k]) >= a.tol_abs + a.tol_rel * fabs(mn[k])) conv = false;
    st.rho0 = rho;
    c0[0] = m0; c0[1] = m1; c0[2] = m2; c0[3] = m3;
    return conv;
  }
}

// Fixed-point iterations from state st until convergence (true) or until the count reaches
// stop_at (false).
template <int LAT, int INTERP, class SRC>
__device__ bool f64_iterate(const SRC& src, const NodeSrc& gsrc, double* __restrict__ fout, int64_t cell, int x, int y,
                            int z, bool zero, F64Iter& st, int stop_at, const StepArgs& a) {
  while (st.niter < stop_at)
    if (f64_once<LAT, INTERP>(src, gsrc, fout, cell, x, y, z, zero, st, a)) return true;
  return false;
}

// starting state of a cell: its previous gauge; an all-zero cell is reset to set_base(1,
// (0,0,0,TLat)) (steps.cu:360-363)
__device__ __forceinline__ bool f64_start(const NodeSrc& gsrc, int64_t cell, const StepArgs& a, F64Iter& st) {
  const int64_t P = a.g.plane;
  st.rho0 = gsrc.m[cell];
  st.c0[0] = gsrc.m[P + cell]; st.c0[1] = gsrc.m[2 * P + cell]; st.c0[2] = gsrc.m[3 * P + cell];
  st.c0[3] = gsrc.m[4 * P + cell];
  const bool zero = st.rho0 == 0 && st.c0[0] == 0 && st.c0[1] == 0 && st.c0[2] == 0 && st.c0[3] == 0;
  if (zero) { st.rho0 = 1; st.c0[3] = a.tlat; }
  st.niter = 0;
  return zero;
}

// BGK towards rho*w_i in the converged gauge (steps.cu:557-564) and the cell's outputs
template <int Q>
__device__ __forceinline__ void f64_finish(double* __restrict__ fout, double* __restrict__ mout,
                                           int* __restrict__ niter_out, int64_t cell, const F64Iter& st,
                                           const StepArgs& a) {
  const int64_t P = a.g.plane;
  for (int i = 0; i < Q; i++) {
    const int64_t o = f64_fidx(Q == 27 ? kD3Q27 : kD1Q5, i, cell, P);
    double f = fout[o];
    const double feq = st.rho0 * a.w[i];
    f += a.dtau * (feq - f);
    fout[o] = f;
  }
  mout[cell] = st.rho0;
  mout[P + cell] = st.c0[0];
  mout[2 * P + cell] = st.c0[1];
  mout[3 * P + cell] = st.c0[2];
  mout[4 * P + cell] = st.c0[3];
  niter_out[cell] = st.niter;
}

// First pass: every cell iterates up to a.pass1_iters times; cells still unconverged are
// appended to the pending list (one atomic per wave) and finished by a second pass over that
// list (k_step_f64_rest or k_step_f64_rest_wide).  A wave runs as long as its slowest cell, and
// a thin layer of cells on shocks and contact discontinuities needs several times the mean
// count.  Each cell's arithmetic and operation order are those of a single pass.
// BUF: node loads through buffer resources (planes up to 59,652,323 cells), else global loads.
// D3Q27: held to two waves per SIMD (256 VGPRs, a few spilled).  Measured on the 384^3
// shock-bubble FP64 step, A/B on one GPU: the grouped WENO5C kernel 2.70 s at one wave per SIMD
// (its registers spilled to AGPRs), 2.58 s at two; Lagrange-4 5.34 s -> 4.99 s.  The D1Q5 kernel
// keeps its registers (held to two waves it would spill 1.7 KB per lane).
template <int LAT, int INTERP, bool BUF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(LAT == kD3Q27 ? 2 : 1))) void k_step_f64(const NodeSrc gsrc, double* __restrict__ fout,
                                                 double* __restrict__ mout, int* __restrict__ niter_out, Pending pend,
                                                 const StepArgs a) {
  using Src = typename std::conditional<BUF, NodeBuf, NodeSrc>::type;
  Src src;
  if constexpr (BUF) src = make_nodebuf<LAT>(gsrc);
  else src = gsrc;
  const Grid& g = a.g;
  int x, y, z;
  bool active = tile_cell(g, xcd_block(blockIdx.x, gridDim.x), threadIdx.x, x, y, z);
  const int64_t cell = active ? cell_index(g, x, y, z) : 0;
  // a switched-off cell is not stepped: its output record and iteration count keep what they
  // held (LBMpull, steps.cu:359)
  if (active && a.cell_set && a.cell_set[cell] == 0) active = false;
  bool pending = false;
  F64Iter st;
  if (active) {
    const bool zero = f64_start(gsrc, cell, a, st);
    const int stop = a.pass1_iters < a.max_iter ? a.pass1_iters : a.max_iter;
    const bool conv = f64_iterate<LAT, INTERP>(src, gsrc, fout, cell, x, y, z, zero, st, stop, a);
    pending = !conv && st.niter < a.max_iter;
    if (!pending) f64_finish<Lat<LAT>::Q>(fout, mout, niter_out, cell, st, a);
  }
  bool any;
  const int64_t e = pending_slot(pend, pending, any);
  if (pending) {
    pend.cell[e] = (int)cell;
    for (int k = 0; k < 4; k++) pend.g[k * pend.cap + e] = st.c0[k];
    pend.res[e] = st.rho0;
    pend.meta[e] = st.niter;
  }
}

// First pass with lane refill: a wave owns `chunk` consecutive 64-cell tiles and its lanes take
// the next unstarted cell of them whenever theirs converges or is set aside (after
// a.pass1_iters iterations, to the pending list), instead of idling until the slowest cell of
// their tile is done.  Each cell's arithmetic is that of k_step_f64.
template <int LAT, int INTERP, bool BUF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(LAT == kD3Q27 ? 2 : 1))) void k_step_f64_refill(const NodeSrc gsrc, double* __restrict__ fout,
                                                        double* __restrict__ mout, int* __restrict__ niter_out,
                                                        Pending pend, const StepArgs a, int chunk) {
  using Src = typename std::conditional<BUF, NodeBuf, NodeSrc>::type;
  Src src;
  if constexpr (BUF) src = make_nodebuf<LAT>(gsrc);
  else src = gsrc;
  const Grid& g = a.g;
  const int lane = threadIdx.x;
  const int64_t ntiles = g.ntiles;
  const int64_t t0 = xcd_block(blockIdx.x, gridDim.x) * chunk;
  const int64_t total = (ntiles - t0 < chunk ? ntiles - t0 : chunk) * 64;  // cell slots of the chunk
  const int stop = a.pass1_iters < a.max_iter ? a.pass1_iters : a.max_iter;
  int64_t next = 0;  // first unstarted slot (the same in every lane)
  bool active = false, zero = false;
  int64_t cell = 0;
  int x = 0, y = 0, z = 0;
  F64Iter st;
  while (true) {
    const uint64_t want = __ballot(!active);
    const int rank = __popcll(want & ((1ull << lane) - 1ull));
    if (!active) {
      const int64_t slot = next + rank;
      if (slot < total && tile_cell(g, t0 + slot / 64, (int)(slot % 64), x, y, z)) {
        cell = cell_index(g, x, y, z);
        zero = f64_start(gsrc, cell, a, st);
        active = true;
      }
    }
    next += __popcll(want);
    if (__ballot(active) == 0 && next >= total) break;
    bool pending = false;
    if (active) {
      // the loop condition of f64_iterate, one iteration per trip
      bool done = st.niter >= stop, conv = false;
      if (!done) {
        conv = f64_once<LAT, INTERP>(src, gsrc, fout, cell, x, y, z, zero, st, a);
        done = conv || st.niter >= stop;
      }
      if (done) {
        pending = !conv && st.niter < a.max_iter;
        if (!pending) f64_finish<Lat<LAT>::Q>(fout, mout, niter_out, cell, st, a);
        active = false;
      }
    }
    bool any;
    const int64_t e = pending_slot(pend, pending, any);
    if (pending) {
      pend.cell[e] = (int)cell;
      for (int k = 0; k < 4; k++) pend.g[k * pend.cap + e] = st.c0[k];
      pend.res[e] = st.rho0;
      pend.meta[e] = st.niter;
    }
  }
}

// Second pass: a fixed grid of waves walks the pending list, every cell resumes its iteration
template <int LAT, int INTERP, bool BUF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(LAT == kD3Q27 ? 2 : 1))) void k_step_f64_rest(const NodeSrc gsrc, double* __restrict__ fout,
                                                      double* __restrict__ mout, int* __restrict__ niter_out,
                                                      Pending pend, const StepArgs a) {
  using Src = typename std::conditional<BUF, NodeBuf, NodeSrc>::type;
  Src src;
  if constexpr (BUF) src = make_nodebuf<LAT>(gsrc);
  else src = gsrc;
  const Grid& g = a.g;
  const int count = *pend.count;
  for (int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x; e - threadIdx.x < count; e += (int64_t)gridDim.x * 64) {
    if (e >= count) continue;
    const int64_t cell = pend.cell[e];
    const int x = (int)(cell % g.nx);
    const int y = (int)((cell / g.nx) % g.ny);
    const int z = (int)(cell / ((int64_t)g.nx * g.ny));
    F64Iter st;
    for (int k = 0; k < 4; k++) st.c0[k] = pend.g[k * pend.cap + e];
    st.rho0 = pend.res[e];
    st.niter = pend.meta[e];
    // the zero-cell start only matters for the first iteration, which the first pass ran
    f64_iterate<LAT, INTERP>(src, gsrc, fout, cell, x, y, z, false, st, a.max_iter, a);
    f64_finish<Lat<LAT>::Q>(fout, mout, niter_out, cell, st, a);
  }
}

// Second pass, one wave per unconverged cell and one lane per target population: each lane
// interpolates its own population with the one-target code, the wave then sums the moments over
// the populations in order from LDS (every lane the same sums), so every value is the one
// f64_once gives, while an iteration of a slow cell costs one population's interpolation instead
// of all of them in one lane.  The waves walk the pending list.  Used for the D3Q27 WENO5 steps
// (384^3 shock-bubble FP64 WENO5C step 2079 -> 1948 ms); the Lagrange-4 step keeps the one-lane
// pass, whose cells take the grouped path (2587 ms against 2712 wide).
template <int LAT, int INTERP, bool BUF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(LAT == kD3Q27 ? 2 : 1))) void k_step_f64_rest_wide(const NodeSrc gsrc, double* __restrict__ fout,
                                                           double* __restrict__ mout, int* __restrict__ niter_out,
                                                           Pending pend, const StepArgs a) {
  using Src = typename std::conditional<BUF, NodeBuf, NodeSrc>::type;
  Src src;
  if constexpr (BUF) src = make_nodebuf<LAT>(gsrc);
  else src = gsrc;
  constexpr int Q = Lat<LAT>::Q;
  constexpr int VAR = INTERP == 2 ? 1 : INTERP == 4 ? 2 : 0;
  __shared__ double sfn[Q];
  const Grid& g = a.g;
  const int64_t P = g.plane;
  const int i = threadIdx.x;
  const bool lane_ok = i < Q;
  const int count = *pend.count;
  for (int64_t e = blockIdx.x; e < count; e += gridDim.x) {
    const int64_t cell = pend.cell[e];
    const int x = (int)(cell % g.nx);
    const int y = (int)((cell / g.nx) % g.ny);
    const int z = (int)(cell / ((int64_t)g.nx * g.ny));
    double c0[4];
    for (int k = 0; k < 4; k++) c0[k] = pend.g[k * pend.cap + e];
    double rho0 = pend.res[e];
    int niter = pend.meta[e];
    const int64_t o = lane_ok ? f64_fidx(LAT, i, cell, P) : 0;
    while (niter < a.max_iter) {
      niter++;
      const double uT[4] = {c0[0], c0[1], c0[2], c0[3]};
      const double s = sqrt(uT[3] / a.tlat);
      bool conv_i = true;
      if (lane_ok) {
        const double vx = s * Lat<LAT>::efc(i, 0) + uT[0];
        const double vy = s * Lat<LAT>::efc(i, 1) + uT[1];
        const double vz = s * Lat<LAT>::efc(i, 2) + uT[2];
        const typename Lat<LAT>::Tgt t = Lat<LAT>::tgt(i, uT);
        const double cx = (double)x - vx, cy = (double)y - vy, cz = (double)z - vz;
        const double fn = INTERP == 1   ? interp_lagr4<LAT, false>(src, cx, cy, cz, t, a.sqrt_tbase)
                          : INTERP == 3 ? interp_lagr4<LAT, true>(src, cx, cy, cz, t, a.sqrt_tbase)
                                        : interp_weno5<LAT, VAR>(src, cx, cy, cz, t, a.sqrt_tbase);
        const double fo = fout[o];  // the previous iterate (the first pass ran at least one)
        if (fabs(fn - fo) >= a.tol_abs + a.tol_rel * fabs(fn)) conv_i = false;
        fout[o] = fn;
        sfn[i] = fn;
      }
      __syncthreads();
      // Cell::calcMoments (structs.cuh:329-330), population order, in every lane
      double rho = 0, m0 = 0, m1 = 0, m2 = 0, m3 = 0;
      for (int j = 0; j < Q; j++) {
        const double vx = s * Lat<LAT>::efc(j, 0) + uT[0];
        const double vy = s * Lat<LAT>::efc(j, 1) + uT[1];
        const double vz = s * Lat<LAT>::efc(j, 2) + uT[2];
        const double fn = sfn[j];
        rho += fn;
        m0 += fn * vx;
        m1 += fn * vy;
        m2 += fn * vz;
        m3 += fn * (vx * vx + vy * vy + vz * vz);
      }
      __syncthreads();  // sfn is rewritten next iteration
      if (rho != 0) { m0 /= rho; m1 /= rho; m2 /= rho; m3 /= rho; }
      m3 = (m3 - m0 * m0 - m1 * m1 - m2 * m2) * a.inv_dim;
      if (m3 < 0) m3 = c0[3];
      const double mn[4] = {m0, m1, m2, m3};
      bool conv = __ballot(!conv_i) == 0;
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (fabs(mn[k] - c0[k]) >= a.tol_abs + a.tol_rel * fabs(mn[k])) conv = false;
      rho0 = rho;
      c0[0] = m0; c0[1] = m1; c0[2] = m2; c0[3] = m3;
      if (conv) break;
    }
    // f64_finish, one population per lane
    if (lane_ok) {
      double f = fout[o];
      const double feq = rho0 * a.w[i];
      f += a.dtau * (feq - f);
      fout[o] = f;
    }
    if (i == 0) {
      mout[cell] = rho0;
      mout[P + cell] = c0[0];
      mout[2 * P + cell] = c0[1];
      mout[3 * P + cell] = c0[2];
      mout[4 * P + cell] = c0[3];
      niter_out[cell] = niter;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// D3Q125: one block per cell, lane = target population (PonDstepD3Q125, steps.cu:151-250)

__global__ __launch_bounds__(128) void k_step_q125_f64(const double* __restrict__ fin, const double* __restrict__ min,
                                                       const unsigned char* __restrict__ okin,
                                                       double* __restrict__ fout, double* __restrict__ mout,
                                                       int* __restrict__ niter_out, const StepArgs a) {
  constexpr int Q = 125;
  __shared__ double sf[27][Q];   // fixed neighbourhood populations (cfix, steps.cu:150)
  __shared__ double sg[27][4];   // their gauges
  __shared__ double snew[Q];     // cnew.f
  __shared__ int sflag[2];
  __shared__ int sok[27];        // node populations finite and below 2^200
  const Grid& g = a.g;
  const int64_t P = g.plane;
  const int64_t cell = xcd_block(blockIdx.x, gridDim.x);
  const int k = threadIdx.x;
  const bool lane_ok = k < Q;
  const int x = (int)(cell % g.nx);
  const int y = (int)((cell / g.nx) % g.ny);
  const int z = (int)(cell / ((int64_t)g.nx * g.ny));
  for (int n = 0; n < 27; n++) {
    // mirror boundaries (steps.cu:163-190): a node beyond a mirrored face is the mirror image of
    // the cell inside, population k read from the mirrored velocity, the axis velocity negated
    bool rx, ry, rz;
    const int px = mirror_wrap(x - 1 + n / 9, g.nx, (g.mirror & 1) != 0, rx);
    const int py = mirror_wrap(y - 1 + (n / 3) % 3, g.ny, (g.mirror & 2) != 0, ry);
    const int pz = mirror_wrap(z - 1 + n % 3, g.nz, (g.mirror & 4) != 0, rz);
    const int64_t c = cell_index(g, px, py, pz);
    const int kq = lane_ok ? k : 0;  // lanes 125..127 hold no population
    const int kx = kq / 25, ky = (kq / 5) % 5, kz = kq % 5;
    const int kl = 25 * (rx ? kMir5[kx] : kx) + 5 * (ry ? kMir5[ky] : ky) + (rz ? kMir5[kz] : kz);
    if (lane_ok) sf[n][k] = fin[f64_fidx(kD3Q125, kl, c, P)];
    if (k < 4) {
      const double v = min[(int64_t)(k + 1) * P + c];
      sg[n][k] = (k == 0 && rx) || (k == 1 && ry) || (k == 2 && rz) ? -v : v;
    }
    if (k == 4) sok[n] = okin[c];
  }
  double rho0 = min[cell], c0[4] = {min[P + cell], min[2 * P + cell], min[3 * P + cell], min[4 * P + cell]};
  const bool zero = rho0 == 0 && c0[0] == 0 && c0[1] == 0 && c0[2] == 0 && c0[3] == 0;
  if (zero) { rho0 = 1; c0[3] = a.tlat; }
  double f0 = lane_ok ? (zero ? a.w[k] * 1.0 : fin[f64_fidx(kD3Q125, k, cell, P)]) : 0.0;
  __syncthreads();

  // per-axis class of target k: index into {0,+ec1,-ec1,+ec2,-ec2}
  const int kk = lane_ok ? k : 0;
  const int X[3] = {kk / 25, (kk / 5) % 5, kk % 5};
  int niter = 0;
  while (niter < a.max_iter) {
    niter++;
    const double uT[4] = {c0[0], c0[1], c0[2], c0[3]};
    const double s = sqrt(uT[3] / a.tlat);
    // target constants (structs.cuh:397-416)
    const double S1 = sqrts_d(uT[3] / 1.0);
    const double e1_to = kEc1 * S1, e2_to = kEc2 * S1;
    double wv[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double en = kA5[X[d]];
      if (en == 0) wv[d] = 1. / (e1_to * e1_to * e2_to * e2_to);
      else if (fabs(en) == kEc1) wv[d] = 1. / (2 * e1_to * e1_to * (e1_to * e1_to - e2_to * e2_to));
      else wv[d] = 1. / (2 * e2_to * e2_to * (e2_to * e2_to - e1_to * e1_to));
    }
    const double W = wv[0] * wv[1] * wv[2];
    const bool wok = fabs(W) < kSafe;
    const double vk[3] = {s * kA5[X[0]] + uT[0], s * kA5[X[1]] + uT[1], s * kA5[X[2]] + uT[2]};
    const double sh[3] = {((double)x - vk[0]) - (double)(x - 1), ((double)y - vk[1]) - (double)(y - 1),
                          ((double)z - vk[2]) - (double)(z - 1)};
    double Fx[3][2], Fy[3][2], Fz[3][2];
    lag3_factors(sh[0], Fx);
    lag3_factors(sh[1], Fy);
    lag3_factors(sh[2], Fz);
    // per-lane offsets of the four factors of the target's basis polynomial, in the cyclic
    // order (X+1..X+4)%5 of structs.cuh:432-434: factor m is b + off[(X+m)%5] with
    // off = {-0.0, -e1_to, +e1_to, -e2_to, +e2_to} (b - c == b + (-c) and b + (-0.0) == b exactly)
    double off[3][4];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int mm = 1; mm <= 4; mm++) {
        const int c = (X[d] + mm) % 5;
        off[d][mm - 1] = c == 0 ? -0.0 : c == 1 ? -e1_to : c == 2 ? e1_to : c == 3 ? -e2_to : e2_to;
      }
    double fpr = 0;
    for (int n = 0; n < 27; n++) {
      const double S0 = sqrts_d(sg[n][3] / 1.0);
      double gb[3][5];
      bool safe = wok && sok[n] != 0;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const double du = sg[n][d] - uT[d];
#pragma unroll
        for (int j = 0; j < 5; j++) {
          const double b = du + kA5[j] * S0;
          gb[d][j] = (b + off[d][0]) * (b + off[d][1]) * (b + off[d][2]) * (b + off[d][3]);
          safe = safe && fabs(gb[d][j]) < kSafe;
        }
      }
      double fnew = 0;
      if (safe) {
        // every term is finite: |W|,|gx|,|gy|,|gz|,|f| < 2^200
#pragma unroll
        for (int jx = 0; jx < 5; jx++)
#pragma unroll
          for (int jy = 0; jy < 5; jy++) {
            const double gxy = gb[0][jx] * gb[1][jy];
#pragma unroll
            for (int jz = 0; jz < 5; jz++) fnew += W * (gxy * gb[2][jz]) * sf[n][25 * jx + 5 * jy + jz];
          }
      } else {
#pragma unroll
        for (int jx = 0; jx < 5; jx++)
#pragma unroll
          for (int jy = 0; jy < 5; jy++) {
            const double gxy = gb[0][jx] * gb[1][jy];
#pragma unroll
            for (int jz = 0; jz < 5; jz++) {
              const double fadd = W * (gxy * gb[2][jz]) * sf[n][25 * jx + 5 * jy + jz];
              if (finite_d(fadd)) fnew += fadd;
            }
          }
      }
      const int ia = n / 9, ib = (n / 3) % 3, ic = n % 3;
      fpr += 1 * Fx[ia][0] * Fx[ia][1] * Fy[ib][0] * Fy[ib][1] * Fz[ic][0] * Fz[ic][1] * fnew;
    }
    if (lane_ok) snew[k] = fpr;
    __syncthreads();
    // Cell::calcMoments on the shared cnew, serial in population order (every lane, same result)
    double rho = 0, m0 = 0, m1 = 0, m2 = 0, m3 = 0;
    for (int i = 0; i < Q; i++) {
      const double fi = snew[i];
      const double vx = s * kA5[i / 25] + uT[0];
      const double vy = s * kA5[(i / 5) % 5] + uT[1];
      const double vz = s * kA5[i % 5] + uT[2];
      rho += fi;
      m0 += fi * vx;
      m1 += fi * vy;
      m2 += fi * vz;
      m3 += fi * (vx * vx + vy * vy + vz * vz);
    }
    if (rho != 0) { m0 /= rho; m1 /= rho; m2 /= rho; m3 /= rho; }
    m3 = (m3 - m0 * m0 - m1 * m1 - m2 * m2) * a.inv_dim;
    if (m3 < 0) m3 = c0[3];
    const double mn[4] = {m0, m1, m2, m3};
    bool conv = true;
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (fabs(mn[q] - c0[q]) >= a.tol_abs + a.tol_rel * fabs(mn[q])) conv = false;
    if (lane_ok && fabs(fpr - f0) >= a.tol_abs + a.tol_rel * fabs(fpr)) conv = false;
    const bool wave_conv = __all(conv);
    if ((threadIdx.x & 63) == 0) sflag[threadIdx.x >> 6] = wave_conv ? 1 : 0;
    __syncthreads();
    const bool all_conv = sflag[0] && sflag[1];
    rho0 = rho;
    c0[0] = m0; c0[1] = m1; c0[2] = m2; c0[3] = m3;
    f0 = fpr;
    __syncthreads();  // sflag and snew are rewritten next iteration
    if (all_conv) break;
  }
  if (lane_ok) {
    const double feq = rho0 * a.w[k];
    f0 += a.dtau * (feq - f0);
    fout[f64_fidx(kD3Q125, k, cell, P)] = f0;
  }
  if (k == 0) {
    mout[cell] = rho0;
    mout[P + cell] = c0[0];
    mout[2 * P + cell] = c0[1];
    mout[3 * P + cell] = c0[2];
    mout[4 * P + cell] = c0[3];
    niter_out[cell] = niter;
  }
}

// ---------------------------------------------------------------------------------------------
// host-side launchers

// per-step node data: sqrt(|T|)*sign(T) and the "populations finite and below 2^200" flag
template <int Q>
__global__ void k_node_prep(const double* __restrict__ f, const double* __restrict__ m, double* __restrict__ P,
                            unsigned char* __restrict__ ok, int64_t n, int64_t plane) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double T = m[4 * plane + c];
  const double Pc = sqrt(fabs(T)) * (double)isign_d(T);
  if (Q == 27) {
    // gauge record (ux, uy, uz, P) of the cell, 32 bytes
    reinterpret_cast<double2*>(P)[2 * c] = make_double2(m[plane + c], m[2 * plane + c]);
    reinterpret_cast<double2*>(P)[2 * c + 1] = make_double2(m[3 * plane + c], Pc);
  } else {
    P[c] = Pc;
  }
  bool good = true;
  for (int i = 0; i < Q; i++)
    good = good && fabs(f[f64_fidx(Q == 27 ? kD3Q27 : Q == 5 ? kD1Q5 : kD3Q125, i, c, plane)]) < kSafe;
  ok[c] = good ? 1 : 0;
}

hipError_t launch_node_prep(int lattice, const double* f, const double* m, double* P, unsigned char* ok, int64_t n,
                            int64_t plane, hipStream_t st) {
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (lattice == kD3Q27)
    hipLaunchKernelGGL(k_node_prep<27>, dim3(blocks), dim3(256), 0, st, f, m, P, ok, n, plane);
  else if (lattice == kD1Q5)
    hipLaunchKernelGGL(k_node_prep<5>, dim3(blocks), dim3(256), 0, st, f, m, P, ok, n, plane);
  else
    hipLaunchKernelGGL(k_node_prep<125>, dim3(blocks), dim3(256), 0, st, f, m, P, ok, n, plane);
  return hipGetLastError();
}

// the second pass: wide for the D3Q27 WENO5 steps, one lane per cell otherwise
template <int LAT, int INTERP, bool BUF>
static void launch_rest(unsigned blocks2, hipStream_t st, const NodeSrc& src, double* fout, double* mout, int* niter,
                        const Pending& pend, const StepArgs& a) {
  if constexpr (LAT == kD3Q27 && (INTERP == 0 || INTERP == 2 || INTERP == 4))
    hipLaunchKernelGGL((k_step_f64_rest_wide<LAT, INTERP, BUF>), dim3(blocks2), dim3(64), 0, st, src, fout, mout, niter,
                       pend, a);
  else
    hipLaunchKernelGGL((k_step_f64_rest<LAT, INTERP, BUF>), dim3(blocks2), dim3(64), 0, st, src, fout, mout, niter,
                       pend, a);
}

hipError_t launch_step_f64(int lattice, int interp, const double* fin, const double* min, const double* Pin,
                           const unsigned char* okin, double* fout, double* mout, int* niter, const Pending& pend,
                           const StepArgs& a, hipStream_t st) {
  if (lattice == kD3Q125) {
    const int64_t blocks = a.g.n;
    hipLaunchKernelGGL(k_step_q125_f64, dim3((unsigned)blocks), dim3(128), 0, st, fin, min, okin, fout, mout, niter,
                       a);
    return hipGetLastError();
  }
  const unsigned blocks = (unsigned)(a.g.ntiles);
  const NodeSrc src{fin, min, Pin, okin, a.g};
  // buffer-resource loads while nine FP64 planes (D1Q5) or four pair planes (D3Q27) fit a 32-bit
  // offset
  const bool buf = (uint64_t)a.g.plane * 8u * (lattice == kD3Q27 ? 8u : (uint64_t)kPlanesPerRsrc64) <= 0xffffffffull &&
                   !a.global_loads;
  hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, st, pend.count);
  const int64_t w2 = (a.g.n + 63) / 64;
  const unsigned blocks2 = (unsigned)(w2 < 16384 ? w2 : 16384);
#define POND_LAUNCH(L, I, B)                                                                                     \
  do {                                                                                                           \
    hipLaunchKernelGGL((k_step_f64<L, I, B>), dim3(blocks), dim3(64), 0, st, src, fout, mout, niter, pend, a);   \
    launch_rest<L, I, B>(blocks2, st, src, fout, mout, niter, pend, a);                                          \
  } while (0)
#define POND_LAUNCH_R(L, I)                                                                                      \
  do {                                                                                                           \
    const unsigned w = (unsigned)(((int64_t)blocks + a.refill_chunk - 1) / a.refill_chunk);                       \
    hipLaunchKernelGGL((k_step_f64_refill<L, I, true>), dim3(w), dim3(64), 0, st, src, fout, mout, niter, pend, a, \
                       a.refill_chunk);                                                                          \
    launch_rest<L, I, true>(blocks2, st, src, fout, mout, niter, pend, a);                                       \
  } while (0)
  // kernel codes: 0 WENO5 as written, 1 Lagrange-4, 2 WENO5 centred, 3 Lagrange-4 with mirror
  // boundaries (its own instantiation, so the periodic kernels carry none of the mirror code),
  // 4 the generic WENO5.  interp is pond_interp (POND_WENO5G = 3).
  const int code = interp == 3 ? 4 : (interp == 1 && a.g.mirror) ? 3 : interp;
  if (lattice == kD3Q27 && buf && a.refill_chunk > 0) {
    if (code == 0) POND_LAUNCH_R(kD3Q27, 0);
    else if (code == 1) POND_LAUNCH_R(kD3Q27, 1);
    else if (code == 3) POND_LAUNCH_R(kD3Q27, 3);
    else if (code == 4) POND_LAUNCH_R(kD3Q27, 4);
    else POND_LAUNCH_R(kD3Q27, 2);
  } else if (lattice == kD3Q27) {
    if (buf) {
      if (code == 0) POND_LAUNCH(kD3Q27, 0, true);
      else if (code == 1) POND_LAUNCH(kD3Q27, 1, true);
      else if (code == 3) POND_LAUNCH(kD3Q27, 3, true);
      else if (code == 4) POND_LAUNCH(kD3Q27, 4, true);
      else POND_LAUNCH(kD3Q27, 2, true);
    } else {
      if (code == 0) POND_LAUNCH(kD3Q27, 0, false);
      else if (code == 1) POND_LAUNCH(kD3Q27, 1, false);
      else if (code == 3) POND_LAUNCH(kD3Q27, 3, false);
      else if (code == 4) POND_LAUNCH(kD3Q27, 4, false);
      else POND_LAUNCH(kD3Q27, 2, false);
    }
  } else {
    // on a line (Ny = Nz = 1) the centred and as-written WENO5 coincide
    if (code == 1) POND_LAUNCH(kD1Q5, 1, false);
    else if (code == 3) POND_LAUNCH(kD1Q5, 3, false);
    else if (code == 4) POND_LAUNCH(kD1Q5, 4, false);
    else POND_LAUNCH(kD1Q5, 0, false);
  }
#undef POND_LAUNCH
#undef POND_LAUNCH_R
  return hipGetLastError();
}

}  // namespace pond
