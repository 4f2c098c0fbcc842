// step_split.hip -- FP32/FP64-split PonD step for D3Q27 on gfx950.
//
// Same algorithm as the plain path (LBMpull, steps.cu:356-588): fixed-point iteration on the
// cell's gauge (u, T), pull streaming of gauge-transferred populations with WENO5 / Lagrange-4
// interpolation, moments, BGK towards rho*w_i.  Precision split:
//   FP32: population storage, the gauge transfer, interpolation, WENO weights, collide;
//         rho and rho*u summed in FP32 with Kahan compensation;
//   FP64: the energy moment sum f|v|^2, the moment solve (u = m/rho, T = (E/rho - |u|^2)/Dim),
//         the stored temperature, and the gauge-convergence residual.
// The gauge transfer is sum-factorised: the target population's basis is a tensor product of
// per-axis quadratic Lagrange polynomials (structs.cuh:525-548), so
//   f_new(k) = sum_jx Lx[jx] sum_jy Ly[jy] sum_jz Lz[jz] f(jx,jy,jz)
// costs 39 FMAs plus 9 basis values instead of 27 five-factor products.
//
// Convergence (FP32-appropriate; the reference's 1e-10 relative test cannot be met by FP32
// populations): the FP64 gauge residual |du_d| <= tol_rel*sqrt(T) + tol_abs and
// |dT| <= tol_rel*T + tol_abs (defaults 1e-7, 1e-12), or, at the FP32 noise floor, a residual
// below 4 tol_rel that no longer halves from one iteration to the next.  Populations are a
// deterministic function of the gauge, so they converge with it.
//
// Starting gauge: the cell's previous gauge, or (learned initializer) a predicted gauge.  The
// fixed-point iteration has the last word: a prediction that is not finite or not positive in T
// is ignored, and if the residual stops falling before the noise floor, or the cell has not
// converged after kLearnedBudget iterations, it restarts from its previous gauge.
#include "pond_device.h"
#include "step_split.h"

namespace pond {

// per-lane constants of one target population in the current gauge
struct TgtF {
  float u1x, u1y, u1z;     // target gauge velocity, rounded to FP32 ...
  float u1xl, u1yl, u1zl;  // ... and the rounding remainder (u1 = hi + lo to ~2^-48)
  float T1;
  float wx, wy, wz;  // 1/(2T1) or -1/T1 per axis
  float cx, cy, cz;  // e_k*sqrt(T1) per axis
  bool kx, ky, kz;
};

__device__ __forceinline__ TgtF make_tgtf(int k, const double* g) {
  TgtF t;
  t.u1x = (float)g[0]; t.u1y = (float)g[1]; t.u1z = (float)g[2];
  t.u1xl = (float)(g[0] - (double)t.u1x);
  t.u1yl = (float)(g[1] - (double)t.u1y);
  t.u1zl = (float)(g[2] - (double)t.u1z);
  const float T1 = (float)g[3];
  t.T1 = T1;
  const float h = 1.0f / (2.0f * T1), m = -1.0f / T1;
  const float sq1 = sqrtf(fabsf(T1)) * (float)((0.0f < T1) - (T1 < 0.0f));
  const int ex = kE27[k][0], ey = kE27[k][1], ez = kE27[k][2];
  t.wx = ex ? h : m; t.wy = ey ? h : m; t.wz = ez ? h : m;
  t.cx = ex * sq1; t.cy = ey * sq1; t.cz = ez * sq1;
  t.kx = ex != 0; t.ky = ey != 0; t.kz = ez != 0;
  return t;
}

// basis values L[j], j = source component -1,0,+1, of one axis (structs.cuh:530-538)
__device__ __forceinline__ void basis3(float D, float sq0, bool kn, float c, float T1, float w, float L[3]) {
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const float A = D - (float)(j - 1) * sq0;
    const float g = kn ? A * (A - c) : __builtin_fmaf(A, A, -T1);
    L[j] = w * g;
  }
}

// basis3 without the per-axis factor w, the axis kind folded into (c, T): a moving axis
// (kn) passes c = e*sqrt(T1), T = 0 and gets A (A - c); a resting one passes c = 0, T = T1 and
// gets A^2 - T1.  fma(A, A - c, -T) is either, rounded as in basis3, without a select between
// the two forms where the kind is only known at run time.  A_j = D - (j - 1) sq0.
__device__ __forceinline__ void basis3u(float D, float sq0, float c, float T, float G[3]) {
  const float A0 = D + sq0, A1 = D, A2 = D - sq0;
  G[0] = __builtin_fmaf(A0, A0 - c, -T);
  G[1] = __builtin_fmaf(A1, A1 - c, -T);
  G[2] = __builtin_fmaf(A2, A2 - c, -T);
}

// Buffer resources: a load names its plane (or quad plane) through the scalar offset and the
// per-lane address is a single 32-bit offset, where a global load needs a 64-bit per-lane
// address (two VGPRs and a 64-bit add per load).  The resource size is exact, so the hardware
// range check returns 0 rather than reading past the array.  Word 3 (0x00020000) is the
// raw-buffer configuration of the gfx9 family.
constexpr uint32_t kRawBufferWord3 = 0x00020000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)(uint32_t)bytes, (int)kRawBufferWord3);
}

// the seven quad planes of the D3Q27 split populations: resource j holds quad planes 2j, 2j+1,
// the quad plane within it is chosen by the scalar offset (plane * 16 bytes each, so up to
// 134,217,727 cells fit the 32-bit offsets)
struct QuadsF {
  __amdgpu_buffer_rsrc_t r[4];
  uint32_t qbytes;  // bytes per quad plane
};

__device__ __forceinline__ QuadsF make_quads(const float* base, int64_t plane) {
  QuadsF q;
  q.qbytes = (uint32_t)(plane * 16);
  for (int j = 0; j < 4; j++) {
    const int cnt = 7 - 2 * j < 2 ? 7 - 2 * j : 2;
    q.r[j] = make_rsrc(base + (int64_t)2 * j * plane * 4, (uint64_t)cnt * q.qbytes);
  }
  return q;
}

// the 27 populations of the cell at byte offset boff of a 4-byte plane (boff = 4 * cell)
__device__ __forceinline__ void load27(const QuadsF& q, uint32_t boff, float f[27]) {
#pragma unroll
  for (int j = 0; j < 7; j++) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(q.r[j >> 1], (int)(boff * 4u), (int)((uint32_t)(j & 1) * q.qbytes), 0);
#pragma unroll
    for (int e = 0; e < 4; e++)
      if (4 * j + e < 27) f[4 * j + e] = __builtin_bit_cast(float, (uint32_t)v[e]);
  }
}

// sum-factorised gauge transfer of the node at byte offset boff to target t.  The gauge
// difference u1 - u0 is formed as (hi - u0) + lo: hi - u0 is exact when the two velocities are
// within a factor of two (Sterbenz), so the difference carries one rounding relative to itself
// rather than one relative to |u1|, which would put a jitter of ulp(|u|) into the velocity
// moment.
__device__ __forceinline__ float transfer27f(const float f[27], float u0x, float u0y, float u0z, float sq0,
                                             const TgtF& t, float sqTb) {
  float Lx[3], Ly[3], Lz[3];
  basis3(((t.u1x - u0x) + t.u1xl) * sqTb, sq0, t.kx, t.cx, t.T1, t.wx, Lx);
  basis3(((t.u1y - u0y) + t.u1yl) * sqTb, sq0, t.ky, t.cy, t.T1, t.wy, Ly);
  basis3(((t.u1z - u0z) + t.u1zl) * sqTb, sq0, t.kz, t.cz, t.T1, t.wz, Lz);
  float out = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float sy = 0.0f;
#pragma unroll
    for (int b = 0; b < 3; b++) {
      float sz = 0.0f;
#pragma unroll
      for (int c = 0; c < 3; c++) sz = __builtin_fmaf(Lz[c], f[kIdx27.v[a][b][c]], sz);
      sy = __builtin_fmaf(Ly[b], sz, sy);
    }
    out = __builtin_fmaf(Lx[a], sy, out);
  }
  return out;
}

// node side of the load state: the population planes and the node gauges, one 16-byte record
// (ux, uy, uz, sqrt(|T|)*sign(T)) per cell written by k_node_gauge before the step, so a node's
// gauge is one load
struct NodeSrcF {
  QuadsF f;
  __amdgpu_buffer_rsrc_t g4;
  Grid g;
};

__device__ __forceinline__ NodeSrcF make_src(const float* f, const float* g4, const Grid& g) {
  return NodeSrcF{make_quads(f, g.plane), make_rsrc(g4, (uint64_t)g.plane * 16u), g};
}

struct Gauge4 {
  float ux, uy, uz, sq;
};

// gauge record of the cell at byte offset boff of the population planes (4 bytes per cell)
__device__ __forceinline__ Gauge4 ldg4(const NodeSrcF& s, uint32_t boff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(s.g4, (int)(boff * 4u), 0, 0);
  return Gauge4{__builtin_bit_cast(float, (uint32_t)v[0]), __builtin_bit_cast(float, (uint32_t)v[1]),
                __builtin_bit_cast(float, (uint32_t)v[2]), __builtin_bit_cast(float, (uint32_t)v[3])};
}

__device__ __forceinline__ float node_xfer(const NodeSrcF& s, int64_t c, const TgtF& t, float sqTb) {
  const uint32_t boff = (uint32_t)c * 4u;
  float f[27];
  load27(s.f, boff, f);
  const Gauge4 n = ldg4(s, boff);
  return transfer27f(f, n.ux, n.uy, n.uz, n.sq, t, sqTb);
}

// node_xfer of the node's mirror image on the axes set in mx, my, mz (Cell::mirrX/Y/Z,
// structs.cuh:552-578: the axis velocity negated, every population exchanged with its mirror
// image).  Exchanging the populations along an axis is the same as reversing that axis's basis
// values (sum_j L[j] f[rev j] = sum_j L[rev j] f[j]), so the populations stay in place and only
// the three basis values of a mirrored axis are swapped end for end.
__device__ float node_xfer_mirrored(const NodeSrcF& s, int64_t c, const TgtF& t, float sqTb, bool mx, bool my,
                                    bool mz) {
  const uint32_t boff = (uint32_t)c * 4u;
  float f[27];
  load27(s.f, boff, f);
  const Gauge4 n = ldg4(s, boff);
  float Lx[3], Ly[3], Lz[3];
  basis3(((t.u1x - (mx ? -n.ux : n.ux)) + t.u1xl) * sqTb, n.sq, t.kx, t.cx, t.T1, t.wx, Lx);
  basis3(((t.u1y - (my ? -n.uy : n.uy)) + t.u1yl) * sqTb, n.sq, t.ky, t.cy, t.T1, t.wy, Ly);
  basis3(((t.u1z - (mz ? -n.uz : n.uz)) + t.u1zl) * sqTb, n.sq, t.kz, t.cz, t.T1, t.wz, Lz);
  auto flip = [](bool m, float L[3]) {
    const float a = L[0];
    L[0] = m ? L[2] : L[0];
    L[2] = m ? a : L[2];
  };
  flip(mx, Lx);
  flip(my, Ly);
  flip(mz, Lz);
  float out = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float sy = 0.0f;
#pragma unroll
    for (int b = 0; b < 3; b++) {
      float sz = 0.0f;
#pragma unroll
      for (int cc = 0; cc < 3; cc++) sz = __builtin_fmaf(Lz[cc], f[kIdx27.v[a][b][cc]], sz);
      sy = __builtin_fmaf(Ly[b], sz, sy);
    }
    out = __builtin_fmaf(Lx[a], sy, out);
  }
  return out;
}

__device__ __forceinline__ void lag3f(float s, float F[3]) {
  // quadratic Lagrange basis on nodes 0,1,2 at s
  F[0] = 0.5f * (s - 1.0f) * (s - 2.0f);
  F[1] = -s * (s - 2.0f);
  F[2] = 0.5f * s * (s - 1.0f);
}

__device__ __forceinline__ void lag4f(float s, float F[4]) {
  // cubic Lagrange basis on nodes 0..3 at s
  const float a = s, b = s - 1.0f, c = s - 2.0f, d = s - 3.0f;
  F[0] = -(b * c * d) * (1.0f / 6.0f);
  F[1] = 0.5f * a * c * d;
  F[2] = -0.5f * a * b * d;
  F[3] = (a * b * c) * (1.0f / 6.0f);
}

// nonlinear WENO5 combination of the three sub-stencil values S at offset c = cx - r from the
// five smoothness samples ub (steps.cu:795-825)
__device__ __forceinline__ float weno5_combine(float c, const float ub[5], const float S[3]) {
  const float d0 = (c * c - 3.f * c + 2.f) * (1.f / 12.f);
  const float d2 = (c * c + 3.f * c + 2.f) * (1.f / 12.f);
  const float d1 = 1.f - d0 - d2;
  // Smoothness indicators in the difference form of Jiang & Shu (J. Comput. Phys. 126, 1996),
  // which expands to exactly the reference's polynomials (steps.cu:801-806).  In FP32 the
  // expanded form cancels terms of size ~25 f^2 down to beta ~ (df)^2, and the rounding of
  // those terms, measured against the 1e-6 regularisation, moves the weights by ~1e-3 from one
  // iteration to the next; the differences carry no such cancellation.
  const float u1 = ub[0], u2 = ub[1], u3 = ub[2], u4 = ub[3], u5 = ub[4];
  const float p0 = u1 - 2.f * u2 + u3, q0 = u1 - 4.f * u2 + 3.f * u3;
  const float p1 = u2 - 2.f * u3 + u4, q1 = u2 - u4;
  const float p2 = u3 - 2.f * u4 + u5, q2 = 3.f * u3 - 4.f * u4 + u5;
  const float b0 = (13.f / 12.f) * p0 * p0 + 0.25f * q0 * q0;
  const float b1 = (13.f / 12.f) * p1 * p1 + 0.25f * q1 * q1;
  const float b2 = (13.f / 12.f) * p2 * p2 + 0.25f * q2 * q2;
  // Weights d_k / (e + b_k)^2, each multiplied by (e + b_min)^2: the same normalised weights,
  // but every factor (e + b_min) / (e + b_k) lies in (0, 1], so large samples (b ~ 1e19 and up,
  // where (e + b)^2 leaves FP32 range and all three weights would round to 0/0) keep finite
  // weights with the largest one equal to d_k.
  const float e = 1e-6f;
  const float em = e + fminf(b0, fminf(b1, b2));
  const float r0 = em / (e + b0), r1 = em / (e + b1), r2 = em / (e + b2);
  const float w0 = d0 * (r0 * r0), w1 = d1 * (r1 * r1), w2 = d2 * (r2 * r2);
  return (w0 * S[0] + w1 * S[1] + w2 * S[2]) / (w0 + w1 + w2);
}


// WENO5 interpolation (propagateWENO<3,5>) in FP32; CENTRED selects the corrected y/z stencil
template <bool CENTRED>
__device__ float interp_weno5f(const NodeSrcF& src, double cx, double cy, double cz, const TgtF& t, float sqTb) {
  const Grid& g = src.g;
  const int r = floor_int(cx + 0.5);
  const int ry = floor_int(cy + 0.5) - (CENTRED ? 1 : 0), rz = floor_int(cz + 0.5) - (CENTRED ? 1 : 0);
  float Fy[3], Fz[3], Fx[3][3];
  lag3f((float)(cy - ry), Fy);
  lag3f((float)(cz - rz), Fz);
#pragma unroll
  for (int k = 0; k < 3; k++) lag3f((float)(cx - (r - 2 + k)), Fx[k]);
  int yy[3], zz[3];
#pragma unroll
  for (int a = 0; a < 3; a++) { yy[a] = wrapi(ry + a, g.ny); zz[a] = wrapi(rz + a, g.nz); }
  const int by = CENTRED ? yy[1] : wrapi(r, g.ny), bz = CENTRED ? zz[1] : wrapi(r, g.nz);
  float S0 = 0.f, S1 = 0.f, S2 = 0.f, ub[5];
  for (int xs = 0; xs < 5; xs++) {
    const int x = wrapi(r - 2 + xs, g.nx);
    float yzsum = 0.f;  // y/z interpolated value of this x slice
    float mid = 0.f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      float zs = 0.f;
#pragma unroll
      for (int b = 0; b < 3; b++) {
        const float T = node_xfer(src, cell_index(g, x, yy[a], zz[b]), t, sqTb);
        if (a == 1 && b == 1) mid = T;
        zs = __builtin_fmaf(Fz[b], T, zs);
      }
      yzsum = __builtin_fmaf(Fy[a], zs, yzsum);
    }
    if (CENTRED) ub[xs] = mid;
    else ub[xs] = node_xfer(src, cell_index(g, x, by, bz), t, sqTb);
    if (xs <= 2) S0 = __builtin_fmaf(Fx[0][xs], yzsum, S0);
    if (xs >= 1 && xs <= 3) S1 = __builtin_fmaf(Fx[1][xs - 1], yzsum, S1);
    if (xs >= 2) S2 = __builtin_fmaf(Fx[2][xs - 2], yzsum, S2);
  }
  const float Sv[3] = {S0, S1, S2};
  return weno5_combine((float)(cx - r), ub, Sv);
}

// MIR: mirror boundaries on the axes of g.mirror (steps.cu:1004-1022)
template <bool MIR>
__device__ float interp_lagr4f(const NodeSrcF& src, double cx, double cy, double cz, const TgtF& t, float sqTb) {
  const Grid& g = src.g;
  const int lx = floor_int(cx) - 1, ly = floor_int(cy) - 1, lz = floor_int(cz) - 1;
  float Fx[4], Fy[4], Fz[4];
  lag4f((float)(cx - lx), Fx);
  lag4f((float)(cy - ly), Fy);
  lag4f((float)(cz - lz), Fz);
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
  float out = 0.f;
  for (int ix = 0; ix < 4; ix++) {
    bool rx = false;
    const int x = MIR ? mirror_wrap(lx + ix, g.nx, (g.mirror & 1) != 0, rx) : wrapi(lx + ix, g.nx);
    float ys = 0.f;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      float zs = 0.f;
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int64_t c = cell_index(g, x, yy[a], zz[b]);
        const float v = MIR && (rx || ry[a] || rz[b]) ? node_xfer_mirrored(src, c, t, sqTb, rx, ry[a], rz[b])
                                                      : node_xfer(src, c, t, sqTb);
        zs = __builtin_fmaf(Fz[b], v, zs);
      }
      ys = __builtin_fmaf(Fy[a], zs, ys);
    }
    out = __builtin_fmaf(Fx[ix], ys, out);
  }
  return out;
}

// WENO5C for the three targets of one (ey, ez) group: ex = -1, 0, +1.  They share the departure
// rows in y and z (same y and z velocity), and their transfer bases differ only along x, so for
// every stencil node the contraction of the 27 source populations over the z and y source
// components, sy[jx] = sum_jy Ly[jy] sum_jz Lz[jz] f(jx,jy,jz), is common to the three:
// 27 loads and 36 FMAs per node, then 3 FMAs and an x basis per target, where three separate
// transfers load the node three times and contract it three times.  The x slices of the three
// five-slice WENO windows are visited once each (their union, at most 5 + |r(-1) - r(+1)|).
// Same interpolation as interp_weno5f<true> per target, with the y/z weights summed in a
// different order.  The x slices are visited two at a time (more loads in flight per lane).
__device__ void interp_group_weno5c(const NodeSrcF& src, int x, int y, int z, double s, const double gc[4], int ey,
                                    int ez, float sqTb, float out[3]) {
  const Grid& g = src.g;
  const float u1x = (float)gc[0], u1y = (float)gc[1], u1z = (float)gc[2];
  const float u1xl = (float)(gc[0] - (double)u1x), u1yl = (float)(gc[1] - (double)u1y),
              u1zl = (float)(gc[2] - (double)u1z);
  const float T1 = (float)gc[3];
  const float h = 1.0f / (2.0f * T1), mm = -1.0f / T1;
  const float sq1 = sqrtf(fabsf(T1)) * (float)((0.0f < T1) - (T1 < 0.0f));
  // y and z: shared rows, Lagrange weights and basis parameters
  const double cy = (double)y - (s * (double)ey + gc[1]), cz = (double)z - (s * (double)ez + gc[2]);
  const int ry = floor_int(cy + 0.5) - 1, rz = floor_int(cz + 0.5) - 1;
  float Fy[3], Fz[3];
  lag3f((float)(cy - ry), Fy);
  lag3f((float)(cz - rz), Fz);
  // row offsets of the stencil rows: node (x, a, b) is cell x + yo[a] + zo[b] (the split
  // path's grids have 32-bit cell indices)
  uint32_t yo[3], zo[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    yo[a] = (uint32_t)wrapi(ry + a, g.ny) * (uint32_t)g.nx;
    zo[a] = (uint32_t)wrapi(rz + a, g.nz) * ((uint32_t)g.nx * (uint32_t)g.ny);
  }
  // y and z basis kinds (basis3u) and factors
  const float cyb = ey ? (float)ey * sq1 : 0.0f, Tyb = ey ? 0.0f : T1;
  const float czb = ez ? (float)ez * sq1 : 0.0f, Tzb = ez ? 0.0f : T1;
  const float wy = ey ? h : mm, wz = ez ? h : mm;
  // x: per target t (ex = t - 1) its window r_t-2 .. r_t+2 and sub-stencil weights
  int r[3];
  float c[3];  // cx - r: sub-stencil k evaluates its quadratic at c + 2 - k
#pragma unroll
  for (int t = 0; t < 3; t++) {
    const double cx = (double)x - (s * (double)(t - 1) + gc[0]);
    r[t] = floor_int(cx + 0.5);
    c[t] = (float)(cx - r[t]);
  }
  const int xlo = min(r[0], min(r[1], r[2])) - 2, xhi = max(r[0], max(r[1], r[2])) + 2;
  // (64-bit difference: for an infinite temperature the windows sit at +-2^30 and the 32-bit
  // difference would overflow)
  if ((int64_t)xhi - (int64_t)xlo > 8) {
    // windows far apart (a gauge with a huge temperature or velocity): target by target, so the
    // loop stays five slices per target whatever the gauge
#pragma unroll 1
    for (int t = 0; t < 3; t++) {
      const int i = kIdx27.v[t][ey + 1][ez + 1];
      const double cx = (double)x - (s * (double)(t - 1) + gc[0]);
      out[t] = interp_weno5f<true>(src, cx, cy, cz, make_tgtf(i, gc), sqTb);
    }
    return;
  }
  // per target and window position j: the middle node's value (smoothness samples) and the
  // slice's y/z sum
  float ub[3][5], yv[3][5];
#pragma unroll
  for (int t = 0; t < 3; t++) {
#pragma unroll
    for (int j = 0; j < 5; j++) ub[t][j] = yv[t][j] = 0.f;
  }
  // contribution of one stencil node (its moments and 27 populations) to the three targets'
  // y/z sums yz (weight wab) and, for the middle node, to their smoothness samples mid
  auto node = [&](float u0x, float u0y, float u0z, float sq0, const float f[27], float wab, bool is_mid, float yz[3],
                  float mid[3]) {
    // basis values without their per-axis factors w: those are constants of the group (y, z)
    // and of the target (x), applied once to the finished window values
    float Ly[3], Lz[3];
    basis3u(((u1y - u0y) + u1yl) * sqTb, sq0, cyb, Tyb, Ly);
    basis3u(((u1z - u0z) + u1zl) * sqTb, sq0, czb, Tzb, Lz);
    float sy[3];
#pragma unroll
    for (int ja = 0; ja < 3; ja++) {
      float acc = 0.f;
#pragma unroll
      for (int jb = 0; jb < 3; jb++) {
        float sz = 0.f;
#pragma unroll
        for (int jc = 0; jc < 3; jc++) sz = __builtin_fmaf(Lz[jc], f[kIdx27.v[ja][jb][jc]], sz);
        acc = __builtin_fmaf(Ly[jb], sz, acc);
      }
      sy[ja] = acc;
    }
    // x basis of the three targets from shared sums: with A_j = Dx - (j - 1) sq0 the moving
    // targets (ex = -1, +1) have g_j = A_j (A_j + sq1), A_j (A_j - sq1) and the resting one
    // A_j^2 - T1, so with P = sum A_j^2 sy_j, Q = sum A_j sy_j, R = sum sy_j the three values
    // are P + sq1 Q, P - T1 R and P - sq1 Q
    const float Dx = ((u1x - u0x) + u1xl) * sqTb;
    const float A0 = Dx + sq0, A1 = Dx, A2 = Dx - sq0;
    const float Q = __builtin_fmaf(A2, sy[2], __builtin_fmaf(A1, sy[1], A0 * sy[0]));
    const float P = __builtin_fmaf(A2, A2 * sy[2], __builtin_fmaf(A1, A1 * sy[1], A0 * (A0 * sy[0])));
    const float R = (sy[0] + sy[1]) + sy[2];
    const float Tt[3] = {__builtin_fmaf(sq1, Q, P), __builtin_fmaf(-T1, R, P), __builtin_fmaf(-sq1, Q, P)};
#pragma unroll
    for (int t = 0; t < 3; t++) {
      yz[t] = __builtin_fmaf(wab, Tt[t], yz[t]);
      if (is_mid) mid[t] = Tt[t];
    }
  };
  // slice xs (its y/z sums yz and middle values mid) enters target t's window at position
  // j = xs - (r_t - 2); the sub-stencil sums are formed once the window is complete
  auto slice = [&](int xs, const float yz[3], const float mid[3]) {
#pragma unroll
    for (int t = 0; t < 3; t++) {
      const int j = xs - (r[t] - 2);
      // (selects with constant indices: the arrays stay in registers)
#pragma unroll
      for (int jj = 0; jj < 5; jj++) {
        const bool at = j == jj;
        ub[t][jj] = at ? mid[t] : ub[t][jj];
        yv[t][jj] = at ? yz[t] : yv[t][jj];
      }
    }
  };
  for (int xs = xlo; xs <= xhi; xs += 2) {
    const int xw = wrapi(xs, g.nx);
    float yz[3] = {0.f, 0.f, 0.f}, mid[3] = {0.f, 0.f, 0.f};
    float yz2[3] = {0.f, 0.f, 0.f}, mid2[3] = {0.f, 0.f, 0.f};
    // slices xs and xs+1 together (the second through the same per-node code)
    const bool two = xs + 1 <= xhi;
    const int xw2 = wrapi(xs + 1, g.nx);
    // one node at a time: its loads are the only ones in flight per lane
#pragma unroll 1
    for (int q = 0; q < 9; q++) {
      const int a = q / 3, b = q - 3 * (q / 3);
      const uint32_t ya = a == 0 ? yo[0] : a == 1 ? yo[1] : yo[2];
      const uint32_t zb = b == 0 ? zo[0] : b == 1 ? zo[1] : zo[2];
      const float fya = a == 0 ? Fy[0] : a == 1 ? Fy[1] : Fy[2];
      const float fzb = b == 0 ? Fz[0] : b == 1 ? Fz[1] : Fz[2];
      const float wab = fya * fzb;
      const uint32_t boff = ((uint32_t)xw + ya + zb) * 4u;
      float f[27];
      load27(src.f, boff, f);
      const Gauge4 n = ldg4(src, boff);
      node(n.ux, n.uy, n.uz, n.sq, f, wab, q == 4, yz, mid);
      if (two) {
        const uint32_t boff2 = ((uint32_t)xw2 + ya + zb) * 4u;
        load27(src.f, boff2, f);
        const Gauge4 n2 = ldg4(src, boff2);
        node(n2.ux, n2.uy, n2.uz, n2.sq, f, wab, q == 4, yz2, mid2);
      }
    }
    slice(xs, yz, mid);
    if (two) slice(xs + 1, yz2, mid2);
  }
  // sub-stencil k (left node r-2+k): quadratic Lagrange weights at c + 2 - k on its nodes
  // m = 0, 1, 2, summed in node order; the window values first take the basis factors
  // w_x(t) w_y w_z left out of the node contributions (the interpolation is linear in them, the
  // WENO weights are not, so they enter before the smoothness indicators)
  const float wyz = wy * wz;
#pragma unroll
  for (int t = 0; t < 3; t++) {
    const float wt = (t != 1 ? h : mm) * wyz;
#pragma unroll
    for (int j = 0; j < 5; j++) {
      ub[t][j] *= wt;
      yv[t][j] *= wt;
    }
    float S[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float sk = c[t] + (float)(2 - k);
      const float w0 = 0.5f * (sk - 1.0f) * (sk - 2.0f), w1 = -sk * (sk - 2.0f), w2 = 0.5f * sk * (sk - 1.0f);
      S[k] = __builtin_fmaf(w2, yv[t][k + 2], __builtin_fmaf(w1, yv[t][k + 1], __builtin_fmaf(w0, yv[t][k], 0.f)));
    }
    out[t] = weno5_combine(c[t], ub[t], S);
  }
}

// WENO5C for the nine targets that share ez (ey, ex = -1, 0, 1): one pass over the union of
// their stencil nodes.  Their departure points share the z rows, and the x windows depend on ex
// only; the y rows of the three ey groups differ by at most a row or two (they coincide wherever
// the thermal speed is below half a cell per step).  So each node of the union (x window union ×
// y row union × 3 z rows) is loaded once for all nine targets instead of once per group, and its
// contraction over the z source components, sz[ja][jb] = sum_jc Lz[jc] f(ja, jb, jc), is formed
// once and shared by the three ey groups; each group whose three rows contain the node then
// contracts over y with its own basis and adds to its three targets as interp_group_weno5c
// does.  The three sub-stencil sums of each target are accumulated slice by slice (a slice's
// y/z sum enters every sub-stencil whose window holds it), so only the five smoothness samples
// of a window are kept.  Each target's value is stored to fout (its population slot of this
// cell).  Windows or rows far apart (absurd gauges) fall back to
// interp_group_weno5c per group, which bounds the loops whatever the gauge.
__device__ void interp_pass_weno5c(const NodeSrcF& src, int x, int y, int z, double s, const double gc[4], int ez,
                                   float sqTb, float* __restrict__ fout, int64_t cell, int64_t P) {
  const Grid& g = src.g;
  const float u1x = (float)gc[0], u1y = (float)gc[1], u1z = (float)gc[2];
  const float u1xl = (float)(gc[0] - (double)u1x), u1yl = (float)(gc[1] - (double)u1y),
              u1zl = (float)(gc[2] - (double)u1z);
  const float T1 = (float)gc[3];
  const float h = 1.0f / (2.0f * T1), mm = -1.0f / T1;
  const float sq1 = sqrtf(fabsf(T1)) * (float)((0.0f < T1) - (T1 < 0.0f));
  // z: the pass's rows, Lagrange weights and basis parameters
  const double cz = (double)z - (s * (double)ez + gc[2]);
  const int rz = floor_int(cz + 0.5) - 1;
  float Fz[3];
  lag3f((float)(cz - rz), Fz);
  const float czb = ez ? (float)ez * sq1 : 0.0f, Tzb = ez ? 0.0f : T1;
  const float wz = ez ? h : mm;
  // y: rows and weights of the three groups ey = e - 1
  int ry[3];
  float Fy[3][3];
  double cyv[3];
#pragma unroll
  for (int e = 0; e < 3; e++) {
    cyv[e] = (double)y - (s * (double)(e - 1) + gc[1]);
    ry[e] = floor_int(cyv[e] + 0.5) - 1;
    lag3f((float)(cyv[e] - ry[e]), Fy[e]);
  }
  // x: per target t (ex = t - 1) its window r_t-2 .. r_t+2 (the same for the three groups)
  int r[3];
  float c[3];
#pragma unroll
  for (int t = 0; t < 3; t++) {
    const double cx = (double)x - (s * (double)(t - 1) + gc[0]);
    r[t] = floor_int(cx + 0.5);
    c[t] = (float)(cx - r[t]);
  }
  const int xlo = min(r[0], min(r[1], r[2])) - 2, xhi = max(r[0], max(r[1], r[2])) + 2;
  const int ylo = min(ry[0], min(ry[1], ry[2])), yhi = max(ry[0], max(ry[1], ry[2])) + 2;
  // (64-bit differences: an infinite temperature puts the windows at +-2^30)
  if ((int64_t)xhi - (int64_t)xlo > 8 || (int64_t)yhi - (int64_t)ylo > 4) {
#pragma unroll 1
    for (int e = 0; e < 3; e++) {
      float f3[3];
      interp_group_weno5c(src, x, y, z, s, gc, e - 1, ez, sqTb, f3);
#pragma unroll
      for (int t = 0; t < 3; t++) fout[split27_fidx(kIdx27.v[t][e][ez + 1], cell, P)] = f3[t];
    }
    return;
  }
  uint32_t zo[3];
#pragma unroll
  for (int b = 0; b < 3; b++) zo[b] = (uint32_t)wrapi(rz + b, g.nz) * ((uint32_t)g.nx * (uint32_t)g.ny);
  // per group e and target t: the three sub-stencil sums, accumulated slice by slice, and the
  // smoothness samples at the five window positions (the middle node of each slice).  The
  // samples live in LDS, one column per lane (the kernels calling this run 64-lane blocks), at
  // the window position the slice has for the target: every position of every target and group
  // is written once (the slices span each target's window, the rows each group's middle row),
  // so nothing needs clearing.
  __shared__ float s_ub[45][64];
  const int lane = (int)threadIdx.x;
  float Sk[3][3][3];
#pragma unroll
  for (int e = 0; e < 3; e++)
#pragma unroll
    for (int t = 0; t < 3; t++)
#pragma unroll
      for (int k = 0; k < 3; k++) Sk[e][t][k] = 0.f;
  // target t's window position of slice xs is j = xs - (r_t - 2); sub-stencil k holds positions
  // k .. k+2, with the quadratic Lagrange weights at sk = c_t + 2 - k.
  //
  // The (slice, row) steps run as one loop whose node loads are one step ahead: node b of the
  // next step (the next row of the slice, or the first row of the next slice) is loaded into
  // the registers of node b of the current step once that node's contributions are added, so a
  // load has the rest of the step (about two nodes of arithmetic) to arrive in.  The loads are
  // unconditional (the step after the last loads three wrapped, unused nodes), so no lane has
  // to keep a node it has used.  The arithmetic and its order are those of a plain slice/row
  // loop.
  float fb[3][27];
  Gauge4 nb[3];
  {
    const uint32_t o0 = (uint32_t)wrapi(xlo, g.nx) + (uint32_t)wrapi(ylo, g.ny) * (uint32_t)g.nx;
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const uint32_t zb = b == 0 ? zo[0] : b == 1 ? zo[1] : zo[2];
      load27(src.f, (o0 + zb) * 4u, fb[b]);
      nb[b] = ldg4(src, (o0 + zb) * 4u);
    }
  }
  float yz[3][3];
#pragma unroll
  for (int e = 0; e < 3; e++)
#pragma unroll
    for (int t = 0; t < 3; t++) yz[e][t] = 0.f;
  int xs = xlo, yr = ylo;
#pragma unroll 1
  while (true) {
    int jt[3];
#pragma unroll
    for (int t = 0; t < 3; t++) jt[t] = xs - (r[t] - 2);
    const bool last_row = yr == yhi;
    const int xn = last_row ? xs + 1 : xs, yn = last_row ? ylo : yr + 1;
    const uint32_t on = (uint32_t)wrapi(xn, g.nx) + (uint32_t)wrapi(yn, g.ny) * (uint32_t)g.nx;
    {
      // the row's place in each group's three rows (outside: the group skips the node)
      int ya[3];
#pragma unroll
      for (int e = 0; e < 3; e++) ya[e] = yr - ry[e];
#pragma unroll
      for (int b = 0; b < 3; b++) {
        const uint32_t zb = b == 0 ? zo[0] : b == 1 ? zo[1] : zo[2];
        const float fzb = b == 0 ? Fz[0] : b == 1 ? Fz[1] : Fz[2];
        const Gauge4 n = nb[b];
        // shared by the three groups: the z basis and contraction, the y and x differences
        float Lz[3];
        basis3u(((u1z - n.uz) + u1zl) * sqTb, n.sq, czb, Tzb, Lz);
        float szz[3][3];
#pragma unroll
        for (int ja = 0; ja < 3; ja++)
#pragma unroll
          for (int jb = 0; jb < 3; jb++) {
            float sz = 0.f;
#pragma unroll
            for (int jc = 0; jc < 3; jc++) sz = __builtin_fmaf(Lz[jc], fb[b][kIdx27.v[ja][jb][jc]], sz);
            szz[ja][jb] = sz;
          }
        // the x basis of the three targets is A_j (A_j - c) or A_j^2 - T1 with A_j = Dx - (j-1) sq0:
        // the x moments R = sum_ja sz, Q = sum_ja A_ja sz, P = sum_ja A_ja^2 sz of the z sums, per
        // source y component jb, are common to every group; each group then contracts them
        // over y and its three targets take P + sq1 Q, P - T1 R and P - sq1 Q
        const float Dy = ((u1y - n.uy) + u1yl) * sqTb;
        const float Dx = ((u1x - n.ux) + u1xl) * sqTb;
        const float A0 = Dx + n.sq, A1 = Dx, A2 = Dx - n.sq;
        const float B0 = A0 * A0, B1 = A1 * A1, B2 = A2 * A2;
        float xr[3], xq[3], xp[3];
#pragma unroll
        for (int jb = 0; jb < 3; jb++) {
          xr[jb] = (szz[0][jb] + szz[1][jb]) + szz[2][jb];
          xq[jb] = __builtin_fmaf(A2, szz[2][jb], __builtin_fmaf(A1, szz[1][jb], A0 * szz[0][jb]));
          xp[jb] = __builtin_fmaf(B2, szz[2][jb], __builtin_fmaf(B1, szz[1][jb], B0 * szz[0][jb]));
        }
        const float Ay0 = Dy + n.sq, Ay1 = Dy, Ay2 = Dy - n.sq;
#pragma unroll
        for (int e = 0; e < 3; e++) {
          const int a = ya[e];
          if (a < 0 || a > 2) continue;
          const int ey = e - 1;
          // basis3u of the group: fma(A, A - c, -T) with (c, T) of a moving or resting y axis
          const float cyb = ey ? (float)ey * sq1 : 0.0f, Tyb = ey ? 0.0f : T1;
          const float Ly0 = __builtin_fmaf(Ay0, Ay0 - cyb, -Tyb), Ly1 = __builtin_fmaf(Ay1, Ay1 - cyb, -Tyb),
                      Ly2 = __builtin_fmaf(Ay2, Ay2 - cyb, -Tyb);
          const float R = __builtin_fmaf(Ly2, xr[2], __builtin_fmaf(Ly1, xr[1], Ly0 * xr[0]));
          const float Q = __builtin_fmaf(Ly2, xq[2], __builtin_fmaf(Ly1, xq[1], Ly0 * xq[0]));
          const float Pp = __builtin_fmaf(Ly2, xp[2], __builtin_fmaf(Ly1, xp[1], Ly0 * xp[0]));
          const float fya = a == 0 ? Fy[e][0] : a == 1 ? Fy[e][1] : Fy[e][2];
          const float wab = fya * fzb;
          const float Tt[3] = {__builtin_fmaf(sq1, Q, Pp), __builtin_fmaf(-T1, R, Pp), __builtin_fmaf(-sq1, Q, Pp)};
#pragma unroll
          for (int t = 0; t < 3; t++) yz[e][t] = __builtin_fmaf(wab, Tt[t], yz[e][t]);
          // the smoothness samples are the middle node's values: the middle z row (b, uniform
          // across the wave) and the group's middle y row
          if (b == 1 && a == 1) {
#pragma unroll
            for (int t = 0; t < 3; t++)
              if (jt[t] >= 0 && jt[t] <= 4) s_ub[(e * 3 + t) * 5 + jt[t]][lane] = Tt[t];
          }
        }
        // node b of this step is used up: load node b of the next step in its place
        load27(src.f, (on + zb) * 4u, fb[b]);
        nb[b] = ldg4(src, (on + zb) * 4u);
      }
    }
    if (last_row) {
      // the slice's y/z sums into the sub-stencil sums whose window holds it
#pragma unroll
      for (int t = 0; t < 3; t++) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const float sk = c[t] + (float)(2 - k);
          const int m = jt[t] - k;  // node of sub-stencil k, 0..2 inside it
          const float wk = m == 0 ? 0.5f * (sk - 1.0f) * (sk - 2.0f) : m == 1 ? -sk * (sk - 2.0f) : 0.5f * sk * (sk - 1.0f);
          const bool in = m >= 0 && m <= 2;
#pragma unroll
          for (int e = 0; e < 3; e++) Sk[e][t][k] = in ? __builtin_fmaf(wk, yz[e][t], Sk[e][t][k]) : Sk[e][t][k];
        }
      }
#pragma unroll
      for (int e = 0; e < 3; e++)
#pragma unroll
        for (int t = 0; t < 3; t++) yz[e][t] = 0.f;
      if (xn > xhi) break;
    }
    xs = xn;
    yr = yn;
  }
  // as in interp_group_weno5c: the basis factors, the sub-stencil sums, the WENO combination
#pragma unroll
  for (int e = 0; e < 3; e++) {
    const float wyz = ((e != 1) ? h : mm) * wz;
#pragma unroll
    for (int t = 0; t < 3; t++) {
      const float wt = (t != 1 ? h : mm) * wyz;
      float ubt[5], S[3];
#pragma unroll
      for (int j = 0; j < 5; j++) ubt[j] = s_ub[(e * 3 + t) * 5 + j][lane] * wt;
#pragma unroll
      for (int k = 0; k < 3; k++) S[k] = Sk[e][t][k] * wt;
      fout[split27_fidx(kIdx27.v[t][e][ez + 1], cell, P)] = weno5_combine(c[t], ubt, S);
    }
  }
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// node gauge record of every cell for the node side of the transfer: (ux, uy, uz) and
// sqrt(|T|)*sign(T) in FP32, 16 bytes per cell
// Node gauge records of a D3Q27 split state, and its non-finite node data.  The reference
// transfer drops every term w*g*f that is NaN or infinite (structs.cuh:542), so a node with
// non-finite data adds only its finite terms to an interpolation instead of turning the whole
// neighbourhood NaN.  With a finite gauge (finite basis values) a term is non-finite exactly
// when its population is, and with a non-finite gauge every term of the node is.  So before the
// step reads the state as nodes, a non-finite population is set to zero, and a node with a
// non-finite gauge gets all-zero populations and a zero gauge record (zero gauge: finite basis
// values, so the node adds nothing).  This rewrites the load state, which the step consumes: its
// cells are updated from their own moments (m, T, unchanged here) and the rewritten populations
// are read only as nodes.  FP32 ranges: a population or gauge beyond FP32 range is already
// infinite in the split state.
__global__ void k_node_gauge(float* __restrict__ f, const float* __restrict__ m, const double* __restrict__ T,
                             float* __restrict__ g4, int64_t n, int64_t plane) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double t = T[c];
  float4 v;
  v.x = m[plane + c];
  v.y = m[2 * plane + c];
  v.z = m[3 * plane + c];
  v.w = (float)(sqrt(fabs(t)) * (double)((0.0 < t) - (t < 0.0)));
  const bool gok = isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w);
  float4 q[7];
  bool ok = gok;
#pragma unroll
  for (int j = 0; j < 7; j++) {
    q[j] = reinterpret_cast<const float4*>(f)[j * plane + c];
    ok = ok && isfinite(q[j].x) && isfinite(q[j].y) && isfinite(q[j].z) && (j == 6 || isfinite(q[j].w));
  }
  if (!ok) {
    auto fix = [gok](float x) { return gok && isfinite(x) ? x : 0.0f; };
#pragma unroll
    for (int j = 0; j < 7; j++) {
      float4 r = make_float4(fix(q[j].x), fix(q[j].y), fix(q[j].z), j == 6 ? q[j].w : fix(q[j].w));
      reinterpret_cast<float4*>(f)[j * plane + c] = r;
    }
    if (!gok) v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  reinterpret_cast<float4*>(g4)[c] = v;
}

// iterations a predicted starting gauge gets to converge before the cell falls back to its
// previous gauge.  Cells that need this many are the ill-conditioned ones (slow contraction at
// contact discontinuities), whose converged state can depend on the starting point.
constexpr int kLearnedBudget = 12;
// SplitArgs::learned_trust: the largest predicted change (in thermal units: du/sqrt(T), dT/T)
// used as a start.  Larger changes sit on shocks and contact layers, where the gauge map can have
// separate fixed points and a start far from the previous gauge may settle on another one than
// the reference reaches.

// Lagrange-4 (propagate<4>, steps.cu:986-1036) for the nine targets that share ez, in one pass
// over the union of their stencil nodes, as interp_pass_weno5c does for WENO5C: the z rows
// (floor(cz)-1 .. +2) are common, the x windows depend on ex only, and the y rows of the three ey
// groups differ by at most a row or two.  Each node of the union is loaded once, its z
// contraction formed once for the three groups; each group whose four rows hold the node
// contracts over y, and each of its targets whose four x slices hold the node adds the node's
// value with its tensor Lagrange weight.  The per-axis basis factors w (1/(2T1) or -1/T1) are
// applied to the finished sums.  MIR: mirror boundaries on the axes of g.mirror (a node beyond
// a mirrored face is the mirror image of the cell inside).  Windows or rows far apart (absurd
// gauges) fall back to interp_lagr4f per target.  Each target's value is stored to fout.
template <bool MIR>
__device__ void interp_pass_lagr4(const NodeSrcF& src, int x, int y, int z, double s, const double gc[4], int ez,
                                  float sqTb, float* __restrict__ fout, int64_t cell, int64_t P) {
  const Grid& g = src.g;
  const float u1x = (float)gc[0], u1y = (float)gc[1], u1z = (float)gc[2];
  const float u1xl = (float)(gc[0] - (double)u1x), u1yl = (float)(gc[1] - (double)u1y),
              u1zl = (float)(gc[2] - (double)u1z);
  const float T1 = (float)gc[3];
  const float h = 1.0f / (2.0f * T1), mm = -1.0f / T1;
  const float sq1 = sqrtf(fabsf(T1)) * (float)((0.0f < T1) - (T1 < 0.0f));
  const double cz = (double)z - (s * (double)ez + gc[2]);
  const int lz = floor_int(cz) - 1;
  float Fz[4];
  lag4f((float)(cz - lz), Fz);
  const float czb = ez ? (float)ez * sq1 : 0.0f, Tzb = ez ? 0.0f : T1;
  const float wz = ez ? h : mm;
  int ly[3], lx[3];
  float Fy[3][4], Fx[3][4];
  double cxv[3], cyv[3];
#pragma unroll
  for (int e = 0; e < 3; e++) {
    cyv[e] = (double)y - (s * (double)(e - 1) + gc[1]);
    ly[e] = floor_int(cyv[e]) - 1;
    lag4f((float)(cyv[e] - ly[e]), Fy[e]);
    cxv[e] = (double)x - (s * (double)(e - 1) + gc[0]);
    lx[e] = floor_int(cxv[e]) - 1;
    lag4f((float)(cxv[e] - lx[e]), Fx[e]);
  }
  const int xlo = min(lx[0], min(lx[1], lx[2])), xhi = max(lx[0], max(lx[1], lx[2])) + 3;
  const int ylo = min(ly[0], min(ly[1], ly[2])), yhi = max(ly[0], max(ly[1], ly[2])) + 3;
  if ((int64_t)xhi - (int64_t)xlo > 6 || (int64_t)yhi - (int64_t)ylo > 6) {
#pragma unroll 1
    for (int q = 0; q < 9; q++) {
      const int e = q / 3, t = q % 3;
      const int i = kIdx27.v[t][e][ez + 1];
      fout[split27_fidx(i, cell, P)] = interp_lagr4f<MIR>(src, cxv[t], cyv[e], cz, make_tgtf(i, gc), sqTb);
    }
    return;
  }
  float acc[3][3];
#pragma unroll
  for (int e = 0; e < 3; e++)
#pragma unroll
    for (int t = 0; t < 3; t++) acc[e][t] = 0.f;
  // byte offset of node (xs, yr, lz + b), its mirror image where it lies beyond a mirrored face
  auto node_off = [&](int xs_, int yr_, int b) -> uint32_t {
    bool rv;
    const int xw = MIR ? mirror_wrap(xs_, g.nx, (g.mirror & 1) != 0, rv) : wrapi(xs_, g.nx);
    const int yw = MIR ? mirror_wrap(yr_, g.ny, (g.mirror & 2) != 0, rv) : wrapi(yr_, g.ny);
    const int zw = MIR ? mirror_wrap(lz + b, g.nz, (g.mirror & 4) != 0, rv) : wrapi(lz + b, g.nz);
    return (uint32_t)cell_index(g, xw, yw, zw) * 4u;
  };
  // The (slice, row) steps run as one loop whose node loads are one step ahead, as in
  // interp_pass_weno5c: node b of the next step is loaded into the registers of node b of the
  // current step once its contributions are added (unconditionally: the step after the last
  // loads four wrapped, unused nodes).  The arithmetic and its order are those of a plain
  // slice/row loop.
  float fb[4][27];
  Gauge4 nb[4];
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const uint32_t boff = node_off(xlo, ylo, b);
    load27(src.f, boff, fb[b]);
    nb[b] = ldg4(src, boff);
  }
  int xs = xlo, yr = ylo;
#pragma unroll 1
  while (true) {
    bool rx = false, ryf = false;
    if (MIR) {
      mirror_wrap(xs, g.nx, (g.mirror & 1) != 0, rx);
      mirror_wrap(yr, g.ny, (g.mirror & 2) != 0, ryf);
    }
    // the slice's Lagrange weight in each target's four x slices (0 outside), the row's in each
    // group's four rows
    float wx[3], wy[3];
#pragma unroll
    for (int t = 0; t < 3; t++) {
      const int m = xs - lx[t];
      wx[t] = m == 0 ? Fx[t][0] : m == 1 ? Fx[t][1] : m == 2 ? Fx[t][2] : m == 3 ? Fx[t][3] : 0.f;
    }
#pragma unroll
    for (int e = 0; e < 3; e++) {
      const int a = yr - ly[e];
      wy[e] = a == 0 ? Fy[e][0] : a == 1 ? Fy[e][1] : a == 2 ? Fy[e][2] : a == 3 ? Fy[e][3] : 0.f;
    }
    const bool last_row = yr == yhi;
    const int xn = last_row ? xs + 1 : xs, yn = last_row ? ylo : yr + 1;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      bool rzf = false;
      if (MIR) mirror_wrap(lz + b, g.nz, (g.mirror & 4) != 0, rzf);
      const float fzb = b == 0 ? Fz[0] : b == 1 ? Fz[1] : b == 2 ? Fz[2] : Fz[3];
      Gauge4 n = nb[b];
      // a node beyond a mirrored face is the mirror image of the cell inside
      // (Cell::mirrX/Y/Z, structs.cuh:552-578): its axis velocity negated, and its populations
      // exchanged with their mirror images, which is the same as reversing that axis's basis
      // values (node_xfer_mirrored)
      if (MIR) {
        n.ux = rx ? -n.ux : n.ux;
        n.uy = ryf ? -n.uy : n.uy;
        n.uz = rzf ? -n.uz : n.uz;
      }
      auto flip = [](bool m, float L[3]) {
        const float a = L[0];
        L[0] = m ? L[2] : L[0];
        L[2] = m ? a : L[2];
      };
      float Lz[3];
      basis3u(((u1z - n.uz) + u1zl) * sqTb, n.sq, czb, Tzb, Lz);
      if (MIR) flip(rzf, Lz);
      float szz[3][3];
#pragma unroll
      for (int ja = 0; ja < 3; ja++)
#pragma unroll
        for (int jb = 0; jb < 3; jb++) {
          float sz = 0.f;
#pragma unroll
          for (int jc = 0; jc < 3; jc++) sz = __builtin_fmaf(Lz[jc], fb[b][kIdx27.v[ja][jb][jc]], sz);
          szz[ja][jb] = sz;
        }
      const float Dy = ((u1y - n.uy) + u1yl) * sqTb;
      const float Dx = ((u1x - n.ux) + u1xl) * sqTb;
      float A0 = Dx + n.sq, A1 = Dx, A2 = Dx - n.sq;
      if (MIR) {
        const float a = A0;
        A0 = rx ? A2 : A0;
        A2 = rx ? a : A2;
      }
      // x moments of the z sums, common to the three groups (as in interp_pass_weno5c)
      const float B0 = A0 * A0, B1 = A1 * A1, B2 = A2 * A2;
      float xr[3], xq[3], xp[3];
#pragma unroll
      for (int jb = 0; jb < 3; jb++) {
        xr[jb] = (szz[0][jb] + szz[1][jb]) + szz[2][jb];
        xq[jb] = __builtin_fmaf(A2, szz[2][jb], __builtin_fmaf(A1, szz[1][jb], A0 * szz[0][jb]));
        xp[jb] = __builtin_fmaf(B2, szz[2][jb], __builtin_fmaf(B1, szz[1][jb], B0 * szz[0][jb]));
      }
#pragma unroll
      for (int e = 0; e < 3; e++) {
        if (wy[e] == 0.f) continue;
        const int ey = e - 1;
        float Ly[3];
        basis3u(Dy, n.sq, ey ? (float)ey * sq1 : 0.0f, ey ? 0.0f : T1, Ly);
        if (MIR) flip(ryf, Ly);
        const float R = __builtin_fmaf(Ly[2], xr[2], __builtin_fmaf(Ly[1], xr[1], Ly[0] * xr[0]));
        const float Q = __builtin_fmaf(Ly[2], xq[2], __builtin_fmaf(Ly[1], xq[1], Ly[0] * xq[0]));
        const float Pp = __builtin_fmaf(Ly[2], xp[2], __builtin_fmaf(Ly[1], xp[1], Ly[0] * xp[0]));
        const float Tt[3] = {__builtin_fmaf(sq1, Q, Pp), __builtin_fmaf(-T1, R, Pp), __builtin_fmaf(-sq1, Q, Pp)};
        const float wyz = wy[e] * fzb;
#pragma unroll
        for (int t = 0; t < 3; t++) acc[e][t] = __builtin_fmaf(wx[t] * wyz, Tt[t], acc[e][t]);
      }
      // node b of this step is used up: load node b of the next step in its place
      const uint32_t boff = node_off(xn, yn, b);
      load27(src.f, boff, fb[b]);
      nb[b] = ldg4(src, boff);
    }
    if (last_row && xn > xhi) break;
    xs = xn;
    yr = yn;
  }
#pragma unroll
  for (int e = 0; e < 3; e++) {
    const float wyz = ((e != 1) ? h : mm) * wz;
#pragma unroll
    for (int t = 0; t < 3; t++)
      fout[split27_fidx(kIdx27.v[t][e][ez + 1], cell, P)] = acc[e][t] * ((t != 1 ? h : mm) * wyz);
  }
}

// Iteration state of one cell, carried from the first pass to the second.
struct SplitIter {
  double gc[4];      // current gauge (u, T)
  double res_prev;   // residual of the previous iteration
  double rho;        // density of the last iterate
  int niter;         // iterations so far, reported
  int base;          // iterations spent on a predicted start that was abandoned
  bool learned;      // started from the predicted gauge and not yet fallen back
  bool contracted;   // a residual has fallen since the current start
};

// previous gauge of the cell (the plain start); an all-zero cell is reset as in steps.cu:360-363
__device__ __forceinline__ void split_prev_gauge(const float* __restrict__ min, const double* __restrict__ Tin,
                                                 int64_t P, int64_t cell, double tlat, double gp[4], double& rho0) {
  rho0 = min[cell];
  gp[0] = min[P + cell]; gp[1] = min[2 * P + cell]; gp[2] = min[3 * P + cell]; gp[3] = Tin[cell];
  if (rho0 == 0 && gp[0] == 0 && gp[1] == 0 && gp[2] == 0 && gp[3] == 0) { rho0 = 1; gp[3] = tlat; }
}

// One fixed-point iteration of one cell: true when it has converged.
template <int INTERP>
__device__ __forceinline__ bool split27_once(const NodeSrcF& src, float* __restrict__ fout, int64_t cell, int x,
                                             int y, int z, const float* __restrict__ min,
                                             const double* __restrict__ Tin, SplitIter& st, const SplitArgs& a) {
  const int64_t P = a.g.plane;
  const float sqTb = (float)a.sqrt_tbase;
  double* gc = st.gc;
  {
    st.niter++;
    const double s = sqrt(gc[3] / a.tlat);
    // moments in the frame of the current gauge: A = sum f, B = sum f xi, C = sum f |xi|^2
    float as = 0.f, ac = 0.f, bxs = 0.f, bxc = 0.f, bys = 0.f, byc = 0.f, bzs = 0.f, bzc = 0.f;
    double cs = 0.0;
    auto add = [&](int ex, int ey, int ez, float fn) {
      fout[split27_fidx(kIdx27.v[ex + 1][ey + 1][ez + 1], cell, P)] = fn;
      kahan_add(as, ac, fn);
      kahan_add(bxs, bxc, (float)ex * fn);
      kahan_add(bys, byc, (float)ey * fn);
      kahan_add(bzs, bzc, (float)ez * fn);
      cs += (double)fn * (double)(ex * ex + ey * ey + ez * ez);
    };
    if constexpr (INTERP == 1 || INTERP == 2 || INTERP == 3) {
      // WENO5C and Lagrange-4: the nine targets of each ez share their node loads and their z
      // contraction (the pass stores the values to fout); the moments are then summed in the
      // order (ey, ez, ex)
#pragma unroll 1
      for (int pz = 0; pz < 3; pz++) {
        if constexpr (INTERP == 2) interp_pass_weno5c(src, x, y, z, s, gc, pz - 1, sqTb, fout, cell, P);
        else interp_pass_lagr4<INTERP == 3>(src, x, y, z, s, gc, pz - 1, sqTb, fout, cell, P);
      }
      for (int gi = 0; gi < 9; gi++) {
        const int ey = gi / 3 - 1, ez = gi % 3 - 1;
#pragma unroll
        for (int ex = -1; ex <= 1; ex++) add(ex, ey, ez, fout[split27_fidx(kIdx27.v[ex + 1][ey + 1][ez + 1], cell, P)]);
      }
    } else {
      for (int i = 0; i < 27; i++) {
        const int ex = kE27[i][0], ey = kE27[i][1], ez = kE27[i][2];
        const double vx = s * (double)ex + gc[0];
        const double vy = s * (double)ey + gc[1];
        const double vz = s * (double)ez + gc[2];
        const TgtF t = make_tgtf(i, gc);
        const double cx = (double)x - vx, cy = (double)y - vy, cz = (double)z - vz;
        // INTERP 3: Lagrange-4 with mirror boundaries
        const float fn = INTERP == 1   ? interp_lagr4f<false>(src, cx, cy, cz, t, sqTb)
                         : INTERP == 3 ? interp_lagr4f<true>(src, cx, cy, cz, t, sqTb)
                                       : interp_weno5f<false>(src, cx, cy, cz, t, sqTb);
        add(ex, ey, ez, fn);
      }
    }
    // FP64 moment solve (Cell::calcMoments, structs.cuh:315-342) restated in the gauge frame:
    // with v = u_g + s xi, sum f v / rho = u_g + s B/rho and
    // (sum f |v|^2/rho - |u|^2)/Dim = s^2 (C - |B|^2/rho) / (rho Dim).  This is the same
    // algebra without the cancellation of |u|^2 against sum f |v|^2 / rho, which at Mach 1
    // and above would amplify the FP32 rounding of the populations several times.
    const double rho = (double)as - (double)ac;
    st.rho = rho;
    const double B[3] = {(double)bxs - (double)bxc, (double)bys - (double)byc, (double)bzs - (double)bzc};
    double m[4];
    if (rho != 0) {
      const double ir = 1.0 / rho;
      for (int k = 0; k < 3; k++) m[k] = gc[k] + s * (B[k] * ir);
      m[3] = s * s * (cs - (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]) * ir) * ir * a.inv_dim;
    } else {  // no division by rho (structs.cuh:329): u = sum f v, T = (sum f |v|^2 - |u|^2)/Dim
      for (int k = 0; k < 3; k++) m[k] = s * B[k];
      m[3] = (2.0 * s * (gc[0] * B[0] + gc[1] * B[1] + gc[2] * B[2]) + s * s * cs - (m[0] * m[0] + m[1] * m[1] + m[2] * m[2])) *
             a.inv_dim;
    }
    if (m[3] < 0) m[3] = gc[3];
    // FP64 gauge residual, scaled by the thermal speed and the temperature
    const double sth = sqrt(fabs(m[3]));
    double res = 0;
    bool conv = true;
    for (int k = 0; k < 4; k++) {
      const double scale = k < 3 ? sth : fabs(m[3]);
      const double d = fabs(m[k] - gc[k]);
      // a NaN difference counts as converged, as in the reference's test (steps.cu:58-75,
      // which fails only on difference >= tolerance), so a cell with non-finite data stops at once
      if (d > a.tol_rel * scale + a.tol_abs) conv = false;
      const double rk = d / (scale + 1e-300);
      if (res == res && !(rk <= res)) res = rk;  // NaN propagates (fmax would drop it)
    }
    // FP32 populations leave a noise floor near the tolerance (interpolation rounding of a few
    // 1e-7 per population): a residual below 4 tol that has stopped falling has reached it.  A
    // residual that still falls, however slowly, keeps iterating down to tol.
    // (A predicted start must first show a falling residual: one that lands in the noise band
    // at once may sit next to a fixed point the previous-gauge start does not reach.)
    if (!conv && st.niter - st.base > 1 && res <= 4.0 * a.tol_rel && res >= st.res_prev &&
        (!st.learned || st.contracted))
      conv = true;
    if (st.niter - st.base > 1 && res < st.res_prev) st.contracted = true;
    // (a NaN iterate from a predicted start is not accepted: the cell falls back)
    if (conv && !(st.learned && res != res)) {
      for (int k = 0; k < 4; k++) gc[k] = m[k];
      return true;
    }
    if (st.learned && (res != res || (st.niter > 1 && !(res < st.res_prev)) || st.niter == kLearnedBudget)) {
      // the predicted start does not contract (or not within the budget): restart from the
      // previous gauge (read again from the step's input, which the step does not write, so
      // no lane carries it through its iterations), so the cell ends where the plain method
      // ends
      double gp[4], rho0;
      split_prev_gauge(min, Tin, P, cell, a.tlat, gp, rho0);
      for (int k = 0; k < 4; k++) gc[k] = gp[k];
      st.learned = false;
      st.contracted = false;
      st.res_prev = 0;
      st.base = st.niter;
      return false;
    }
    st.res_prev = res;
    for (int k = 0; k < 4; k++) gc[k] = m[k];
  }
  return false;
}

// Fixed-point iterations of one cell from state st until it converges (true), its count
// reaches stop_at, or its count since its current start reaches the cap (false).  A cell that
// abandons a predicted start gets the full cap from its previous gauge, so a cell that does
// not converge ends on the same iterate as without the prediction.  The iterate's populations
// are left in fout.
template <int INTERP>
__device__ bool split27_iterate(const NodeSrcF& src, float* __restrict__ fout, int64_t cell, int x, int y, int z,
                                const float* __restrict__ min, const double* __restrict__ Tin, SplitIter& st,
                                int stop_at, const SplitArgs& a) {
  while (st.niter < stop_at && st.niter - st.base < a.max_iter)
    if (split27_once<INTERP>(src, fout, cell, x, y, z, min, Tin, st, a)) return true;
  return false;
}

// BGK in FP32 (steps.cu:557-564) on the iterate left in fout, and the cell's outputs
__device__ __forceinline__ void split27_finish(float* __restrict__ fout, float* __restrict__ mout,
                                               double* __restrict__ Tout, int* __restrict__ niter_out, int64_t cell,
                                               const SplitIter& st, const SplitArgs& a) {
  const int64_t P = a.g.plane;
  const float rf = (float)st.rho, dt = (float)a.dtau;
  for (int i = 0; i < 27; i++) {
    const int64_t o = split27_fidx(i, cell, P);
    float f = fout[o];
    f = __builtin_fmaf(dt, rf * a.w[i] - f, f);
    fout[o] = f;
  }
  mout[cell] = rf;
  mout[P + cell] = (float)st.gc[0];
  mout[2 * P + cell] = (float)st.gc[1];
  mout[3 * P + cell] = (float)st.gc[2];
  Tout[cell] = st.gc[3];
  niter_out[cell] = st.niter;
}

// A predicted start is taken only in cells whose previous step took kLearnedPrevMin ..
// kLearnedPrevMax iterations (step_split.h).  Where the gauge map contracts slowly (cells that
// needed more, on shocks and contact layers) it can have a second fixed point near the first,
// and a start from elsewhere may settle there: on the 384^3 shock-bubble, one step from the same
// FP64 state, ungated predicted starts left 13-189 cells per step 1e-4 to 2e-3 (relative, u and
// T) away from the FP64 step, all in cells whose FP64 step needed 9-10 iterations; with an upper
// bound of 8 previous iterations 76-78 cells were still up to 1.7e-3 off; with 4 or 6 every cell
// stayed within the plain start's 2e-7.  On the 64^3 shock-bubble (sharper fronts per cell) an
// upper bound of 6 left cells 5.8e-5 off at step 12, and 4 kept every cell within 3e-7 through
// step 12.  Below: a cell that took one or two iterations gains
// nothing (a prediction lands within the 1e-7 tolerance of the fixed point only by chance, so
// such a cell takes as many iterations from it; windows from 1, 2 or 3 gave the same counts).

// Starting state of a cell: its previous gauge, or the predicted one where a prediction is
// given, finite, not all zero, has T > 0, lies within the trust radius and the cell's previous
// step took kLearnedPrevMin .. kLearnedPrevMax iterations (niter_prev: the counts of the previous step,
// which this step overwrites cell by cell after reading).
__device__ __forceinline__ void split27_start(const float* __restrict__ min, const double* __restrict__ Tin,
                                              const float* __restrict__ guess, const int* __restrict__ niter_prev,
                                              int64_t cell, SplitIter& st, const SplitArgs& a) {
  const int64_t P = a.g.plane;
  double gp[4], rho0;
  split_prev_gauge(min, Tin, P, cell, a.tlat, gp, rho0);
  for (int k = 0; k < 4; k++) st.gc[k] = gp[k];
  st.res_prev = 0;
  st.rho = rho0;
  st.niter = 0;
  st.base = 0;
  st.learned = false;
  st.contracted = false;
  // (the previous count is tested first, so the prediction is read only where it may be taken)
  if (guess && niter_prev[cell] >= kLearnedPrevMin && niter_prev[cell] <= kLearnedPrevMax) {
    // decode the predicted increment: (du/sqrt(T), dT/T) relative to the previous gauge
    const double sq = sqrt(gp[3]);
    const float p0 = guess[cell], p1 = guess[P + cell], p2 = guess[2 * P + cell], p3 = guess[3 * P + cell];
    const double q[4] = {gp[0] + sq * p0, gp[1] + sq * p1, gp[2] + sq * p2, gp[3] * (1.0 + (double)p3)};
    const float pmax = fmaxf(fmaxf(fabsf(p0), fabsf(p1)), fmaxf(fabsf(p2), fabsf(p3)));
    // (a prediction of no change at all is the previous-gauge start, with its stopping rules)
    if (isfinite(q[0]) && isfinite(q[1]) && isfinite(q[2]) && isfinite(q[3]) && q[3] > 0 && pmax <= a.learned_trust &&
        pmax > 0.0f) {
      for (int k = 0; k < 4; k++) st.gc[k] = q[k];
      st.learned = true;
    }
  }
}

// First pass: every cell iterates up to a.pass1_iters times.  A wave runs as long as its
// slowest cell, and on contact discontinuities and shocks a thin layer of cells needs several
// times the mean, so cells still unconverged after the first pass are appended to a list
// (wave-aggregated: one atomic per wave) and finished by k_step_split27_rest, whose waves
// hold only such cells.  Each cell's arithmetic is the same as in a single pass.
// INTERP: 0 WENO5 as written, 1 Lagrange-4, 2 WENO5C with the targets of an (ey, ez) group
// sharing their node work.
template <int INTERP>
__device__ __forceinline__ void split27_pass1(const float* __restrict__ fin, const float* __restrict__ min,
                                              const double* __restrict__ Tin, const float* __restrict__ g4in,
                                              float* __restrict__ fout, float* __restrict__ mout,
                                              double* __restrict__ Tout, int* __restrict__ niter_out,
                                              const float* __restrict__ guess, const SplitPending& pend,
                                              const SplitArgs& a) {
  const Grid& g = a.g;
  int x, y, z;
  const bool live = tile_cell(g, xcd_block(blockIdx.x, gridDim.x), threadIdx.x, x, y, z);
  bool pending = false;
  SplitIter st;
  const int64_t cell = live ? cell_index(g, x, y, z) : 0;
  if (live) {
    split27_start(min, Tin, guess, niter_out, cell, st, a);
    const NodeSrcF src = make_src(fin, g4in, g);
    const bool conv = split27_iterate<INTERP>(src, fout, cell, x, y, z, min, Tin, st, a.pass1_iters, a);
    pending = !conv && st.niter - st.base < a.max_iter;
    if (!pending) split27_finish(fout, mout, Tout, niter_out, cell, st, a);
  }
  // append the pending cells of this wave to the list
  bool any;
  const int64_t e = pending_slot(pend, pending, any);
  if (pending) {
    pend.cell[e] = (int)cell;
    for (int k = 0; k < 4; k++) pend.g[k * pend.cap + e] = st.gc[k];
    pend.res[e] = st.res_prev;
    pend.meta[e] = st.niter | (st.base << 8) | (st.learned ? 0x10000 : 0) | (st.contracted ? 0x20000 : 0);
  }
}

// The D3Q27 split kernels are held to at least two waves per SIMD (256 registers): the WENO5C
// and Lagrange-4 passes, with their node loads one step ahead, would otherwise take a few more
// registers and drop to one wave, which leaves the load latency unhidden; held to two, they keep
// a few rarely used values in scratch instead.
#define POND_SPLIT27_KERNEL __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
template <int INTERP>
POND_SPLIT27_KERNEL void k_step_split27(const float* __restrict__ fin, const float* __restrict__ min,
                                                     const double* __restrict__ Tin, const float* __restrict__ g4in,
                                                     float* __restrict__ fout, float* __restrict__ mout,
                                                     double* __restrict__ Tout, int* __restrict__ niter_out,
                                                     const float* __restrict__ guess, SplitPending pend,
                                                     const SplitArgs a) {
  split27_pass1<INTERP>(fin, min, Tin, g4in, fout, mout, Tout, niter_out, guess, pend, a);
}


// Persistent refill waves (below) are built but not used: on the 384^3 shock-bubble they took
// the WENO5C step from 157.8 to 155.2 ms, but the 256^3 vortex from 66.2 to 78.7 ms, the 384^3
// Taylor-Green from 179.3 to 208.6 ms and the Lagrange-4 shock-bubble from 224 to 294 ms (a
// wave handing over from one row to the next spans two distant rows, and the fetches from
// beyond L2 double), so every kernel keeps one chunk per workgroup.
constexpr bool kRefillPersistent[4] = {false, false, false, false};

// Single pass with lane refill.  Cells are dealt in chunks of `chunk` consecutive 64-cell tiles
// (a whole x row by default).  Whenever lanes of a wave finish their cells, they take the next
// unstarted cells of its chunk in order (one ballot per iteration), so a lane is idle only when
// no unstarted cell is left.  On the shock-bubble the mean gauge iterations per cell are about
// 1.3 while the slowest cell of a 64-cell tile needs 3.5-5, so lanes of the one-tile-per-wave
// step sit idle most of the time.
//
// Each workgroup owns one chunk, in the XCD-contiguous order of xcd_block.  Where
// kRefillPersistent is set (for no interpolator at present, see there) the waves are persistent
// instead: the grid is what the GPU holds at once, and a wave whose chunk has no unstarted cell
// left claims the next chunk while its lanes still finish the old one, so the tail of a chunk
// (the lanes of its last, slowest cells) is filled from the next one.  With one chunk per wave,
// 30-40 % of the lane-iterations of the 384^3 shock-bubble step are such tails.
// The chunks are split into eight contiguous slabs, one per XCD (workgroups are dealt to the XCDs
// in turn, so a wave's XCD is its block index mod 8), and the waves of XCD k claim the chunks of
// slab k in order from the counter claim[k] (zeroed before the launch), so the waves resident on
// one XCD work on neighbouring rows and share their stencil halo in its L2.  Each cell's
// arithmetic is the same as in k_step_split27 (bitwise the same results, whichever wave runs
// it); the active cells of a wave stay within a window of one or two chunks, so node loads
// remain close together.
template <int INTERP>
__device__ __forceinline__ void split27_refill(const float* __restrict__ fin, const float* __restrict__ min,
                                               const double* __restrict__ Tin, const float* __restrict__ g4in,
                                               float* __restrict__ fout, float* __restrict__ mout,
                                               double* __restrict__ Tout, int* __restrict__ niter_out,
                                               const float* __restrict__ guess, const SplitArgs& a, int chunk,
                                               int* __restrict__ claim) {
  const Grid& g = a.g;
  const int lane = threadIdx.x;
  const int64_t ntiles = g.ntiles;
  const int64_t nchunks = (ntiles + chunk - 1) / chunk;
  constexpr int kXcd = 8;
  const int xk = (int)(blockIdx.x % kXcd);
  const int64_t q = nchunks / kXcd, rq = nchunks % kXcd;
  const int64_t c_lo = xk * q + (xk < rq ? xk : rq), c_hi = c_lo + q + (xk < rq ? 1 : 0);
  const NodeSrcF src = make_src(fin, g4in, g);
  int64_t t0 = 0, total = 0;  // first tile and cell slots of the current chunk
  int64_t next = 0;           // first unstarted slot (the same in every lane)
  bool more = true;           // the slab may still hold unclaimed chunks
  if constexpr (!kRefillPersistent[INTERP]) {
    // one chunk per workgroup, in the XCD-contiguous order of xcd_block
    t0 = xcd_block(blockIdx.x, gridDim.x) * chunk;
    total = (ntiles - t0 < chunk ? ntiles - t0 : chunk) * 64;
    more = false;
  }
  bool active = false;
  int64_t cell = 0;
  int x = 0, y = 0, z = 0;
  SplitIter st;
  while (true) {
    const uint64_t want = __ballot(!active);
    if (kRefillPersistent[INTERP] && want != 0 && next >= total && more) {
      // lanes are free and the chunk has no unstarted cell: claim the slab's next chunk
      int idx = 0;
      if (lane == 0) idx = atomicAdd(&claim[xk], 1);
      idx = __shfl(idx, 0);
      const int64_t c = c_lo + idx;
      if (c < c_hi) {
        t0 = c * chunk;
        total = (ntiles - t0 < chunk ? ntiles - t0 : chunk) * 64;
        next = 0;
      } else {
        more = false;
      }
    }
    // lanes without a cell take the next slots, in lane order; a slot outside the grid (ragged
    // last tile) is used up without a cell
    const int rank = __popcll(want & ((1ull << lane) - 1ull));
    if (!active) {
      const int64_t slot = next + rank;
      if (slot < total && tile_cell(g, t0 + slot / 64, (int)(slot % 64), x, y, z)) {
        cell = cell_index(g, x, y, z);
        split27_start(min, Tin, guess, niter_out, cell, st, a);
        active = true;
      }
    }
    next += __popcll(want);
    if (__ballot(active) == 0 && next >= total && !more) break;
    if (active) {
      // the loop condition of split27_iterate, one iteration per trip
      bool done = st.niter - st.base >= a.max_iter;
      if (!done)
        done = split27_once<INTERP>(src, fout, cell, x, y, z, min, Tin, st, a) || st.niter - st.base >= a.max_iter;
      if (done) {
        split27_finish(fout, mout, Tout, niter_out, cell, st, a);
        active = false;
      }
    }
  }
}

template <int INTERP>
POND_SPLIT27_KERNEL void k_step_split27_refill(
    const float* __restrict__ fin, const float* __restrict__ min, const double* __restrict__ Tin,
    const float* __restrict__ g4in, float* __restrict__ fout, float* __restrict__ mout, double* __restrict__ Tout,
    int* __restrict__ niter_out, const float* __restrict__ guess, const SplitArgs a, int chunk, int* claim) {
  split27_refill<INTERP>(fin, min, Tin, g4in, fout, mout, Tout, niter_out, guess, a, chunk, claim);
}


// Second pass: a fixed grid of waves walks the pending list; every cell resumes its iteration
// where the first pass left it.
template <int INTERP>
__device__ __forceinline__ void split27_rest(const float* __restrict__ fin, const float* __restrict__ min,
                                             const double* __restrict__ Tin, const float* __restrict__ g4in,
                                             float* __restrict__ fout, float* __restrict__ mout,
                                             double* __restrict__ Tout, int* __restrict__ niter_out,
                                             const SplitPending& pend, const SplitArgs& a) {
  const Grid& g = a.g;
  const int count = *pend.count;
  const NodeSrcF src = make_src(fin, g4in, g);
  for (int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x; e - threadIdx.x < count; e += (int64_t)gridDim.x * 64) {
    if (e >= count) continue;
    const int64_t cell = pend.cell[e];
    const int x = (int)(cell % g.nx);
    const int y = (int)((cell / g.nx) % g.ny);
    const int z = (int)(cell / ((int64_t)g.nx * g.ny));
    double gp[4], rho0;
    split_prev_gauge(min, Tin, g.plane, cell, a.tlat, gp, rho0);
    SplitIter st;
    for (int k = 0; k < 4; k++) st.gc[k] = pend.g[k * pend.cap + e];
    st.res_prev = pend.res[e];
    st.rho = rho0;
    const int meta = pend.meta[e];
    st.niter = meta & 0xff;
    st.base = (meta >> 8) & 0xff;
    st.learned = (meta & 0x10000) != 0;
    st.contracted = (meta & 0x20000) != 0;
    split27_iterate<INTERP>(src, fout, cell, x, y, z, min, Tin, st, 0x7fffffff, a);
    split27_finish(fout, mout, Tout, niter_out, cell, st, a);
  }
}

template <int INTERP>
POND_SPLIT27_KERNEL void k_step_split27_rest(const float* __restrict__ fin, const float* __restrict__ min,
                                                          const double* __restrict__ Tin, const float* __restrict__ g4in,
                                                          float* __restrict__ fout, float* __restrict__ mout,
                                                          double* __restrict__ Tout, int* __restrict__ niter_out,
                                                          SplitPending pend, const SplitArgs a) {
  split27_rest<INTERP>(fin, min, Tin, g4in, fout, mout, Tout, niter_out, pend, a);
}


__global__ void k_zero_count(int* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0;
}

__global__ void k_zero_ints(int* p, int n) {
  if (blockIdx.x == 0 && (int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// Workgroups of the persistent refill kernel: 16 64-lane blocks per CU, twice what two waves per
// SIMD hold (blocks beyond those resident start when others have ended, find their slab used up
// and end), a multiple of 8 (one share per XCD slab; at least 8 so that every slab has waves),
// never more than the chunks, and one per chunk when there are fewer than 8.
static unsigned refill_blocks(int64_t nchunks) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    hipDeviceProp_t prop;
    cus = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
           prop.multiProcessorCount > 0)
              ? prop.multiProcessorCount
              : 256;
  }
  if (nchunks < 8) return (unsigned)nchunks;
  const int64_t resident = (int64_t)16 * cus;
  int64_t b = resident < nchunks ? resident : nchunks;
  b -= b % 8;
  return (unsigned)(b < 8 ? 8 : b);
}

// ---------------------------------------------------------------------------------------------
// D3Q125 split step: one 128-lane block per cell, lane = target population (PonDstepD3Q125,
// steps.cu:151-250), the fixed 3x3x3 neighbourhood staged in LDS as FP32.
//
// Transfer of node n to target k (structs.cuh:396-524): per axis the target's quartic Lagrange
// basis on the target nodes {0, +-ec1 S1, +-ec2 S1} evaluated at the five source velocities
// (u0 - u1) + a_j S0, so f_new(k) = sum_jx Lx sum_jy Ly sum_jz Lz f(jx,jy,jz): 155 FMAs instead
// of 125 four-factor products.  Interpolation on the fixed stencil coord-1..coord+1 with the
// quadratic Lagrange weights, also sum-factorised.  Moments in the frame of the current gauge
// (A = sum f, B = sum f xi, C = sum f |xi|^2) are block reductions in FP64; the moment solve,
// the residual and the stopping rules are those of the D3Q27 split step.
constexpr double kA5d[5] = {0.0, kEc1, -kEc1, kEc2, -kEc2};

// target node m of one axis, {0, +e1, -e1, +e2, -e2}[m], as selects (m varies per lane)
__device__ __forceinline__ double node5(int m, double e1, double e2) {
  return m == 0 ? 0.0 : m == 1 ? e1 : m == 2 ? -e1 : m == 3 ? e2 : -e2;
}

__device__ __forceinline__ double block_sum128(double v, double* red, int slot) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[2 * slot + (threadIdx.x >> 6)] = v;
  __syncthreads();
  const double r = red[2 * slot] + red[2 * slot + 1];
  return r;
}

__global__ __launch_bounds__(128) void k_step_split125(const float* __restrict__ fin, const float* __restrict__ min,
                                                       const double* __restrict__ Tin, float* __restrict__ fout,
                                                       float* __restrict__ mout, double* __restrict__ Tout,
                                                       int* __restrict__ niter_out, const float* __restrict__ guess,
                                                       const SplitArgs a) {
  constexpr int Q = 125;
  __shared__ float sf[27][Q];   // neighbourhood populations
  __shared__ float su[27][3];   // neighbourhood gauge velocities
  __shared__ float ss0[27];     // neighbourhood sqrt(|T|)*sign(T)
  __shared__ double red[10];    // per-wave partial sums of the five moments
  __shared__ float sL[2][3][5][5];  // a node's basis values L[d][target index][source index], by node parity
  __shared__ float sTz[Q], sTy[Q];  // a node's z sums and y sums
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
    // non-finite data as in k_node_gauge (the reference drops non-finite transfer terms,
    // structs.cuh:516): a non-finite population stages as zero, a node with a non-finite gauge
    // as all zeros with a zero gauge
    const float ux = rx ? -min[P + c] : min[P + c], uy = ry ? -min[2 * P + c] : min[2 * P + c],
                uz = rz ? -min[3 * P + c] : min[3 * P + c];
    const double t = Tin[c];
    const float s0 = (float)(sqrt(fabs(t)) * (double)((0.0 < t) - (t < 0.0)));
    const bool gok = isfinite(ux) && isfinite(uy) && isfinite(uz) && isfinite(s0);
    if (lane_ok) {
      const float v = fin[c * kSplit125Floats + kl];
      sf[n][k] = gok && isfinite(v) ? v : 0.0f;
    }
    if (k < 3) su[n][k] = gok ? (k == 0 ? ux : k == 1 ? uy : uz) : 0.0f;
    if (k == 3) ss0[n] = gok ? s0 : 0.0f;
  }
  double rho0 = min[cell];
  double gp[4] = {min[P + cell], min[2 * P + cell], min[3 * P + cell], Tin[cell]};
  if (rho0 == 0 && gp[0] == 0 && gp[1] == 0 && gp[2] == 0 && gp[3] == 0) { rho0 = 1; gp[3] = a.tlat; }
  double gc[4] = {gp[0], gp[1], gp[2], gp[3]};
  bool learned = false;
  if (guess) {
    const double st = sqrt(gp[3]);
    const double q[4] = {gp[0] + st * guess[cell], gp[1] + st * guess[P + cell], gp[2] + st * guess[2 * P + cell],
                         gp[3] * (1.0 + (double)guess[3 * P + cell])};
    const float pmax = fmaxf(fmaxf(fabsf(guess[cell]), fabsf(guess[P + cell])),
                             fmaxf(fabsf(guess[2 * P + cell]), fabsf(guess[3 * P + cell])));
    if (isfinite(q[0]) && isfinite(q[1]) && isfinite(q[2]) && isfinite(q[3]) && q[3] > 0 && pmax <= a.learned_trust &&
        pmax > 0.0f && niter_out[cell] >= kLearnedPrevMin && niter_out[cell] <= kLearnedPrevMax) {
      for (int d = 0; d < 4; d++) gc[d] = q[d];
      learned = true;
    }
  }
  __syncthreads();
  const int kk = lane_ok ? k : 0;
  const int X[3] = {kk / 25, (kk / 5) % 5, kk % 5};
  const double xa[3] = {node5(X[0], kEc1, kEc2), node5(X[1], kEc1, kEc2), node5(X[2], kEc1, kEc2)};
  const double xi2 = xa[0] * xa[0] + xa[1] * xa[1] + xa[2] * xa[2];
  int niter = 0, base = 0;  // base: iterations spent on an abandoned predicted start
  bool contracted = false;   // a residual has fallen since the current start
  double res_prev = 0, rho = rho0;
  float fn = 0.f;
  while (niter - base < a.max_iter) {
    niter++;
    const double s = sqrt(gc[3] / a.tlat);
    // target basis constants (structs.cuh:397-416)
    const double S1 = sqrt(fabs(gc[3])) * (double)((0.0 < gc[3]) - (gc[3] < 0.0));
    const double e1 = kEc1 * S1, e2 = kEc2 * S1;
    // departure point on the fixed stencil coord-1..coord+1 (steps.cu:163,207-208)
    float Fx[3], Fy[3], Fz[3];
    {
      const double vx = s * xa[0] + gc[0], vy = s * xa[1] + gc[1], vz = s * xa[2] + gc[2];
      lag3f((float)(((double)x - vx) - (double)(x - 1)), Fx);
      lag3f((float)(((double)y - vy) - (double)(y - 1)), Fy);
      lag3f((float)(((double)z - vz) - (double)(z - 1)), Fz);
    }
    const float u1[3] = {(float)gc[0], (float)gc[1], (float)gc[2]};
    const float u1l[3] = {(float)(gc[0] - (double)u1[0]), (float)(gc[1] - (double)u1[1]), (float)(gc[2] - (double)u1[2])};
    // Transfer of each node to every target, contracted one axis at a time across the block:
    // the basis values L[d][kd][j] of the node (75, one per lane), the z sums tz(jx, jy, kz), the
    // y sums ty(jx, ky, kz), then each target's x sum, every value formed once and read by the
    // targets that share it.  A target's value is the same sequence of operations as contracting
    // its own 125 terms (z inner, then y, then x).
    const int tb = lane_ok ? k : 0;
    const int bd = tb / 25, bk = (tb / 5) % 5, bj = tb % 5;  // basis entry of lanes 0..74
    float bw = 0.f, bn[4] = {0.f, 0.f, 0.f, 0.f};
    {
      const double xn = node5(bk, e1, e2);
      double den = 1.0;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const double v = node5(q < bk ? q : q + 1, e1, e2);
        den *= xn - v;
        bn[q] = (float)v;
      }
      bw = (float)(1.0 / den);
    }
    float acc = 0.f, accy = 0.f, accz = 0.f;
    // basis values of node n (lanes 0..74) into sL[n & 1]
    auto basis = [&](int n) {
      if (k < 75) {
        const float du = (su[n][bd] - u1[bd]) - u1l[bd];
        const float bv = du + (float)kA5d[bj] * ss0[n];
        sL[n & 1][bd][bk][bj] = bw * ((bv - bn[0]) * (bv - bn[1]) * (bv - bn[2]) * (bv - bn[3]));
      }
    };
    // z sums of node n: tz(jx, jy, kz) at index 25 jx + 5 jy + kz
    auto zsums = [&](int n) {
      if (lane_ok) {
        float tz = 0.f;
#pragma unroll
        for (int jz = 0; jz < 5; jz++) tz = __builtin_fmaf(sL[n & 1][2][bj][jz], sf[n][25 * bd + 5 * bk + jz], tz);
        sTz[k] = tz;
      }
    };
    // The three phases of consecutive nodes overlap, two barriers per node: the y sums of node
    // n with the basis values of node n + 1 (the other sL slot), then the x sums of node n with
    // the z sums of node n + 1 (sTz, whose node-n values the y sums have read).  Each value is
    // the same sequence of operations as with the phases of one node at a time.
    basis(0);
    __syncthreads();
    zsums(0);
    __syncthreads();
#pragma unroll 1
    for (int n = 0; n < 27; n++) {
      const int pb = n & 1;
      if (lane_ok) {  // ty(jx, ky, kz) at index 25 jx + 5 ky + kz
        float ty = 0.f;
#pragma unroll
        for (int jy = 0; jy < 5; jy++) ty = __builtin_fmaf(sL[pb][1][bk][jy], sTz[25 * bd + 5 * jy + bj], ty);
        sTy[k] = ty;
      }
      if (n + 1 < 27) basis(n + 1);
      __syncthreads();
      float tx = 0.f;
      if (lane_ok) {
#pragma unroll
        for (int jx = 0; jx < 5; jx++) tx = __builtin_fmaf(sL[pb][0][X[0]][jx], sTy[25 * jx + 5 * X[1] + X[2]], tx);
      }
      if (n + 1 < 27) zsums(n + 1);
      // (the stencil sums of the fixed 3x3x3 neighbourhood: z inner, then y, then x)
      const int ia = n / 9, ib = (n / 3) % 3, ic = n % 3;
      accz = __builtin_fmaf(Fz[ic], tx, accz);
      if (ic == 2) {
        accy = __builtin_fmaf(Fy[ib], accz, accy);
        accz = 0.f;
        if (ib == 2) {
          acc = __builtin_fmaf(Fx[ia], accy, acc);
          accy = 0.f;
        }
      }
      __syncthreads();
    }
    fn = lane_ok ? acc : 0.f;
    // moments in the frame of the current gauge, FP64 block sums
    const double fd = (double)fn;
    const double A = block_sum128(fd, red, 0);
    const double Bx = block_sum128(fd * xa[0], red, 1);
    const double By = block_sum128(fd * xa[1], red, 2);
    const double Bz = block_sum128(fd * xa[2], red, 3);
    const double C = block_sum128(fd * xi2, red, 4);
    __syncthreads();  // red is rewritten next iteration
    rho = A;
    const double B[3] = {Bx, By, Bz};
    double m[4];
    if (rho != 0) {
      const double ir = 1.0 / rho;
      for (int d = 0; d < 3; d++) m[d] = gc[d] + s * (B[d] * ir);
      m[3] = s * s * (C - (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]) * ir) * ir * a.inv_dim;
    } else {
      for (int d = 0; d < 3; d++) m[d] = s * B[d];
      m[3] = (2.0 * s * (gc[0] * B[0] + gc[1] * B[1] + gc[2] * B[2]) + s * s * C - (m[0] * m[0] + m[1] * m[1] + m[2] * m[2])) *
             a.inv_dim;
    }
    if (m[3] < 0) m[3] = gc[3];
    const double st = sqrt(fabs(m[3]));
    double res = 0;
    bool conv = true;
    for (int d = 0; d < 4; d++) {
      const double scale = d < 3 ? st : fabs(m[3]);
      const double dd = fabs(m[d] - gc[d]);
      if (dd > a.tol_rel * scale + a.tol_abs) conv = false;  // NaN: converged, as in the D3Q27 step
      const double rk = dd / (scale + 1e-300);
      if (res == res && !(rk <= res)) res = rk;  // NaN propagates (fmax would drop it)
    }
    if (!conv && niter - base > 1 && res <= 4.0 * a.tol_rel && res >= res_prev && (!learned || contracted)) conv = true;
    if (niter - base > 1 && res < res_prev) contracted = true;
    // (unlike the D3Q27 step, no fallback from a predicted start on a NaN residual: with that
    // rule the 256^3 D3Q125 turbulence took 4.58 mean iterations instead of 2.35, although a
    // probe counted no NaN residual there; the cause is not found)
    if (conv) {
      for (int d = 0; d < 4; d++) gc[d] = m[d];
      break;
    }
    if (learned && ((niter > 1 && !(res < res_prev)) || niter == kLearnedBudget)) {
      for (int d = 0; d < 4; d++) gc[d] = gp[d];
      learned = false;
      contracted = false;
      res_prev = 0;
      base = niter;
      continue;
    }
    res_prev = res;
    for (int d = 0; d < 4; d++) gc[d] = m[d];
  }
  const float rf = (float)rho;
  if (lane_ok) fout[cell * kSplit125Floats + k] = __builtin_fmaf((float)a.dtau, rf * a.w[k] - fn, fn);
  if (k == 0) {
    mout[cell] = rf;
    mout[P + cell] = (float)gc[0];
    mout[2 * P + cell] = (float)gc[1];
    mout[3 * P + cell] = (float)gc[2];
    Tout[cell] = gc[3];
    niter_out[cell] = niter;
  }
}

hipError_t launch_node_gauge(float* f, const float* m, const double* T, float* g4, int64_t n, int64_t plane,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_node_gauge, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, f, m, T, g4, n, plane);
  return hipGetLastError();
}

hipError_t launch_step_split(int lattice, int interp, const SplitPtrs& in, const SplitPtrs& out, const float* g4,
                             int* niter, const float* guess, const SplitPending& pend, const SplitArgs& a,
                             hipStream_t st) {
  if (lattice == kD3Q125) {
    hipLaunchKernelGGL(k_step_split125, dim3((unsigned)a.g.n), dim3(128), 0, st, in.f, in.m, in.T, out.f, out.m, out.T,
                       niter, guess, a);
    return hipGetLastError();
  }
  if (lattice != kD3Q27) return hipErrorInvalidValue;
  const unsigned blocks = (unsigned)(a.g.ntiles);
  hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, st, pend.count);
  // second-pass grid: enough waves to fill the chip several times, never more than the cells
  const int64_t w2 = (a.g.n + 63) / 64;
  const unsigned blocks2 = (unsigned)(w2 < 16384 ? w2 : 16384);
#define POND_LAUNCH_S(K, I)                                                                                       \
  do {                                                                                                            \
    hipLaunchKernelGGL((K), dim3(blocks), dim3(64), 0, st, in.f, in.m, in.T, g4, out.f, out.m, out.T, niter, guess, \
                       pend, a);                                                                                  \
    hipLaunchKernelGGL((k_step_split27_rest<I>), dim3(blocks2), dim3(64), 0, st, in.f, in.m, in.T, g4, out.f,     \
                       out.m, out.T, niter, pend, a);                                                             \
  } while (0)
  if (interp != 0 && a.refill_chunk > 0 && a.pass1_iters >= a.max_iter + kLearnedBudget) {
    const int64_t nchunks = ((int64_t)blocks + a.refill_chunk - 1) / a.refill_chunk;
    // the slabs' chunk counters, in the pending list's header (its count is the first int)
    int* claim = pend.count + 16;
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, claim, 8);
#define POND_LAUNCH_R(I)                                                                                              \
  hipLaunchKernelGGL(k_step_split27_refill<I>, dim3(kRefillPersistent[I] ? refill_blocks(nchunks) : (unsigned)nchunks), \
                     dim3(64), 0, st, in.f, in.m, in.T, g4,                                                              \
                     out.f, out.m, out.T, niter, guess, a, a.refill_chunk, claim)
    if (interp == 2) POND_LAUNCH_R(2);
    else if (a.g.mirror) POND_LAUNCH_R(3);
    else POND_LAUNCH_R(1);
#undef POND_LAUNCH_R
    return hipGetLastError();
  }
  if (interp == 1 && a.g.mirror) POND_LAUNCH_S(k_step_split27<3>, 3);
  else if (interp == 1) POND_LAUNCH_S(k_step_split27<1>, 1);
  else if (interp == 2) POND_LAUNCH_S(k_step_split27<2>, 2);
  else POND_LAUNCH_S(k_step_split27<0>, 0);
#undef POND_LAUNCH_S
  return hipGetLastError();
}

}  // namespace pond
