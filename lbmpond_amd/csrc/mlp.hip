// mlp.hip -- learned gauge initializer: a 135-64-64-4 ReLU network in BF16 with FP32
// accumulation on the CDNA matrix cores (v_mfma_f32_16x16x16_bf16), no reference counterpart.
//
// Input of a cell: the previous step's (rho, u, T) on its 3x3x3 neighbourhood, as
// (rho_n/rho_c - 1, (u_n - u_c)/sqrt(T_c), T_n/T_c - 1) for the 27 neighbours (x outer, z inner),
// each feature type divided by a scale fitted on the training set.  Output: the converged
// gauge of the step relative to the previous one, (du/sqrt(T_c), dT/T_c), times fitted scales.
//
// MFMA operand layout of v_mfma_f32_16x16x16_bf16 (lane l, element v = 0..3):
//   A[i][k]: i = l%16, k = 4*(l/16)+v      B[k][j]: j = l%16, k = 4*(l/16)+v
//   D[i][j]: j = l%16, i = 4*(l/16)+v
// Inference computes H^T = W^T X^T per 16-cell tile (cells on the j axis): the D tile of one
// layer holds, per lane, 4 consecutive hidden units of one cell, which is exactly the B operand
// of the next layer, so activations never leave registers.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "mlp.h"
#include "pond_device.h"

namespace pond {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

// round to nearest even in one instruction (v_cvt_pk_bf16_f32); the same as f2bf for every
// finite value
__device__ __forceinline__ short f2bf_hw(float f) { return (short)__builtin_bit_cast(unsigned short, (__bf16)f); }

// four values as one bf16 MFMA operand, two conversions per instruction (v_cvt_pk_bf16_f32)
typedef float f2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf2_t __attribute__((ext_vector_type(2)));
struct U32x2 {
  unsigned int lo, hi;
};
__device__ __forceinline__ short4_t pack_bf16x4(float a, float b, float c, float d) {
  const bf2_t lo = __builtin_convertvector((f2_t){a, b}, bf2_t), hi = __builtin_convertvector((f2_t){c, d}, bf2_t);
  return __builtin_bit_cast(short4_t, (U32x2{__builtin_bit_cast(unsigned int, lo), __builtin_bit_cast(unsigned int, hi)}));
}

__device__ __forceinline__ float4_t mfma16(short4_t a, short4_t b, float4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------------------ feature gather

// raw feature f (0..134) of cell (x,y,z) from a split-layout state
__device__ __forceinline__ float raw_feature(const MlpState& s, int x, int y, int z, int f, float rc, float uxc,
                                            float uyc, float uzc, double Tc, float isT) {
  const int nb = f / 5, type = f - 5 * (f / 5);
  const int dx = nb / 9 - 1, dy = (nb / 3) % 3 - 1, dz = nb % 3 - 1;
  const Grid& g = s.g;
  const int64_t c = cell_index(g, wrapi(x + dx, g.nx), wrapi(y + dy, g.ny), wrapi(z + dz, g.nz));
  const int64_t P = g.plane;
  switch (type) {
    case 0: return s.m[c] / rc - 1.0f;
    case 1: return (s.m[P + c] - uxc) * isT;
    case 2: return (s.m[2 * P + c] - uyc) * isT;
    case 3: return (s.m[3 * P + c] - uzc) * isT;
    default: return (float)(s.T[c] / Tc - 1.0);
  }
}

// ------------------------------------------------------------------------ fused inference

// cells of the neighbourhood of a 64-cell row segment: 66 x 3 x 3
constexpr int kStage = 66 * 9;

// orders this wave's LDS writes before its following reads (waves stage independently)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Waves loop over 16-cell tiles.  Features are taken from the state (mode 0) or from a
// normalized feature matrix (mode 1, for checking the network arithmetic).
// Input feature held in slot k = 16*st + 4*q + e of the first layer's operands (st: K step,
// q: lane group, e: element), or -1 for the 9 unused slots.  The 135 features (neighbour nb,
// type: rho, ux, uy, uz, T) are laid out so that the type depends on (st, e) only: every lane
// of a wave builds the same kind of feature at each unrolled (st, e), for different
// neighbours, with no divergent branch.  The order of the sum over features inside the MFMA is
// all this changes.
__device__ __forceinline__ int slot_feature(int st, int q, int e) {
  const int j = st * 4 + e;  // 0..35; seven (st, e) pairs per type
  const int type = j / 7, nb = (j % 7) * 4 + q;
  return (j < 35 && nb < 27) ? nb * 5 + type : -1;
}

// Two waves per SIMD (256 VGPRs, a few spilled): the weights held in registers take most of
// them, and with one wave per SIMD nothing hides the LDS and MFMA latencies.  Measured on the
// 384^3 shock-bubble: 14.9 ms per step at one wave per SIMD, 8.4 ms at two.
template <int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_mlp_infer(const MlpWeightsBF w, MlpState s, const float* __restrict__ Xin,
                                                  int64_t ncell, float* __restrict__ out, int64_t out_plane) {
  const int l = threadIdx.x & 63;
  const int q = l >> 4, col = l & 15;
  // weights into registers: A operands W^T[i][k] = W[k][i]
  short4_t a1[4][9], a2[4][4], a3[4];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int st = 0; st < 9; st++) {
      short4_t v;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int f = slot_feature(st, q, e);
        v[e] = f >= 0 ? (short)w.W1[f * kMlpH + 16 * t + col] : (short)0;
      }
      a1[t][st] = v;
    }
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int st = 0; st < 4; st++) {
      short4_t v;
#pragma unroll
      for (int e = 0; e < 4; e++) v[e] = (short)w.W2[(16 * st + 4 * q + e) * kMlpH + 16 * t + col];
      a2[t][st] = v;
    }
#pragma unroll
  for (int st = 0; st < 4; st++) {
    short4_t v;
#pragma unroll
    for (int e = 0; e < 4; e++) v[e] = (short)w.W3[(16 * st + 4 * q + e) * kMlpOutPad + col];
    a3[st] = v;
  }
  float b1[4][4], b2[4][4], b3[4];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int e = 0; e < 4; e++) {
      b1[t][e] = w.b1[16 * t + 4 * q + e];
      b2[t][e] = w.b2[16 * t + 4 * q + e];
    }
#pragma unroll
  for (int e = 0; e < 4; e++) b3[e] = (4 * q + e) < kMlpOut ? w.b3[4 * q + e] : 0.0f;

  const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t nwave = (int64_t)gridDim.x * (blockDim.x >> 6);
  // MODE 0: a wave takes 64 consecutive cells of one x-row at a time (a row has ceil(nx/64)
  // such segments); the 3x3x3 neighbourhoods of its cells lie in 66 x 3 x 3 cells, whose
  // (rho, u, T) it stages in LDS once, then it runs the network on four 16-cell tiles built
  // from there.  Features are formed in FP32 (they enter the network as bf16).
  __shared__ float lm[4][kStage][5];
  const int wv = threadIdx.x >> 6;
  const Grid& g = s.g;
  const int nseg = MODE == 0 ? (g.nx + 63) / 64 : 0;
  const int64_t nunit = MODE == 0 ? (int64_t)nseg * g.ny * g.nz : (ncell + 15) / 16;
  for (int64_t unit = wave; unit < nunit; unit += nwave) {
    int xs0 = 0, y = 0, z = 0;
    uint64_t gate = ~0ull;  // 16-cell tiles to evaluate, one bit per cell
    if (MODE == 0) {
      xs0 = 64 * (int)(unit % nseg);
      const int64_t row = unit / nseg;
      y = (int)(row % g.ny);
      z = (int)(row / g.ny);
      if (s.niter) {
        // cells of the segment that will take a predicted start (one per lane, uniform result)
        const int xl = xs0 + l;
        bool want = false;
        if (xl < g.nx) {
          const int ni = s.niter[cell_index(g, xl, y, z)];
          want = ni >= s.gate_lo && ni <= s.gate_hi;
        }
        gate = __ballot(want);
        if (gate == 0) continue;
      }
      wave_sync();  // the previous segment's reads of this wave's LDS are done
      // Entries e = l + 64 i.  The loop is unrolled and its indices wrap without branches
      // (wrap2: one period either way, which the y and z rows always are and the x positions
      // are on rows of 65 cells or more), so a lane's loads of all its entries can be in
      // flight together; shorter rows take wrapi, whose slow path branches.
      auto wrap2 = [](int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); };
      auto stage = [&](auto wrapx) {
#pragma unroll
        for (int i = 0; i < (kStage + 63) / 64; i++) {
          const int e = l + 64 * i;
          if (e < kStage) {
            const int xx = e % 66, r = e / 66;
            const int64_t c = cell_index(g, wrapx(xs0 - 1 + xx), wrap2(y + r / 3 - 1, g.ny), wrap2(z + r % 3 - 1, g.nz));
#pragma unroll
            for (int k = 0; k < 4; k++) lm[wv][e][k] = s.m[(int64_t)k * g.plane + c];
            lm[wv][e][4] = (float)s.T[c];
          }
        }
      };
      if (g.nx >= 65) stage([&](int v) { return wrap2(v, g.nx); });
      else stage([&](int v) { return wrapi(v, g.nx); });
      wave_sync();
    }
#pragma unroll 1
  for (int sub = 0; sub < (MODE == 0 ? 4 : 1); sub++) {
    if (MODE == 0 && ((gate >> (16 * sub)) & 0xFFFFull) == 0) continue;
    int64_t cell;
    bool live;
    // B operands of layer 1: features 16*st + 4*q + e of this lane's cell
    short4_t xb[9];
    if (MODE == 0) {
      const int x = xs0 + 16 * sub + col;
      live = x < g.nx;
      cell = cell_index(g, live ? x : 0, y, z);
      const int ec = 4 * 66 + 16 * sub + col + 1;  // the cell itself: dx = dy = dz = 0
      const float rc = lm[wv][ec][0], rci = 1.0f / rc, uxc = lm[wv][ec][1], uyc = lm[wv][ec][2], uzc = lm[wv][ec][3];
      const float Tc = lm[wv][ec][4], Tci = 1.0f / Tc;
      const float isT = 1.0f / sqrtf(lm[wv][ec][4]);
      const int lane_base = 66 * q + 16 * sub + col;
      // per feature type: reference value and factor (the cell's scale times the fitted one)
      const float ref[5] = {rc, uxc, uyc, uzc, Tc};
      const float mul[5] = {rci * w.in_inv[0], isT * w.in_inv[1], isT * w.in_inv[2], isT * w.in_inv[3],
                            Tci * w.in_inv[4]};
#pragma unroll
      for (int st = 0; st < 9; st++) {
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int j = st * 4 + e;
          const int type = j / 7, m = j % 7;  // slot_feature: type fixed per (st, e)
          // raw_feature of neighbour nb = 4m + q, read from the staged neighbourhood at entry
          // (nb mod 9) * 66 + 16 sub + col + nb / 9 = 66 nb - 593 (nb / 9) + 16 sub + col.
          // With q in 0..3, nb / 9 is k0 = 4m / 9 or k0 + 1 (from q >= t), so the entry is a
          // per-lane base plus one of two constants.  Every lane reads (slots 27, 28 land on
          // staged entries too) and the unused values are zeroed afterwards, so the reads
          // carry no branch and a lane's LDS reads can be in flight together.
          const int k0 = (4 * m) / 9, t = 9 * (k0 + 1) - 4 * m;  // constants once unrolled
          const int c0 = 264 * m - 593 * k0;
          const int nb = 4 * m + q;
          const float cv = lm[wv][lane_base + (q >= t ? c0 - 593 : c0)][type < 5 ? type : 0];
          // (relative differences as (v - v_c) / v_c: exactly zero where a neighbour equals
          // the cell, so a uniform neighbourhood has all-zero features)
          const int ty = type < 5 ? type : 0;
          const float raw = (cv - ref[ty]) * mul[ty];
          v[e] = (j < 35 && live && nb < 27) ? raw : 0.0f;
        }
        xb[st] = pack_bf16x4(v[0], v[1], v[2], v[3]);
      }
    } else {
      cell = unit * 16 + col;
      live = cell < ncell;
#pragma unroll
      for (int st = 0; st < 9; st++) {
        short4_t v;
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int f = slot_feature(st, q, e);
          v[e] = (short)f2bf(live && f >= 0 ? Xin[cell * kMlpIn + f] : 0.0f);
        }
        xb[st] = v;
      }
    }
    // layer 1
    short4_t h1[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      float4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int st = 0; st < 9; st++) acc = mfma16(a1[t][st], xb[st], acc);
      h1[t] = pack_bf16x4(fmaxf(acc[0] + b1[t][0], 0.0f), fmaxf(acc[1] + b1[t][1], 0.0f),
                          fmaxf(acc[2] + b1[t][2], 0.0f), fmaxf(acc[3] + b1[t][3], 0.0f));
    }
    // layer 2
    short4_t h2[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      float4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int st = 0; st < 4; st++) acc = mfma16(a2[t][st], h1[st], acc);
      h2[t] = pack_bf16x4(fmaxf(acc[0] + b2[t][0], 0.0f), fmaxf(acc[1] + b2[t][1], 0.0f),
                          fmaxf(acc[2] + b2[t][2], 0.0f), fmaxf(acc[3] + b2[t][3], 0.0f));
    }
    // layer 3: rows 0..3 of the output tile live in lanes 0..15 (q == 0)
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int st = 0; st < 4; st++) acc = mfma16(a3[st], h2[st], acc);
    if (q == 0 && live) {
#pragma unroll
      for (int e = 0; e < kMlpOut; e++) out[e * out_plane + cell] = (acc[e] + b3[e]) * w.out_scale[e];
    }
  }
  }
}

// ------------------------------------------------------------------------ training pieces

// generic tile GEMM C[M][N] (+)= sum_k A(i,k) B(k,j), bf16 operands, FP32 accumulation.
// A(i,k) = A[i*sai + k*sak], B(k,j) = B[k*sbk + j*sbj]; one wave per 16x16 tile of C.
__global__ __launch_bounds__(64) void k_gemm_bf16(const unsigned short* __restrict__ A, int64_t sai, int64_t sak,
                                                  const unsigned short* __restrict__ B, int64_t sbk, int64_t sbj,
                                                  float* __restrict__ C, int64_t ldc, int M, int N, int K, int accum) {
  const int l = threadIdx.x;
  const int q = l >> 4, col = l & 15;
  const int ti = blockIdx.x, tj = blockIdx.y;
  float4_t acc = {0.f, 0.f, 0.f, 0.f};
  const int i = 16 * ti + col;  // A row held by this lane
  const int j = 16 * tj + col;  // B column held by this lane
  for (int k0 = 0; k0 < K; k0 += 16) {
    short4_t av, bv;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int k = k0 + 4 * q + e;
      av[e] = (short)((i < M && k < K) ? A[i * sai + k * sak] : 0);
      bv[e] = (short)((j < N && k < K) ? B[k * sbk + j * sbj] : 0);
    }
    acc = mfma16(av, bv, acc);
  }
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int r = 16 * ti + 4 * q + e;
    if (r < M && j < N) {
      float* p = C + (int64_t)r * ldc + j;
      *p = accum ? *p + acc[e] : acc[e];
    }
  }
}

// Z = Zacc + bias (FP32), A = bf16(relu(Z)) ; rows x cols, row stride ld
__global__ void k_bias_relu(const float* __restrict__ Zacc, const float* __restrict__ bias, float* __restrict__ Z,
                            unsigned short* __restrict__ Ab, int rows, int cols, int ld, int relu) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)rows * ld) return;
  const int c = (int)(t % ld);
  if (c >= cols) {
    Ab[t] = 0;
    Z[t] = 0.f;
    return;
  }
  const float z = Zacc[t] + bias[c];
  Z[t] = z;
  Ab[t] = f2bf(relu ? fmaxf(z, 0.f) : z);
}

// dY = 2 (Y - T)/B (FP32, padded to ld columns), dYb = bf16(dY), per-row squared error
__global__ void k_out_grad(const float* __restrict__ Y, const float* __restrict__ T, float* __restrict__ dY,
                           unsigned short* __restrict__ dYb, float* __restrict__ sqerr, int B, int nout, int ld) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float se = 0.f;
  for (int c = 0; c < ld; c++) {
    float d = 0.f;
    if (c < nout) {
      const float e = Y[(int64_t)b * ld + c] - T[(int64_t)b * nout + c];
      se += e * e;
      d = 2.0f * e / (float)B;
    }
    dY[(int64_t)b * ld + c] = d;
    dYb[(int64_t)b * ld + c] = f2bf(d);
  }
  sqerr[b] = se;
}

// dZ = G * (Z > 0), dZb = bf16(dZ)
__global__ void k_relu_grad(const float* __restrict__ G, const float* __restrict__ Z, float* __restrict__ dZ,
                            unsigned short* __restrict__ dZb, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float d = Z[t] > 0.f ? G[t] : 0.f;
  dZ[t] = d;
  dZb[t] = f2bf(d);
}

// column sums of a rows x cols matrix (row stride ld) into out[cols] (bias gradients)
// column sums of a rows x cols matrix (stride ld): one 256-thread block per column
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ M, int rows, int cols, int ld,
                                                float* __restrict__ out) {
  __shared__ float sh[256];
  const int c = blockIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int r = threadIdx.x; r < rows; r += 256) s += M[(int64_t)r * ld + c];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sh[0];
}

// sum of n values, one 256-thread block
__global__ __launch_bounds__(256) void k_sum(const float* __restrict__ v, int n, float* __restrict__ out) {
  __shared__ float sh[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) s += v[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// FP32 -> bf16 with row padding: src rows x cols (stride src_ld) -> dst rows x ld
__global__ void k_to_bf16(const float* __restrict__ src, int rows, int cols, int src_ld, unsigned short* __restrict__ dst,
                          int ld) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)rows * ld) return;
  const int r = (int)(t / ld), c = (int)(t % ld);
  dst[t] = c < cols ? f2bf(src[(int64_t)r * src_ld + c]) : 0;
}

// Adam on the FP32 master parameters
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  // the biases stay where they are (zero after pond_mlp_init): a bias-free ReLU network is
  // positively homogeneous, net(a x) = a net(x) for a >= 0
  if ((t >= kOffB1 && t < kOffW2) || (t >= kOffB2 && t < kOffW3) || t >= kOffB3) return;
  const float gi = g[t];
  const float mi = b1 * m[t] + (1.f - b1) * gi;
  const float vi = b2 * v[t] + (1.f - b2) * gi * gi;
  m[t] = mi;
  v[t] = vi;
  p[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// gather a minibatch: X (normalized features, B x kMlpIn) and T (normalized targets) from the
// dataset rows idx[b].
//  * Augmentation: row b is seen through one of the 48 symmetries of the cubic lattice (an
//    axis permutation with sign flips, chosen by a hash of (seed, b)).  New axis a is old axis
//    perm[a] times sign[a]: the neighbour at new offset d' is the old neighbour at
//    d_perm[a] = sign[a] d'_a, and velocity component a is sign[a] times old component perm[a].
//  * Relative error: the pair is scaled by 1 / sqrt(max_k T_k^2 + rel_eps^2).  The network is
//    positively homogeneous, so this turns the squared error into a relative squared error and
//    leaves the prediction for the unscaled features unchanged.
__global__ void k_gather_batch(const float* __restrict__ Xall, const float* __restrict__ Yall,
                               const int* __restrict__ idx, int B, const float* __restrict__ in_inv,
                               const float* __restrict__ out_inv, uint64_t seed, int augment, float rel_eps,
                               float* __restrict__ X, float* __restrict__ T) {
  const int b = blockIdx.x;
  if (b >= B) return;
  const int64_t r = idx[b];
  int perm[3] = {0, 1, 2}, sgn[3] = {1, 1, 1};
  if (augment) {
    uint64_t h = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(b + 1);
    h = (h ^ (h >> 30)) * 0xbf58476d1ce4e5b9ull;
    h = (h ^ (h >> 27)) * 0x94d049bb133111ebull;
    h ^= h >> 31;
    const int sym = (int)(h % 48u);
    const int pi = sym >> 3;  // 6 permutations
    constexpr int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (int a = 0; a < 3; a++) {
      perm[a] = kPerm[pi][a];
      sgn[a] = (sym >> a) & 1 ? -1 : 1;
    }
  }
  float t[kMlpOut];
  float tmax = 0.f;
  for (int c = 0; c < kMlpOut; c++) {
    const int src = c < 3 ? perm[c] : c;
    t[c] = Yall[r * kMlpOut + src] * out_inv[c] * (c < 3 ? (float)sgn[c] : 1.f);
    tmax = fmaxf(tmax, fabsf(t[c]));
  }
  const float sc = 1.0f / sqrtf(tmax * tmax + rel_eps * rel_eps);
  for (int f = threadIdx.x; f < kMlpIn; f += blockDim.x) {
    const int nb = f / 5, type = f - 5 * nb;
    const int dn[3] = {nb / 9 - 1, (nb / 3) % 3 - 1, nb % 3 - 1};
    int d[3];
    for (int a = 0; a < 3; a++) d[perm[a]] = sgn[a] * dn[a];
    const int onb = 9 * (d[0] + 1) + 3 * (d[1] + 1) + (d[2] + 1);
    float sgv = 1.f;
    int otype = type;
    if (type >= 1 && type <= 3) {
      otype = 1 + perm[type - 1];
      sgv = (float)sgn[type - 1];
    }
    X[(int64_t)b * kMlpIn + f] = Xall[r * kMlpIn + 5 * onb + otype] * sgv * in_inv[type] * sc;
  }
  if (threadIdx.x < kMlpOut) T[(int64_t)b * kMlpOut + threadIdx.x] = t[threadIdx.x] * sc;
}

// dataset collection: raw features of sampled cells from the previous state, raw targets
// (du/sqrt(T_c), dT/T_c) from the stepped state
__global__ void k_collect(MlpState prev, MlpState next, int64_t nsamp, uint64_t seed, float* __restrict__ X,
                          float* __restrict__ Y) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nsamp) return;
  // splitmix64 of (seed, b) picks the cell
  uint64_t h = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(b + 1);
  h = (h ^ (h >> 30)) * 0xbf58476d1ce4e5b9ull;
  h = (h ^ (h >> 27)) * 0x94d049bb133111ebull;
  h ^= h >> 31;
  const Grid& g = prev.g;
  const int64_t cell = (int64_t)(h % (uint64_t)g.n);
  const int x = (int)(cell % g.nx), y = (int)((cell / g.nx) % g.ny), z = (int)(cell / ((int64_t)g.nx * g.ny));
  const int64_t P = g.plane;
  const float rc = prev.m[cell], uxc = prev.m[P + cell], uyc = prev.m[2 * P + cell], uzc = prev.m[3 * P + cell];
  const double Tc = prev.T[cell];
  const float isT = (float)(1.0 / sqrt(Tc));
  for (int f = 0; f < kMlpIn; f++) X[b * kMlpIn + f] = raw_feature(prev, x, y, z, f, rc, uxc, uyc, uzc, Tc, isT);
  Y[b * kMlpOut + 0] = (next.m[P + cell] - uxc) * isT;
  Y[b * kMlpOut + 1] = (next.m[2 * P + cell] - uyc) * isT;
  Y[b * kMlpOut + 2] = (next.m[3 * P + cell] - uzc) * isT;
  Y[b * kMlpOut + 3] = (float)(next.T[cell] / Tc - 1.0);
}

// per-type second moments of the raw dataset (for the input / output scales)
__global__ void k_moments2(const float* __restrict__ X, const float* __restrict__ Y, int64_t n, double* __restrict__ acc) {
  __shared__ double sh[9][64];
  double loc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    for (int f = 0; f < kMlpIn; f++) {
      const double v = X[r * kMlpIn + f];
      loc[f % 5] += v * v;
    }
    for (int o = 0; o < kMlpOut; o++) {
      const double v = Y[r * kMlpOut + o];
      loc[5 + o] += v * v;
    }
  }
  for (int k = 0; k < 9; k++) sh[k][threadIdx.x] = loc[k];
  __syncthreads();
  if (threadIdx.x < 9) {
    double s = 0;
    for (int i = 0; i < 64; i++) s += sh[threadIdx.x][i];
    atomicAdd(&acc[threadIdx.x], s);
  }
}

// ------------------------------------------------------------------------ launchers

hipError_t mlp_launch_infer(const MlpWeightsBF& w, const MlpState& s, float* out, int64_t out_plane, hipStream_t st) {
  const int64_t nseg = (int64_t)((s.g.nx + 63) / 64) * s.g.ny * s.g.nz;
  const int64_t waves = nseg < 16384 ? nseg : 16384;
  const unsigned blocks = (unsigned)((waves + 3) / 4);
  hipLaunchKernelGGL(k_mlp_infer<0>, dim3(blocks), dim3(256), 0, st, w, s, (const float*)nullptr, s.g.n, out, out_plane);
  return hipGetLastError();
}

hipError_t mlp_launch_infer_matrix(const MlpWeightsBF& w, const float* X, int64_t n, float* out, hipStream_t st) {
  MlpState s{};
  const int64_t ntile = (n + 15) / 16;
  const unsigned blocks = (unsigned)((ntile + 3) / 4);
  hipLaunchKernelGGL(k_mlp_infer<1>, dim3(blocks), dim3(256), 0, st, w, s, X, n, out, n);
  return hipGetLastError();
}

hipError_t mlp_gemm(const unsigned short* A, int64_t sai, int64_t sak, const unsigned short* B, int64_t sbk,
                    int64_t sbj, float* C, int64_t ldc, int M, int N, int K, int accum, hipStream_t st) {
  hipLaunchKernelGGL(k_gemm_bf16, dim3((M + 15) / 16, (N + 15) / 16), dim3(64), 0, st, A, sai, sak, B, sbk, sbj, C,
                     ldc, M, N, K, accum);
  return hipGetLastError();
}

static unsigned nb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t mlp_bias_relu(const float* Zacc, const float* bias, float* Z, unsigned short* Ab, int rows, int cols, int ld,
                         int relu, hipStream_t st) {
  hipLaunchKernelGGL(k_bias_relu, dim3(nb((int64_t)rows * ld, 256)), dim3(256), 0, st, Zacc, bias, Z, Ab, rows, cols,
                     ld, relu);
  return hipGetLastError();
}

hipError_t mlp_out_grad(const float* Y, const float* T, float* dY, unsigned short* dYb, float* sqerr, int B, int nout,
                        int ld, hipStream_t st) {
  hipLaunchKernelGGL(k_out_grad, dim3(nb(B, 256)), dim3(256), 0, st, Y, T, dY, dYb, sqerr, B, nout, ld);
  return hipGetLastError();
}

hipError_t mlp_relu_grad(const float* G, const float* Z, float* dZ, unsigned short* dZb, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_relu_grad, dim3(nb(n, 256)), dim3(256), 0, st, G, Z, dZ, dZb, n);
  return hipGetLastError();
}

hipError_t mlp_colsum(const float* M, int rows, int cols, int ld, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_colsum, dim3((unsigned)cols), dim3(256), 0, st, M, rows, cols, ld, out);
  return hipGetLastError();
}

hipError_t mlp_sum(const float* v, int n, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, st, v, n, out);
  return hipGetLastError();
}

hipError_t mlp_to_bf16(const float* src, int rows, int cols, int src_ld, unsigned short* dst, int ld, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16, dim3(nb((int64_t)rows * ld, 256)), dim3(256), 0, st, src, rows, cols, src_ld, dst, ld);
  return hipGetLastError();
}

hipError_t mlp_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                    float c1, float c2, hipStream_t st) {
  hipLaunchKernelGGL(k_adam, dim3(nb(n, 256)), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t mlp_gather_batch(const float* Xall, const float* Yall, const int* idx, int B, const float* in_inv,
                            const float* out_inv, uint64_t seed, int augment, float rel_eps, float* X, float* T,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_gather_batch, dim3(B), dim3(64), 0, st, Xall, Yall, idx, B, in_inv, out_inv, seed, augment,
                     rel_eps, X, T);
  return hipGetLastError();
}

hipError_t mlp_collect(const MlpState& prev, const MlpState& next, int64_t nsamp, uint64_t seed, float* X, float* Y,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_collect, dim3(nb(nsamp, 64)), dim3(64), 0, st, prev, next, nsamp, seed, X, Y);
  return hipGetLastError();
}

hipError_t mlp_moments2(const float* X, const float* Y, int64_t n, double* acc, hipStream_t st) {
  const int64_t blocks = n < 64 * 1024 ? nb(n, 64) : 1024;
  hipLaunchKernelGGL(k_moments2, dim3((unsigned)(blocks > 0 ? blocks : 1)), dim3(64), 0, st, X, Y, n, acc);
  return hipGetLastError();
}

}  // namespace pond
