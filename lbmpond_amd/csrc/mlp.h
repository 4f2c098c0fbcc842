// mlp.h -- shapes and launchers of the learned gauge initializer (135-64-64-4, BF16 MFMA).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "pond_device.h"

namespace pond {

constexpr int kMlpIn = 135, kMlpInPad = 144, kMlpH = 64, kMlpOut = 4, kMlpOutPad = 16;
// packed FP32 parameters: W1[135][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4]
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpIn * kMlpH, kOffW2 = kOffB1 + kMlpH, kOffB2 = kOffW2 + kMlpH * kMlpH,
              kOffW3 = kOffB2 + kMlpH, kOffB3 = kOffW3 + kMlpH * kMlpOut, kMlpParams = kOffB3 + kMlpOut;

// device view of the network for the kernels: bf16 weights (W1 rows padded to 144, W3 columns
// padded to 16), FP32 biases, feature and output scales
struct MlpWeightsBF {
  const unsigned short* W1;
  const unsigned short* W2;
  const unsigned short* W3;
  const float* b1;
  const float* b2;
  const float* b3;
  float in_inv[5];     // 1/scale per feature type (rho, ux, uy, uz, T)
  float out_scale[4];  // decoded output = network output * out_scale
};

// split-layout state the features are read from
struct MlpState {
  Grid g;
  const float* m;   // rho, ux, uy, uz planes
  const double* T;  // temperature plane
  // when set, only 16-cell tiles holding a cell whose count lies in [gate_lo, gate_hi] are
  // evaluated (the step takes a predicted start nowhere else); the others' outputs are not written
  const int* niter = nullptr;
  int gate_lo = 0, gate_hi = 0;
};

hipError_t mlp_launch_infer(const MlpWeightsBF& w, const MlpState& s, float* out, int64_t out_plane, hipStream_t st);
hipError_t mlp_launch_infer_matrix(const MlpWeightsBF& w, const float* X, int64_t n, float* out, hipStream_t st);
hipError_t mlp_gemm(const unsigned short* A, int64_t sai, int64_t sak, const unsigned short* B, int64_t sbk,
                    int64_t sbj, float* C, int64_t ldc, int M, int N, int K, int accum, hipStream_t st);
hipError_t mlp_bias_relu(const float* Zacc, const float* bias, float* Z, unsigned short* Ab, int rows, int cols, int ld,
                         int relu, hipStream_t st);
hipError_t mlp_out_grad(const float* Y, const float* T, float* dY, unsigned short* dYb, float* sqerr, int B, int nout,
                        int ld, hipStream_t st);
hipError_t mlp_relu_grad(const float* G, const float* Z, float* dZ, unsigned short* dZb, int64_t n, hipStream_t st);
hipError_t mlp_colsum(const float* M, int rows, int cols, int ld, float* out, hipStream_t st);
hipError_t mlp_sum(const float* v, int n, float* out, hipStream_t st);
hipError_t mlp_to_bf16(const float* src, int rows, int cols, int src_ld, unsigned short* dst, int ld, hipStream_t st);
hipError_t mlp_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                    float c1, float c2, hipStream_t st);
hipError_t mlp_gather_batch(const float* Xall, const float* Yall, const int* idx, int B, const float* in_inv,
                            const float* out_inv, uint64_t seed, int augment, float rel_eps, float* X, float* T,
                            hipStream_t st);
hipError_t mlp_collect(const MlpState& prev, const MlpState& next, int64_t nsamp, uint64_t seed, float* X, float* Y,
                       hipStream_t st);
hipError_t mlp_moments2(const float* X, const float* Y, int64_t n, double* acc, hipStream_t st);

}  // namespace pond
