// mlp_api.hip -- host side of the learned gauge initializer: dataset collection from the
// solver's own steps, training with the BF16 MFMA kernels (Adam on FP32 master weights),
// parameter I/O, and the forward/gradient entry points used to check the kernels.
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>
#include "pond_ctx.h"

using namespace pond;

namespace {

// relative-error floor of the training loss, in units of the RMS target (k_gather_batch)
constexpr float kRelEps = 0.1f;
constexpr double kPi = 3.14159265358979323846;

#define MLP_HIP(ctx, expr)                                                        \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);             \
      return POND_ERR_HIP;                                                        \
    }                                                                             \
  } while (0)

template <typename T>
hipError_t dalloc(pond_ctx* ctx, T** p, size_t n) {
  hipError_t e = hipMalloc(p, n * sizeof(T));
  if (e == hipSuccess) ctx->bytes += (int64_t)(n * sizeof(T));
  return e;
}

pond_status mlp_alloc(pond_ctx* ctx) {
  MlpCtx& m = ctx->mlp;
  if (m.params) return POND_OK;
  MLP_HIP(ctx, dalloc(ctx, &m.params, kMlpParams));
  MLP_HIP(ctx, dalloc(ctx, &m.grad, kMlpParams));
  MLP_HIP(ctx, dalloc(ctx, &m.adam_m, kMlpParams));
  MLP_HIP(ctx, dalloc(ctx, &m.adam_v, kMlpParams));
  MLP_HIP(ctx, dalloc(ctx, &m.W1b, (size_t)kMlpInPad * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.W2b, (size_t)kMlpH * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.W3b, (size_t)kMlpH * kMlpOutPad));
  MLP_HIP(ctx, dalloc(ctx, &m.gw3tmp, (size_t)kMlpH * kMlpOutPad));
  MLP_HIP(ctx, dalloc(ctx, &m.mom, (size_t)16));
  MLP_HIP(ctx, hipMemsetAsync(m.W1b, 0, (size_t)kMlpInPad * kMlpH * 2, ctx->stream));
  MLP_HIP(ctx, hipMemsetAsync(m.W3b, 0, (size_t)kMlpH * kMlpOutPad * 2, ctx->stream));
  MLP_HIP(ctx, hipMemsetAsync(m.adam_m, 0, (size_t)kMlpParams * 4, ctx->stream));
  MLP_HIP(ctx, hipMemsetAsync(m.adam_v, 0, (size_t)kMlpParams * 4, ctx->stream));
  return POND_OK;
}

pond_status mlp_batch_alloc(pond_ctx* ctx, int B) {
  MlpCtx& m = ctx->mlp;
  if (m.bcap >= B) return POND_OK;
  float** fbufs[] = {&m.bx, &m.bt, &m.z1acc, &m.z1, &m.z2acc, &m.z2, &m.yacc, &m.y, &m.dy, &m.g2, &m.dz2, &m.g1,
                     &m.dz1, &m.sqerr, &m.loss};
  for (auto p : fbufs)
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  unsigned short** hbufs[] = {&m.xb, &m.a1b, &m.a2b, &m.dyb, &m.dz2b, &m.dz1b};
  for (auto p : hbufs)
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  if (m.idx) { (void)hipFree(m.idx); m.idx = nullptr; }
  const size_t b = (size_t)B;
  MLP_HIP(ctx, dalloc(ctx, &m.bx, b * kMlpIn));
  MLP_HIP(ctx, dalloc(ctx, &m.bt, b * kMlpOut));
  MLP_HIP(ctx, dalloc(ctx, &m.z1acc, b * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.z1, b * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.z2acc, b * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.z2, b * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.yacc, b * kMlpOutPad));
  MLP_HIP(ctx, dalloc(ctx, &m.y, b * kMlpOutPad));
  MLP_HIP(ctx, dalloc(ctx, &m.dy, b * kMlpOutPad));
  MLP_HIP(ctx, dalloc(ctx, &m.g2, b * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.dz2, b * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.g1, b * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.dz1, b * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.sqerr, b));
  MLP_HIP(ctx, dalloc(ctx, &m.loss, 4));
  MLP_HIP(ctx, dalloc(ctx, &m.xb, b * kMlpInPad));
  MLP_HIP(ctx, dalloc(ctx, &m.a1b, b * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.a2b, b * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.dyb, b * kMlpOutPad));
  MLP_HIP(ctx, dalloc(ctx, &m.dz2b, b * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.dz1b, b * kMlpH));
  MLP_HIP(ctx, dalloc(ctx, &m.idx, b));
  m.bcap = B;
  return POND_OK;
}

// bf16 working copies of the master weights
pond_status mlp_refresh_bf16(pond_ctx* ctx) {
  MlpCtx& m = ctx->mlp;
  MLP_HIP(ctx, mlp_to_bf16(m.params + kOffW1, kMlpIn, kMlpH, kMlpH, m.W1b, kMlpH, ctx->stream));
  MLP_HIP(ctx, mlp_to_bf16(m.params + kOffW2, kMlpH, kMlpH, kMlpH, m.W2b, kMlpH, ctx->stream));
  MLP_HIP(ctx, mlp_to_bf16(m.params + kOffW3, kMlpH, kMlpOut, kMlpOut, m.W3b, kMlpOutPad, ctx->stream));
  return POND_OK;
}

// forward + backward of one minibatch held in m.bx (normalized X) / m.bt (normalized T);
// gradients land in m.grad, the summed squared error in m.loss[0].  Training holds the biases
// at zero and skips their gradients (bias_grad = false); the gradient entry point computes all.
pond_status mlp_fwd_bwd(pond_ctx* ctx, int B, bool bias_grad) {
  MlpCtx& m = ctx->mlp;
  hipStream_t st = ctx->stream;
  const float* P = m.params;
  MLP_HIP(ctx, mlp_to_bf16(m.bx, B, kMlpIn, kMlpIn, m.xb, kMlpInPad, st));
  // forward
  MLP_HIP(ctx, mlp_gemm(m.xb, kMlpInPad, 1, m.W1b, kMlpH, 1, m.z1acc, kMlpH, B, kMlpH, kMlpInPad, 0, st));
  MLP_HIP(ctx, mlp_bias_relu(m.z1acc, P + kOffB1, m.z1, m.a1b, B, kMlpH, kMlpH, 1, st));
  MLP_HIP(ctx, mlp_gemm(m.a1b, kMlpH, 1, m.W2b, kMlpH, 1, m.z2acc, kMlpH, B, kMlpH, kMlpH, 0, st));
  MLP_HIP(ctx, mlp_bias_relu(m.z2acc, P + kOffB2, m.z2, m.a2b, B, kMlpH, kMlpH, 1, st));
  MLP_HIP(ctx, mlp_gemm(m.a2b, kMlpH, 1, m.W3b, kMlpOutPad, 1, m.yacc, kMlpOutPad, B, kMlpOutPad, kMlpH, 0, st));
  // y = yacc + b3 (no activation); the padded columns get bias 0 via cols = kMlpOut
  MLP_HIP(ctx, mlp_bias_relu(m.yacc, P + kOffB3, m.y, m.dyb, B, kMlpOut, kMlpOutPad, 0, st));
  // backward
  MLP_HIP(ctx, mlp_out_grad(m.y, m.bt, m.dy, m.dyb, m.sqerr, B, kMlpOut, kMlpOutPad, st));
  if (bias_grad) MLP_HIP(ctx, mlp_colsum(m.dy, B, kMlpOut, kMlpOutPad, m.grad + kOffB3, st));
  MLP_HIP(ctx, mlp_gemm(m.a2b, 1, kMlpH, m.dyb, kMlpOutPad, 1, m.gw3tmp, kMlpOutPad, kMlpH, kMlpOutPad, B, 0, st));
  // pack the 64x16 padded W3 gradient into the 64x4 parameter block
  MLP_HIP(ctx, hipMemcpy2DAsync(m.grad + kOffW3, kMlpOut * sizeof(float), m.gw3tmp, kMlpOutPad * sizeof(float),
                                kMlpOut * sizeof(float), kMlpH, hipMemcpyDeviceToDevice, st));
  MLP_HIP(ctx, mlp_gemm(m.dyb, kMlpOutPad, 1, m.W3b, 1, kMlpOutPad, m.g2, kMlpH, B, kMlpH, kMlpOutPad, 0, st));
  MLP_HIP(ctx, mlp_relu_grad(m.g2, m.z2, m.dz2, m.dz2b, (int64_t)B * kMlpH, st));
  if (bias_grad) MLP_HIP(ctx, mlp_colsum(m.dz2, B, kMlpH, kMlpH, m.grad + kOffB2, st));
  MLP_HIP(ctx, mlp_gemm(m.a1b, 1, kMlpH, m.dz2b, kMlpH, 1, m.grad + kOffW2, kMlpH, kMlpH, kMlpH, B, 0, st));
  MLP_HIP(ctx, mlp_gemm(m.dz2b, kMlpH, 1, m.W2b, 1, kMlpH, m.g1, kMlpH, B, kMlpH, kMlpH, 0, st));
  MLP_HIP(ctx, mlp_relu_grad(m.g1, m.z1, m.dz1, m.dz1b, (int64_t)B * kMlpH, st));
  if (bias_grad) MLP_HIP(ctx, mlp_colsum(m.dz1, B, kMlpH, kMlpH, m.grad + kOffB1, st));
  MLP_HIP(ctx, mlp_gemm(m.xb, 1, kMlpInPad, m.dz1b, kMlpH, 1, m.grad + kOffW1, kMlpH, kMlpIn, kMlpH, B, 0, st));
  MLP_HIP(ctx, mlp_sum(m.sqerr, B, m.loss, st));
  return POND_OK;
}

MlpState state_view(pond_ctx* ctx, int b) { return MlpState{ctx->g, ctx->sp[b].m, ctx->sp[b].T}; }

// ---- device scratch of pond_mlp_forward_noalloc: X rows, then the output planes, then the
// output rows, for at most kNoAllocRows rows.  One copy per GPU, held until the process ends.
constexpr int64_t kNoAllocRows = 16384;
constexpr size_t kNoAllocFloats = (size_t)kNoAllocRows * (kMlpIn + 2 * kMlpOut);
static __device__ __align__(256) float scratch[kNoAllocFloats];
constexpr int kMaxDevices = 64;
float* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

// the device address of this GPU's scratch array (looked up once per device)
hipError_t scratch_ptr(float** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[dev] = static_cast<float*>(ptr);
  }
  *out = g_scratch_dev[dev];
  return hipSuccess;
}

// output planes [o][b] -> rows [b][o]
__global__ void k_planes_to_rows(const float* __restrict__ planes, int64_t n, float* __restrict__ rows) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * kMlpOut) return;
  const int64_t b = i / kMlpOut;
  const int o = (int)(i % kMlpOut);
  rows[i] = planes[(int64_t)o * n + b];
}

}  // namespace

// device view of the network used by the step driver
MlpWeightsBF pond_mlp_weights(const pond_ctx* ctx) {
  const MlpCtx& m = ctx->mlp;
  MlpWeightsBF w;
  w.W1 = m.W1b;
  w.W2 = m.W2b;
  w.W3 = m.W3b;
  w.b1 = m.params + kOffB1;
  w.b2 = m.params + kOffB2;
  w.b3 = m.params + kOffB3;
  for (int k = 0; k < 5; k++) w.in_inv[k] = m.in_inv[k];
  for (int k = 0; k < 4; k++) w.out_scale[k] = m.out_scale[k];
  return w;
}

void pond_mlp_free(pond_ctx* ctx) {
  MlpCtx& m = ctx->mlp;
  void* ptrs[] = {m.params, m.grad, m.adam_m, m.adam_v, m.W1b, m.W2b, m.W3b, m.gw3tmp, m.mom, m.X, m.Y, m.bx, m.bt,
                  m.z1acc, m.z1, m.z2acc, m.z2, m.yacc, m.y, m.dy, m.g2, m.dz2, m.g1, m.dz1, m.sqerr, m.loss, m.xb,
                  m.a1b, m.a2b, m.dyb, m.dz2b, m.dz1b, m.idx};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  m = MlpCtx();
}

extern "C" {

int64_t pond_mlp_num_params(void) { return kMlpParams; }

pond_status pond_mlp_init(pond_ctx* ctx, uint64_t seed) {
  if (!ctx) return POND_ERR_INVALID;
  pond_status s = mlp_alloc(ctx);
  if (s != POND_OK) return s;
  MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));  // no step may still read the network
  // He-uniform weights and zero biases from a seeded generator; training keeps the biases at zero
  std::vector<float> p(kMlpParams, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int fan_in, int n) {
    const float a = std::sqrt(6.0f / (float)fan_in);
    std::uniform_real_distribution<float> u(-a, a);
    for (int i = 0; i < n; i++) p[off + i] = u(rng);
  };
  fill(kOffW1, kMlpIn, kMlpIn * kMlpH);
  fill(kOffW2, kMlpH, kMlpH * kMlpH);
  fill(kOffW3, kMlpH, kMlpH * kMlpOut);
  for (int i = 0; i < kMlpH * kMlpOut; i++) p[kOffW3 + i] *= 0.1f;
  MLP_HIP(ctx, hipMemcpy(ctx->mlp.params, p.data(), (size_t)kMlpParams * 4, hipMemcpyHostToDevice));
  MLP_HIP(ctx, hipMemset(ctx->mlp.adam_m, 0, (size_t)kMlpParams * 4));
  MLP_HIP(ctx, hipMemset(ctx->mlp.adam_v, 0, (size_t)kMlpParams * 4));
  ctx->mlp.adam_t = 0;
  ctx->mlp.trained = false;
  s = mlp_refresh_bf16(ctx);
  if (s != POND_OK) return s;
  MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return POND_OK;
}

pond_status pond_mlp_collect(pond_ctx* ctx, int64_t nsamples, uint64_t seed) {
  if (!ctx) return POND_ERR_INVALID;
  if (!ctx->split) { ctx->err = "pond_mlp_collect: needs a POND_SPLIT context"; return POND_ERR_INVALID; }
  if (ctx->step < 1) { ctx->err = "pond_mlp_collect: needs at least one step taken"; return POND_ERR_STATE; }
  if (nsamples <= 0) return POND_OK;
  MlpCtx& m = ctx->mlp;
  if (m.count + nsamples > m.cap) {
    const int64_t cap = std::max<int64_t>(m.count + nsamples, 2 * m.cap);
    float *X = nullptr, *Y = nullptr;
    MLP_HIP(ctx, dalloc(ctx, &X, (size_t)cap * kMlpIn));
    MLP_HIP(ctx, dalloc(ctx, &Y, (size_t)cap * kMlpOut));
    if (m.count) {
      MLP_HIP(ctx, hipMemcpyAsync(X, m.X, (size_t)m.count * kMlpIn * 4, hipMemcpyDeviceToDevice, ctx->stream));
      MLP_HIP(ctx, hipMemcpyAsync(Y, m.Y, (size_t)m.count * kMlpOut * 4, hipMemcpyDeviceToDevice, ctx->stream));
    }
    MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (m.X) (void)hipFree(m.X);
    if (m.Y) (void)hipFree(m.Y);
    m.X = X;
    m.Y = Y;
    m.cap = cap;
  }
  // features from the state before the last step (buffer 1-cur), targets from the state after it
  MLP_HIP(ctx, mlp_collect(state_view(ctx, 1 - ctx->cur), state_view(ctx, ctx->cur), nsamples, seed,
                           m.X + m.count * kMlpIn, m.Y + m.count * kMlpOut, ctx->stream));
  MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  m.count += nsamples;
  return POND_OK;
}

pond_status pond_mlp_dataset_size(pond_ctx* ctx, int64_t* n) {
  if (!ctx || !n) return POND_ERR_INVALID;
  *n = ctx->mlp.count;
  return POND_OK;
}

pond_status pond_mlp_train(pond_ctx* ctx, int epochs, int batch, float lr, uint64_t seed, float* epoch_loss) {
  if (!ctx || epochs < 1 || batch < 16 || batch % 16 != 0 || !(lr > 0)) return POND_ERR_INVALID;
  MlpCtx& m = ctx->mlp;
  if (m.count < batch) { ctx->err = "pond_mlp_train: dataset smaller than one batch"; return POND_ERR_STATE; }
  pond_status s = mlp_alloc(ctx);
  if (s != POND_OK) return s;
  if (!m.initialized) {
    s = pond_mlp_init(ctx, seed);
    if (s != POND_OK) return s;
    m.initialized = true;
  }
  s = mlp_batch_alloc(ctx, batch);
  if (s != POND_OK) return s;
  // feature / output scales from the dataset's second moments
  MLP_HIP(ctx, hipMemsetAsync(m.mom, 0, 16 * sizeof(double), ctx->stream));
  MLP_HIP(ctx, mlp_moments2(m.X, m.Y, m.count, m.mom, ctx->stream));
  double mom[16];
  // the context stream is non-blocking: copies go through it, then wait for it
  MLP_HIP(ctx, hipMemcpyAsync(mom, m.mom, 16 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // the three velocity components share one scale (the lattice symmetries used in training
  // exchange them)
  double msi[5], mso[4];
  for (int k = 0; k < 5; k++) msi[k] = mom[k] / ((double)m.count * 27.0);
  for (int o = 0; o < 4; o++) mso[o] = mom[5 + o] / (double)m.count;
  msi[1] = msi[2] = msi[3] = (msi[1] + msi[2] + msi[3]) / 3.0;
  mso[0] = mso[1] = mso[2] = (mso[0] + mso[1] + mso[2]) / 3.0;
  for (int k = 0; k < 5; k++) m.in_inv[k] = msi[k] > 0 ? (float)(1.0 / std::sqrt(msi[k])) : 1.0f;
  for (int o = 0; o < 4; o++) {
    m.out_scale[o] = mso[o] > 0 ? (float)std::sqrt(mso[o]) : 1.0f;
    m.out_inv[o] = 1.0f / m.out_scale[o];
  }
  float *din = nullptr, *dout = nullptr;
  MLP_HIP(ctx, hipMalloc(&din, 5 * sizeof(float)));
  MLP_HIP(ctx, hipMalloc(&dout, 4 * sizeof(float)));
  MLP_HIP(ctx, hipMemcpy(din, m.in_inv, 5 * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP(ctx, hipMemcpy(dout, m.out_inv, 4 * sizeof(float), hipMemcpyHostToDevice));
  std::mt19937_64 rng(seed ^ 0x5bd1e995ull);
  std::vector<int> perm((size_t)m.count);
  for (int64_t i = 0; i < m.count; i++) perm[i] = (int)i;
  const int64_t nb = m.count / batch;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const double total = (double)epochs * (double)nb;  // cosine decay of the step size over the run
  pond_status st = POND_OK;
  for (int ep = 0; ep < epochs && st == POND_OK; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    double lsum = 0;
    int lcount = 0;
    for (int64_t bi = 0; bi < nb && st == POND_OK; bi++) {
      if (hipMemcpyAsync(m.idx, perm.data() + bi * batch, (size_t)batch * sizeof(int), hipMemcpyHostToDevice,
                         ctx->stream) != hipSuccess ||
          mlp_gather_batch(m.X, m.Y, m.idx, batch, din, dout, seed + 0x2545f4914f6cdd1dull * (uint64_t)(ep * nb + bi + 1),
                           1, kRelEps, m.bx, m.bt, ctx->stream) != hipSuccess) {
        st = POND_ERR_HIP;
        ctx->err = "pond_mlp_train: batch gather failed";
        break;
      }
      st = mlp_fwd_bwd(ctx, batch, false);
      if (st != POND_OK) break;
      m.adam_t++;
      const float c1 = 1.0f - std::pow(b1, (float)m.adam_t), c2 = 1.0f - std::pow(b2, (float)m.adam_t);
      const float lr_t = (float)(0.5 * (double)lr * (1.0 + std::cos(kPi * (double)(ep * nb + bi) / total)));
      if (mlp_adam(m.params, m.grad, m.adam_m, m.adam_v, kMlpParams, lr_t, b1, b2, eps, c1, c2, ctx->stream) !=
          hipSuccess) {
        st = POND_ERR_HIP;
        ctx->err = "pond_mlp_train: adam failed";
        break;
      }
      st = mlp_refresh_bf16(ctx);
      if (st != POND_OK) break;
      if ((bi & 63) == 63 || bi == nb - 1) {
        float l = 0;
        if (hipMemcpyAsync(&l, m.loss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
          st = POND_ERR_HIP;
          break;
        }
        lsum += l / (float)batch;
        lcount++;
      }
    }
    if (epoch_loss) epoch_loss[ep] = lcount ? (float)(lsum / lcount) : 0.0f;
  }
  (void)hipFree(din);
  (void)hipFree(dout);
  if (st != POND_OK) return st;
  MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  m.trained = true;
  return POND_OK;
}

pond_status pond_mlp_get_params(pond_ctx* ctx, float* params, float* scales) {
  if (!ctx) return POND_ERR_INVALID;
  MlpCtx& m = ctx->mlp;
  if (!m.params) { ctx->err = "pond_mlp_get_params: no network"; return POND_ERR_STATE; }
  MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (params) MLP_HIP(ctx, hipMemcpy(params, m.params, (size_t)kMlpParams * 4, hipMemcpyDeviceToHost));
  if (scales) {
    for (int k = 0; k < 5; k++) scales[k] = m.in_inv[k];
    for (int k = 0; k < 4; k++) scales[5 + k] = m.out_scale[k];
  }
  return POND_OK;
}

pond_status pond_mlp_set_params(pond_ctx* ctx, const float* params, const float* scales) {
  if (!ctx || !params) return POND_ERR_INVALID;
  pond_status s = mlp_alloc(ctx);
  if (s != POND_OK) return s;
  MlpCtx& m = ctx->mlp;
  MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));  // no step may still read the network
  MLP_HIP(ctx, hipMemcpy(m.params, params, (size_t)kMlpParams * 4, hipMemcpyHostToDevice));
  if (scales) {
    for (int k = 0; k < 5; k++) m.in_inv[k] = scales[k];
    for (int k = 0; k < 4; k++) {
      m.out_scale[k] = scales[5 + k];
      m.out_inv[k] = scales[5 + k] != 0 ? 1.0f / scales[5 + k] : 1.0f;
    }
  }
  s = mlp_refresh_bf16(ctx);
  if (s != POND_OK) return s;
  MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  m.initialized = true;
  m.trained = true;
  return POND_OK;
}

pond_status pond_mlp_predict(pond_ctx* ctx, float* out) {
  if (!ctx || !out) return POND_ERR_INVALID;
  MlpCtx& m = ctx->mlp;
  if (!m.params) { ctx->err = "pond_mlp_predict: no network"; return POND_ERR_STATE; }
  if (!ctx->split || !ctx->has_state) { ctx->err = "pond_mlp_predict: needs a POND_SPLIT context with a state"; return POND_ERR_STATE; }
  const Grid& g = ctx->g;
  float* d = nullptr;
  MLP_HIP(ctx, hipMalloc(&d, (size_t)4 * g.plane * sizeof(float)));
  hipError_t e = mlp_launch_infer(pond_mlp_weights(ctx), state_view(ctx, ctx->cur), d, g.plane, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  for (int k = 0; k < 4 && e == hipSuccess; k++)
    e = hipMemcpy(out + (size_t)k * g.n, d + (size_t)k * g.plane, (size_t)g.n * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) {
    ctx->err = std::string("pond_mlp_predict: ") + hipGetErrorString(e);
    return POND_ERR_HIP;
  }
  return POND_OK;
}

pond_status pond_mlp_forward(pond_ctx* ctx, const float* X, int64_t B, float* Y) {
  if (!ctx || !X || !Y || B < 1) return POND_ERR_INVALID;
  MlpCtx& m = ctx->mlp;
  if (!m.params) { ctx->err = "pond_mlp_forward: no network"; return POND_ERR_STATE; }
  float *dX = nullptr, *dY = nullptr;
  MLP_HIP(ctx, hipMalloc(&dX, (size_t)B * kMlpIn * 4));
  MLP_HIP(ctx, hipMalloc(&dY, (size_t)B * kMlpOut * 4));
  std::vector<float> tmp((size_t)B * kMlpOut);
  hipError_t e = hipMemcpy(dX, X, (size_t)B * kMlpIn * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = mlp_launch_infer_matrix(pond_mlp_weights(ctx), dX, B, dY, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(tmp.data(), dY, tmp.size() * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dX);
  (void)hipFree(dY);
  if (e != hipSuccess) {
    ctx->err = std::string("pond_mlp_forward: ") + hipGetErrorString(e);
    return POND_ERR_HIP;
  }
  // output planes [o][b] -> rows [b][o]
  for (int64_t b = 0; b < B; b++)
    for (int o = 0; o < kMlpOut; o++) Y[b * kMlpOut + o] = tmp[(size_t)o * B + b];
  return POND_OK;
}

pond_status pond_mlp_grad(pond_ctx* ctx, const float* X, const float* T, int64_t B, float* grad, float* loss) {
  if (!ctx || !X || !T || !grad || B < 16 || B % 16 != 0) return POND_ERR_INVALID;
  MlpCtx& m = ctx->mlp;
  if (!m.params) { ctx->err = "pond_mlp_grad: no network"; return POND_ERR_STATE; }
  pond_status s = mlp_batch_alloc(ctx, (int)B);
  if (s != POND_OK) return s;
  MLP_HIP(ctx, hipMemcpy(m.bx, X, (size_t)B * kMlpIn * 4, hipMemcpyHostToDevice));
  MLP_HIP(ctx, hipMemcpy(m.bt, T, (size_t)B * kMlpOut * 4, hipMemcpyHostToDevice));
  s = mlp_fwd_bwd(ctx, (int)B, true);
  if (s != POND_OK) return s;
  MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  MLP_HIP(ctx, hipMemcpy(grad, m.grad, (size_t)kMlpParams * 4, hipMemcpyDeviceToHost));
  if (loss) {
    float l = 0;
    MLP_HIP(ctx, hipMemcpy(&l, m.loss, sizeof(float), hipMemcpyDeviceToHost));
    *loss = l / (float)B;
  }
  return POND_OK;
}

// pond_mlp_forward without any allocation: the device staging of X and Y lives in a static
// device array (kNoAllocRows rows at most), shared by every call on the same GPU, so calls that
// overlap on one GPU race.  Run with the context's device current.
pond_status pond_mlp_forward_noalloc(pond_ctx* ctx, const float* X, int64_t B, float* Y) {
  if (!ctx || !X || !Y || B < 1) return POND_ERR_INVALID;
  if (B > kNoAllocRows) {
    ctx->err = "pond_mlp_forward_noalloc: B = " + std::to_string(B) + " exceeds the maximum of " +
               std::to_string(kNoAllocRows) + " rows";
    return POND_ERR_INVALID;
  }
  MlpCtx& m = ctx->mlp;
  if (!m.params) { ctx->err = "pond_mlp_forward_noalloc: no network"; return POND_ERR_STATE; }
  float* base = nullptr;
  MLP_HIP(ctx, scratch_ptr(&base));
  float* dX = base;
  float* dP = base + (size_t)kNoAllocRows * kMlpIn;
  float* dR = dP + (size_t)kNoAllocRows * kMlpOut;
  MLP_HIP(ctx, hipMemcpyAsync(dX, X, (size_t)B * kMlpIn * 4, hipMemcpyHostToDevice, ctx->stream));
  MLP_HIP(ctx, mlp_launch_infer_matrix(pond_mlp_weights(ctx), dX, B, dP, ctx->stream));
  const int64_t nv = B * kMlpOut;
  hipLaunchKernelGGL(k_planes_to_rows, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, ctx->stream, dP, B, dR);
  MLP_HIP(ctx, hipGetLastError());
  MLP_HIP(ctx, hipMemcpyAsync(Y, dR, (size_t)nv * 4, hipMemcpyDeviceToHost, ctx->stream));
  MLP_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return POND_OK;
}

}  // extern "C"
