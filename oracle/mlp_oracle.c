/*
 * mlp_oracle.c -- CPU restatement of the learned gauge initializer's network.
 *
 * TEST INFRASTRUCTURE ONLY (checker of the BF16 MFMA forward/backward kernels); no
 * counterpart in the reference.
 *
 * Network 135 -> 64 -> 64 -> 4, ReLU, parameters packed in FP32:
 *   W1[135][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4]   (13124 values; W[k][j] = input k
 *   to unit j).
 * Arithmetic of the GPU kernels (mode bf16 = 1): operands of every product are rounded to
 * bfloat16 (round to nearest even) -- inputs, weights, hidden activations, and the output
 * and hidden gradients fed to the next product -- and products are accumulated in FP32/FP64;
 * biases, pre-activations and loss are FP32.  Mode bf16 = 0 is the same network in plain
 * FP64, used to pin the backward formulas by finite differences.
 * Loss = (1/B) sum_b sum_o (y_bo - t_bo)^2.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NIN 135
#define NH 64
#define NOUT 4
#define NPARAM (NIN * NH + NH + NH * NH + NH + NH * NOUT + NOUT)

static double bf(double v, int on) {
  if (!on) return v;
  float f = (float)v;
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return f; /* inf / nan unchanged */
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return f;
}

static double f32(double v, int on) { return on ? (double)(float)v : v; }

int orc_mlp_nparam(void) { return NPARAM; }

typedef struct {
  const float *W1, *b1, *W2, *b2, *W3, *b3;
} net;

static net unpack(const float* p) {
  net n;
  n.W1 = p;
  n.b1 = n.W1 + NIN * NH;
  n.W2 = n.b1 + NH;
  n.b2 = n.W2 + NH * NH;
  n.W3 = n.b2 + NH;
  n.b3 = n.W3 + NH * NOUT;
  return n;
}

/* forward of one sample; keeps pre-activations and rounded activations for the backward */
static void fwd(const net* n, const float* x, int on, double* z1, double* a1, double* z2, double* a2, double* y) {
  for (int j = 0; j < NH; j++) {
    double s = 0;
    for (int k = 0; k < NIN; k++) s += bf(x[k], on) * bf(n->W1[k * NH + j], on);
    z1[j] = f32(f32(s, on) + n->b1[j], on);
    a1[j] = bf(z1[j] > 0 ? z1[j] : 0, on);
  }
  for (int j = 0; j < NH; j++) {
    double s = 0;
    for (int k = 0; k < NH; k++) s += a1[k] * bf(n->W2[k * NH + j], on);
    z2[j] = f32(f32(s, on) + n->b2[j], on);
    a2[j] = bf(z2[j] > 0 ? z2[j] : 0, on);
  }
  for (int o = 0; o < NOUT; o++) {
    double s = 0;
    for (int k = 0; k < NH; k++) s += a2[k] * bf(n->W3[k * NOUT + o], on);
    y[o] = f32(f32(s, on) + n->b3[o], on);
  }
}

void orc_mlp_forward(const float* params, int64_t B, const float* X, float* Y, int on) {
  net n = unpack(params);
  for (int64_t b = 0; b < B; b++) {
    double z1[NH], a1[NH], z2[NH], a2[NH], y[NOUT];
    fwd(&n, X + b * NIN, on, z1, a1, z2, a2, y);
    for (int o = 0; o < NOUT; o++) Y[b * NOUT + o] = (float)y[o];
  }
}

/* gradient of the mean-squared loss w.r.t. all parameters (packed like params) */
void orc_mlp_grad(const float* params, int64_t B, const float* X, const float* T, double* grad, double* loss,
                  int on) {
  net n = unpack(params);
  memset(grad, 0, sizeof(double) * NPARAM);
  double* gW1 = grad;
  double* gb1 = gW1 + NIN * NH;
  double* gW2 = gb1 + NH;
  double* gb2 = gW2 + NH * NH;
  double* gW3 = gb2 + NH;
  double* gb3 = gW3 + NH * NOUT;
  double L = 0;
  for (int64_t b = 0; b < B; b++) {
    const float* x = X + b * NIN;
    double z1[NH], a1[NH], z2[NH], a2[NH], y[NOUT];
    fwd(&n, x, on, z1, a1, z2, a2, y);
    double dy[NOUT], dyb[NOUT];
    for (int o = 0; o < NOUT; o++) {
      const double d = y[o] - T[b * NOUT + o];
      L += d * d;
      dy[o] = f32(2.0 * d / (double)B, on);
      dyb[o] = bf(dy[o], on);
      gb3[o] += dy[o];
    }
    double dz2[NH], dz2b[NH];
    for (int k = 0; k < NH; k++) {
      double s = 0;
      for (int o = 0; o < NOUT; o++) {
        gW3[k * NOUT + o] += a2[k] * dyb[o];
        s += dyb[o] * bf(n.W3[k * NOUT + o], on);
      }
      dz2[k] = z2[k] > 0 ? f32(s, on) : 0.0;
      dz2b[k] = bf(dz2[k], on);
      gb2[k] += dz2[k];
    }
    double dz1[NH], dz1b[NH];
    for (int k = 0; k < NH; k++) {
      double s = 0;
      for (int j = 0; j < NH; j++) {
        gW2[k * NH + j] += a1[k] * dz2b[j];
        s += dz2b[j] * bf(n.W2[k * NH + j], on);
      }
      dz1[k] = z1[k] > 0 ? f32(s, on) : 0.0;
      dz1b[k] = bf(dz1[k], on);
      gb1[k] += dz1[k];
    }
    for (int k = 0; k < NIN; k++)
      for (int j = 0; j < NH; j++) gW1[k * NH + j] += bf(x[k], on) * dz1b[j];
  }
  *loss = L / (double)B;
}
