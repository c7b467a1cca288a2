/*
 * CPU oracle for the BC7 candidate-ranking MLP (64 -> 256 -> 256 -> 512).
 *
 * TEST INFRASTRUCTURE ONLY.  The network has no counterpart in the
 * reference; this file is the plain-C statement of what the HIP kernels
 * compute, and the tests check the kernels against it.  Pinned by a
 * finite-difference check of its own gradients (exact arithmetic mode) and by
 * closed forms for softmax cross-entropy and Adam.
 *
 * Numerics mirror the GPU: with round_bf16 set, inputs, weights and the
 * hidden activations are rounded to bfloat16 (round to nearest even) where
 * the kernels feed them to bf16 matrix products, and every product
 * accumulates in double here (fp32 on the GPU, so comparisons use a
 * tolerance).  Parameters are one flat float array:
 *   W1[64][256] b1[256] W2[256][256] b2[256] W3[256][512] b3[512]
 * with W[in][out].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mlp_oracle.h"

float ora_bf16_round(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) return x; /* inf / nan unchanged */
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  u &= 0xFFFF0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

static float rb(float x, int on) { return on ? ora_bf16_round(x) : x; }

/* Logit index = mode * 64 + slot.  Slots: the partition shape for modes
 * 0-3 and 7 (mode 0 has 16), rotation * 2 + index mode for mode 4, the
 * rotation for mode 5, and slot 0 for mode 6. */
int ora_mlp_valid(int logit) {
  const int mode = logit >> 6, slot = logit & 63;
  switch (mode) {
    case 0: return slot < 16;
    case 4: return slot < 8;
    case 5: return slot < 4;
    case 6: return slot == 0;
    default: return 1;
  }
}

static void layer(const float *in, int B, int nin, int nout, const float *W, const float *b,
                  float *out, int relu, int round_in, int round_w) {
  for (int r = 0; r < B; r++) {
    for (int o = 0; o < nout; o++) {
      double acc = 0.0;
      for (int i = 0; i < nin; i++)
        acc += (double)rb(in[r * nin + i], round_in) * (double)rb(W[i * nout + o], round_w);
      float v = (float)(acc + (double)b[o]);
      if (relu && v < 0.0f) v = 0.0f;
      out[r * nout + o] = v;
    }
  }
}

void ora_mlp_features(const uint32_t *px, int nblocks, float *x) {
  for (int b = 0; b < nblocks; b++)
    for (int i = 0; i < 16; i++)
      for (int c = 0; c < 4; c++)
        x[b * 64 + i * 4 + c] =
            ora_bf16_round((float)((px[b * 16 + i] >> (8 * c)) & 0xFF) / 255.0f);
}

void ora_mlp_forward(const float *params, const float *x, int B, int round_bf16, float *h1,
                     float *h2, float *logits) {
  const float *W1 = params, *b1 = W1 + ORA_MLP_IN * ORA_MLP_H;
  const float *W2 = b1 + ORA_MLP_H, *b2 = W2 + ORA_MLP_H * ORA_MLP_H;
  const float *W3 = b2 + ORA_MLP_H, *b3 = W3 + ORA_MLP_H * ORA_MLP_OUT;
  layer(x, B, ORA_MLP_IN, ORA_MLP_H, W1, b1, h1, 1, round_bf16, round_bf16);
  layer(h1, B, ORA_MLP_H, ORA_MLP_H, W2, b2, h2, 1, round_bf16, round_bf16);
  layer(h2, B, ORA_MLP_H, ORA_MLP_OUT, W3, b3, logits, 0, round_bf16, round_bf16);
}

/* Mean softmax cross-entropy over the valid logits, and its gradient
 * dlogits = (softmax - onehot) / B (zero on invalid logits). */
double ora_mlp_xent(const float *logits, const int32_t *labels, int B, float *dlogits) {
  double loss = 0.0;
  for (int r = 0; r < B; r++) {
    const float *z = logits + r * ORA_MLP_OUT;
    double mx = -INFINITY;
    for (int j = 0; j < ORA_MLP_OUT; j++)
      if (ora_mlp_valid(j) && z[j] > mx) mx = z[j];
    double s = 0.0;
    for (int j = 0; j < ORA_MLP_OUT; j++)
      if (ora_mlp_valid(j)) s += exp((double)z[j] - mx);
    loss += -(((double)z[labels[r]] - mx) - log(s));
    if (dlogits) {
      for (int j = 0; j < ORA_MLP_OUT; j++) {
        double p = ora_mlp_valid(j) ? exp((double)z[j] - mx) / s : 0.0;
        if (j == labels[r]) p -= 1.0;
        dlogits[r * ORA_MLP_OUT + j] = (float)(p / B);
      }
    }
  }
  return loss / B;
}

/* dW[i][o] = sum_r in[r][i] * dy[r][o]; db[o] = sum_r dy[r][o];
 * din[r][i] = sum_o dy[r][o] * W[i][o] (masked by relu'(in) if requested). */
static void layer_back(const float *in, const float *dy, int B, int nin, int nout,
                       const float *W, float *dW, float *db, float *din, const float *relu_of,
                       int round_on) {
  for (int i = 0; i < nin; i++)
    for (int o = 0; o < nout; o++) {
      double acc = 0.0;
      for (int r = 0; r < B; r++)
        acc += (double)rb(in[r * nin + i], round_on) * (double)rb(dy[r * nout + o], round_on);
      dW[i * nout + o] = (float)acc;
    }
  for (int o = 0; o < nout; o++) {
    double acc = 0.0;
    for (int r = 0; r < B; r++) acc += (double)dy[r * nout + o];
    db[o] = (float)acc;
  }
  if (!din) return;
  for (int r = 0; r < B; r++)
    for (int i = 0; i < nin; i++) {
      double acc = 0.0;
      for (int o = 0; o < nout; o++)
        acc += (double)rb(dy[r * nout + o], round_on) * (double)rb(W[i * nout + o], round_on);
      float v = (float)acc;
      if (relu_of && relu_of[r * nin + i] <= 0.0f) v = 0.0f;
      din[r * nin + i] = v;
    }
}

double ora_mlp_loss_grad(const float *params, const float *x, const int32_t *labels, int B,
                         int round_bf16, float *grad) {
  float *h1 = malloc(sizeof(float) * B * ORA_MLP_H);
  float *h2 = malloc(sizeof(float) * B * ORA_MLP_H);
  float *lg = malloc(sizeof(float) * B * ORA_MLP_OUT);
  float *dl = malloc(sizeof(float) * B * ORA_MLP_OUT);
  float *d2 = malloc(sizeof(float) * B * ORA_MLP_H);
  float *d1 = malloc(sizeof(float) * B * ORA_MLP_H);
  ora_mlp_forward(params, x, B, round_bf16, h1, h2, lg);
  const double loss = ora_mlp_xent(lg, labels, B, dl);
  if (grad) {
    const float *W2 = params + ORA_MLP_IN * ORA_MLP_H + ORA_MLP_H;
    const float *W3 = W2 + ORA_MLP_H * ORA_MLP_H + ORA_MLP_H;
    float *gW1 = grad, *gb1 = gW1 + ORA_MLP_IN * ORA_MLP_H;
    float *gW2 = gb1 + ORA_MLP_H, *gb2 = gW2 + ORA_MLP_H * ORA_MLP_H;
    float *gW3 = gb2 + ORA_MLP_H, *gb3 = gW3 + ORA_MLP_H * ORA_MLP_OUT;
    layer_back(h2, dl, B, ORA_MLP_H, ORA_MLP_OUT, W3, gW3, gb3, d2, h2, round_bf16);
    layer_back(h1, d2, B, ORA_MLP_H, ORA_MLP_H, W2, gW2, gb2, d1, h1, round_bf16);
    layer_back(x, d1, B, ORA_MLP_IN, ORA_MLP_H, params, gW1, gb1, NULL, NULL, round_bf16);
  }
  free(h1);
  free(h2);
  free(lg);
  free(dl);
  free(d2);
  free(d1);
  return loss;
}

/* Adam (Kingma & Ba, 2015) with bias correction, step t >= 1. */
void ora_adam_step(float *w, float *m, float *v, const float *g, int n, int t, float lr,
                   float beta1, float beta2, float eps) {
  const double c1 = 1.0 - pow((double)beta1, t), c2 = 1.0 - pow((double)beta2, t);
  for (int i = 0; i < n; i++) {
    m[i] = beta1 * m[i] + (1.0f - beta1) * g[i];
    v[i] = beta2 * v[i] + (1.0f - beta2) * g[i] * g[i];
    const double mh = m[i] / c1, vh = v[i] / c2;
    w[i] = (float)((double)w[i] - (double)lr * mh / (sqrt(vh) + (double)eps));
  }
}
