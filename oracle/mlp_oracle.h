/*
 * CPU oracle for the BC7 candidate-ranking MLP.  TEST INFRASTRUCTURE ONLY:
 * only tests/ and the benchmark's checks load it.  See mlp_oracle.c.
 */
#ifndef FASTC_MLP_ORACLE_H_
#define FASTC_MLP_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_MLP_IN = 64, ORA_MLP_H = 256, ORA_MLP_OUT = 512 };
#define ORA_MLP_NPARAMS (64 * 256 + 256 + 256 * 256 + 256 + 256 * 512 + 512)

float ora_bf16_round(float x);
int ora_mlp_valid(int logit);
void ora_mlp_features(const uint32_t *px, int nblocks, float *x);
void ora_mlp_forward(const float *params, const float *x, int B, int round_bf16, float *h1,
                     float *h2, float *logits);
double ora_mlp_xent(const float *logits, const int32_t *labels, int B, float *dlogits);
double ora_mlp_loss_grad(const float *params, const float *x, const int32_t *labels, int B,
                         int round_bf16, float *grad);
void ora_adam_step(float *w, float *m, float *v, const float *g, int n, int t, float lr,
                   float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif

#endif
