/*
 * Learned start-point proposer: a small multilayer perceptron trained on the
 * GPU with hand-written BF16 matrix-core (MFMA) kernels, FP32 accumulation
 * and FP32 master weights (Adam).  It learns (parameters -> NLL, gradient)
 * from samples the likelihood itself produces and ranks random fit starting
 * points; every proposal is then refit with the FP64 likelihood.
 *
 * No counterpart in the reference: its fits start from random points
 * (src/execute.cpp:194-284, RooTDalitz/Setup.h:323-474) without a model of
 * the likelihood surface.
 *
 * Network: n_in -> [Linear(H) -> GELU] x n_layers -> Linear(n_out); GELU is
 * 0.5 z (1 + erf(z / sqrt 2)).  Loss: mean over the batch of
 * sum_j out_weight[j] (out_j - y_j)^2.
 * Parameter vector layout (pwa_mlp_get_params): for each layer in order, W
 * (fan_in x fan_out, row-major) then b (fan_out).
 * All pointers are host pointers; status codes as in pwa_amp.h.
 */
#ifndef PWA_MLP_H
#define PWA_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pwa_mlp pwa_mlp;

int pwa_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out, uint64_t seed, int device,
                   pwa_mlp** out);
int pwa_mlp_destroy(pwa_mlp* m);
const char* pwa_mlp_last_error(const pwa_mlp* m);

/* number of trainable parameters */
int pwa_mlp_n_params(const pwa_mlp* m, int64_t* n);
int pwa_mlp_get_params(const pwa_mlp* m, float* w);
int pwa_mlp_set_params(pwa_mlp* m, const float* w);

/* training set: x[n * n_in], y[n * n_out] (row-major), out_weight[n_out] */
int pwa_mlp_set_data(pwa_mlp* m, int64_t n, const float* x, const float* y, const float* out_weight);

/* loss and its gradient with respect to every parameter (grad may be NULL)
   on training samples [first, first + batch), without an update */
int pwa_mlp_loss_grad(pwa_mlp* m, int64_t first, int32_t batch, float* loss, float* grad);

/* Adam (beta 0.9 / 0.999, eps 1e-8) over shuffled minibatches; mean batch
   loss of the last epoch in *loss */
int pwa_mlp_train(pwa_mlp* m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* loss);

/* y[n * n_out] = network(x[n * n_in]) */
int pwa_mlp_predict(pwa_mlp* m, int64_t n, const float* x, float* y);

#ifdef __cplusplus
}
#endif

#endif
