/*
 * rx_mlp.h -- the learned warm start's network (no reference counterpart).
 *
 * 16 -> 64 -> 64 -> 1 with SiLU (inputs: the 12 warm-start features of RX_ARRAY_FEATURES plus 4 zero pads),
 * evaluated and trained with bf16 MFMA kernels that accumulate in FP32.  Inputs are standardised with
 * (x - mu) / sd inside the kernels and the output is mapped back with y * out_scale + out_bias.
 * Parameters (FP32 master copy) in this order: W1[64][16], b1[64], W2[64][64], b2[64], w3[64], b3.
 * Entry points returning int give 0 on success.
 */
#ifndef RX_MLP_H_
#define RX_MLP_H_

#ifdef __cplusplus
extern "C" {
#endif

int rxMLP_NumParams(void);
/* Returns NULL (and records the error) when a device allocation fails. */
void* rxMLP_Create(unsigned seed);
void rxMLP_Destroy(void* mlp);
int rxMLP_SetParams(void* mlp, const float* params);
int rxMLP_GetParams(void* mlp, float* params);
/* in_mu, in_sd: 12 values (features 12..15 are padding and always 0). */
int rxMLP_SetNormalization(void* mlp, const float* in_mu, const float* in_sd, float out_scale, float out_bias);
/* Y[i] = network(X[i]) for n rows of 16 floats (host arrays). */
int rxMLP_Forward(void* mlp, const float* X, int n, float* Y);
/* Gradients of 0.5 mean_i (y_i - T_i)^2 over n rows (T in standardised units), parameter order above. */
int rxMLP_Gradients(void* mlp, const float* X, const float* T, int n, float* grads, float* loss);
/* Adam training on the GPU (minibatches of `batch` rows, cosine learning-rate schedule); loss_history (may be
 * NULL) receives the last minibatch loss of each epoch. */
int rxMLP_Train(void* mlp, const float* X, const float* T, int n, int epochs, int batch, float lr, unsigned seed,
                float* loss_history);

#ifdef __cplusplus
}
#endif

#endif /* RX_MLP_H_ */
