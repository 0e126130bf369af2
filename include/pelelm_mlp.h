/*
 * pelelm_mlp -- the learned Newton predictor of the MI355X reactor, C ABI.
 *
 * No counterpart in the reference: PeleLM's reactor (Source/PeleLM.cpp:5752-5758) has
 * no learned component.  This is the network the reactor's chosen form uses as the
 * initial iterate of each BDF Newton iteration, and its GPU trainer:
 *   n_in -> 128 -> 128 -> 128 -> n_out, GELU on the hidden layers, BF16 weights and
 *   activations on the matrix cores with FP32 accumulation, FP32 master weights,
 *   mean-squared-error loss, Adam (beta1 0.9, beta2 0.999, eps 1e-8).
 * Device arrays are row-major [rows][features] FP32; work is enqueued on the given
 * HIP stream.  Return codes: 0 ok, < 0 error (mi_mlp_last_error describes it).
 */
#ifndef PELELM_MLP_H
#define PELELM_MLP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MI_MLP_OK 0
#define MI_MLP_EINVAL -1
#define MI_MLP_EHIP -3
#define MI_MLP_ENOMEM -5

/* New network with He-uniform weights drawn from `seed`; 1 <= n_in, n_out <= 64. */
int mi_mlp_create(int n_in, int n_out, unsigned long long seed, void** mlp);

/* Number of parameters; the flat order is, per layer l = 0..3, W_l[fan_out][fan_in]
   row-major followed by b_l[fan_out]. */
int mi_mlp_nparams(void* mlp);
int mi_mlp_get_params(void* mlp, float* host, int n);
int mi_mlp_set_params(void* mlp, const float* host, int n);

/* Feature normalisation the reactor's predictor applies: x_n = (x - in_mean) / in_std,
   y = out_mean + out_std * network(x_n).  Host arrays of n_in / n_out doubles. */
int mi_mlp_set_normalisation(void* mlp, const double* in_mean, const double* in_std, const double* out_mean,
                             const double* out_std);

/* Y[B][n_out] = network(X[B][n_in]) on normalised features (device arrays). */
int mi_mlp_forward(void* mlp, const float* X, float* Y, long B, void* hip_stream);

/* Gradient of L = sum (network(X) - Yt)^2 / (B n_out) with respect to every parameter
   (flat order above) into the device array grad[nparams]; *loss (host, nullable)
   receives L and then the call synchronises the stream. */
int mi_mlp_grad(void* mlp, const float* X, const float* Yt, long B, float* grad, float* loss, void* hip_stream);

/* One Adam step on that loss with learning rate lr. */
int mi_mlp_train_step(void* mlp, const float* X, const float* Yt, long B, float lr, float* loss, void* hip_stream);

/* Attach a network to a reactor; a null network detaches it.  Each BDF Newton
   attempt then starts from the network's prediction of the state over the attempted
   step, with the Nordsieck prediction kept as the error-test reference and as the
   fallback when the Newton iteration fails.  Shape: (nspec + 3) inputs (T, log p,
   log10 Y_k, log h) and (nspec + 1) outputs (dT, dY_k), normalised as set by
   mi_mlp_set_normalisation.  Needs mi.schedule = queue.  The network is not owned. */
int mi_chem_attach_predictor(void* reactor, void* mlp);

const char* mi_mlp_last_error(void* mlp);
int mi_mlp_close(void* mlp);

#ifdef __cplusplus
}
#endif
#endif
