/*
 * alamo_unet.h -- C-ABI of the learned coarse-level solver of
 * libalamo_elastic.so: a 3D U-Net on a nodal grid (four levels, channels
 * 16-32-64-128, 3x3x3 convolutions, full-weighting down / trilinear up
 * sampling, skip connections by concatenation), bfloat16 weights and
 * activations on the gfx950 matrix cores with float32 accumulation, trained
 * by its own forward and backward kernels.
 *
 * The reference has no learned component; nothing here replaces a reference
 * interface.  The network supplies the initial guess and a coarse-level
 * correction of the elastic solve (alamo_el_set_learned in alamo_elastic.h).
 *
 * Conventions follow alamo_elastic.h: int status (0 ok, negative errors,
 * alamo_el_last_error() describes the last one), device pointers unless a
 * name ends in _host, every launch on the stream passed.  Fields are float32
 * component-major: v[c * N + node], node index i fastest.
 *
 * Parameters are one float32 vector, per convolution in network order
 * (encoder level 0..3: conv A, conv B; decoder level 2..0: conv A, conv B;
 * output conv): W[Cout][27][Cin] then bias[Cout]; tap t = (dz+1)*9 + (dy+1)*3
 * + (dx+1) reads the input node (x+dx, y+dy, z+dz).
 */
#ifndef ALAMO_UNET_H
#define ALAMO_UNET_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct alamo_nn_s* alamo_nn_t;

/* nodes[3] = (nx, ny, nz) of the finest network level; every nodes[a] - 1
 * must be divisible by 8 (three nodal coarsenings).  n_in input and n_out
 * output components. */
int alamo_nn_create(const int nodes[3], int n_in, int n_out, alamo_nn_t* out);
int alamo_nn_destroy(alamo_nn_t nn);
int alamo_nn_num_params(alamo_nn_t nn, long* n);
int alamo_nn_shape(alamo_nn_t nn, int nodes[3], int* n_in, int* n_out);
/* He-normal weights and zero biases from a seeded host generator. */
int alamo_nn_init_params(alamo_nn_t nn, unsigned long long seed, hipStream_t stream);
int alamo_nn_set_params(alamo_nn_t nn, const float* params, hipStream_t stream);
int alamo_nn_get_params(alamo_nn_t nn, float* params, hipStream_t stream);

/* out[n_out][N] = net(in[n_in][N]); keeps the activations for
 * alamo_nn_backward. */
int alamo_nn_forward(alamo_nn_t nn, const float* in, float* out, hipStream_t stream);
/* After alamo_nn_forward: gradients of <out, dout> with respect to the
 * parameters (dparams, num_params floats) and, when din is not NULL, the input
 * (din[n_in][N]). */
int alamo_nn_backward(alamo_nn_t nn, const float* dout, float* dparams, float* din, hipStream_t stream);
/* After alamo_nn_forward: the activations it kept for convolution `conv`
 * (0 .. 14, in parameter order), bf16 values widened to float32 -- its input
 * x[cin][nodes of its level] and, when y is not NULL, its output after the
 * ReLU y[cout][nodes] (ALAMO_EL_EINVAL for the output layer, which has no
 * ReLU and keeps no bf16 output).  No counterpart in the reference (the
 * network is this build's own): it lets a test run a backward pass on the
 * device's own forward tape, so that the comparison is not blurred by ReLU
 * masks that bf16 rounding flipped in the forward pass. */
int alamo_nn_tape(alamo_nn_t nn, int conv, float* x, float* y, hipStream_t stream);
/* loss = mean((out - target)^2) over n_out * N values; dout (may be NULL)
 * receives its gradient 2 (out - target) / (n_out N). */
int alamo_nn_mse(alamo_nn_t nn, const float* out, const float* target, float* dout, double* loss_host,
                 hipStream_t stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) of the parameters with
 * gradient dparams; the moment estimates live in the handle. */
int alamo_nn_adam(alamo_nn_t nn, const float* dparams, float lr, hipStream_t stream);
/* forward, mse, backward and adam in one call; loss_host gets the loss
 * before the update. */
int alamo_nn_train_step(alamo_nn_t nn, const float* in, const float* target, float lr, double* loss_host,
                        hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
