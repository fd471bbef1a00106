/*
 * flip3d_unet.h - C ABI of the learned coarse-level pressure solver.
 *
 * No counterpart in the reference: the reference solves the pressure system
 * with PCG only (src/solver.cpp:183-218).  This network is the coarse solve of
 * the FLIP_PRECOND_LEARNED two-grid preconditioner: a 4-level 3-D U-Net
 * (channels 16-32-64-64, BF16 weights and activations, FP32 accumulation on
 * the MFMA units) that maps a right-hand side b and the cell flags / level set
 * (A, L) of an R^3 pressure system (R a multiple of 8, spacing 1/R) to an
 * approximate solution x of the 7-point system the solver builds.
 *
 * Training: flip_unet_train_step takes S systems (A int8, L, b, x as flat
 * R^3 arrays, [i][j][k] order, concatenated) whose x the plain GPU solver
 * produced, and takes one Adam step on the loss
 *   mean over samples of |y - t|_A^2 / |t|_A^2
 * in the unit-spacing operator A, t the target in network units.
 * Calls return 0 or a negative FLIP_E_* code (message: flip_gpu_last_error).
 */
#ifndef FLIP3D_UNET_H
#define FLIP3D_UNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct flip_unet flip_unet;

/* A network for R^3 inputs and up to `capacity` (1..64) samples per call,
 * weights He-initialised from `seed`. */
int flip_unet_create(int R, int capacity, uint64_t seed, flip_unet **out);
int flip_unet_destroy(flip_unet *net);
int64_t flip_unet_num_params(flip_unet *net);
/* fp32 master weights, layer order: 7 convs (weights [27][cout][cin], then
 * bias [cout]) and the head (16 weights, 1 bias) */
int flip_unet_get_params(flip_unet *net, float *w);
int flip_unet_set_params(flip_unet *net, const float *w);
/* Raw network on given input features (S * R^3 * 8 floats, channels last)
 * and output mask (S * R^3); out: S * R^3 in network units. */
int flip_unet_forward_features(flip_unet *net, int S, const float *features, const float *mask, float *out);
/* Gradient of a loss with dL/dout = gout (S * R^3) after the last forward
 * call; grads: one float per parameter. */
int flip_unet_backward(flip_unet *net, int S, const float *gout, float *grads);
/* One Adam step on S systems; loss: mean relative A-norm error before the step. */
int flip_unet_train_step(flip_unet *net, int S, const int8_t *A, const float *L, const float *b, const float *x,
                         float lr, double *loss);
/* x ~ A^-1 b for S systems. */
int flip_unet_predict(flip_unet *net, int S, const int8_t *A, const float *L, const float *b, float *x,
                      double *reserved);
/* predict() for one system (S = 1) through the solver's inference kernels:
 * the same network with its upsample-adds and head fused into the
 * convolutions.  The result equals predict()'s bit for bit; no counterpart
 * in the reference (the learned solver is this library's own). */
int flip_unet_predict_fused(flip_unet *net, const int8_t *A, const float *L, const float *b, float *x);
/* The network as the learned solver runs it on one system (S = 1), with
 * sigma from per-256-cell partials of b: path 0 through separate kernels
 * (features, seven convolutions, three poolings), path 1 through the seven
 * layer kernels that stage each brick's halo in LDS and fuse the features,
 * poolings and head.  The two paths agree bit for bit; no counterpart in the
 * reference (the learned solver is this library's own). */
int flip_unet_solve_system(flip_unet *net, int path, const int8_t *A, const float *L, const float *b, float *x);
/* Weights file: "FLIPUNET", int32 version 1, int32 count, count floats. */
int flip_unet_save(flip_unet *net, const char *path);
int flip_unet_load(flip_unet *net, const char *path);

#ifdef __cplusplus
}
#endif

#endif
