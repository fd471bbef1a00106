// unet_internal.hpp - the learned coarse-level solver: a 4-level 3-D U-Net
// (channels 16-32-64-64) on R^3 cubes, BF16 weights and activations with
// FP32 accumulation, trained and run by the kernels of k_conv.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

// ---- k_conv.hip launchers ----
void unet_conv3_fwd(int cin, int cout, const uint16_t *X, const uint16_t *W, const float *bias, uint16_t *Y, float *Yf,
                    int R, long nvox, bool relu, hipStream_t s);
// conv(up(U) + X) with leaky ReLU, the upsampled sum formed on the fly (the
// bf16 result equals conv of k_up_add's output)
void unet_conv3_fwd_up(int cin, int cout, const uint16_t *U, const uint16_t *X, const uint16_t *W, const float *bias,
                       uint16_t *Y, int R, long nvox, hipStream_t s);
// last layer conv(up(U) + X) (16 -> 16) with the head and the unit change
// fused: x = mask (hw . y + hb) sigma scale
void unet_conv3_fwd_up_head(const uint16_t *U, const uint16_t *X, const uint16_t *W, const float *bias,
                            const float *hw, const float *hb, const float *mask, const float *sigma, float scale,
                            float *x, int R, long nvox, hipStream_t s);
void unet_conv3_wgrad(int cin, int cout, const uint16_t *X, const uint16_t *dY, float *partial, float *dW, int R,
                      long nvox, hipStream_t s);
int unet_wgrad_chunks(long nvox);
void unet_bias_grad(const uint16_t *dY, int C, long nvox, float *partial, float *db, hipStream_t s);
void unet_weight_flip(const uint16_t *W, int cout, int cin, uint16_t *Wt, hipStream_t s);
void unet_relu_back(long n, const uint16_t *Y, uint16_t *dY, hipStream_t s);
void unet_pool2(const uint16_t *in, int R, int C, long nout, uint16_t *out, hipStream_t s);
void unet_pool2_back(const uint16_t *dout, int R, int C, long nin, uint16_t *din, hipStream_t s);
void unet_up_add(const uint16_t *coarse, const uint16_t *skip, int R, int C, long nfine, uint16_t *out, hipStream_t s);
void unet_up_back(const uint16_t *dout, int R, int C, long ncoarse, uint16_t *dcoarse, hipStream_t s);
void unet_add(long n, uint16_t *a, const uint16_t *b, hipStream_t s);
void unet_head_fwd(const uint16_t *X, int C, const float *w, const float *b, const float *mask, long nvox, float *out,
                   hipStream_t s);
int unet_head_blocks(long nvox);
void unet_head_back(const uint16_t *X, int C, const float *w, const float *mask, const float *g, long nvox, uint16_t *dX,
                    float *partial, float *dwb, hipStream_t s);
// Adam step with the gradient clipped to global norm `clip` (partial: 256
// doubles of scratch, gscale: 1 float of scratch)
void unet_adam(long n, float *w, const float *g, float *m, float *v, float lr, int t, float clip, double *partial,
               float *gscale, uint16_t *wbf, hipStream_t s);
void unet_to_bf16(long n, const float *a, uint16_t *b, hipStream_t s);
void unet_to_f32(long n, const uint16_t *a, float *b, hipStream_t s);

// ---- the network ----
constexpr int UNET_FEATURES = 8;  // input channels (5 used, padded to 8)
constexpr float UNET_LEAK = 0.1f;  // negative slope of the leaky ReLU
constexpr int UNET_NCONV = 7;
constexpr float UNET_OUT_UNIT = 64.0f;  // the network's output unit: x = out UNET_OUT_UNIT sigma / R^2
constexpr float UNET_CLIP = 1.0f;  // global gradient-norm clip of the training step

struct UNetConv {
    int cin, cout, level;  // level: spatial side R >> level
    long w_off, b_off;     // offsets into the parameter vector
};

struct UNet {
    int R = 0;        // input cube side
    int cap = 0;      // batch capacity (samples)
    hipStream_t stream = nullptr;
    UNetConv conv[UNET_NCONV];
    long head_w = 0, head_b = 0;  // head: 16 weights + 1 bias
    long nparam = 0;
    int step = 0;  // Adam step count
    // parameters: fp32 master, gradient, Adam moments, bf16 copy, flipped bf16 copies
    float *pw = nullptr, *pg = nullptr, *pm = nullptr, *pv = nullptr;
    uint16_t *pwb = nullptr, *pwt = nullptr;
    // activations (bf16, channels-last)
    uint16_t *F0 = nullptr, *e0 = nullptr, *p0 = nullptr, *e1 = nullptr, *p1 = nullptr, *e2 = nullptr, *p2 = nullptr,
             *e3 = nullptr, *u2 = nullptr, *d2 = nullptr, *u1 = nullptr, *d1 = nullptr, *u0 = nullptr, *d0 = nullptr;
    float *out = nullptr, *mask = nullptr;
    // gradient scratch
    uint16_t *g0 = nullptr, *g1 = nullptr, *g2 = nullptr, *gA = nullptr, *gB = nullptr, *gC = nullptr;
    float *gout = nullptr, *partial = nullptr;
    long partial_len = 0;

    long vox(int level, int samples) const
    {
        long r = R >> level;
        return (long)samples * r * r * r;
    }
};

UNet *unet_new(int R, int cap, hipStream_t s);
void unet_free(UNet *u);
void unet_init_weights(UNet *u, uint64_t seed);
// forward on u->F0 / u->mask (samples <= cap); result in u->out
void unet_forward(UNet *u, int samples);
// backward of a loss with dL/dout in u->gout; fills u->pg
void unet_backward(UNet *u, int samples);
void unet_refresh_bf16(UNet *u);
// one R^3 system on the device: x ~ A^-1 r by the network (sigma: 1 float,
// part: 2 * unet_partials_count doubles of scratch)
void unet_solve(UNet *u, const int8_t *A, const float *L, const float *r, float *x, float *sigma, double *part);
// the same with the right-hand side's per-256-cell partials (sum r^2 over
// FLUID cells, FLUID count) already in part, as the producer of r wrote them
void unet_solve_partials(UNet *u, const int8_t *A, const float *L, const float *r, const double *part, float *x,
                         float *sigma);
int unet_partials_count(const UNet *u);
// the same forward in seven layer kernels (k_infer.hip): features built while
// staging the first layer, poolings and the head in the layers' epilogues;
// every kernel returns at once while *stop is nonzero (stop may be null).
// Equals unet_solve_partials bit for bit.
void unet_infer_layers(UNet *u, const int8_t *A, const float *L, const float *b, const double *part, int nparts,
                       const int32_t *stop, float *x, float *sigma);
bool unet_read_weights(const char *path, long nparam, std::vector<float> &w, std::string &err);
