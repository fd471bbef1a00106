// unet.hip - the learned coarse-level solver: network orchestration,
// training step (A-norm loss, Adam) and its C ABI.
//
// Network (R = input side, channels-last bf16, every conv 3x3x3 + bias +
// leaky ReLU with slope 0.1):
//   e0 = conv(F0: 8->16)            R        p0 = pool(e0)   R/2
//   e1 = conv(p0: 16->32)           R/2      p1 = pool(e1)   R/4
//   e2 = conv(p1: 32->64)           R/4      p2 = pool(e2)   R/8
//   e3 = conv(p2: 64->64)           R/8
//   d2 = conv(up(e3) + e2: 64->32)  R/4
//   d1 = conv(up(d2) + e1: 32->16)  R/2
//   d0 = conv(up(d1) + e0: 16->16)  R
//   out = mask * (w . d0 + b)       R   (fp32 head, 1x1x1)
// Input features per cell: r/sigma on FLUID cells, FLUID, AIR and WALL
// indicators, clamp(L, -2, 2) off walls, three zero pads; sigma is the RMS of
// the right-hand side over FLUID cells.  The output is the solution in the
// unit x * R^2 / (64 sigma).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <random>
#include <string>
#include <vector>

#include "../../include/flip3d_unet.h"
#include "flip_internal.hpp"
#include "unet_internal.hpp"

namespace {

constexpr float OUT_UNIT = UNET_OUT_UNIT;

inline unsigned nb(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

template <class T>
void dmalloc(T **p, long n)
{
    FLIP_CHECK(hipMalloc((void **)p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
}

template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

__device__ __forceinline__ float bf2f_d(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint16_t f2bf_d(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// per-sample sum of b^2 and FLUID count: grid (blocks, samples)
__global__ void k_rhs_stats(const int8_t *__restrict__ A, const float *__restrict__ b, long per, double *partial)
{
    __shared__ double s0[256], s1[256];
    const int smp = blockIdx.y;
    double a = 0.0, c = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < per; i += (long)gridDim.x * blockDim.x) {
        long g = (long)smp * per + i;
        if (A[g] == FLIP_FLUID) {
            a += (double)b[g] * b[g];
            c += 1.0;
        }
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s0[threadIdx.x] += s0[threadIdx.x + o];
            s1[threadIdx.x] += s1[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[((long)smp * gridDim.x + blockIdx.x) * 2] = s0[0];
        partial[((long)smp * gridDim.x + blockIdx.x) * 2 + 1] = s1[0];
    }
}

// sigma[s] = sqrt(sum b^2 / count) (1 when there is no FLUID cell or b = 0)
__global__ void k_rhs_sigma(const double *__restrict__ partial, int blocks, int samples, float *sigma)
{
    int s = threadIdx.x;
    if (s >= samples) return;
    double a = 0.0, c = 0.0;
    for (int k = 0; k < blocks; ++k) {
        a += partial[((long)s * blocks + k) * 2];
        c += partial[((long)s * blocks + k) * 2 + 1];
    }
    double v = c > 0.0 ? sqrt(a / c) : 0.0;
    sigma[s] = v > 0.0 ? (float)v : 1.0f;
}

__global__ void k_features(const int8_t *__restrict__ A, const float *__restrict__ L, const float *__restrict__ b,
                           const float *__restrict__ sigma, long per, long n, uint16_t *F, float *mask)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int s = (int)(g / per);
    int a = A[g];
    float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (a == FLIP_FLUID) f[0] = b[g] / sigma[s];
    f[1] = a == FLIP_FLUID;
    f[2] = a == FLIP_AIR;
    f[3] = a == FLIP_WALL;
    if (a != FLIP_WALL) f[4] = fminf(fmaxf(L[g], -2.0f), 2.0f);
    for (int c = 0; c < 8; ++c) F[g * 8 + c] = f2bf_d(f[c]);
    mask[g] = a == FLIP_FLUID ? 1.0f : 0.0f;
}

// target in network units: t = x R^2 / (64 sigma) on FLUID cells
__global__ void k_target(const int8_t *__restrict__ A, const float *__restrict__ x, const float *__restrict__ sigma,
                         long per, long n, float scale, float *t)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    t[g] = A[g] == FLIP_FLUID ? x[g] * scale / sigma[g / per] : 0.0f;
}

// network output back to the solution's unit: x = out 64 sigma / R^2
__global__ void k_unscale(const float *__restrict__ out, const float *__restrict__ sigma, long per, long n, float scale,
                          float *x)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    x[g] = out[g] * sigma[g / per] * scale;
}

// unit-spacing 7-point operator of the solver (ghost-fluid AIR, Neumann WALL
// and domain edge) on each R^3 sample: y = 6x - sum x_ref on FLUID, 0 elsewhere
__global__ void k_grid_ax(const int8_t *__restrict__ A, const float *__restrict__ L, const float *__restrict__ x, int R,
                          long n, float *y)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    if (A[g] != FLIP_FLUID) {
        y[g] = 0.0f;
        return;
    }
    int k = (int)(g % R), j = (int)((g / R) % R), i = (int)((g / ((long)R * R)) % R);
    long base = g - (((long)i * R + j) * R + k);
    float xs = x[g], Ls = L[g];
    float s = 6.0f * xs;
    const int d[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int q = 0; q < 6; ++q) {
        int a = i + d[q][0], bb = j + d[q][1], c = k + d[q][2];
        if (a < 0 || a >= R || bb < 0 || bb >= R || c < 0 || c >= R) {
            s -= xs;
            continue;
        }
        long h = base + ((long)a * R + bb) * R + c;
        int f = A[h];
        if (f == FLIP_FLUID)
            s -= x[h];
        else if (f == FLIP_WALL)
            s -= xs;
        else
            s -= L[h] / fminf(1.0e-6f, Ls) * xs;
    }
    y[g] = s;
}

// per-sample partial dots u.v and w.z: grid (blocks, samples)
__global__ void k_dots2(const float *__restrict__ u, const float *__restrict__ v, const float *__restrict__ w,
                        const float *__restrict__ z, long per, double *partial)
{
    __shared__ double s0[256], s1[256];
    const int smp = blockIdx.y;
    double a = 0.0, c = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < per; i += (long)gridDim.x * blockDim.x) {
        long g = (long)smp * per + i;
        a += (double)u[g] * v[g];
        c += (double)w[g] * z[g];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s0[threadIdx.x] += s0[threadIdx.x + o];
            s1[threadIdx.x] += s1[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[((long)smp * gridDim.x + blockIdx.x) * 2] = s0[0];
        partial[((long)smp * gridDim.x + blockIdx.x) * 2 + 1] = s1[0];
    }
}

// loss[s] = e.Ae / t.At; coef[s] = 2 / (t.At * samples)
__global__ void k_loss_fin(const double *__restrict__ partial, int blocks, int samples, double *loss, float *coef)
{
    int s = threadIdx.x;
    if (s >= samples) return;
    double a = 0.0, c = 0.0;
    for (int k = 0; k < blocks; ++k) {
        a += partial[((long)s * blocks + k) * 2];
        c += partial[((long)s * blocks + k) * 2 + 1];
    }
    double den = c > 0.0 ? c : 1.0;
    loss[s] = a / den;
    coef[s] = (float)(2.0 / (den * samples));
}

__global__ void k_sub(long n, const float *__restrict__ a, const float *__restrict__ b, float *c)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i] - b[i];
}

__global__ void k_scale_per_sample(long n, long per, const float *__restrict__ coef, const float *__restrict__ a,
                                   float *out)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = coef[i / per] * a[i];
}

struct Trainer {
    UNet *net = nullptr;
    int8_t *A = nullptr;
    float *L = nullptr, *b = nullptr, *x = nullptr, *t = nullptr, *err = nullptr, *Aerr = nullptr, *At = nullptr;
    float *sigma = nullptr, *coef = nullptr;
    double *part = nullptr, *loss = nullptr;
    double *spart = nullptr;  // per-256-cell right-hand-side partials of one system (solver inference)
    double *gpart = nullptr;  // gradient-norm partials (256)
    float *gscale = nullptr;  // clip factor of the last step
    int blocks = 64;
};

struct flip_unet_impl {
    Trainer tr;
    hipStream_t stream = nullptr;
};

int fail(int code, const std::string &m)
{
    flip_internal_fail(code, m.c_str());
    return code;
}

template <class F>
int guarded_u(F &&f)
{
    try {
        return f();
    } catch (const flip_hip_error &e) {
        return fail(FLIP_E_HIP, e.what);
    } catch (const flip_arg_error &e) {
        return fail(e.code, e.what);
    } catch (const std::exception &e) {
        return fail(FLIP_E_STATE, e.what());
    }
}

}  // namespace

// ------------------------------------------------------------ network
UNet *unet_new(int R, int cap, hipStream_t s)
{
    if (R < 8 || R % 8 != 0) throw flip_arg_error{FLIP_E_ARG, "U-Net side must be a positive multiple of 8"};
    if (cap < 1) throw flip_arg_error{FLIP_E_ARG, "U-Net batch capacity must be >= 1"};
    UNet *u = new UNet();
    u->R = R;
    u->cap = cap;
    u->stream = s;
    const int spec[UNET_NCONV][3] = {{UNET_FEATURES, 16, 0}, {16, 32, 1}, {32, 64, 2}, {64, 64, 3},
                                     {64, 32, 2},            {32, 16, 1}, {16, 16, 0}};
    long off = 0;
    for (int i = 0; i < UNET_NCONV; ++i) {
        UNetConv &c = u->conv[i];
        c.cin = spec[i][0];
        c.cout = spec[i][1];
        c.level = spec[i][2];
        c.w_off = off;
        off += 27L * c.cin * c.cout;
        c.b_off = off;
        off += c.cout;
    }
    u->head_w = off;
    off += 16;
    u->head_b = off;
    off += 1;
    u->nparam = off;
    try {
        dmalloc(&u->pw, off);
        dmalloc(&u->pg, off);
        dmalloc(&u->pm, off);
        dmalloc(&u->pv, off);
        dmalloc(&u->pwb, off);
        dmalloc(&u->pwt, off);
        FLIP_CHECK(hipMemsetAsync(u->pm, 0, sizeof(float) * off, s));
        FLIP_CHECK(hipMemsetAsync(u->pv, 0, sizeof(float) * off, s));
        const long v0 = u->vox(0, cap), v1 = u->vox(1, cap), v2 = u->vox(2, cap), v3 = u->vox(3, cap);
        dmalloc(&u->F0, v0 * UNET_FEATURES);
        dmalloc(&u->e0, v0 * 16);
        dmalloc(&u->p0, v1 * 16);
        dmalloc(&u->e1, v1 * 32);
        dmalloc(&u->p1, v2 * 32);
        dmalloc(&u->e2, v2 * 64);
        dmalloc(&u->p2, v3 * 64);
        dmalloc(&u->e3, v3 * 64);
        dmalloc(&u->u2, v2 * 64);
        dmalloc(&u->d2, v2 * 32);
        dmalloc(&u->u1, v1 * 32);
        dmalloc(&u->d1, v1 * 16);
        dmalloc(&u->u0, v0 * 16);
        dmalloc(&u->d0, v0 * 16);
        dmalloc(&u->out, v0);
        dmalloc(&u->mask, v0);
        long gmax = std::max(std::max(v0 * 16, v1 * 32), std::max(v2 * 64, v3 * 64));
        dmalloc(&u->g0, gmax);
        dmalloc(&u->g1, gmax);
        dmalloc(&u->g2, gmax);
        dmalloc(&u->gA, gmax);
        dmalloc(&u->gB, gmax);
        dmalloc(&u->gC, gmax);
        dmalloc(&u->gout, v0);
        long pl = (long)unet_wgrad_chunks(v0) * 27 * 64 * 64;
        pl = std::max(pl, (long)unet_head_blocks(v0) * 17);
        u->partial_len = pl;
        dmalloc(&u->partial, pl);
    } catch (...) {
        unet_free(u);
        throw;
    }
    return u;
}

void unet_free(UNet *u)
{
    if (!u) return;
    dfree(u->pw);
    dfree(u->pg);
    dfree(u->pm);
    dfree(u->pv);
    dfree(u->pwb);
    dfree(u->pwt);
    uint16_t **bufs[] = {&u->F0, &u->e0, &u->p0, &u->e1, &u->p1, &u->e2, &u->p2, &u->e3, &u->u2, &u->d2,
                         &u->u1, &u->d1, &u->u0, &u->d0, &u->g0, &u->g1, &u->g2, &u->gA, &u->gB, &u->gC};
    for (auto b : bufs) dfree(*b);
    dfree(u->out);
    dfree(u->mask);
    dfree(u->gout);
    dfree(u->partial);
    delete u;
}

void unet_refresh_bf16(UNet *u) { unet_to_bf16(u->nparam, u->pw, u->pwb, u->stream); }

// He-normal conv weights, zero biases, zero head (the untrained network
// outputs 0 and the loss starts at 1)
void unet_init_weights(UNet *u, uint64_t seed)
{
    std::vector<float> w(u->nparam, 0.0f);
    std::mt19937_64 g(seed);
    for (int i = 0; i < UNET_NCONV; ++i) {
        const UNetConv &c = u->conv[i];
        std::normal_distribution<float> nd(0.0f, sqrtf(2.0f / (27.0f * (float)(i == 0 ? 5 : c.cin))));
        for (long k = 0; k < 27L * c.cin * c.cout; ++k) {
            long ci = k % c.cin;
            w[c.w_off + k] = (i == 0 && ci >= 5) ? 0.0f : nd(g);
        }
    }
    FLIP_CHECK(hipMemcpy(u->pw, w.data(), sizeof(float) * u->nparam, hipMemcpyHostToDevice));
    u->step = 0;
    FLIP_CHECK(hipMemsetAsync(u->pm, 0, sizeof(float) * u->nparam, u->stream));
    FLIP_CHECK(hipMemsetAsync(u->pv, 0, sizeof(float) * u->nparam, u->stream));
    unet_refresh_bf16(u);
}

void unet_forward(UNet *u, int S)
{
    hipStream_t s = u->stream;
    const int R0 = u->R, R1 = R0 / 2, R2 = R0 / 4, R3 = R0 / 8;
    const long v0 = u->vox(0, S), v1 = u->vox(1, S), v2 = u->vox(2, S), v3 = u->vox(3, S);
    auto W = [&](int i) { return u->pwb + u->conv[i].w_off; };
    auto B = [&](int i) { return u->pw + u->conv[i].b_off; };
    unet_conv3_fwd(UNET_FEATURES, 16, u->F0, W(0), B(0), u->e0, nullptr, R0, v0, true, s);
    unet_pool2(u->e0, R0, 16, v1, u->p0, s);
    unet_conv3_fwd(16, 32, u->p0, W(1), B(1), u->e1, nullptr, R1, v1, true, s);
    unet_pool2(u->e1, R1, 32, v2, u->p1, s);
    unet_conv3_fwd(32, 64, u->p1, W(2), B(2), u->e2, nullptr, R2, v2, true, s);
    unet_pool2(u->e2, R2, 64, v3, u->p2, s);
    unet_conv3_fwd(64, 64, u->p2, W(3), B(3), u->e3, nullptr, R3, v3, true, s);
    unet_up_add(u->e3, u->e2, R2, 64, v2, u->u2, s);
    unet_conv3_fwd(64, 32, u->u2, W(4), B(4), u->d2, nullptr, R2, v2, true, s);
    unet_up_add(u->d2, u->e1, R1, 32, v1, u->u1, s);
    unet_conv3_fwd(32, 16, u->u1, W(5), B(5), u->d1, nullptr, R1, v1, true, s);
    unet_up_add(u->d1, u->e0, R0, 16, v0, u->u0, s);
    unet_conv3_fwd(16, 16, u->u0, W(6), B(6), u->d0, nullptr, R0, v0, true, s);
    unet_head_fwd(u->d0, 16, u->pw + u->head_w, u->pw + u->head_b, u->mask, v0, u->out, s);
}

void unet_backward(UNet *u, int S)
{
    hipStream_t s = u->stream;
    const int R0 = u->R, R1 = R0 / 2, R2 = R0 / 4, R3 = R0 / 8;
    const long v0 = u->vox(0, S), v1 = u->vox(1, S), v2 = u->vox(2, S), v3 = u->vox(3, S);
    float *G = u->pg;
    auto Wt = [&](int i) { return u->pwt + u->conv[i].w_off; };
    for (int i = 1; i < UNET_NCONV; ++i)
        unet_weight_flip(u->pwb + u->conv[i].w_off, u->conv[i].cout, u->conv[i].cin, Wt(i), s);
    auto grads = [&](int i, const uint16_t *X, const uint16_t *dY, int R, long nv) {
        const UNetConv &c = u->conv[i];
        unet_conv3_wgrad(c.cin, c.cout, X, dY, u->partial, G + c.w_off, R, nv, s);
        unet_bias_grad(dY, c.cout, nv, u->partial, G + c.b_off, s);
    };
    // head -> d0
    uint16_t *dd0 = u->g0;
    unet_head_back(u->d0, 16, u->pw + u->head_w, u->mask, u->gout, v0, dd0, u->partial, G + u->head_w, s);
    unet_relu_back(v0 * 16, u->d0, dd0, s);
    grads(6, u->u0, dd0, R0, v0);
    uint16_t *de0 = u->gA;  // du0: gradient of u0 = up(d1) + e0
    unet_conv3_fwd(16, 16, dd0, Wt(6), nullptr, de0, nullptr, R0, v0, false, s);
    uint16_t *dd1 = u->g1;
    unet_up_back(de0, R0, 16, v1, dd1, s);
    unet_relu_back(v1 * 16, u->d1, dd1, s);
    grads(5, u->u1, dd1, R1, v1);
    uint16_t *de1 = u->gB;
    unet_conv3_fwd(16, 32, dd1, Wt(5), nullptr, de1, nullptr, R1, v1, false, s);
    uint16_t *dd2 = u->g2;
    unet_up_back(de1, R1, 32, v2, dd2, s);
    unet_relu_back(v2 * 32, u->d2, dd2, s);
    grads(4, u->u2, dd2, R2, v2);
    uint16_t *de2 = u->gC;
    unet_conv3_fwd(32, 64, dd2, Wt(4), nullptr, de2, nullptr, R2, v2, false, s);
    uint16_t *de3 = u->g0;  // dd0 no longer needed
    unet_up_back(de2, R2, 64, v3, de3, s);
    unet_relu_back(v3 * 64, u->e3, de3, s);
    grads(3, u->p2, de3, R3, v3);
    uint16_t *dp2 = u->g1;
    unet_conv3_fwd(64, 64, de3, Wt(3), nullptr, dp2, nullptr, R3, v3, false, s);
    uint16_t *tmp = u->g2;
    unet_pool2_back(dp2, R2, 64, v2, tmp, s);
    unet_add(v2 * 64, de2, tmp, s);
    unet_relu_back(v2 * 64, u->e2, de2, s);
    grads(2, u->p1, de2, R2, v2);
    uint16_t *dp1 = u->g0;
    unet_conv3_fwd(64, 32, de2, Wt(2), nullptr, dp1, nullptr, R2, v2, false, s);
    unet_pool2_back(dp1, R1, 32, v1, tmp, s);
    unet_add(v1 * 32, de1, tmp, s);
    unet_relu_back(v1 * 32, u->e1, de1, s);
    grads(1, u->p0, de1, R1, v1);
    uint16_t *dp0 = u->g0;
    unet_conv3_fwd(32, 16, de1, Wt(1), nullptr, dp0, nullptr, R1, v1, false, s);
    unet_pool2_back(dp0, R0, 16, v0, tmp, s);
    unet_add(v0 * 16, de0, tmp, s);
    unet_relu_back(v0 * 16, u->e0, de0, s);
    grads(0, u->F0, de0, R0, v0);
}

// ----------------------------------------------------- solver-side inference
// features from (A, L, r) of one R^3 grid, forward, solution in x
// Features of one sample with sigma from per-block partials (sum b^2, FLUID
// count) that the producer of b wrote; every block sums them in the same
// order, block 0 stores sigma for the head.
__global__ __launch_bounds__(256) void k_features_p(const int8_t *__restrict__ A, const float *__restrict__ L,
                                                    const float *__restrict__ b, const double *__restrict__ part,
                                                    int nparts, long n, uint16_t *F, float *mask, float *sigma_out)
{
    __shared__ double s0[4], s1[4];
    __shared__ float sig;
    double a = 0.0, c = 0.0;
    for (int k = threadIdx.x; k < nparts; k += 256) {
        a += part[2 * k];
        c += part[2 * k + 1];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        c += __shfl_xor(c, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s0[threadIdx.x >> 6] = a;
        s1[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double ta = s0[0] + s0[1] + s0[2] + s0[3], tc = s1[0] + s1[1] + s1[2] + s1[3];
        const double v = tc > 0.0 ? sqrt(ta / tc) : 0.0;
        sig = v > 0.0 ? (float)v : 1.0f;
        if (blockIdx.x == 0) sigma_out[0] = sig;
    }
    __syncthreads();
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    int av = A[g];
    float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (av == FLIP_FLUID) f[0] = b[g] / sig;
    f[1] = av == FLIP_FLUID;
    f[2] = av == FLIP_AIR;
    f[3] = av == FLIP_WALL;
    if (av != FLIP_WALL) f[4] = fminf(fmaxf(L[g], -2.0f), 2.0f);
    for (int ch = 0; ch < 8; ++ch) F[g * 8 + ch] = f2bf_d(f[ch]);
    mask[g] = av == FLIP_FLUID ? 1.0f : 0.0f;
}

// per-block partials (sum b^2 over FLUID cells, FLUID count), blocks of 256
__global__ void k_rhs_partials(const int8_t *__restrict__ A, const float *__restrict__ b, long n, double *part)
{
    __shared__ double s0[4], s1[4];
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, c = 0.0;
    if (g < n && A[g] == FLIP_FLUID) {
        a = (double)b[g] * b[g];
        c = 1.0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        c += __shfl_xor(c, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s0[threadIdx.x >> 6] = a;
        s1[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0] + s0[1] + s0[2] + s0[3];
        part[2 * blockIdx.x + 1] = s1[0] + s1[1] + s1[2] + s1[3];
    }
}

// Inference forward of one sample: the decoder's upsample-adds are formed
// inside the convolutions, and the last convolution applies the head and
// writes the solution directly.  (Forming the encoder's poolings inside the
// next convolution's operand loads was slower: each pooled voxel is then
// averaged once per tap, 27 times.)
static void unet_infer(UNet *u, const float *sigma, float *x)
{
    hipStream_t s = u->stream;
    const int R0 = u->R, R1 = R0 / 2, R2 = R0 / 4, R3 = R0 / 8;
    const long v0 = u->vox(0, 1), v1 = u->vox(1, 1), v2 = u->vox(2, 1), v3 = u->vox(3, 1);
    auto W = [&](int i) { return u->pwb + u->conv[i].w_off; };
    auto B = [&](int i) { return u->pw + u->conv[i].b_off; };
    unet_conv3_fwd(UNET_FEATURES, 16, u->F0, W(0), B(0), u->e0, nullptr, R0, v0, true, s);
    unet_pool2(u->e0, R0, 16, v1, u->p0, s);
    unet_conv3_fwd(16, 32, u->p0, W(1), B(1), u->e1, nullptr, R1, v1, true, s);
    unet_pool2(u->e1, R1, 32, v2, u->p1, s);
    unet_conv3_fwd(32, 64, u->p1, W(2), B(2), u->e2, nullptr, R2, v2, true, s);
    unet_pool2(u->e2, R2, 64, v3, u->p2, s);
    unet_conv3_fwd(64, 64, u->p2, W(3), B(3), u->e3, nullptr, R3, v3, true, s);
    unet_conv3_fwd_up(64, 32, u->e3, u->e2, W(4), B(4), u->d2, R2, v2, s);
    unet_conv3_fwd_up(32, 16, u->d2, u->e1, W(5), B(5), u->d1, R1, v1, s);
    const float scale = OUT_UNIT / ((float)u->R * (float)u->R);
    unet_conv3_fwd_up_head(u->d1, u->e0, W(6), B(6), u->pw + u->head_w, u->pw + u->head_b, u->mask, sigma, scale, x,
                           R0, v0, s);
}

int unet_partials_count(const UNet *u) { return (int)nb(u->vox(0, 1)); }

void unet_solve_partials(UNet *u, const int8_t *A, const float *L, const float *r, const double *part, float *x,
                         float *sigma)
{
    const long n = u->vox(0, 1);
    hipLaunchKernelGGL(k_features_p, dim3(nb(n)), dim3(256), 0, u->stream, A, L, r, part, unet_partials_count(u), n,
                       u->F0, u->mask, sigma);
    unet_infer(u, sigma, x);
}

void unet_solve(UNet *u, const int8_t *A, const float *L, const float *r, float *x, float *sigma, double *part)
{
    const long n = u->vox(0, 1);
    hipLaunchKernelGGL(k_rhs_partials, dim3(nb(n)), dim3(256), 0, u->stream, A, r, n, part);
    unet_solve_partials(u, A, L, r, part, x, sigma);
}

// ------------------------------------------------------------------ C ABI
struct flip_unet {
    flip_unet_impl impl;
};

namespace {

void trainer_alloc(Trainer &t, UNet *u)
{
    const long n = u->vox(0, u->cap);
    dmalloc(&t.A, n);
    dmalloc(&t.L, n);
    dmalloc(&t.b, n);
    dmalloc(&t.x, n);
    dmalloc(&t.t, n);
    dmalloc(&t.err, n);
    dmalloc(&t.Aerr, n);
    dmalloc(&t.At, n);
    dmalloc(&t.sigma, u->cap);
    dmalloc(&t.coef, u->cap);
    dmalloc(&t.part, 2L * t.blocks * u->cap);
    dmalloc(&t.spart, 2L * unet_partials_count(u));
    dmalloc(&t.loss, u->cap);
    dmalloc(&t.gpart, 256);
    dmalloc(&t.gscale, 1);
}

void trainer_free(Trainer &t)
{
    dfree(t.A);
    dfree(t.L);
    dfree(t.b);
    dfree(t.x);
    dfree(t.t);
    dfree(t.err);
    dfree(t.Aerr);
    dfree(t.At);
    dfree(t.sigma);
    dfree(t.coef);
    dfree(t.part);
    dfree(t.spart);
    dfree(t.loss);
    dfree(t.gpart);
    dfree(t.gscale);
    if (t.net) unet_free(t.net);
    t.net = nullptr;
}

// upload S systems and build features, mask and (if x) the target
void stage_systems(Trainer &t, int S, const int8_t *A, const float *L, const float *b, const float *x)
{
    UNet *u = t.net;
    hipStream_t s = u->stream;
    const long per = u->vox(0, 1), n = per * S;
    FLIP_CHECK(hipMemcpyAsync(t.A, A, n, hipMemcpyHostToDevice, s));
    FLIP_CHECK(hipMemcpyAsync(t.L, L, n * sizeof(float), hipMemcpyHostToDevice, s));
    FLIP_CHECK(hipMemcpyAsync(t.b, b, n * sizeof(float), hipMemcpyHostToDevice, s));
    if (x) FLIP_CHECK(hipMemcpyAsync(t.x, x, n * sizeof(float), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_rhs_stats, dim3(t.blocks, S), dim3(256), 0, s, t.A, t.b, per, t.part);
    hipLaunchKernelGGL(k_rhs_sigma, dim3(1), dim3(64), 0, s, t.part, t.blocks, S, t.sigma);
    hipLaunchKernelGGL(k_features, dim3(nb(n)), dim3(256), 0, s, t.A, t.L, t.b, t.sigma, per, n, u->F0, u->mask);
    if (x) {
        const float scale = ((float)u->R * (float)u->R) / OUT_UNIT;
        hipLaunchKernelGGL(k_target, dim3(nb(n)), dim3(256), 0, s, t.A, t.x, t.sigma, per, n, scale, t.t);
    }
}

}  // namespace

extern "C" {

int flip_unet_create(int R, int capacity, uint64_t seed, flip_unet **out)
{
    return guarded_u([&]() -> int {
        if (!out) return fail(FLIP_E_ARG, "null output");
        if (capacity < 1 || capacity > 64) return fail(FLIP_E_ARG, "batch capacity must be 1..64");
        flip_unet *h = new flip_unet();
        try {
            FLIP_CHECK(hipStreamCreate(&h->impl.stream));
            h->impl.tr.net = unet_new(R, capacity, h->impl.stream);
            trainer_alloc(h->impl.tr, h->impl.tr.net);
            unet_init_weights(h->impl.tr.net, seed);
            FLIP_CHECK(hipStreamSynchronize(h->impl.stream));
        } catch (...) {
            trainer_free(h->impl.tr);
            if (h->impl.stream) (void)hipStreamDestroy(h->impl.stream);
            delete h;
            throw;
        }
        *out = h;
        return FLIP_OK;
    });
}

int flip_unet_destroy(flip_unet *h)
{
    if (!h) return FLIP_OK;
    (void)hipStreamSynchronize(h->impl.stream);
    trainer_free(h->impl.tr);
    (void)hipStreamDestroy(h->impl.stream);
    delete h;
    return FLIP_OK;
}

int64_t flip_unet_num_params(flip_unet *h) { return h ? h->impl.tr.net->nparam : -1; }

int flip_unet_get_params(flip_unet *h, float *w)
{
    return guarded_u([&]() -> int {
        if (!h || !w) return fail(FLIP_E_ARG, "null argument");
        UNet *u = h->impl.tr.net;
        FLIP_CHECK(hipStreamSynchronize(u->stream));
        FLIP_CHECK(hipMemcpy(w, u->pw, sizeof(float) * u->nparam, hipMemcpyDeviceToHost));
        return FLIP_OK;
    });
}

int flip_unet_set_params(flip_unet *h, const float *w)
{
    return guarded_u([&]() -> int {
        if (!h || !w) return fail(FLIP_E_ARG, "null argument");
        UNet *u = h->impl.tr.net;
        FLIP_CHECK(hipStreamSynchronize(u->stream));
        FLIP_CHECK(hipMemcpy(u->pw, w, sizeof(float) * u->nparam, hipMemcpyHostToDevice));
        unet_refresh_bf16(u);
        FLIP_CHECK(hipStreamSynchronize(u->stream));
        return FLIP_OK;
    });
}

int flip_unet_forward_features(flip_unet *h, int S, const float *features, const float *mask, float *out)
{
    return guarded_u([&]() -> int {
        if (!h || !features || !mask || !out) return fail(FLIP_E_ARG, "null argument");
        UNet *u = h->impl.tr.net;
        if (S < 1 || S > u->cap) return fail(FLIP_E_ARG, "sample count out of range");
        const long n = u->vox(0, S);
        std::vector<float> f(features, features + n * UNET_FEATURES);
        float *tmp = nullptr;
        dmalloc(&tmp, n * UNET_FEATURES);
        FLIP_CHECK(hipMemcpyAsync(tmp, f.data(), sizeof(float) * n * UNET_FEATURES, hipMemcpyHostToDevice, u->stream));
        unet_to_bf16(n * UNET_FEATURES, tmp, u->F0, u->stream);
        FLIP_CHECK(hipMemcpyAsync(u->mask, mask, sizeof(float) * n, hipMemcpyHostToDevice, u->stream));
        unet_forward(u, S);
        FLIP_CHECK(hipMemcpyAsync(out, u->out, sizeof(float) * n, hipMemcpyDeviceToHost, u->stream));
        FLIP_CHECK(hipStreamSynchronize(u->stream));
        (void)hipFree(tmp);
        return FLIP_OK;
    });
}

int flip_unet_backward(flip_unet *h, int S, const float *gout, float *grads)
{
    return guarded_u([&]() -> int {
        if (!h || !gout || !grads) return fail(FLIP_E_ARG, "null argument");
        UNet *u = h->impl.tr.net;
        if (S < 1 || S > u->cap) return fail(FLIP_E_ARG, "sample count out of range");
        const long n = u->vox(0, S);
        FLIP_CHECK(hipMemcpyAsync(u->gout, gout, sizeof(float) * n, hipMemcpyHostToDevice, u->stream));
        unet_backward(u, S);
        FLIP_CHECK(hipMemcpyAsync(grads, u->pg, sizeof(float) * u->nparam, hipMemcpyDeviceToHost, u->stream));
        FLIP_CHECK(hipStreamSynchronize(u->stream));
        return FLIP_OK;
    });
}

int flip_unet_train_step(flip_unet *h, int S, const int8_t *A, const float *L, const float *b, const float *x,
                         float lr, double *loss)
{
    return guarded_u([&]() -> int {
        if (!h || !A || !L || !b || !x) return fail(FLIP_E_ARG, "null argument");
        Trainer &t = h->impl.tr;
        UNet *u = t.net;
        if (S < 1 || S > u->cap) return fail(FLIP_E_ARG, "sample count out of range");
        hipStream_t s = u->stream;
        const long per = u->vox(0, 1), n = per * S;
        stage_systems(t, S, A, L, b, x);
        unet_forward(u, S);
        hipLaunchKernelGGL(k_sub, dim3(nb(n)), dim3(256), 0, s, n, u->out, t.t, t.err);
        hipLaunchKernelGGL(k_grid_ax, dim3(nb(n)), dim3(256), 0, s, t.A, t.L, t.err, u->R, n, t.Aerr);
        hipLaunchKernelGGL(k_grid_ax, dim3(nb(n)), dim3(256), 0, s, t.A, t.L, t.t, u->R, n, t.At);
        hipLaunchKernelGGL(k_dots2, dim3(t.blocks, S), dim3(256), 0, s, t.err, t.Aerr, t.t, t.At, per, t.part);
        hipLaunchKernelGGL(k_loss_fin, dim3(1), dim3(64), 0, s, t.part, t.blocks, S, t.loss, t.coef);
        hipLaunchKernelGGL(k_scale_per_sample, dim3(nb(n)), dim3(256), 0, s, n, per, t.coef, t.Aerr, u->gout);
        unet_backward(u, S);
        u->step++;
        unet_adam(u->nparam, u->pw, u->pg, u->pm, u->pv, lr, u->step, UNET_CLIP, t.gpart, t.gscale, u->pwb, s);
        std::vector<double> ls(S);
        FLIP_CHECK(hipMemcpyAsync(ls.data(), t.loss, sizeof(double) * S, hipMemcpyDeviceToHost, s));
        FLIP_CHECK(hipStreamSynchronize(s));
        double m = 0.0;
        for (double v : ls) m += v;
        if (loss) *loss = m / S;
        return FLIP_OK;
    });
}

int flip_unet_predict(flip_unet *h, int S, const int8_t *A, const float *L, const float *b, float *x, double *loss_vs)
{
    return guarded_u([&]() -> int {
        if (!h || !A || !L || !b || !x) return fail(FLIP_E_ARG, "null argument");
        Trainer &t = h->impl.tr;
        UNet *u = t.net;
        if (S < 1 || S > u->cap) return fail(FLIP_E_ARG, "sample count out of range");
        hipStream_t s = u->stream;
        const long per = u->vox(0, 1), n = per * S;
        stage_systems(t, S, A, L, b, nullptr);
        unet_forward(u, S);
        const float scale = OUT_UNIT / ((float)u->R * (float)u->R);
        hipLaunchKernelGGL(k_unscale, dim3(nb(n)), dim3(256), 0, s, u->out, t.sigma, per, n, scale, t.err);
        FLIP_CHECK(hipMemcpyAsync(x, t.err, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        FLIP_CHECK(hipStreamSynchronize(s));
        (void)loss_vs;
        return FLIP_OK;
    });
}

int flip_unet_predict_fused(flip_unet *h, const int8_t *A, const float *L, const float *b, float *x)
{
    return guarded_u([&]() -> int {
        if (!h || !A || !L || !b || !x) return fail(FLIP_E_ARG, "null argument");
        Trainer &t = h->impl.tr;
        UNet *u = t.net;
        hipStream_t s = u->stream;
        const long n = u->vox(0, 1);
        stage_systems(t, 1, A, L, b, nullptr);
        unet_infer(u, t.sigma, t.err);
        FLIP_CHECK(hipMemcpyAsync(x, t.err, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        FLIP_CHECK(hipStreamSynchronize(s));
        return FLIP_OK;
    });
}

int flip_unet_solve_system(flip_unet *h, int path, const int8_t *A, const float *L, const float *b, float *x)
{
    return guarded_u([&]() -> int {
        if (!h || !A || !L || !b || !x) return fail(FLIP_E_ARG, "null argument");
        if (path != 0 && path != 1) return fail(FLIP_E_ARG, "path must be 0 or 1");
        Trainer &t = h->impl.tr;
        UNet *u = t.net;
        hipStream_t s = u->stream;
        const long n = u->vox(0, 1);
        FLIP_CHECK(hipMemcpyAsync(t.A, A, n, hipMemcpyHostToDevice, s));
        FLIP_CHECK(hipMemcpyAsync(t.L, L, n * sizeof(float), hipMemcpyHostToDevice, s));
        FLIP_CHECK(hipMemcpyAsync(t.b, b, n * sizeof(float), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_rhs_partials, dim3(nb(n)), dim3(256), 0, s, t.A, t.b, n, t.spart);
        if (path == 0)
            unet_solve_partials(u, t.A, t.L, t.b, t.spart, t.err, t.sigma);
        else
            unet_infer_layers(u, t.A, t.L, t.b, t.spart, unet_partials_count(u), nullptr, t.err, t.sigma);
        FLIP_CHECK(hipMemcpyAsync(x, t.err, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        FLIP_CHECK(hipStreamSynchronize(s));
        return FLIP_OK;
    });
}

int flip_unet_save(flip_unet *h, const char *path)
{
    return guarded_u([&]() -> int {
        if (!h || !path) return fail(FLIP_E_ARG, "null argument");
        UNet *u = h->impl.tr.net;
        std::vector<float> w(u->nparam);
        FLIP_CHECK(hipStreamSynchronize(u->stream));
        FLIP_CHECK(hipMemcpy(w.data(), u->pw, sizeof(float) * u->nparam, hipMemcpyDeviceToHost));
        FILE *f = fopen(path, "wb");
        if (!f) return fail(FLIP_E_ARG, "cannot open weights file for writing");
        const char magic[8] = {'F', 'L', 'I', 'P', 'U', 'N', 'E', 'T'};
        int32_t hdr[2] = {1, (int32_t)u->nparam};
        bool ok = fwrite(magic, 8, 1, f) == 1 && fwrite(hdr, sizeof(hdr), 1, f) == 1 &&
                  fwrite(w.data(), sizeof(float), w.size(), f) == w.size();
        fclose(f);
        return ok ? FLIP_OK : fail(FLIP_E_ARG, "short write of weights file");
    });
}

}  // extern "C"

// weights file -> host vector (checks magic, version and size)
bool unet_read_weights(const char *path, long nparam, std::vector<float> &w, std::string &err)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        err = std::string("cannot open U-Net weights file ") + path;
        return false;
    }
    char magic[8];
    int32_t hdr[2];
    bool ok = fread(magic, 8, 1, f) == 1 && memcmp(magic, "FLIPUNET", 8) == 0 && fread(hdr, sizeof(hdr), 1, f) == 1 &&
              hdr[0] == 1 && hdr[1] == nparam;
    if (ok) {
        w.resize(nparam);
        ok = fread(w.data(), sizeof(float), nparam, f) == (size_t)nparam;
    }
    fclose(f);
    if (!ok) err = std::string("bad U-Net weights file ") + path;
    return ok;
}

extern "C" int flip_unet_load(flip_unet *h, const char *path)
{
    return guarded_u([&]() -> int {
        if (!h || !path) return fail(FLIP_E_ARG, "null argument");
        UNet *u = h->impl.tr.net;
        std::vector<float> w;
        std::string err;
        if (!unet_read_weights(path, u->nparam, w, err)) return fail(FLIP_E_ARG, err);
        FLIP_CHECK(hipMemcpy(u->pw, w.data(), sizeof(float) * u->nparam, hipMemcpyHostToDevice));
        unet_refresh_bf16(u);
        FLIP_CHECK(hipStreamSynchronize(u->stream));
        return FLIP_OK;
    });
}
