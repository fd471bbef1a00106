// k_conv.hip - BF16 MFMA 3-D convolution kernels (gfx950) for the learned
// coarse-level solver, with FP32 accumulation.
//
// Activations are channels-last: element (s, z, y, x, c) of a batch of S
// cubes of side R sits at ((((s*R + z)*R + y)*R + x)*C + c), bf16.  Weights of
// a 3x3x3 layer are [27][Cout][Cin] bf16 (tap = (dz+1)*9 + (dy+1)*3 + (dx+1)),
// so the eight input channels one MFMA lane needs are one 16-byte load, as
// are the eight weights it needs.
//
// A 3x3x3 "same" convolution is an implicit GEMM: M = voxels, N = Cout,
// K = 27*Cin.  One wave owns MT tiles of 16 voxels and all Cout columns and
// steps K by 32 with v_mfma_f32_16x16x32_bf16: lane l feeds row l%16 of A
// (a voxel) and column l%16 of B (an output channel) with the eight K values
// of chunk 4*step + l/16, where a chunk is eight consecutive input channels
// of one tap.  Any K order is correct as long as A and B use the same one.
// The 16x16 result sits in lane l as rows 4*(l/16)+i, column l%16.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "flip_internal.hpp"
#include "unet_internal.hpp"

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

struct Vox {
    int s, z, y, x;
};

__device__ __forceinline__ Vox vox_of(long v, int R)
{
    Vox o;
    o.x = (int)(v % R);
    long t = v / R;
    o.y = (int)(t % R);
    t /= R;
    o.z = (int)(t % R);
    o.s = (int)(t / R);
    return o;
}

__device__ __forceinline__ long vox_shift(const Vox &o, int tap, int R)
{
    int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    int z = o.z + dz, y = o.y + dy, x = o.x + dx;
    if ((unsigned)z >= (unsigned)R || (unsigned)y >= (unsigned)R || (unsigned)x >= (unsigned)R) return -1;
    return (((long)o.s * R + z) * R + y) * R + x;
}

__device__ __forceinline__ bf16x8 load8(const uint16_t *p)
{
    uint4 u = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ bf16x8 zero8()
{
    uint4 u = make_uint4(0, 0, 0, 0);
    return __builtin_bit_cast(bf16x8, u);
}

// ------------------------------------------------------------ forward conv
// Y = act(conv3(X, W) + bias), act = leaky ReLU (slope UNET_LEAK) when relu
// is set; writes bf16 Y (if Y) and fp32 Yf (if Yf).
template <int CIN, int COUT, int MT>
__global__ __launch_bounds__(256) void k_conv3_fwd(const uint16_t *__restrict__ X, const uint16_t *__restrict__ Wt,
                                                   const float *__restrict__ bias, uint16_t *Y, float *Yf, int R,
                                                   long nvox, int relu)
{
    static_assert(CIN % 8 == 0 && COUT % 16 == 0, "channel counts");
    constexpr int NT = COUT / 16;
    constexpr int CPT = CIN / 8;  // chunks per tap
    constexpr int NCH = 27 * CPT;
    constexpr int STEPS = (NCH + 3) / 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, grp = lane >> 4;
    const long base = ((long)blockIdx.x * 4 + wave) * (16 * MT);
    if (base >= nvox) return;

    Vox vo[MT];
    bool valid[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        long v = base + m * 16 + row;
        valid[m] = v < nvox;
        vo[m] = vox_of(valid[m] ? v : 0, R);
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int st = 0; st < STEPS; ++st) {
        const int q = st * 4 + grp;
        const bool live = q < NCH;
        const int tap = live ? q / CPT : 0;
        const int c0 = live ? (q % CPT) * 8 : 0;
        bf16x8 a[MT], b[NT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            long sv = (live && valid[m]) ? vox_shift(vo[m], tap, R) : -1;
            a[m] = sv >= 0 ? load8(X + sv * CIN + c0) : zero8();
        }
#pragma unroll
        for (int n = 0; n < NT; ++n)
            b[n] = live ? load8(Wt + ((long)tap * COUT + n * 16 + row) * CIN + c0) : zero8();
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int n = 0; n < NT; ++n) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[n], acc[m][n], 0, 0, 0);
    }
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const int co = n * 16 + row;
            const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                long v = base + m * 16 + grp * 4 + i;
                if (v >= nvox) continue;
                float val = acc[m][n][i] + bv;
                if (relu && val < 0.0f) val *= UNET_LEAK;
                if (Y) Y[v * COUT + co] = f2bf(val);
                if (Yf) Yf[v * COUT + co] = val;
            }
        }
}

// Split-K variant for small layers: one block per (16-voxel tile, 16-channel
// output tile); its KS waves take every KS-th K step (the next step's
// operands are loaded while the current MFMA runs) and are summed through
// LDS.  This gives layers of a few hundred voxels (the bottom of the U-Net)
// KS * NT times the waves of k_conv3_fwd, and each wave a K loop of
// STEPS / KS steps.
// Input modes (the fused form rounds to bf16 exactly as the stored
// intermediate would be rounded, so results equal the unfused network):
//   IN_PLAIN  X
//   IN_UP     up(U) + X: nearest upsampling of the (R/2)^3 tensor U (k_up_add)
// HEAD: the layer is the last one (16 output channels, one tile); instead of
// storing it, the block applies the 1x1x1 head and writes the solution
// x = mask (w . y + b) sigma scale (k_head_fwd followed by the unit change).
enum { IN_PLAIN = 0, IN_UP = 1 };

struct HeadArgs {
    const float *w, *b, *mask, *sigma;
    float scale;
    float *x;
};

__device__ __forceinline__ uint32_t add_bf16x2(uint32_t u, uint32_t x)
{
    const uint16_t lo = f2bf(bf2f((uint16_t)(u & 0xffff)) + bf2f((uint16_t)(x & 0xffff)));
    const uint16_t hi = f2bf(bf2f((uint16_t)(u >> 16)) + bf2f((uint16_t)(x >> 16)));
    return (uint32_t)lo | ((uint32_t)hi << 16);
}

template <int CIN, int COUT, int MODE, bool HEAD, int KS>
__global__ __launch_bounds__(64 * KS) void k_conv3_fwd_sk(const uint16_t *__restrict__ X, const uint16_t *__restrict__ U,
                                                      const uint16_t *__restrict__ Wt, const float *__restrict__ bias,
                                                      uint16_t *Y, float *Yf, int R, long nvox, int relu, HeadArgs hd)
{
    static_assert(CIN % 8 == 0 && COUT % 16 == 0, "channel counts");
    static_assert(!HEAD || COUT == 16, "the head follows a 16-channel layer");
    constexpr int CPT = CIN / 8;
    constexpr int NCH = 27 * CPT;
    constexpr int STEPS = (NCH + 3) / 4;
    __shared__ f32x4 red[KS - 1][64];
    __shared__ float hy[16][17];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, grp = lane >> 4;
    const long base = (long)blockIdx.x * 16;
    const int n0 = blockIdx.y * 16;
    const long v = base + row;
    const bool valid = v < nvox;
    const Vox vo = vox_of(valid ? v : 0, R);
    auto load_a = [&](int st) -> bf16x8 {
        const int q = st * 4 + grp;
        if (st >= STEPS || q >= NCH || !valid) return zero8();
        const int tap = q / CPT, c0 = (q % CPT) * 8;
        const long sv = vox_shift(vo, tap, R);
        if (sv < 0) return zero8();
        if (MODE == IN_PLAIN) return load8(X + sv * CIN + c0);
        const int z = vo.z + tap / 9 - 1, y = vo.y + (tap / 3) % 3 - 1, x = vo.x + tap % 3 - 1;
        const int Rc = R >> 1;
        const long p = (((long)vo.s * Rc + (z >> 1)) * Rc + (y >> 1)) * Rc + (x >> 1);
        const uint4 xs = *reinterpret_cast<const uint4 *>(X + sv * CIN + c0);
        const uint4 us = *reinterpret_cast<const uint4 *>(U + p * CIN + c0);
        return __builtin_bit_cast(bf16x8, make_uint4(add_bf16x2(us.x, xs.x), add_bf16x2(us.y, xs.y),
                                                     add_bf16x2(us.z, xs.z), add_bf16x2(us.w, xs.w)));
    };
    auto load_b = [&](int st) -> bf16x8 {
        const int q = st * 4 + grp;
        if (st >= STEPS || q >= NCH) return zero8();
        return load8(Wt + ((long)(q / CPT) * COUT + n0 + row) * CIN + (q % CPT) * 8);
    };
    // A wave has only a few K steps (STEPS / KS, about four): all of their
    // operands are loaded before the first MFMA, so the wave waits on memory
    // once, and the steps are then summed in the same order as a loop would.
    constexpr int NS = (STEPS + KS - 1) / KS;
    bf16x8 a[NS], b[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        a[i] = load_a(wave + i * KS);
        b[i] = load_b(wave + i * KS);
    }
    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < NS; ++i)
        if (wave + i * KS < STEPS) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[i], acc, 0, 0, 0);
    if (wave > 0) red[wave - 1][lane] = acc;
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int w = 0; w < KS - 1; ++w) {
            const f32x4 o = red[w][lane];
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[i] += o[i];
        }
        const int co = n0 + row;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long vv = base + grp * 4 + i;
            float val = acc[i] + bv;
            if (relu && val < 0.0f) val *= UNET_LEAK;
            if (HEAD) {
                hy[grp * 4 + i][co] = bf2f(f2bf(val));
                continue;
            }
            if (vv >= nvox) continue;
            if (Y) Y[vv * COUT + co] = f2bf(val);
            if (Yf) Yf[vv * COUT + co] = val;
        }
    }
    if (!HEAD) return;
    __syncthreads();
    if (threadIdx.x < 16) {
        const long vv = base + threadIdx.x;
        if (vv < nvox) {
            float sh = hd.b[0];
            for (int c = 0; c < 16; ++c) sh += hd.w[c] * hy[threadIdx.x][c];
            const float out = hd.mask[vv] * sh;
            hd.x[vv] = out * hd.sigma[0] * hd.scale;
        }
    }
}

// ------------------------------------------------------- weight gradient
// partial[(chunk*27 + tap)][co][ci] = sum over the chunk's voxels v of
// dY[v][co] * X[v + tap][ci].  grid (chunks, 27), 4 waves per block; each
// wave strides over the chunk 32 voxels at a time.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv3_wgrad(const uint16_t *__restrict__ X, const uint16_t *__restrict__ dY,
                                                     float *partial, int R, long nvox, int vox_per_chunk)
{
    constexpr int MT = COUT / 16;
    constexpr int NT = (CIN + 15) / 16;
    __shared__ float red[4][MT * NT * 16 * 16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, grp = lane >> 4;
    const int tap = blockIdx.y;
    const long c_begin = (long)blockIdx.x * vox_per_chunk;
    const long c_end = min(nvox, c_begin + vox_per_chunk);
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    for (long kb = c_begin + wave * 32; kb < c_end; kb += 4 * 32) {
        long src[8];
        long vv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            long v = kb + grp * 8 + j;
            vv[j] = v < c_end ? v : -1;
            src[j] = vv[j] >= 0 ? vox_shift(vox_of(v, R), tap, R) : -1;
        }
        bf16x8 a[MT], b[NT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            uint16_t t[8];
            const int co = m * 16 + row;
#pragma unroll
            for (int j = 0; j < 8; ++j) t[j] = vv[j] >= 0 ? dY[vv[j] * COUT + co] : (uint16_t)0;
            uint4 u = make_uint4(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16),
                                 t[4] | ((uint32_t)t[5] << 16), t[6] | ((uint32_t)t[7] << 16));
            a[m] = __builtin_bit_cast(bf16x8, u);
        }
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            uint16_t t[8];
            const int ci = n * 16 + row;
#pragma unroll
            for (int j = 0; j < 8; ++j) t[j] = (src[j] >= 0 && ci < CIN) ? X[src[j] * CIN + ci] : (uint16_t)0;
            uint4 u = make_uint4(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16),
                                 t[4] | ((uint32_t)t[5] << 16), t[6] | ((uint32_t)t[7] << 16));
            b[n] = __builtin_bit_cast(bf16x8, u);
        }
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int n = 0; n < NT; ++n) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[n], acc[m][n], 0, 0, 0);
    }
    // D(co, ci): lane holds rows co = 4*grp + i, column ci = row
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n)
#pragma unroll
            for (int i = 0; i < 4; ++i) red[wave][((m * NT + n) * 16 + grp * 4 + i) * 16 + row] = acc[m][n][i];
    __syncthreads();
    float *out = partial + ((long)blockIdx.x * 27 + tap) * COUT * CIN;
    for (int e = threadIdx.x; e < MT * NT * 256; e += blockDim.x) {
        int tile = e / 256, rr = (e / 16) % 16, cc = e % 16;
        int m = tile / NT, n = tile % NT;
        int co = m * 16 + rr, ci = n * 16 + cc;
        if (ci >= CIN) continue;
        float s = red[0][e] + red[1][e] + red[2][e] + red[3][e];
        out[co * CIN + ci] = s;
    }
}

// dW[i] = sum over chunks of partial[chunk][i] (fixed order)
__global__ void k_reduce_chunks(const float *__restrict__ partial, int chunks, long len, float *out)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    float s = 0.0f;
    for (int c = 0; c < chunks; ++c) s += partial[(long)c * len + i];
    out[i] = s;
}

// db[co] partials: block b sums dY over its voxel chunk
__global__ void k_bias_grad(const uint16_t *__restrict__ dY, int C, long nvox, int vox_per_chunk, float *partial)
{
    __shared__ float red[256];
    const int co = threadIdx.x % C, sub = threadIdx.x / C, nsub = blockDim.x / C;
    const long b0 = (long)blockIdx.x * vox_per_chunk, b1 = min(nvox, b0 + vox_per_chunk);
    float s = 0.0f;
    if (sub < nsub)
        for (long v = b0 + sub; v < b1; v += nsub) s += bf2f(dY[v * C + co]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < C) {
        float t = 0.0f;
        for (int k = 0; k < nsub; ++k) t += red[k * C + threadIdx.x];
        partial[(long)blockIdx.x * C + threadIdx.x] = t;
    }
}

// ---------------------------------------------------------- elementwise
// flipped, transposed weights for the data gradient: Wt[t][ci][co] = W[26-t][co][ci]
__global__ void k_weight_flip(const uint16_t *__restrict__ W, int cout, int cin, uint16_t *Wt)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = 27L * cout * cin;
    if (i >= n) return;
    int t = (int)(i / ((long)cout * cin));
    int r = (int)(i % ((long)cout * cin));
    int ci = r / cout, co = r % cout;  // Wt layout [t][ci][co]
    Wt[i] = W[((long)(26 - t) * cout + co) * cin + ci];
}

// dY <- dY * act'(Y): 1 where Y > 0, UNET_LEAK elsewhere (Y and the
// activation's input have the same sign)
__global__ void k_relu_back(long n, const uint16_t *__restrict__ Y, uint16_t *dY)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(bf2f(Y[i]) > 0.0f)) dY[i] = f2bf(bf2f(dY[i]) * UNET_LEAK);
}

// 2x2x2 average pooling: out (R/2 side) from in (R side), C channels
__global__ void k_pool2(const uint16_t *__restrict__ in, int R, int C, long nout, uint16_t *out)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nout * C) return;
    int c = (int)(i % C);
    long v = i / C;
    int Rc = R / 2;
    Vox o = vox_of(v, Rc);
    float s = 0.0f;
    for (int d = 0; d < 8; ++d) {
        long f = (((long)o.s * R + 2 * o.z + (d >> 2)) * R + 2 * o.y + ((d >> 1) & 1)) * R + 2 * o.x + (d & 1);
        s += bf2f(in[f * C + c]);
    }
    out[i] = f2bf(s * 0.125f);
}

// backward of k_pool2: din[f] = dout[parent(f)] / 8
__global__ void k_pool2_back(const uint16_t *__restrict__ dout, int R, int C, long nin, uint16_t *din)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nin * C) return;
    int c = (int)(i % C);
    Vox o = vox_of(i / C, R);
    int Rc = R / 2;
    long p = (((long)o.s * Rc + o.z / 2) * Rc + o.y / 2) * Rc + o.x / 2;
    din[i] = f2bf(bf2f(dout[p * C + c]) * 0.125f);
}

// out[f] = coarse[parent(f)] + skip[f] (nearest upsampling, R = fine side)
__global__ void k_up_add(const uint16_t *__restrict__ coarse, const uint16_t *__restrict__ skip, int R, int C,
                         long nfine, uint16_t *out)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nfine * C) return;
    int c = (int)(i % C);
    Vox o = vox_of(i / C, R);
    int Rc = R / 2;
    long p = (((long)o.s * Rc + o.z / 2) * Rc + o.y / 2) * Rc + o.x / 2;
    out[i] = f2bf(bf2f(coarse[p * C + c]) + bf2f(skip[i]));
}

// backward of the upsampling half: dcoarse[p] = sum of the 8 children's dout
__global__ void k_up_back(const uint16_t *__restrict__ dout, int R, int C, long ncoarse, uint16_t *dcoarse)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncoarse * C) return;
    int c = (int)(i % C);
    int Rc = R / 2;
    Vox o = vox_of(i / C, Rc);
    float s = 0.0f;
    for (int d = 0; d < 8; ++d) {
        long f = (((long)o.s * R + 2 * o.z + (d >> 2)) * R + 2 * o.y + ((d >> 1) & 1)) * R + 2 * o.x + (d & 1);
        s += bf2f(dout[f * C + c]);
    }
    dcoarse[i] = f2bf(s);
}

// a += b (bf16)
__global__ void k_add_bf16(long n, uint16_t *a, const uint16_t *__restrict__ b)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

// 1x1x1 output layer: out[v] = mask[v] * (sum_c w[c] x[v][c] + b)
__global__ void k_head_fwd(const uint16_t *__restrict__ X, int C, const float *__restrict__ w, const float *__restrict__ b,
                           const float *__restrict__ mask, long nvox, float *out)
{
    long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    float s = b[0];
    for (int c = 0; c < C; ++c) s += w[c] * bf2f(X[v * C + c]);
    out[v] = mask[v] * s;
}

// head backward: dX[v][c] = mask g w[c]; partial dw/db per block
__global__ void k_head_back(const uint16_t *__restrict__ X, int C, const float *__restrict__ w,
                            const float *__restrict__ mask, const float *__restrict__ g, long nvox, uint16_t *dX,
                            float *partial)
{
    __shared__ float red[256][17];
    long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float gm = v < nvox ? g[v] * mask[v] : 0.0f;
    for (int c = 0; c < C; ++c) {
        float xv = v < nvox ? bf2f(X[v * C + c]) : 0.0f;
        if (v < nvox) dX[v * C + c] = f2bf(gm * w[c]);
        red[threadIdx.x][c] = gm * xv;
    }
    red[threadIdx.x][C] = gm;
    __syncthreads();
    if (threadIdx.x <= (unsigned)C) {
        float s = 0.0f;
        for (int t = 0; t < (int)blockDim.x; ++t) s += red[t][threadIdx.x];
        partial[(long)blockIdx.x * (C + 1) + threadIdx.x] = s;
    }
}

// Adam on fp32 master weights, then the bf16 working copy
__global__ void k_adam(long n, float *w, const float *__restrict__ g, float *m, float *v, float lr, float b1, float b2,
                       float eps, float c1, float c2, const float *__restrict__ gscale, uint16_t *wbf)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i] * gscale[0];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wbf) wbf[i] = f2bf(wi);
}

// per-block sums of g^2 (gradient norm for clipping)
__global__ void k_sumsq(long n, const float *__restrict__ g, double *partial)
{
    __shared__ double lds[4];
    double a = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        a += (double)g[i] * (double)g[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}

// gscale = min(1, clip / |g|) from the per-block sums
__global__ void k_clip_scale(const double *__restrict__ partial, int blocks, float clip, float *gscale)
{
    if (threadIdx.x != 0) return;
    double a = 0.0;
    for (int b = 0; b < blocks; ++b) a += partial[b];
    double nrm = sqrt(a);
    gscale[0] = (nrm > (double)clip && nrm == nrm) ? (float)((double)clip / nrm) : (nrm == nrm ? 1.0f : 0.0f);
}

__global__ void k_to_bf16(long n, const float *__restrict__ a, uint16_t *b)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

__global__ void k_to_f32(long n, const uint16_t *__restrict__ a, float *b)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = bf2f(a[i]);
}

inline unsigned nb(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// waves per block of the split-K kernel: about four K steps per wave,
// between 4 and 16 waves
constexpr int sk_waves(int cin)
{
    return ((27 * cin / 8 + 3) / 4) >= 48 ? 16 : (((27 * cin / 8 + 3) / 4) >= 24 ? 8 : 4);
}

template <int CIN, int COUT>
void fwd_t(const uint16_t *X, const uint16_t *W, const float *bias, uint16_t *Y, float *Yf, int R, long nvox, int relu,
           hipStream_t s)
{
    constexpr int MT = COUT >= 64 ? 2 : 4;
    long per_block = 4L * 16 * MT;
    const long waves = (nvox + 16 * MT - 1) / (16 * MT);
    if (waves < 2048) {
        // too few voxel tiles to fill the chip: split K and the output channels
        constexpr int KS = sk_waves(CIN);
        hipLaunchKernelGGL((k_conv3_fwd_sk<CIN, COUT, IN_PLAIN, false, KS>), dim3(nb(nvox, 16), COUT / 16),
                           dim3(64 * KS), 0, s, X, (const uint16_t *)nullptr, W, bias, Y, Yf, R, nvox, relu,
                           HeadArgs{});
        return;
    }
    hipLaunchKernelGGL((k_conv3_fwd<CIN, COUT, MT>), dim3(nb(nvox, (int)per_block)), dim3(256), 0, s, X, W, bias, Y,
                       Yf, R, nvox, relu);
}

template <int CIN, int COUT>
void wgrad_t(const uint16_t *X, const uint16_t *dY, float *partial, int chunks, int R, long nvox, int vpc,
             hipStream_t s)
{
    hipLaunchKernelGGL((k_conv3_wgrad<CIN, COUT>), dim3(chunks, 27), dim3(256), 0, s, X, dY, partial, R, nvox, vpc);
}

}  // namespace

// ------------------------------------------------------------ launchers
void unet_conv3_fwd(int cin, int cout, const uint16_t *X, const uint16_t *W, const float *bias, uint16_t *Y, float *Yf,
                    int R, long nvox, bool relu, hipStream_t s)
{
#define CASE(a, b)                                                  \
    if (cin == a && cout == b) {                                    \
        fwd_t<a, b>(X, W, bias, Y, Yf, R, nvox, relu ? 1 : 0, s);   \
        return;                                                     \
    }
    CASE(8, 16) CASE(16, 16) CASE(16, 32) CASE(32, 32) CASE(32, 64) CASE(64, 64) CASE(64, 32) CASE(32, 16)
#undef CASE
    throw flip_arg_error{FLIP_E_ARG, "unsupported conv channel pair"};
}

void unet_conv3_fwd_up(int cin, int cout, const uint16_t *U, const uint16_t *X, const uint16_t *W, const float *bias,
                       uint16_t *Y, int R, long nvox, hipStream_t s)
{
#define CASE(a, b)                                                                                                   \
    if (cin == a && cout == b) {                                                                                     \
        constexpr int KS = sk_waves(a);                                                                              \
        hipLaunchKernelGGL((k_conv3_fwd_sk<a, b, IN_UP, false, KS>), dim3(nb(nvox, 16), b / 16), dim3(64 * KS), 0, s, \
                           X, U, W, bias, Y, (float *)nullptr, R, nvox, 1, HeadArgs{});                              \
        return;                                                                                                      \
    }
    CASE(64, 32) CASE(32, 16) CASE(16, 16)
#undef CASE
    throw flip_arg_error{FLIP_E_ARG, "unsupported conv channel pair"};
}

void unet_conv3_fwd_up_head(const uint16_t *U, const uint16_t *X, const uint16_t *W, const float *bias,
                            const float *hw, const float *hb, const float *mask, const float *sigma, float scale,
                            float *x, int R, long nvox, hipStream_t s)
{
    HeadArgs hd{hw, hb, mask, sigma, scale, x};
    constexpr int KS = sk_waves(16);
    hipLaunchKernelGGL((k_conv3_fwd_sk<16, 16, IN_UP, true, KS>), dim3(nb(nvox, 16), 1), dim3(64 * KS), 0, s, X, U, W,
                       bias, (uint16_t *)nullptr, (float *)nullptr, R, nvox, 1, hd);
}

void unet_conv3_wgrad(int cin, int cout, const uint16_t *X, const uint16_t *dY, float *partial, float *dW, int R,
                      long nvox, hipStream_t s)
{
    int vpc = 8192;
    int chunks = (int)((nvox + vpc - 1) / vpc);
#define CASE(a, b)                                                  \
    if (cin == a && cout == b) {                                    \
        wgrad_t<a, b>(X, dY, partial, chunks, R, nvox, vpc, s);     \
        goto reduce;                                                \
    }
    CASE(8, 16) CASE(16, 16) CASE(16, 32) CASE(32, 32) CASE(32, 64) CASE(64, 64) CASE(64, 32) CASE(32, 16)
#undef CASE
    throw flip_arg_error{FLIP_E_ARG, "unsupported conv channel pair"};
reduce:
    long len = 27L * cout * cin;
    hipLaunchKernelGGL(k_reduce_chunks, dim3(nb(len)), dim3(256), 0, s, partial, chunks, len, dW);
}

int unet_wgrad_chunks(long nvox) { return (int)((nvox + 8191) / 8192); }

void unet_bias_grad(const uint16_t *dY, int C, long nvox, float *partial, float *db, hipStream_t s)
{
    int vpc = 8192;
    int chunks = (int)((nvox + vpc - 1) / vpc);
    int threads = (256 / C) * C;
    hipLaunchKernelGGL(k_bias_grad, dim3(chunks), dim3(threads), 0, s, dY, C, nvox, vpc, partial);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(nb(C)), dim3(256), 0, s, partial, chunks, (long)C, db);
}

void unet_weight_flip(const uint16_t *W, int cout, int cin, uint16_t *Wt, hipStream_t s)
{
    long n = 27L * cout * cin;
    hipLaunchKernelGGL(k_weight_flip, dim3(nb(n)), dim3(256), 0, s, W, cout, cin, Wt);
}

void unet_relu_back(long n, const uint16_t *Y, uint16_t *dY, hipStream_t s)
{
    hipLaunchKernelGGL(k_relu_back, dim3(nb(n)), dim3(256), 0, s, n, Y, dY);
}

void unet_pool2(const uint16_t *in, int R, int C, long nout, uint16_t *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_pool2, dim3(nb(nout * C)), dim3(256), 0, s, in, R, C, nout, out);
}

void unet_pool2_back(const uint16_t *dout, int R, int C, long nin, uint16_t *din, hipStream_t s)
{
    hipLaunchKernelGGL(k_pool2_back, dim3(nb(nin * C)), dim3(256), 0, s, dout, R, C, nin, din);
}

void unet_up_add(const uint16_t *coarse, const uint16_t *skip, int R, int C, long nfine, uint16_t *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_up_add, dim3(nb(nfine * C)), dim3(256), 0, s, coarse, skip, R, C, nfine, out);
}

void unet_up_back(const uint16_t *dout, int R, int C, long ncoarse, uint16_t *dcoarse, hipStream_t s)
{
    hipLaunchKernelGGL(k_up_back, dim3(nb(ncoarse * C)), dim3(256), 0, s, dout, R, C, ncoarse, dcoarse);
}

void unet_add(long n, uint16_t *a, const uint16_t *b, hipStream_t s)
{
    hipLaunchKernelGGL(k_add_bf16, dim3(nb(n)), dim3(256), 0, s, n, a, b);
}

void unet_head_fwd(const uint16_t *X, int C, const float *w, const float *b, const float *mask, long nvox, float *out,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_head_fwd, dim3(nb(nvox)), dim3(256), 0, s, X, C, w, b, mask, nvox, out);
}

int unet_head_blocks(long nvox) { return (int)nb(nvox); }

void unet_head_back(const uint16_t *X, int C, const float *w, const float *mask, const float *g, long nvox, uint16_t *dX,
                    float *partial, float *dwb, hipStream_t s)
{
    int blocks = (int)nb(nvox);
    hipLaunchKernelGGL(k_head_back, dim3(blocks), dim3(256), 0, s, X, C, w, mask, g, nvox, dX, partial);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(256), 0, s, partial, blocks, (long)(C + 1), dwb);
}

void unet_adam(long n, float *w, const float *g, float *m, float *v, float lr, int t, float clip, double *partial,
               float *gscale, uint16_t *wbf, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    const int blocks = 256;
    hipLaunchKernelGGL(k_sumsq, dim3(blocks), dim3(256), 0, s, n, g, partial);
    hipLaunchKernelGGL(k_clip_scale, dim3(1), dim3(64), 0, s, partial, blocks, clip, gscale);
    hipLaunchKernelGGL(k_adam, dim3(nb(n)), dim3(256), 0, s, n, w, g, m, v, lr, b1, b2, 1e-8f, c1, c2, gscale, wbf);
}

void unet_to_bf16(long n, const float *a, uint16_t *b, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(nb(n)), dim3(256), 0, s, n, a, b);
}

void unet_to_f32(long n, const uint16_t *a, float *b, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_f32, dim3(nb(n)), dim3(256), 0, s, n, a, b);
}
