// k_infer.hip - the learned solver's network inference as seven layer
// kernels, one per convolution (gfx950, v_mfma_f32_16x16x32_bf16).
//
// Each workgroup owns one brick of output voxels and 16 * NTW output
// channels.  It stages the brick's input halo (the brick plus one voxel on
// every side, zero outside the cube) and its slice of the layer's weights in
// LDS with one round of coalesced global loads, then runs the implicit GEMM
// (M = brick voxels, N = output channels, K = 27 * Cin) from LDS.  What the
// separate kernels of k_conv.hip do around the convolutions is formed while
// staging or in the epilogue:
//   - the first layer builds its input features from (A, L, b) with sigma
//     from the right-hand side's per-block partials (k_features_p);
//   - the encoder layers also write the 2x2x2 average pool of their output
//     (k_pool2), every cube being inside one brick;
//   - the decoder layers form up(U) + X while staging (the bf16 sum that
//     k_up_add would store);
//   - the last layer applies the 1x1x1 head and writes the solution in the
//     system's units.
// The arithmetic is that of the separate kernels, in the same order: the
// K steps are summed in the split-K kernel's chains (chain j takes steps j,
// j + KS, j + 2 KS, ... and the chains are added in order j = 0, 1, ...), so
// the result equals unet_infer's bit for bit.
//
// Every kernel returns at once when the solver's stop flag is set, so a
// V-cycle after convergence costs only the launches.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "flip_internal.hpp"
#include "unet_internal.hpp"

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even, NaN kept quiet (the convolutions' rounding)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// round to nearest even (the features' rounding)
__device__ __forceinline__ uint16_t f2bf_feat(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ uint32_t add_bf16x2(uint32_t u, uint32_t x)
{
    const uint16_t lo = f2bf(bf2f((uint16_t)(u & 0xffff)) + bf2f((uint16_t)(x & 0xffff)));
    const uint16_t hi = f2bf(bf2f((uint16_t)(u >> 16)) + bf2f((uint16_t)(x >> 16)));
    return (uint32_t)lo | ((uint32_t)hi << 16);
}

__device__ __forceinline__ bf16x8 lds8(const uint16_t *p) { return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(p)); }

typedef __attribute__((address_space(3))) void *lds_ptr;

__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

enum { SRC_FEAT = 0, SRC_PLAIN = 1, SRC_UP = 2 };
enum { DST_STORE = 0, DST_POOL = 1, DST_HEAD = 2 };

struct LayerArgs {
    const int32_t *stop;  // the solver's stop flag (null: always run)
    int R;                // cube side at this layer's level
    const uint16_t *X;    // input [R^3][CIN] (SRC_PLAIN; SRC_UP: the skip tensor)
    const uint16_t *U;    // SRC_UP: the coarse tensor [(R/2)^3][CIN]
    const uint16_t *W;    // weights [27][COUT][CIN]
    const float *bias;
    uint16_t *Y;  // output [R^3][COUT] (DST_STORE, DST_POOL)
    uint16_t *P;  // DST_POOL: pooled output [(R/2)^3][COUT]
    // SRC_FEAT: the system and the right-hand side's partials (sum b^2 over
    // FLUID cells, FLUID count) per 256-cell block
    const int8_t *A;
    const float *L, *b;
    const double *part;
    int nparts;
    float *sigma;  // SRC_FEAT writes it, DST_HEAD reads it
    // DST_HEAD
    const float *hw, *hb;
    float scale;
    float *x;
};

// One layer.  Grid: (bricks, COUT / (16 NTW)), 256 threads.  A unit is one
// 16-voxel row tile times one 16-channel column tile; with four or more
// units a wave owns whole units (all KS chains), with fewer the waves of a
// unit split its chains and add them through LDS in chain order.
template <int CIN, int COUT, int SRC, int DST, int KS, int BZ, int BY, int BX, int NTW>
__global__ __launch_bounds__(256) void k_unet_layer(LayerArgs a)
{
    constexpr int CPT = CIN / 8, NCH = 27 * CPT, STEPS = (NCH + 3) / 4;
    constexpr int SWZ = CPT - 1;  // chunk swizzle mask (CPT is a power of two)
    constexpr int HZ = BZ + 2, HY = BY + 2, HX = BX + 2, HV = HZ * HY * HX;
    // the parents of the halo (SRC_UP): z in [z0 - 1, z0 + BZ] has parents
    // z0 / 2 - 1 .. z0 / 2 + BZ / 2
    constexpr int UZ = BZ / 2 + 2, UY = BY / 2 + 2, UX = BX / 2 + 2, UV = UZ * UY * UX;
    constexpr int NV = BZ * BY * BX, MT = NV / 16, NCO = 16 * NTW, UNITS = MT * NTW;
    constexpr int WPU = UNITS >= 4 ? 1 : 4 / UNITS;  // waves per unit
    constexpr int UPW = UNITS >= 4 ? UNITS / 4 : 1;  // units per wave
    constexpr int CPW = KS / WPU;                    // chains per wave
    constexpr int TPC = (STEPS + KS - 1) / KS;       // steps per chain
    static_assert(CIN % 8 == 0 && COUT % NCO == 0 && NV % 16 == 0, "layer shape");
    static_assert((CPT & (CPT - 1)) == 0, "chunks per voxel: a power of two");
    static_assert(SRC != SRC_UP || (BZ % 2 == 0 && BY % 2 == 0 && BX % 2 == 0), "upsampled bricks");
    static_assert(UNITS % 4 == 0 || 4 % UNITS == 0, "units per block");
    static_assert(KS % WPU == 0, "chains per wave");
    static_assert(SRC != SRC_FEAT || CIN == 8, "features are eight channels");
    static_assert(DST != DST_POOL || (BZ % 2 == 0 && BY % 2 == 0 && BX % 2 == 0), "pooled bricks");
    static_assert(DST != DST_HEAD || COUT == 16, "the head follows a 16-channel layer");

    // LDS holds 16-byte chunks (8 channels): the halo's chunk ch of voxel hv
    // at slot hv * CPT + (ch ^ (hv & SWZ)), the weights' chunk ch of row
    // (tap, co) at slot (tap * NCO + co) * CPT + (ch ^ (co & SWZ)); the XOR
    // spreads the 16 rows an MFMA operand reads over the LDS banks
    __shared__ __attribute__((aligned(16))) uint16_t sx[HV * CIN];
    __shared__ __attribute__((aligned(16))) uint16_t su[SRC == SRC_UP ? UV * CIN : 8];
    __shared__ __attribute__((aligned(16))) uint16_t sw[27 * NCO * CIN];
    __shared__ f32x4 red[WPU > 1 ? UNITS * KS : 1][64];
    __shared__ __attribute__((aligned(16))) uint16_t so[DST == DST_POOL ? NV * NCO : 8];
    __shared__ float hy[DST == DST_HEAD ? NV : 1][17];
    __shared__ double s0[4], s1[4];
    __shared__ float s_sig;

    if (a.stop && *a.stop) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row = lane & 15, grp = lane >> 4;
    const int R = a.R;
    const int nbx = (R + BX - 1) / BX, nby = (R + BY - 1) / BY;
    const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nby, bz = blockIdx.x / (nbx * nby);
    const int z0 = bz * BZ, y0 = by * BY, x0 = bx * BX;
    const int n0 = blockIdx.y * NCO;

    constexpr int NX = HV * CPT;
    constexpr int IX = (NX + 255) / 256;
    // halo element it of this thread (element e = slot e): whether it exists,
    // its voxel's cube index (clamped into the cube), whether the voxel is
    // inside the cube, and the chunk it holds
    auto halo = [&](int it, bool &has, long &g, bool &in, int &ch) {
        const int e0 = tid + it * 256, e = e0 < NX ? e0 : NX - 1;
        has = e0 < NX;
        const int hv = e / CPT;
        ch = (e - hv * CPT) ^ (hv & SWZ);
        const int hx = hv % HX, hyy = (hv / HX) % HY, hz = hv / (HX * HY);
        const int z = z0 - 1 + hz, y = y0 - 1 + hyy, x = x0 - 1 + hx;
        in = (unsigned)z < (unsigned)R && (unsigned)y < (unsigned)R && (unsigned)x < (unsigned)R;
        const int zc = min(max(z, 0), R - 1), yc = min(max(y, 0), R - 1), xc = min(max(x, 0), R - 1);
        g = ((long)zc * R + yc) * R + xc;
    };
    // wave-uniform LDS address of a 64-slot run (direct loads write lane l's
    // 16 bytes at that address + 16 l).  A macro: the host pass of this
    // compiler drops the kernel's launch stub when a lambda returns an LDS
    // pointer.
#define RUN_OF(base, it) ((lds_ptr)((base) + ((it) * 256 + wave * 64) * 8))

    // ---- staging: direct global -> LDS loads (global_load_lds_dwordx4) for
    // the weights and the halos, so the block waits on global memory once
    {
        constexpr int NW = 27 * NCO * CPT;
#pragma unroll
        for (int it = 0; it < (NW + 255) / 256; ++it) {
            const int e = tid + it * 256;
            if (e < NW) {
                const int tap = e / (NCO * CPT), r = e - tap * (NCO * CPT), co = r / CPT, ch = (r - co * CPT) ^ (co & SWZ);
                __builtin_amdgcn_global_load_lds(a.W + ((long)(tap * COUT + n0 + co) * CIN + ch * 8), RUN_OF(sw, it), 16,
                                                 0, 0);
            }
        }
    }
    if constexpr (SRC != SRC_FEAT) {
#pragma unroll
        for (int it = 0; it < IX; ++it) {
            bool has, in;
            long g;
            int ch;
            halo(it, has, g, in, ch);
            if (has && in)
                __builtin_amdgcn_global_load_lds(a.X + g * CIN + ch * 8, RUN_OF(sx, it), 16, 0, 0);
            else if (has)
                *reinterpret_cast<uint4 *>(sx + (tid + it * 256) * 8) = make_uint4(0, 0, 0, 0);
        }
    }
    if constexpr (SRC == SRC_UP) {
        constexpr int NU = UV * CPT;
        const int Rc = R >> 1;
#pragma unroll
        for (int it = 0; it < (NU + 255) / 256; ++it) {
            const int e = tid + it * 256;
            if (e < NU) {
                const int pv = e / CPT, ch = (e - pv * CPT) ^ (pv & SWZ);
                const int px = pv % UX, py = (pv / UX) % UY, pz = pv / (UX * UY);
                const int z = z0 / 2 - 1 + pz, y = y0 / 2 - 1 + py, x = x0 / 2 - 1 + px;
                if ((unsigned)z < (unsigned)Rc && (unsigned)y < (unsigned)Rc && (unsigned)x < (unsigned)Rc)
                    __builtin_amdgcn_global_load_lds(a.U + (((long)z * Rc + y) * Rc + x) * CIN + ch * 8,
                                                     RUN_OF(su, it), 16, 0, 0);
                else
                    *reinterpret_cast<uint4 *>(su + e * 8) = make_uint4(0, 0, 0, 0);
            }
        }
    }
    if constexpr (SRC == SRC_FEAT) {
        // the halo's (A, b, L) are loaded before sigma is formed
        int av[IX];
        float bg[IX], lg[IX];
#pragma unroll
        for (int it = 0; it < IX; ++it) {
            bool has, in;
            long g;
            int ch;
            halo(it, has, g, in, ch);
            av[it] = a.A[g];
            bg[it] = a.b[g];
            lg[it] = a.L[g];
        }
        // sigma = RMS of b over FLUID cells, summed as k_features_p sums it
        double pa = 0.0, pc = 0.0;
        for (int k = tid; k < a.nparts; k += 256) {
            pa += a.part[2 * k];
            pc += a.part[2 * k + 1];
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            pa += __shfl_xor(pa, o, 64);
            pc += __shfl_xor(pc, o, 64);
        }
        if (lane == 0) {
            s0[wave] = pa;
            s1[wave] = pc;
        }
        __syncthreads();
        if (tid == 0) {
            const double ta = s0[0] + s0[1] + s0[2] + s0[3], tc = s1[0] + s1[1] + s1[2] + s1[3];
            const double v = tc > 0.0 ? sqrt(ta / tc) : 0.0;
            const float sg = v > 0.0 ? (float)v : 1.0f;
            s_sig = sg;
            if (blockIdx.x == 0 && blockIdx.y == 0) a.sigma[0] = sg;
        }
        __syncthreads();
        const float sig = s_sig;
#pragma unroll
        for (int it = 0; it < IX; ++it) {
            bool has, in;
            long g;
            int ch;
            halo(it, has, g, in, ch);
            const int A = av[it];
            const float f0 = A == FLIP_FLUID ? bg[it] / sig : 0.0f;
            const float f1 = A == FLIP_FLUID, f2 = A == FLIP_AIR, f3 = A == FLIP_WALL;
            const float f4 = A != FLIP_WALL ? fminf(fmaxf(lg[it], -2.0f), 2.0f) : 0.0f;
            uint4 v = make_uint4((uint32_t)f2bf_feat(f0) | ((uint32_t)f2bf_feat(f1) << 16),
                                 (uint32_t)f2bf_feat(f2) | ((uint32_t)f2bf_feat(f3) << 16), (uint32_t)f2bf_feat(f4),
                                 0u);
            if (!in) v = make_uint4(0, 0, 0, 0);
            if (has) *reinterpret_cast<uint4 *>(sx + (tid + it * 256) * 8) = v;
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    if constexpr (SRC == SRC_UP) {
        // up(U) + X in place, rounded to bf16 as k_up_add rounds it
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IX; ++it) {
            bool has, in;
            long g;
            int ch;
            halo(it, has, g, in, ch);
            if (!(has && in)) continue;
            const int e = tid + it * 256, hv = e / CPT;
            const int hx = hv % HX, hyy = (hv / HX) % HY, hz = hv / (HX * HY);
            // parent offsets inside the parents' halo: (z0 - 1 + hz) / 2 - (z0 / 2 - 1)
            const int pv = (((hz + 1) >> 1) * UY + ((hyy + 1) >> 1)) * UX + ((hx + 1) >> 1);
            uint4 *xp = reinterpret_cast<uint4 *>(sx + e * 8);
            const uint4 xs = *xp;
            const uint4 us = *reinterpret_cast<const uint4 *>(su + (pv * CPT + (ch ^ (pv & SWZ))) * 8);
            *xp = make_uint4(add_bf16x2(us.x, xs.x), add_bf16x2(us.y, xs.y), add_bf16x2(us.z, xs.z),
                             add_bf16x2(us.w, xs.w));
        }
    }
    __syncthreads();

    // ---- implicit GEMM from LDS
    const int sub = WPU == 1 ? 0 : wave % WPU;
    int uid[UPW], lz[UPW], ly[UPW], lx[UPW];
    bool rv[UPW];
#pragma unroll
    for (int uu = 0; uu < UPW; ++uu) {
        uid[uu] = WPU == 1 ? wave * UPW + uu : wave / WPU;
        const int lv = (uid[uu] % MT) * 16 + row;
        lz[uu] = lv / (BY * BX);
        ly[uu] = (lv / BX) % BY;
        lx[uu] = lv % BX;
        rv[uu] = z0 + lz[uu] < R && y0 + ly[uu] < R && x0 + lx[uu] < R;
    }
    f32x4 acc[UPW][CPW];
#pragma unroll
    for (int uu = 0; uu < UPW; ++uu)
#pragma unroll
        for (int lc = 0; lc < CPW; ++lc) acc[uu][lc] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < TPC; ++t)
#pragma unroll
        for (int lc = 0; lc < CPW; ++lc) {
            const int st = sub + lc * WPU + t * KS;  // step t of chain sub + lc * WPU
            if (st >= STEPS) continue;
            const int q = st * 4 + grp;
            const bool live = q < NCH;
            const int tap = live ? q / CPT : 0, cc = live ? q % CPT : 0;
            const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
#pragma unroll
            for (int uu = 0; uu < UPW; ++uu) {
                const int nt = uid[uu] / MT;
                const int hv = ((lz[uu] + dz) * HY + ly[uu] + dy) * HX + lx[uu] + dx, wr = tap * NCO + nt * 16 + row;
                const bf16x8 av = (live && rv[uu]) ? lds8(sx + (hv * CPT + (cc ^ (hv & SWZ))) * 8) : zero8();
                const bf16x8 bv = live ? lds8(sw + (wr * CPT + (cc ^ (wr & SWZ))) * 8) : zero8();
                acc[uu][lc] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[uu][lc], 0, 0, 0);
            }
        }

    // ---- chains in order, bias, leaky ReLU, bf16
    f32x4 tot[UPW];
    bool owner = true;
    if constexpr (WPU == 1) {
#pragma unroll
        for (int uu = 0; uu < UPW; ++uu) {
            tot[uu] = acc[uu][0];
#pragma unroll
            for (int lc = 1; lc < CPW; ++lc)
#pragma unroll
                for (int i = 0; i < 4; ++i) tot[uu][i] += acc[uu][lc][i];
        }
    } else {
#pragma unroll
        for (int lc = 0; lc < CPW; ++lc) red[uid[0] * KS + sub + lc * WPU][lane] = acc[0][lc];
        __syncthreads();
        owner = sub == 0;
        tot[0] = red[uid[0] * KS][lane];
#pragma unroll
        for (int j = 1; j < KS; ++j) {
            const f32x4 o = red[uid[0] * KS + j][lane];
#pragma unroll
            for (int i = 0; i < 4; ++i) tot[0][i] += o[i];
        }
    }
    if (owner) {
#pragma unroll
        for (int uu = 0; uu < UPW; ++uu) {
            const int m = uid[uu] % MT, nt = uid[uu] / MT;
            const int co = nt * 16 + row;
            const float bv = a.bias[n0 + co];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int lv = m * 16 + grp * 4 + i;
                float val = tot[uu][i] + bv;
                if (val < 0.0f) val *= UNET_LEAK;
                if constexpr (DST == DST_HEAD) {
                    hy[lv][co] = bf2f(f2bf(val));
                } else {
                    const int z = z0 + lv / (BY * BX), y = y0 + (lv / BX) % BY, x = x0 + lv % BX;
                    const uint16_t h = f2bf(val);
                    if (z < R && y < R && x < R) a.Y[(((long)z * R + y) * R + x) * COUT + n0 + co] = h;
                    if constexpr (DST == DST_POOL) so[lv * NCO + co] = h;
                }
            }
        }
    }
    if constexpr (DST == DST_POOL) {
        // 2x2x2 average of the brick's bf16 outputs, in k_pool2's order
        __syncthreads();
        constexpr int PX = BX / 2, PY = BY / 2, PZ = BZ / 2;
        const int Rh = R >> 1;
        for (int e = tid; e < PZ * PY * PX * NCO; e += 256) {
            const int c = e % NCO, pv = e / NCO;
            const int px = pv % PX, py = (pv / PX) % PY, pz = pv / (PX * PY);
            const int gz = z0 / 2 + pz, gy = y0 / 2 + py, gx = x0 / 2 + px;
            if (gz >= Rh || gy >= Rh || gx >= Rh) continue;
            float s = 0.0f;
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                const int cz = 2 * pz + (d >> 2), cy = 2 * py + ((d >> 1) & 1), cx = 2 * px + (d & 1);
                s += bf2f(so[((cz * BY + cy) * BX + cx) * NCO + c]);
            }
            a.P[(((long)gz * Rh + gy) * Rh + gx) * COUT + n0 + c] = f2bf(s * 0.125f);
        }
    }
    if constexpr (DST == DST_HEAD) {
        // x = mask (w . y + b) sigma scale, as the split-K kernel's head
        __syncthreads();
        for (int t = tid; t < NV; t += 256) {
            const int z = z0 + t / (BY * BX), y = y0 + (t / BX) % BY, x = x0 + t % BX;
            if (z >= R || y >= R || x >= R) continue;
            const long g = ((long)z * R + y) * R + x;
            float sh = a.hb[0];
            for (int c = 0; c < 16; ++c) sh += a.hw[c] * hy[t][c];
            const float mask = a.A[g] == FLIP_FLUID ? 1.0f : 0.0f;
            const float out = mask * sh;
            a.x[g] = out * a.sigma[0] * a.scale;
        }
    }
}

#undef RUN_OF

template <int CIN, int COUT, int SRC, int DST, int KS, int BZ, int BY, int BX, int NTW>
void launch_layer(const LayerArgs &a, hipStream_t s)
{
    const int R = a.R;
    const unsigned bricks = (unsigned)(((R + BZ - 1) / BZ) * ((R + BY - 1) / BY) * ((R + BX - 1) / BX));
    hipLaunchKernelGGL((k_unet_layer<CIN, COUT, SRC, DST, KS, BZ, BY, BX, NTW>), dim3(bricks, COUT / (16 * NTW)),
                       dim3(256), 0, s, a);
}

}  // namespace

// The network's forward for one system in seven launches (see the file
// comment); sigma and the solution as unet_solve_partials writes them.
void unet_infer_layers(UNet *u, const int8_t *A, const float *L, const float *b, const double *part, int nparts,
                       const int32_t *stop, float *x, float *sigma)
{
    hipStream_t s = u->stream;
    const int R0 = u->R, R1 = R0 / 2, R2 = R0 / 4, R3 = R0 / 8;
    auto W = [&](int i) { return u->pwb + u->conv[i].w_off; };
    auto B = [&](int i) { return u->pw + u->conv[i].b_off; };
    LayerArgs a{};
    a.stop = stop;
    // encoder: features, conv 8 -> 16 and pool at R0
    a.R = R0;
    a.W = W(0);
    a.bias = B(0);
    a.Y = u->e0;
    a.P = u->p0;
    a.A = A;
    a.L = L;
    a.b = b;
    a.part = part;
    a.nparts = nparts;
    a.sigma = sigma;
    launch_layer<8, 16, SRC_FEAT, DST_POOL, 4, 4, 4, 8, 1>(a, s);
    // conv 16 -> 32 and pool at R1
    a.R = R1;
    a.X = u->p0;
    a.W = W(1);
    a.bias = B(1);
    a.Y = u->e1;
    a.P = u->p1;
    launch_layer<16, 32, SRC_PLAIN, DST_POOL, 4, 2, 4, 4, 2>(a, s);
    // conv 32 -> 64 and pool at R2
    a.R = R2;
    a.X = u->p1;
    a.W = W(2);
    a.bias = B(2);
    a.Y = u->e2;
    a.P = u->p2;
    launch_layer<32, 64, SRC_PLAIN, DST_POOL, 8, 2, 2, 4, 1>(a, s);
    // bottom: conv 64 -> 64 at R3
    a.R = R3;
    a.X = u->p2;
    a.W = W(3);
    a.bias = B(3);
    a.Y = u->e3;
    a.P = nullptr;
    launch_layer<64, 64, SRC_PLAIN, DST_STORE, 16, 1, 4, 4, 1>(a, s);
    // decoder: conv(up(e3) + e2) 64 -> 32 at R2
    a.R = R2;
    a.X = u->e2;
    a.U = u->e3;
    a.W = W(4);
    a.bias = B(4);
    a.Y = u->d2;
    launch_layer<64, 32, SRC_UP, DST_STORE, 16, 2, 2, 4, 1>(a, s);
    // conv(up(d2) + e1) 32 -> 16 at R1
    a.R = R1;
    a.X = u->e1;
    a.U = u->d2;
    a.W = W(5);
    a.bias = B(5);
    a.Y = u->d1;
    launch_layer<32, 16, SRC_UP, DST_STORE, 8, 2, 4, 4, 1>(a, s);
    // conv(up(d1) + e0) 16 -> 16 at R0 with the head
    a.R = R0;
    a.X = u->e0;
    a.U = u->d1;
    a.W = W(6);
    a.bias = B(6);
    a.Y = nullptr;
    a.hw = u->pw + u->head_w;
    a.hb = u->pw + u->head_b;
    a.scale = UNET_OUT_UNIT / ((float)R0 * (float)R0);
    a.x = x;
    launch_layer<16, 16, SRC_UP, DST_HEAD, 4, 4, 4, 8, 1>(a, s);
}
