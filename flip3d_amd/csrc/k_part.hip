// k_part.hip - particle-side kernels of the FLIP step on cell-sorted storage.
//
// Particles are stored sorted by (cell, creation index), so the particles of
// the cells (si, sj, klo..khi) form ONE contiguous range
// [cell_start[c(si,sj,klo)], cell_start[c(si,sj,khi)+1]) whose order is exactly
// the reference's window order (sorter.cpp:42-64: si -> sj -> sk, then bucket
// order).  A 3x3x3 window is 9 ranges, a 2x4x4 P2G window 8 ranges.
// pos.w carries the mass with the sign bit set for WALL particles, so
// neighbour loops need no type load.
#include "flip_device.hpp"
#include "flip_internal.hpp"

#include <algorithm>
#include <cstring>
#include <cmath>
#include <unordered_map>

using namespace flipdev;

namespace {

constexpr int TB = 256;
inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ bool is_wall_w(float w) { return __builtin_signbit(w); }
__device__ __forceinline__ bool finite3(float a, float b, float c)
{
    return __builtin_isfinite(a) && __builtin_isfinite(b) && __builtin_isfinite(c);
}
__device__ __forceinline__ float mass_of(float w) { return __builtin_fabsf(w); }

// ------------------------------------------------------------- density
// flip3D.cpp:166-190 computeDensity (+ add_ExtForce when gdt != 0 is fused
// by the step driver through a separate call).
// POW2: h*h is a power of two, so r2 / (h*h) == r2 * (1 / (h*h)) exactly
// (scaling by a power of two rounds nothing while the result stays finite).
// FORCE: add_ExtForce fused in (flip3D.cpp:870-875, u_y += gdt for every
// particle); it reads and writes only velocities, which density does not use.
struct P3 {
    float x, y, z;
};

// The unit-mass P2G records of particle i: (x, y, z, u_c) per component with
// the position scaled by n and clamped to [0, n] (mapper.cpp:27-29); a WALL
// particle is placed at infinity, outside every face's support.
__device__ __forceinline__ void p2g_records_wall(int64_t i, float4 *px, float4 *py, float4 *pz)
{
    const float inf = __builtin_inff();
    const float4 far = make_float4(inf, inf, inf, 0.0f);
    px[i] = far;
    py[i] = far;
    pz[i] = far;
}

__device__ __forceinline__ void p2g_records(int64_t i, int n, const float4 &p, const float4 &v, float4 *px,
                                            float4 *py, float4 *pz)
{
    const float g = (float)n;
    const float x = fmaxf(fminf(g * p.x, g), 0.0f);
    const float y = fmaxf(fminf(g * p.y, g), 0.0f);
    const float z = fmaxf(fminf(g * p.z, g), 0.0f);
    px[i] = make_float4(x, y, z, v.x);
    py[i] = make_float4(x, y, z, v.y);
    pz[i] = make_float4(x, y, z, v.z);
}

template <bool POW2, bool FORCE>
__global__ __launch_bounds__(TB) void k_density(int64_t np, int n, const float4 *__restrict__ pos,
                                                float4 *__restrict__ vel, const int32_t *__restrict__ cs, float h,
                                                float inv_hh, float max_dens, const double *__restrict__ dtv,
                                                double gravity)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double gdt = FORCE ? -dtv[0] * gravity : 0.0;  // add_ExtForce's -DT * GRAVITY
    float4 p = pos[i];
    if (is_wall_w(p.w)) {
        if (FORCE) {
            float4 v = vel[i];
            v.y = (float)((double)v.y + gdt);
            v.w = 1.0f;
            vel[i] = v;
        } else {
            vel[i].w = 1.0f;
        }
        return;
    }
    int ci = cell_of(n, p.x), cj = cell_of(n, p.y), ck = cell_of(n, p.z);
    int klo = max(ck - 1, 0), khi = min(ck + 1, n - 1);
    // The window's nine rows (si, sj) in the reference's order; their bounds
    // are loaded together (rows outside the grid are empty) and each row is
    // read eight particles at a time, so a wave waits on memory fewer times.
    int rlo[9], rhi[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int si = ci - 1 + r / 3, sj = cj - 1 + r % 3;
        const bool ok = si >= 0 && si <= n - 1 && sj >= 0 && sj <= n - 1;
        const long c0 = ok ? cidx(n, si, sj, klo) : cidx(n, ci, cj, klo);
        const int b = cs[c0], e = cs[c0 + (khi - klo) + 1];
        rlo[r] = ok ? b : 0;
        rhi[r] = ok ? e : 0;
    }
    float wsum = 0.0f;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        for_range8(pos, rlo[r], rhi[r], [&](const float4 &q, int) {
            if (is_wall_w(q.w)) return;
            const float r2 = length2(q.x, q.y, q.z, p.x, p.y, p.z);
            float w = q.w * (POW2 ? smooth_kernel_pow2(r2, inv_hh) : smooth_kernel(r2, h));
            wsum += w;
        });
    }
    if (FORCE) {
        float4 v = vel[i];
        v.y = (float)((double)v.y + gdt);
        v.w = wsum / max_dens;
        vel[i] = v;
    } else {
        vel[i].w = wsum / max_dens;
    }
}

// k_density<true, FORCE> for unit masses while the storage is sorted: the
// window is read from the sort's 12-byte records xyz3, where a WALL particle
// sits at +inf.  Its weight is then exactly +0, which leaves the sum
// unchanged (it starts at +0 and never becomes -0), as skipping it does; a
// unit mass times the kernel is the kernel.  The storage is sorted by cell, so the block's 256 particles lie in a few (i, j)
// cell rows ("groups", k ascending inside each).  For every group the block
// copies the union of its particles' windows, nine rows (i+-1, j+-1) over the
// cells [kmin-1, kmax+1], from the sort's records into LDS with coalesced
// loads; every particle then reads its own window rows out of LDS in the
// reference's order, so the sums are those of k_density bit for bit.  The
// lanes of one cell read the same LDS words together (a broadcast), where
// each had issued its own global gather.  A block with more than DL_G groups
// or more than DL_CAP staged records takes the global loop.
constexpr int DL_CAP = 2560;  // staged records per block (30 KiB)
constexpr int DL_G = 3;

template <bool FORCE>
__global__ __launch_bounds__(TB) void k_density_lds(int64_t np, int n, const float4 *__restrict__ pos,
                                                    float4 *__restrict__ vel, const int32_t *__restrict__ cs,
                                                    float inv_hh, float max_dens, const double *__restrict__ dtv,
                                                    double gravity, const P3 *__restrict__ xyz3, float4 *px,
                                                    float4 *py, float4 *pz, float *dw)
{
    __shared__ P3 st[DL_CAP];
    __shared__ int skey[TB];
    __shared__ int wcnt[TB / 64];
    __shared__ int gkey[DL_G], gk0[DL_G], gk1[DL_G];
    __shared__ int rgb[DL_G * 9], rlb[DL_G * 9], rlen[DL_G * 9];
    __shared__ int s_ngrp, s_total;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int64_t i = (int64_t)blockIdx.x * TB + t;
    const bool valid = i < np;
    const double gdt = FORCE ? -dtv[0] * gravity : 0.0;
    float4 p = valid ? pos[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int ci = cell_of(n, p.x), cj = cell_of(n, p.y), ck = cell_of(n, p.z);
    const int key = ci * n + cj;
    skey[t] = valid ? key : -1;
    __syncthreads();
    // groups: runs of equal (i, j) keys (nondecreasing along the storage)
    const bool first = valid && (t == 0 || skey[t - 1] != key);
    const bool last = valid && (t == TB - 1 || skey[t + 1] != key);
    const uint64_t bal = __ballot(first);
    if (lane == 0) wcnt[wv] = __popcll(bal);
    __syncthreads();
    int g = __popcll(bal & ((2ull << lane) - 1)) - 1;  // starts up to and including this lane
    for (int w = 0; w < wv; ++w) g += wcnt[w];
    if (t == 0) {
        int s = 0;
        for (int w = 0; w < TB / 64; ++w) s += wcnt[w];
        s_ngrp = s;
    }
    if (first && g < DL_G) {
        gkey[g] = key;
        gk0[g] = ck;
    }
    if (last && g < DL_G) gk1[g] = ck;
    __syncthreads();
    const int ngrp = s_ngrp;
    bool staged = ngrp <= DL_G;
    if (staged) {
        if (t < ngrp * 9) {
            const int gg = t / 9, r = t % 9;
            const int gi = gkey[gg] / n, gj = gkey[gg] % n;
            const int si = gi - 1 + r / 3, sj = gj - 1 + r % 3;
            int b = 0, e = 0;
            if (si >= 0 && si <= n - 1 && sj >= 0 && sj <= n - 1) {
                const int klo = max(gk0[gg] - 1, 0), khi = min(gk1[gg] + 1, n - 1);
                const long c0 = cidx(n, si, sj, klo);
                b = cs[c0];
                e = cs[c0 + (khi - klo) + 1];
            }
            rgb[t] = b;
            rlen[t] = e - b;
        }
        __syncthreads();
        if (t == 0) {
            int s = 0;
            for (int q = 0; q < ngrp * 9; ++q) {
                rlb[q] = s;
                s += rlen[q];
            }
            s_total = s;
        }
        __syncthreads();
        staged = s_total <= DL_CAP;
        if (staged) {
            for (int q = 0; q < ngrp * 9; ++q) {
                const int b = rgb[q], l = rlen[q], o = rlb[q];
                for (int e = t; e < l; e += TB) st[o + e] = xyz3[b + e];
            }
            __syncthreads();
        }
    }
    if (!valid) return;
    if (is_wall_w(p.w)) {
        float4 v = vel[i];
        if (FORCE) v.y = (float)((double)v.y + gdt);
        v.w = 1.0f;
        vel[i] = v;
        if (px) p2g_records_wall(i, px, py, pz);
        if (dw) dw[i] = -1.0f;
        return;
    }
    const int klo = max(ck - 1, 0), khi = min(ck + 1, n - 1);
    float wsum = 0.0f;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int si = ci - 1 + r / 3, sj = cj - 1 + r % 3;
        if (si < 0 || si > n - 1 || sj < 0 || sj > n - 1) continue;
        const long c0 = cidx(n, si, sj, klo);
        const int b = cs[c0], e = cs[c0 + (khi - klo) + 1];
        if (staged) {
            const int q = g * 9 + r;
            for_range8(st + (rlb[q] + (b - rgb[q])), 0, e - b, [&](const P3 &s, int) {
                wsum += smooth_kernel_pow2(length2(s.x, s.y, s.z, p.x, p.y, p.z), inv_hh);
            });
        } else {
            for_range8(xyz3, b, e, [&](const P3 &s, int) {
                wsum += smooth_kernel_pow2(length2(s.x, s.y, s.z, p.x, p.y, p.z), inv_hh);
            });
        }
    }
    float4 v = vel[i];
    if (FORCE) v.y = (float)((double)v.y + gdt);
    v.w = wsum / max_dens;
    vel[i] = v;
    if (px) p2g_records(i, n, p, v, px, py, pz);
    if (dw) dw[i] = v.w;
}

// flip3D.cpp:870-875 add_ExtForce: u_y = (float)((double)u_y + gdt), gdt = -dt*g
__global__ void k_ext_force(int64_t np, float4 *vel, const double *__restrict__ dtv, double gravity)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double gdt = -dtv[0] * gravity;
    vel[i].y = (float)((double)vel[i].y + gdt);
}

// --------------------------------------------------------- flags / level set
// sorter.cpp:70-93 markWater + levelset (flip3D.cpp:824-826 recomputes the
// same L): WALL if the bucket holds a wall particle (L = 1.0), else
// L = (float)(0.2*n0 - accm) with accm the float sum of dens in bucket order.
// The block's cells hold one contiguous range of particles: it is first
// copied to LDS with coalesced loads (the wall bit and dens of each), and
// each thread then walks its bucket there; a range over FLAGS_CAP reads
// global memory instead.
constexpr int FLAGS_TB = 256, FLAGS_CAP = 4096;

__global__ __launch_bounds__(FLAGS_TB) void k_flags(long C, const int32_t *__restrict__ cs,
                                                    const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                                    const float *__restrict__ dw, int8_t *A, float *L, double c0,
                                                    int32_t *fflag, float *press)
{
    __shared__ float s_dens[FLAGS_CAP];
    __shared__ uint8_t s_wall[FLAGS_CAP];
    const long cb = (long)blockIdx.x * FLAGS_TB;
    const long ce = cb + FLAGS_TB < C ? cb + FLAGS_TB : C;
    const int pb = cs[cb], pe = cs[ce];
    const bool staged = pe - pb <= FLAGS_CAP;
    if (staged) {
        for (int k = threadIdx.x; k < pe - pb; k += FLAGS_TB) {
            if (dw) {  // the density pass's copy: the density, or -1 for WALL
                const float d = dw[pb + k];
                s_dens[k] = d;
                s_wall[k] = d < 0.0f ? 1 : 0;
            } else {
                s_dens[k] = vel[pb + k].w;
                s_wall[k] = is_wall_w(pos[pb + k].w) ? 1 : 0;
            }
        }
    }
    __syncthreads();
    long c = cb + threadIdx.x;
    if (c >= C) return;
    int b = cs[c], e = cs[c + 1];
    float accm = 0.0f;
    bool wall = false;
    if (staged) {
        for (int a = b - pb; a < e - pb; ++a) {
            if (s_wall[a]) {
                wall = true;
                break;
            }
            accm += s_dens[a];
        }
    } else {
        for (int a = b; a < e; ++a) {
            if (is_wall_w(pos[a].w)) {
                wall = true;
                break;
            }
            accm += vel[a].w;
        }
    }
    int8_t a = WALL;
    if (wall) {
        L[c] = 1.0f;
    } else {
        float l = (float)(c0 - (double)accm);
        L[c] = l;
        a = l < 0.0f ? FLUID : AIR;
    }
    A[c] = a;
    if (fflag) {  // the PCG setup's FLUID flags (k_fluid_flag)
        fflag[c] = a == FLUID ? 1 : 0;
        if (a != FLUID) press[c] = 0.0f;
    }
}

// ------------------------------------------------------------------- P2G
// P2G input: clamp(n*p, 0, n) per axis (mapper.cpp:34-36) for FLUID particles;
// others are placed infinitely far so every weight test rejects them.
__global__ void k_p2g_prep(int64_t np, int n, const float4 *__restrict__ pos, float4 *spos)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float4 p = pos[i];
    if (is_wall_w(p.w)) {
        const float inf = __builtin_inff();
        spos[i] = make_float4(inf, inf, inf, 0.0f);
        return;
    }
    float g = (float)n;
    float x = fmaxf(fminf(g * p.x, g), 0.0f);
    float y = fmaxf(fminf(g * p.y, g), 0.0f);
    float z = fmaxf(fminf(g * p.z, g), 0.0f);
    spos[i] = make_float4(x, y, z, p.w);
}

__device__ __forceinline__ int wall_chk(int8_t a) { return a == WALL ? 1 : -1; }

// Unit-mass P2G input: per face component, (clamped x, y, z, u_c) in one
// float4 (w = m * sharp = sharp when m == 1, which every reference particle has).
__global__ void k_p2g_prep_unit(int64_t np, int n, const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                float4 *px, float4 *py, float4 *pz)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float4 p = pos[i];
    if (is_wall_w(p.w)) {
        p2g_records_wall(i, px, py, pz);
        return;
    }
    p2g_records(i, n, p, vel[i], px, py, pz);
}

// mapper.cpp:17-88 window sum of one face with unit masses and packed
// (x, y, z, u_c) input read from `src` (global memory, or a block's LDS copy
// through loff).  The window is the one described at k_p2g.
template <int COMP>
__device__ __forceinline__ void p2g_face_coord(int i, int j, int k, float &fx, float &fy, float &fz)
{
    fx = COMP == 0 ? (float)i : (float)(i + 0.5);
    fy = COMP == 1 ? (float)j : (float)(j + 0.5);
    fz = COMP == 2 ? (float)k : (float)(k + 0.5);
}

// Branch-free: four window particles' weights are formed side by side (four
// independent division chains per wave), then added in window order.  A
// particle outside the support adds +0, which leaves both sums unchanged:
// they start at +0 and a float sum of values that is -0 needs every term
// -0, so a sum never becomes -0 and x + 0 == x.
__device__ __forceinline__ void p2g_term(const float4 &q, float fx, float fy, float fz, float hh, float &tw,
                                         float &tx)
{
    const float r2 = length2(q.x, q.y, q.z, fx, fy, fz);
    const bool in = r2 < hh;
    const float w = sharp_kernel_in(r2, hh);
    tw = in ? w : 0.0f;
    tx = in ? w * q.w : 0.0f;
}

__device__ __forceinline__ void p2g_accum(const float4 *__restrict__ src, int b, int e, float fx, float fy, float fz,
                                          float hh, float &sumw, float &sumx)
{
    int a = b;
    for (; a + 3 < e; a += 4) {
        float tw[4], tx[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) p2g_term(src[a + u], fx, fy, fz, hh, tw[u], tx[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            sumx += tx[u];
            sumw += tw[u];
        }
    }
    for (; a < e; ++a) {
        float tw, tx;
        p2g_term(src[a], fx, fy, fz, hh, tw, tx);
        sumx += tx;
        sumw += tw;
    }
}

// u_save = raw value (save_grid, flip3D.cpp:971) and u = value with
// enforce_boundary (flip3D.cpp:783-805)
template <int COMP>
__device__ __forceinline__ void p2g_store(int n, int i, int j, int k, float sumw, float sumx,
                                          const int8_t *__restrict__ A, float *__restrict__ ug, float *__restrict__ us)
{
    const int ni = n + (COMP == 0), nj = n + (COMP == 1), nk = n + (COMP == 2);
    const long f = ((long)i * nj + j) * nk + k;
    float val = sumw != 0.0f ? sumx / sumw : 0.0f;
    us[f] = val;
    int a0 = COMP == 0 ? i : (COMP == 1 ? j : k);
    int na = COMP == 0 ? ni : (COMP == 1 ? nj : nk);
    bool zero = a0 == 0 || a0 == na - 1;
    if (!zero) {
        long c1 = cidx(n, i, j, k);
        long c0 = COMP == 0 ? cidx(n, i - 1, j, k) : (COMP == 1 ? cidx(n, i, j - 1, k) : cidx(n, i, j, k - 1));
        zero = wall_chk(A[c1]) * wall_chk(A[c0]) < 0;
    }
    ug[f] = zero ? 0.0f : val;
}

template <int COMP>
__device__ __forceinline__ void p2g_face_global(int n, int i, int j, int k, const float4 *__restrict__ pk,
                                                const int32_t *__restrict__ cs, float hh, float &sumw, float &sumx)
{
    float fx, fy, fz;
    p2g_face_coord<COMP>(i, j, k, fx, fy, fz);
    const int ihi = COMP == 0 ? 0 : 1, jhi = COMP == 1 ? 0 : 1, khi_off = COMP == 2 ? 0 : 1;
    int klo = max(k - 1, 0), khi = min(k + khi_off, n - 1);
    sumw = 0.0f;
    sumx = 0.0f;
    if (klo > khi) return;
    for (int si = i - 1; si <= i + ihi; ++si) {
        if (si < 0 || si > n - 1) continue;
        for (int sj = j - 1; sj <= j + jhi; ++sj) {
            if (sj < 0 || sj > n - 1) continue;
            int b = cs[cidx(n, si, sj, klo)], e = cs[cidx(n, si, sj, khi) + 1];
            p2g_accum(pk, b, e, fx, fy, fz, hh, sumw, sumx);
        }
    }
}

// mapper.cpp:17-88 with unit masses and packed input, over bricks of
// P2G_BI x P2G_BJ x P2G_BK faces (i, j, k; k fastest, one wave per i layer).  The block first copies the particles of
// the brick's whole cell window into LDS: one contiguous range per (si, sj)
// row, rows back to back in row order, with loff[row][c] the LDS offset of
// the row's c-th cell from sklo.  Every face then walks its own window out
// of LDS in the reference's order, so the sums are those of p2g_face_global
// bit for bit.  A brick whose window holds more than P2G_CAP particles reads
// global memory instead.
typedef __attribute__((address_space(3))) void *lds_ptr;
constexpr int P2G_BI = 4, P2G_BJ = 8, P2G_BK = 8;
constexpr int P2G_TB = P2G_BI * P2G_BJ * P2G_BK;
constexpr int P2G_ROWS = (P2G_BI + 2) * (P2G_BJ + 2);
constexpr int P2G_RK = P2G_BK + 2;
constexpr int P2G_CAP = 4608;  // 72 KiB of float4: two blocks per CU

template <int COMP>
__host__ __device__ __forceinline__ void p2g_bricks(int n, int &nbi, int &nbj, int &nbk)
{
    nbi = (n + (COMP == 0) + P2G_BI - 1) / P2G_BI;
    nbj = (n + (COMP == 1) + P2G_BJ - 1) / P2G_BJ;
    nbk = (n + (COMP == 2) + P2G_BK - 1) / P2G_BK;
}

struct P2GLds {
    float4 sp[P2G_CAP];
    int loff[P2G_ROWS][P2G_RK + 1];
    int rbase[P2G_ROWS];
    int total;
};

// one brick of component COMP (brick = the block's brick index)
template <int COMP>
__device__ __forceinline__ void p2g_brick(int n, const float4 *__restrict__ pk, const int32_t *__restrict__ cs,
                                          const int8_t *__restrict__ A, float *__restrict__ ug,
                                          float *__restrict__ us, float hh, int brick, P2GLds &S)
{
    float4 *sp = S.sp;
    auto &loff = S.loff;
    int *rbase = S.rbase;
    int &total = S.total;
    const int ni = n + (COMP == 0), nj = n + (COMP == 1), nk = n + (COMP == 2);
    const int ihi = COMP == 0 ? 0 : 1, jhi = COMP == 1 ? 0 : 1, khi_off = COMP == 2 ? 0 : 1;
    int nbi, nbj, nbk;
    p2g_bricks<COMP>(n, nbi, nbj, nbk);
    const int bk = brick % nbk, bj = (brick / nbk) % nbj, bi = brick / (nbk * nbj);
    const int i0 = bi * P2G_BI, j0 = bj * P2G_BJ, k0 = bk * P2G_BK;
    // rows (si, sj) = (i0 - 1 + r / RJ, j0 - 1 + r % RJ); cells sklo..skhi
    const int RI = P2G_BI + 1 + ihi, RJ = P2G_BJ + 1 + jhi;
    const int sklo = max(k0 - 1, 0), skhi = min(k0 + P2G_BK - 1 + khi_off, n - 1);
    const int nkc = skhi - sklo + 1;  // 1..P2G_RK
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int idx = tid; idx < RI * RJ * (nkc + 1); idx += P2G_TB) {
        const int r = idx / (nkc + 1), c = idx - r * (nkc + 1);
        const int si = i0 - 1 + r / RJ, sj = j0 - 1 + r % RJ;
        const bool ok = si >= 0 && si <= n - 1 && sj >= 0 && sj <= n - 1;
        loff[r][c] = ok ? cs[cidx(n, si, sj, sklo) + c] : 0;
    }
    __syncthreads();
    if (wave == 0) {
        // exclusive scan of the row counts (RI * RJ <= 60 rows, one per lane)
        int cnt = lane < RI * RJ ? loff[lane][nkc] - loff[lane][0] : 0;
        int inc = cnt;
        for (int d = 1; d < 64; d <<= 1) {
            int v = __shfl_up(inc, d, 64);
            if (lane >= d) inc += v;
        }
        if (lane < RI * RJ) rbase[lane] = inc - cnt;
        if (lane == 63) total = inc;
    }
    __syncthreads();
    const bool staged = total <= P2G_CAP;
    if (staged) {
        // Direct global -> LDS copies (global_load_lds_dwordx4: lane l of the
        // wave writes the 16 bytes at the wave-uniform LDS address + 16 l), so
        // a wave issues all of its rows' loads without waiting on any.
        for (int r = wave; r < RI * RJ; r += P2G_TB / 64) {
            const int g0 = loff[r][0], cnt = loff[r][nkc] - g0, base = rbase[r];
            for (int e0 = 0; e0 < cnt; e0 += 64)
                if (e0 + lane < cnt)
                    __builtin_amdgcn_global_load_lds(pk + g0 + e0 + lane, (lds_ptr)(sp + base + e0), 16, 0, 0);
            // every lane has read loff[r][*] above; rebase the row to LDS offsets
            if (lane <= nkc) loff[r][lane] = loff[r][lane] - g0 + base;
        }
        __builtin_amdgcn_s_waitcnt(0);
    }
    __syncthreads();
    const int i = i0 + (tid >> 6), j = j0 + ((tid >> 3) & 7), k = k0 + (tid & 7);
    if (i >= ni || j >= nj || k >= nk) return;
    float sumw, sumx;
    if (!staged) {
        p2g_face_global<COMP>(n, i, j, k, pk, cs, hh, sumw, sumx);
    } else {
        float fx, fy, fz;
        p2g_face_coord<COMP>(i, j, k, fx, fy, fz);
        const int klo = max(k - 1, 0), khi = min(k + khi_off, n - 1);
        sumw = 0.0f;
        sumx = 0.0f;
        if (klo <= khi) {
            for (int si = i - 1; si <= i + ihi; ++si) {
                if (si < 0 || si > n - 1) continue;
                for (int sj = j - 1; sj <= j + jhi; ++sj) {
                    if (sj < 0 || sj > n - 1) continue;
                    const int r = (si - (i0 - 1)) * RJ + (sj - (j0 - 1));
                    p2g_accum(sp, loff[r][klo - sklo], loff[r][khi - sklo + 1], fx, fy, fz, hh, sumw, sumx);
                }
            }
        }
    }
    p2g_store<COMP>(n, i, j, k, sumw, sumx, A, ug, us);
}

// the three components in one launch: nb0 bricks of X faces, then nb1 of Y
// faces, then those of Z faces
__global__ __launch_bounds__(P2G_TB) void k_p2g_lds3(int n, const float4 *__restrict__ p0,
                                                     const float4 *__restrict__ p1, const float4 *__restrict__ p2,
                                                     const int32_t *__restrict__ cs, const int8_t *__restrict__ A,
                                                     float *u0, float *u1, float *u2, float *s0, float *s1, float *s2,
                                                     float hh, int nb0, int nb1)
{
    __shared__ P2GLds S;
    const int b = blockIdx.x;
    if (b < nb0)
        p2g_brick<0>(n, p0, cs, A, u0, s0, hh, b, S);
    else if (b < nb0 + nb1)
        p2g_brick<1>(n, p1, cs, A, u1, s1, hh, b - nb0, S);
    else
        p2g_brick<2>(n, p2, cs, A, u2, s2, hh, b - nb0 - nb1, S);
}

// mapper.cpp:17-88 one face component per launch.  COMP selects the face
// family: X faces (i, j+.5, k+.5) over window [i-1,i]x[j-2,j+1]x[k-2,k+1],
// and the analogous Y/Z windows.  Writes u_save = raw value (save_grid,
// flip3D.cpp:971) and u = value with enforce_boundary (flip3D.cpp:783-805).
template <int COMP>
__global__ __launch_bounds__(TB) void k_p2g(int n, const float4 *__restrict__ spos, const float4 *__restrict__ vel,
                                            const int32_t *__restrict__ cs, const int8_t *__restrict__ A,
                                            float *__restrict__ ug, float *__restrict__ us, float hh)
{
    const int ni = n + (COMP == 0), nj = n + (COMP == 1), nk = n + (COMP == 2);
    long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= (long)ni * nj * nk) return;
    int k = (int)(f % nk);
    int j = (int)((f / nk) % nj);
    int i = (int)(f / ((long)nk * nj));
    float fx = COMP == 0 ? (float)i : (float)(i + 0.5);
    float fy = COMP == 1 ? (float)j : (float)(j + 0.5);
    float fz = COMP == 2 ? (float)k : (float)(k + 0.5);
    // The reference window is [i-2, i+1] on the two transverse axes, but a
    // particle in the first row (cell i-2, clamped coordinate < i-1) is at
    // least 1.5 from the face centre i+0.5, so r2 >= 2.25 > RE^2 and its weight
    // is exactly 0: those rows are skipped.  Along the normal axis the window
    // is [i-1, i] as in the reference.
    const int ilo_off = COMP == 0 ? 1 : 1, jlo_off = 1, klo_off = 1;
    const int ihi_off = COMP == 0 ? 0 : 1, jhi_off = COMP == 1 ? 0 : 1, khi_off = COMP == 2 ? 0 : 1;
    int klo = max(k - klo_off, 0), khi = min(k + khi_off, n - 1);
    float sumw = 0.0f, sumx = 0.0f;
    if (klo <= khi) {
        for (int si = i - ilo_off; si <= i + ihi_off; ++si) {
            if (si < 0 || si > n - 1) continue;
            for (int sj = j - jlo_off; sj <= j + jhi_off; ++sj) {
                if (sj < 0 || sj > n - 1) continue;
                int b = cs[cidx(n, si, sj, klo)], e = cs[cidx(n, si, sj, khi) + 1];
                for (int a = b; a < e; ++a) {
                    float4 q = spos[a];
                    float r2 = length2(q.x, q.y, q.z, fx, fy, fz);
                    if (!(r2 < hh)) continue;  // weight is exactly 0 (sharp_kernel_hh)
                    float w = q.w * sharp_kernel_hh(r2, hh);
                    float4 v = vel[a];
                    float uc = COMP == 0 ? v.x : (COMP == 1 ? v.y : v.z);
                    sumx += w * uc;
                    sumw += w;
                }
            }
        }
    }
    float val = sumw != 0.0f ? sumx / sumw : 0.0f;
    us[f] = val;
    int a0 = COMP == 0 ? i : (COMP == 1 ? j : k);
    int na = COMP == 0 ? ni : (COMP == 1 ? nj : nk);
    bool zero = a0 == 0 || a0 == na - 1;
    if (!zero) {
        long c1 = cidx(n, i, j, k);
        long c0 = COMP == 0 ? cidx(n, i - 1, j, k) : (COMP == 1 ? cidx(n, i, j - 1, k) : cidx(n, i, j, k - 1));
        zero = wall_chk(A[c1]) * wall_chk(A[c0]) < 0;
    }
    ug[f] = zero ? 0.0f : val;
}

// ------------------------------------------------------------------- G2P
// mapper.cpp:96-106 linear(), including its k = min(z, h-2) clamp; a read
// past the last k layer (grid[1] rows, mapper.cpp:102) returns 0.
template <int COMP>
__device__ __forceinline__ float grid_at(const float *q, int n, int i, int j, int k, int d)
{
    if (k >= d) return 0.0f;
    long f = COMP == 0 ? fidx0(n, i, j, k) : (COMP == 1 ? fidx1(n, i, j, k) : fidx2(n, i, j, k));
    return q[f];
}

template <int COMP>
__device__ __forceinline__ float linear(const float *__restrict__ q, int n, float x, float y, float z)
{
    const int w = n + (COMP == 0), h = n + (COMP == 1), d = n + (COMP == 2);
    x = fmaxf(0.0f, fminf((float)w, x));
    y = fmaxf(0.0f, fminf((float)h, y));
    z = fmaxf(0.0f, fminf((float)d, z));
    int i = (int)(x > (float)(w - 2) ? (float)(w - 2) : x);
    int j = (int)(y > (float)(h - 2) ? (float)(h - 2) : y);
    int k = (int)(z > (float)(h - 2) ? (float)(h - 2) : z);
    float ix = (float)(i + 1) - x, xi = x - (float)i;
    float jy = (float)(j + 1) - y, yj = y - (float)j;
    float kz = (float)(k + 1) - z, zk = z - (float)k;
    float lo = (ix * grid_at<COMP>(q, n, i, j, k, d) + xi * grid_at<COMP>(q, n, i + 1, j, k, d)) * jy +
               (ix * grid_at<COMP>(q, n, i, j + 1, k, d) + xi * grid_at<COMP>(q, n, i + 1, j + 1, k, d)) * yj;
    float hi = (ix * grid_at<COMP>(q, n, i, j, k + 1, d) + xi * grid_at<COMP>(q, n, i + 1, j, k + 1, d)) * jy +
               (ix * grid_at<COMP>(q, n, i, j + 1, k + 1, d) + xi * grid_at<COMP>(q, n, i + 1, j + 1, k + 1, d)) * yj;
    return kz * lo + zk * hi;
}

// linear() of the two grids interleaved in q (.x and .y), same arithmetic
template <int COMP>
__device__ __forceinline__ float2 linear2(const float2 *__restrict__ q, int n, float x, float y, float z)
{
    const int w = n + (COMP == 0), h = n + (COMP == 1), d = n + (COMP == 2);
    x = fmaxf(0.0f, fminf((float)w, x));
    y = fmaxf(0.0f, fminf((float)h, y));
    z = fmaxf(0.0f, fminf((float)d, z));
    int i = (int)(x > (float)(w - 2) ? (float)(w - 2) : x);
    int j = (int)(y > (float)(h - 2) ? (float)(h - 2) : y);
    int k = (int)(z > (float)(h - 2) ? (float)(h - 2) : z);
    float ix = (float)(i + 1) - x, xi = x - (float)i;
    float jy = (float)(j + 1) - y, yj = y - (float)j;
    float kz = (float)(k + 1) - z, zk = z - (float)k;
    auto at = [&](int a, int b, int c) -> float2 {
        if (c >= d) return make_float2(0.0f, 0.0f);
        return q[COMP == 0 ? fidx0(n, a, b, c) : (COMP == 1 ? fidx1(n, a, b, c) : fidx2(n, a, b, c))];
    };
    const float2 q000 = at(i, j, k), q100 = at(i + 1, j, k), q010 = at(i, j + 1, k), q110 = at(i + 1, j + 1, k);
    const float2 q001 = at(i, j, k + 1), q101 = at(i + 1, j, k + 1), q011 = at(i, j + 1, k + 1),
                 q111 = at(i + 1, j + 1, k + 1);
    float2 r;
    {
        float lo = (ix * q000.x + xi * q100.x) * jy + (ix * q010.x + xi * q110.x) * yj;
        float hi = (ix * q001.x + xi * q101.x) * jy + (ix * q011.x + xi * q111.x) * yj;
        r.x = kz * lo + zk * hi;
    }
    {
        float lo = (ix * q000.y + xi * q100.y) * jy + (ix * q010.y + xi * q110.y) * yj;
        float hi = (ix * q001.y + xi * q101.y) * jy + (ix * q011.y + xi * q111.y) * yj;
        r.y = kz * lo + zk * hi;
    }
    return r;
}

// mapper.cpp:108-112 fetchVelocity; gn*p - 0.5 in double rounds like the
// float subtraction (one rounding of an exact value).
__device__ __forceinline__ float3 fetch_velocity(const float *g0, const float *g1, const float *g2, int n, float px,
                                                 float py, float pz)
{
    float gx = (float)n * px, gy = (float)n * py, gz = (float)n * pz;
    float3 u;
    u.x = linear<0>(g0, n, gx, gy - 0.5f, gz - 0.5f);
    u.y = linear<1>(g1, n, gx - 0.5f, gy, gz - 0.5f);
    u.z = linear<2>(g2, n, gx - 0.5f, gy - 0.5f, gz);
    return u;
}

// flip3D.cpp:1071-1097 FLIP/PIC: t = G2P(du) + u; u = (1-a)*G2P(u_grid) + a*t
// in double, rounded to float.  Wall velocities are not read by the step and
// are left as they are.
__global__ __launch_bounds__(TB) void k_g2p_flip(int64_t np, int n, const float4 *__restrict__ pos,
                                                 float4 *__restrict__ vel, const float *__restrict__ u0,
                                                 const float *__restrict__ u1, const float *__restrict__ u2,
                                                 const float *__restrict__ d0, const float *__restrict__ d1,
                                                 const float *__restrict__ d2, double c_pic, double c_flip)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float4 p = pos[i];
    if (is_wall_w(p.w)) return;
    float4 v = vel[i];
    float3 du = fetch_velocity(d0, d1, d2, n, p.x, p.y, p.z);
    float tx = du.x + v.x, ty = du.y + v.y, tz = du.z + v.z;
    float3 pic = fetch_velocity(u0, u1, u2, n, p.x, p.y, p.z);
    v.x = (float)(c_pic * (double)pic.x + c_flip * (double)tx);
    v.y = (float)(c_pic * (double)pic.y + c_flip * (double)ty);
    v.z = (float)(c_pic * (double)pic.z + c_flip * (double)tz);
    vel[i] = v;
}

// k_g2p_flip reading the extrapolation's interleaved (u, u - u_save) faces
__global__ __launch_bounds__(TB) void k_g2p_flip2(int64_t np, int n, const float4 *__restrict__ pos,
                                                  float4 *__restrict__ vel, const float2 *__restrict__ q0,
                                                  const float2 *__restrict__ q1, const float2 *__restrict__ q2,
                                                  double c_pic, double c_flip)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float4 p = pos[i];
    if (is_wall_w(p.w)) return;
    float4 v = vel[i];
    const float gx = (float)n * p.x, gy = (float)n * p.y, gz = (float)n * p.z;
    const float2 a = linear2<0>(q0, n, gx, gy - 0.5f, gz - 0.5f);
    const float2 b = linear2<1>(q1, n, gx - 0.5f, gy, gz - 0.5f);
    const float2 c = linear2<2>(q2, n, gx - 0.5f, gy - 0.5f, gz);
    float tx = a.y + v.x, ty = b.y + v.y, tz = c.y + v.z;
    v.x = (float)(c_pic * (double)a.x + c_flip * (double)tx);
    v.y = (float)(c_pic * (double)b.x + c_flip * (double)ty);
    v.z = (float)(c_pic * (double)c.x + c_flip * (double)tz);
    vel[i] = v;
}

__global__ void k_g2p_points(int64_t m, int n, const float4 *__restrict__ p, const float *__restrict__ u0,
                             const float *__restrict__ u1, const float *__restrict__ u2, float4 *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    float4 q = p[i];
    float3 u = fetch_velocity(u0, u1, u2, n, q.x, q.y, q.z);
    out[i] = make_float4(u.x, u.y, u.z, 0.0f);
}

// ---------------------------------------------------------------- advect
// flip3D.cpp:933-955 stuck test of a moved FLUID particle: in a WALL cell, or
// thin (dens < 0.04) next to a wall in y.
__device__ __forceinline__ bool is_stuck(int n, const float4 &p, float dens, const int8_t *__restrict__ A)
{
    bool rem = A[cidx(n, cell_of_b(n, p.x), cell_of_b(n, p.y), cell_of_b(n, p.z))] == WALL;
    float g = (float)n;
    int ii = (int)fmin((double)(n - 3), fmax(2.0, (double)(p.x * g)));
    int jj = (int)fmin((double)(n - 3), fmax(2.0, (double)(p.y * g)));
    int kk = (int)fmin((double)(n - 3), fmax(2.0, (double)(p.z * g)));
    // for n < 5 the reference indexes out of range; keep the window inside
    ii = min(max(ii, 0), n - 1);
    jj = min(max(jj, 0), n - 1);
    kk = min(max(kk, 0), n - 1);
    int jm = jj - 1 > 0 ? jj - 1 : 0, jp = jj + 1 < n - 1 ? jj + 1 : n - 1;
    if ((double)dens < 0.04 && (A[cidx(n, ii, jm, kk)] == WALL || A[cidx(n, ii, jp, kk)] == WALL)) rem = true;
    return rem;
}

// flip3D.cpp:877-930: p += dt*G2P(u) (double), clamp to [1/N, 1-1/N], then
// push out of wall particles within 1.5*DENSITY/N, in window order, updating
// p in place and projecting out the normal velocity.  Wall particles never
// move, so the static wall bins are the reference's post-advect wall buckets.
// With slist, the stuck particles' creation indices are appended to it
// (unordered; the reposition plan orders them).
__global__ __launch_bounds__(TB) void k_advect(int64_t np, int n, float4 *__restrict__ pos, float4 *__restrict__ vel,
                                               const float *__restrict__ u0, const float *__restrict__ u1,
                                               const float *__restrict__ u2, const double *__restrict__ dtv,
                                               float rlo, double re,
                                               const int32_t *__restrict__ wst, const int8_t *__restrict__ wnear,
                                               const float4 *__restrict__ wpos, const float4 *__restrict__ wnrm,
                                               int32_t *nonfinite, const int32_t *__restrict__ meta,
                                               const int8_t *__restrict__ A, int32_t *slist, int32_t *scount)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float4 p = pos[i];
    if (is_wall_w(p.w)) return;
    float4 v = vel[i];
    float3 u = fetch_velocity(u0, u1, u2, n, p.x, p.y, p.z);
    const double dt = dtv[0];
    // the clamp below maps a NaN position into the domain, so the check
    // looks at the inputs of the move
    if (!finite3(u.x, u.y, u.z) || !finite3(v.x, v.y, v.z) || !finite3(p.x, p.y, p.z)) atomicOr(nonfinite, 1);
    p.x = (float)((double)p.x + dt * (double)u.x);
    p.y = (float)((double)p.y + dt * (double)u.y);
    p.z = (float)((double)p.z + dt * (double)u.z);
    double lo = (double)rlo, hi = 1.0 - (double)rlo;
    p.x = (float)fmax(lo, fmin(hi, (double)p.x));
    p.y = (float)fmax(lo, fmin(hi, (double)p.y));
    p.z = (float)fmax(lo, fmin(hi, (double)p.z));
    int ci = cell_of_b(n, p.x), cj = cell_of_b(n, p.y), ck = cell_of_b(n, p.z);
    int klo = max(ck - 1, 0), khi = min(ck + 1, n - 1);
    // a window with no wall particle pushes nothing
    const bool walls = wnear[cidx(n, ci, cj, ck)] != 0;
    for (int si = ci - 1; walls && si <= ci + 1; ++si) {
        if (si < 0 || si > n - 1) continue;
        for (int sj = cj - 1; sj <= cj + 1; ++sj) {
            if (sj < 0 || sj > n - 1) continue;
            int b = wst[cidx(n, si, sj, klo)], e = wst[cidx(n, si, sj, khi) + 1];
            for (int a = b; a < e; ++a) {
                float4 q = wpos[a];
                float dist = length3(p.x, p.y, p.z, q.x, q.y, q.z);
                if ((double)dist < re) {
                    float4 nr = wnrm[a];
                    float nx = nr.x, ny = nr.y, nz = nr.z;
                    if (nx == 0.0f && ny == 0.0f && nz == 0.0f && dist != 0.0f) {
                        nx = (p.x - q.x) / dist;
                        ny = (p.y - q.y) / dist;
                        nz = (p.z - q.z) / dist;
                    }
                    double sh = re - (double)dist;
                    p.x = (float)((double)p.x + sh * (double)nx);
                    p.y = (float)((double)p.y + sh * (double)ny);
                    p.z = (float)((double)p.z + sh * (double)nz);
                    float dot = v.x * nx + v.y * ny + v.z * nz;
                    v.x -= dot * nx;
                    v.y -= dot * ny;
                    v.z -= dot * nz;
                }
            }
        }
    }
    pos[i] = p;
    vel[i] = v;
    // with slist: the stuck test (flip3D.cpp:933-955) on the moved particle
    if (slist && is_stuck(n, p, v.w, A)) {
        const int32_t at = atomicAdd(scount, 1);
        slist[at] = meta_id(meta[i]);
    }
}

// flip3D.cpp:414-421 deep-water cells: FLUID with every in-range neighbour
// FLUID (nothing to do when no particle is stuck: mcount[0] == 0)
__global__ void k_deep(long C, int n, const int8_t *__restrict__ A, int32_t *flag, const int32_t *__restrict__ mcount)
{
    if (mcount[0] == 0) return;
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const Split3 s3 = split3(c, n, n);
    const int i = s3.i, j = s3.j, k = s3.k;
    bool d = A[c] == FLUID;
    if (d && i > 0 && A[c - (long)n * n] != FLUID) d = false;
    if (d && i < n - 1 && A[c + (long)n * n] != FLUID) d = false;
    if (d && j > 0 && A[c - n] != FLUID) d = false;
    if (d && j < n - 1 && A[c + n] != FLUID) d = false;
    if (d && k > 0 && A[c - 1] != FLUID) d = false;
    if (d && k < n - 1 && A[c + 1] != FLUID) d = false;
    flag[c] = d ? 1 : 0;
}

// the first mcount[0] deep cells in scan order
__global__ void k_deep_compact(long C, const int32_t *__restrict__ flag, const int32_t *__restrict__ at,
                               const int32_t *__restrict__ mcount, int32_t *out)
{
    const int limit = mcount[0];
    if (limit == 0) return;
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    if (flag[c] && at[c] < limit) out[at[c]] = (int32_t)c;
}

// write the positions of the mcount[0] particles given by creation index
__global__ void k_set_pos_by_id(const int32_t *__restrict__ mcount, const int32_t *__restrict__ ids,
                                const float *__restrict__ xyz, const int32_t *__restrict__ inv, float4 *pos)
{
    const int m = mcount[0];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x) {
        int32_t s = inv[ids[t]];
        float4 p = pos[s];
        p.x = xyz[3 * t];
        p.y = xyz[3 * t + 1];
        p.z = xyz[3 * t + 2];
        pos[s] = p;
    }
}

// one glibc rand() draw from a TYPE_3 state held in s (31 words, indices f
// and r = f + 28 mod 31)
__device__ __forceinline__ int rand_step(uint32_t *s, int &f, int &r)
{
    const uint32_t v = s[f] + s[r];
    s[f] = v;
    f = f == 30 ? 0 : f + 1;
    r = r == 30 ? 0 : r + 1;
    return (int)(v >> 1);
}

// flip3D.cpp:405-437 the reposition plan, in one block: the stuck list in
// creation order (the reference lists particle indices in order), the deep
// cells repeated until there are as many as stuck particles (:409-429), the
// libstdc++ random_shuffle of them with the reference's rand() (:432), three
// rand() % 101 per particle for its position in its cell (:436-441), all
// drawn in the reference's order from the device's generator (canonical
// order at rest: f = 0, r = 28).  Then the dependency levels of the
// resample (k_resample_levels): level(t) = 1 + the highest level of an
// earlier listed particle within t's support (a few parallel rounds, then
// list order for long chains), with the level offsets and the list ordered
// by level.  ev[EV_MOVED] and ev[EV_NLEV] get the counts
// (both 0 when nothing is stuck or there is no deep water: the reference
// then moves nothing).
constexpr int RP_TB = 1024, RP_LDS = 8192, RP_ROUNDS = 8;
__global__ __launch_bounds__(RP_TB) void k_repo_plan(int32_t *ev, const int32_t *__restrict__ list,
                                                     const int32_t *__restrict__ nd_src,
                                                     const int32_t *__restrict__ deep, int32_t *ids,
                                                     int32_t *waters, int32_t *rng, float *xyz, int32_t *level,
                                                     int32_t *loff, int32_t *fill, int32_t *order, int n)
{
    __shared__ int32_t sw[RP_LDS];
    __shared__ uint32_t st[31];
    __shared__ int s_more, s_max;
    const int tid = threadIdx.x;
    const int m = ev[EV_STUCK];
    const int nd = m > 0 ? nd_src[0] : 0;
    if (m == 0 || nd == 0) {
        if (tid == 0) {
            ev[EV_MOVED] = 0;
            ev[EV_NLEV] = 0;
        }
        return;
    }
    const int take = min(nd, m);
    const bool in_lds = m <= RP_LDS;
    int32_t *w = in_lds ? sw : waters;
    for (int t = tid; t < m; t += RP_TB) {
        const int32_t id = list[t];
        int rank = 0;
        for (int u = 0; u < m; ++u) rank += list[u] < id;  // creation indices are unique
        ids[rank] = id;
        w[t] = deep[t % take];
    }
    __threadfence_block();
    __syncthreads();
    const float h = (float)(1.0 / n);
    if (tid == 0) {
        for (int j = 0; j < 31; ++j) st[j] = (uint32_t)rng[j];
        int f = 0, r = 28;
        // libstdc++ random_shuffle: swap(a[i], a[rand() % (i + 1)]) for i >= 1
        for (int i = 1; i < m; ++i) {
            const int j = rand_step(st, f, r) % (i + 1);
            if (i != j) {
                const int32_t x = w[i];
                w[i] = w[j];
                w[j] = x;
            }
        }
        for (int t = 0; t < m; ++t) {
            const int c = w[t];
            const int wi = c / (n * n), wj = (c / n) % n, wk = c % n;
            const int r0 = rand_step(st, f, r) % 101;
            xyz[3 * t + 0] = (float)((double)h * (wi + 0.25 + 0.5 * r0 / 100));
            const int r1 = rand_step(st, f, r) % 101;
            xyz[3 * t + 1] = (float)((double)h * (wj + 0.25 + 0.5 * r1 / 100));
            const int r2 = rand_step(st, f, r) % 101;
            xyz[3 * t + 2] = (float)((double)h * (wk + 0.25 + 0.5 * r2 / 100));
        }
        for (int j = 0; j < 31; ++j) rng[j] = (int32_t)st[(f + j) % 31];
        rng[31] = 0;
        rng[32] = 28;
    }
    for (int t = tid; t < m; t += RP_TB) level[t] = -1;
    __threadfence_block();
    __syncthreads();
    // levels: a few rounds in which every particle whose earlier neighbours
    // are all placed gets placed (the earliest unplaced one always is), then
    // whatever is left (long chains) one particle at a time in list order
    const float hh = h * h;
    for (int round = 0; round < RP_ROUNDS; ++round) {
        if (tid == 0) s_more = 0;
        __syncthreads();
        for (int t = tid; t < m; t += RP_TB) {
            if (level[t] >= 0) continue;
            const float px = xyz[3 * t], py = xyz[3 * t + 1], pz = xyz[3 * t + 2];
            int mx = -1;
            bool ready = true;
            for (int u = 0; u < t && ready; ++u) {
                const float dx = px - xyz[3 * u], dy = py - xyz[3 * u + 1], dz = pz - xyz[3 * u + 2];
                if (dx * dx + dy * dy + dz * dz < hh) {
                    const int lu = __atomic_load_n(&level[u], __ATOMIC_RELAXED);
                    if (lu < 0) ready = false;
                    else mx = max(mx, lu);
                }
            }
            if (ready) __atomic_store_n(&level[t], mx + 1, __ATOMIC_RELAXED);
            else s_more = 1;
        }
        __threadfence_block();
        __syncthreads();
        if (!s_more) break;
    }
    if (s_more) {
        for (int t = 0; t < m; ++t) {
            if (level[t] >= 0) continue;  // the same for every thread after the last barrier
            const float px = xyz[3 * t], py = xyz[3 * t + 1], pz = xyz[3 * t + 2];
            int mx = -1;
            for (int u = tid; u < t; u += RP_TB) {
                const float dx = px - xyz[3 * u], dy = py - xyz[3 * u + 1], dz = pz - xyz[3 * u + 2];
                if (dx * dx + dy * dy + dz * dz < hh) mx = max(mx, level[u]);
            }
            if (tid == 0) s_max = -1;
            __syncthreads();
            atomicMax(&s_max, mx);
            __syncthreads();
            if (tid == 0) level[t] = s_max + 1;
            __threadfence_block();
            __syncthreads();
        }
    }
    if (tid == 0) s_max = 0;
    __syncthreads();
    for (int t = tid; t < m; t += RP_TB) atomicMax(&s_max, level[t]);
    __syncthreads();
    const int nlev = s_max + 1;
    // level offsets: counts, then an exclusive scan over nlev + 1 entries
    // (chunk per thread, the chunk sums scanned in LDS)
    for (int l = tid; l <= nlev; l += RP_TB) loff[l] = 0;
    __threadfence_block();
    __syncthreads();
    for (int t = tid; t < m; t += RP_TB) atomicAdd(&loff[level[t]], 1);
    __threadfence_block();
    __syncthreads();
    const int chunk = (nlev + 1 + RP_TB - 1) / RP_TB, l0 = tid * chunk, l1 = min(l0 + chunk, nlev + 1);
    int csum = 0;
    for (int l = l0; l < l1; ++l) csum += loff[l];
    sw[tid] = csum;  // the shuffle's LDS copy is no longer needed
    __syncthreads();
    for (int d = 1; d < RP_TB; d <<= 1) {
        const int v = tid >= d ? sw[tid - d] : 0;
        __syncthreads();
        sw[tid] += v;
        __syncthreads();
    }
    int run = sw[tid] - csum;
    for (int l = l0; l < l1; ++l) {
        const int cnt = loff[l];
        loff[l] = run;
        if (l < nlev) fill[l] = run;
        run += cnt;
    }
    __threadfence_block();
    __syncthreads();
    // each particle at its level's next slot: the order inside a level is
    // free, its particles never read each other
    for (int t = tid; t < m; t += RP_TB) order[atomicAdd(&fill[level[t]], 1)] = t;
    if (tid == 0) {
        ev[EV_MOVED] = m;
        ev[EV_NLEV] = nlev;
    }
}

// zero slots [lo, hi) of the step's event counts (a kernel, so that steps
// can be graph-captured)
__global__ void k_clear_events(int32_t *ev, int lo, int hi)
{
    const int t = lo + (int)threadIdx.x;
    if (t < hi) ev[t] = 0;
}

// flip3D.cpp:445-450: resample each repositioned particle in list order with
// re = h = 1/N (corrector.cpp:15-47), u starting from 0, each particle's
// result visible to the later ones (the reference's loop order).  A listed
// particle reads another listed particle's velocity exactly when the two lie
// within the kernel support (length2 < h^2, a symmetric test), so the list
// splits into dependency levels: level(t) = 1 + the highest level of an
// earlier listed particle within t's support (the host forms them from the
// new positions).  Particles of one level never read each other, and every
// particle they read from the list is on a lower level (earlier in the list,
// new value) or a higher one (later in the list, old value), exactly as in
// the serial loop (k_repo_plan forms the levels).  One block walks the levels; within a level each wave
// resamples one particle: its lanes form the window's terms (w u, w) in
// parallel into LDS, in window order, and lane 0 adds them up in that order,
// so the sums are the reference's serial ones (a skipped particle stores +0
// terms, which leave the sums unchanged: they start at +0 and never become
// -0).
constexpr int RL_WAVES = 16, RL_CAP = 384;

__global__ __launch_bounds__(64 * RL_WAVES) void k_resample_levels(const int32_t *__restrict__ nlev_p,
                                                                   const int32_t *__restrict__ loff,
                                                                   const int32_t *__restrict__ order,
                                                                   const int32_t *__restrict__ ids,
                                                                   const int32_t *__restrict__ inv, int n,
                                                                   const float4 *__restrict__ pos, float4 *vel,
                                                                   const int32_t *__restrict__ cs, float hh)
{
    __shared__ float4 terms[RL_WAVES][RL_CAP];
    __shared__ unsigned long long present[RL_WAVES][RL_CAP / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nlev = nlev_p[0];
    for (int l = 0; l < nlev; ++l) {
        for (int x = loff[l] + wave; x < loff[l + 1]; x += RL_WAVES) {
            const int32_t slot = inv[ids[order[x]]];
            const float4 p = pos[slot];
            const int ci = cell_of(n, p.x), cj = cell_of(n, p.y), ck = cell_of(n, p.z);
            const int klo = max(ck - 1, 0), khi = min(ck + 1, n - 1);
            int rb[9], rl[9], total = 0;
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                const int si = ci - 1 + r / 3, sj = cj - 1 + r % 3;
                const bool ok = si >= 0 && si <= n - 1 && sj >= 0 && sj <= n - 1;
                const long c0 = ok ? cidx(n, si, sj, klo) : 0;
                const int b = cs[c0], e = cs[c0 + (ok ? khi - klo : 0) + 1];
                rb[r] = b;
                rl[r] = ok ? e - b : 0;
                total += rl[r];
            }
            float ax = 0.0f, ay = 0.0f, az = 0.0f, wsum = 0.0f;
            for (int base = 0; base < total; base += RL_CAP) {
                const int cnt = min(RL_CAP, total - base);
                for (int k0 = 0; k0 < cnt; k0 += 64) {
                    const int k = k0 + lane;
                    bool has = false;
                    if (k < cnt) {
                    // window position base + k -> storage slot (rows in window order)
                    int rem = base + k, a = -1;
#pragma unroll
                    for (int r = 0; r < 9; ++r) {
                        if (a < 0) {
                            if (rem < rl[r]) a = rb[r] + rem;
                            else rem -= rl[r];
                        }
                    }
                    float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    const float4 q = pos[a];
                    const float d2 = length2(p.x, p.y, p.z, q.x, q.y, q.z);
                    if (!is_wall_w(q.w) && d2 < hh) {
                        const float w = q.w * sharp_kernel_hh(d2, hh);
                        const float4 u = vel[a];
                        t = make_float4(w * u.x, w * u.y, w * u.z, w);
                        has = true;
                    }
                    terms[wave][k] = t;
                    }
                    const unsigned long long m = __ballot(has);
                    if (lane == 0) present[wave][k0 / 64] = m;
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) {
                    // the terms in window order, zero terms (outside the
                    // support, or WALL) skipped
                    for (int k0 = 0; k0 < cnt; k0 += 64) {
                        for (unsigned long long left = present[wave][k0 / 64]; left; left &= left - 1) {
                            const float4 t = terms[wave][k0 + __builtin_ctzll(left)];
                            ax += t.x;
                            ay += t.y;
                            az += t.z;
                            wsum += t.w;
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
            if (lane == 0) {
                float4 v = vel[slot];
                if (wsum != 0.0f) {
                    v.x = ax / wsum;
                    v.y = ay / wsum;
                    v.z = az / wsum;
                } else {
                    v.x = v.y = v.z = 0.0f;
                }
                vel[slot] = v;
            }
        }
        __threadfence_block();
        __syncthreads();
    }
}

// ------------------------------------------------------------- corrector
// Conservative culling of a 3x3x3 neighbour window (the corrector's radius is
// half a cell, so most of the window is out of reach).  For a particle at
// g*p = (gx, gy, gz) with window centre cell (ci, cj, ck), the squared
// distance in cell units from g*p to the lower and upper neighbour cells
// along x is at least max(gx - ci - 1e-3, 0)^2 and max(ci + 1 - gx - 1e-3,
// 0)^2; the margin covers the float rounding of every particle's cell
// assignment.  A centre past the last cell (the corrector's clamp to g
// allows ci == n) culls nothing below it, since the last cell reaches to
// infinity.  A row or cell whose bound reaches r2 g^2 1.001 (r2: squared
// radius in world units) holds only particles beyond the radius, whose
// contributions are exactly zero and skipped by the callers anyway, so the
// particles left are visited in the same order with the same sums.
struct WinCull {
    float xl, xh, yl, yh, zl, zh, thr;
    __device__ __forceinline__ static float sq0(float v) { return v > 0.0f ? v * v : 0.0f; }
    __device__ __forceinline__ WinCull(int n, float gx, float gy, float gz, int ci, int cj, int ck, float r2)
    {
        const float m = 1e-3f, g = (float)n;
        xl = ci < n ? sq0(gx - (float)ci - m) : 0.0f;
        yl = cj < n ? sq0(gy - (float)cj - m) : 0.0f;
        zl = ck < n ? sq0(gz - (float)ck - m) : 0.0f;
        xh = sq0((float)(ci + 1) - gx - m);
        yh = sq0((float)(cj + 1) - gy - m);
        zh = sq0((float)(ck + 1) - gz - m);
        thr = r2 * g * g * 1.001f;
    }
    // bound of row (si, sj); false when the whole row is out of reach
    __device__ __forceinline__ bool row(int ci, int cj, int si, int sj, float &dij) const
    {
        float dx = si < ci ? xl : (si > ci ? xh : 0.0f);
        float dy = sj < cj ? yl : (sj > cj ? yh : 0.0f);
        dij = dx + dy;
        return dij < thr;
    }
    // cells [klo, khi] of a row within reach (klo > khi: none)
    __device__ __forceinline__ void cells(int n, int ck, float dij, int &klo, int &khi) const
    {
        klo = (ck >= 1 && dij + zl < thr) ? ck - 1 : ck;
        khi = (ck + 1 <= n - 1 && dij + zh < thr) ? ck + 1 : min(ck, n - 1);
    }
};

// Bounds of the nine window rows (si, sj) = (ci - 1 + r / 3, cj - 1 + r % 3),
// in the reference's visiting order, loaded together: rows outside the grid
// or out of reach (wc.row), and rows whose cells in reach are none, are
// empty.
__device__ __forceinline__ void window_rows(int n, const WinCull &wc, int ci, int cj, int ck,
                                            const int32_t *__restrict__ cs, int rlo[9], int rhi[9])
{
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int si = ci - 1 + r / 3, sj = cj - 1 + r % 3;
        float dij;
        int klo, khi;
        const bool reach = wc.row(ci, cj, si, sj, dij);
        wc.cells(n, ck, dij, klo, khi);
        const bool ok = si >= 0 && si <= n - 1 && sj >= 0 && sj <= n - 1 && reach && klo <= khi;
        const long c0 = ok ? cidx(n, si, sj, klo) : 0;
        const int b = cs[c0], e = cs[c0 + (ok ? khi - klo : 0) + 1];
        rlo[r] = ok ? b : 0;
        rhi[r] = ok ? e : 0;
    }
}

// Block-local regrouping of a block's B particles by the octant of the cell
// that centres their window.  With the corrector's half-cell radius the
// octant (which half of the cell along each axis) fixes the neighbour rows
// and cells within reach, so a wave whose lanes share an octant walks only
// those.  Group 8 collects idle threads.  Stable within a group (index
// order); built from wave ballots and one wave scan.  Returns the
// block-relative particle index this thread handles.  The mapping changes
// only which thread computes a particle, never what it computes.
template <int B>
__device__ __forceinline__ int octant_regroup(int grp, int32_t *s_off, int32_t *s_idx)
{
    constexpr int W = B / 64, NG = 9, E = (NG * W + 63) / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int rank = 0;
    for (int o = 0; o < NG; ++o) {
        const unsigned long long m = __ballot(grp == o);
        if (grp == o) rank = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) s_off[o * W + wave] = __popcll(m);
    }
    __syncthreads();
    if (wave == 0) {
        // exclusive scan of the (group, wave) counts in that order
        int v[E], sum = 0;
        for (int e = 0; e < E; ++e) {
            const int j = lane * E + e;
            v[e] = j < NG * W ? s_off[j] : 0;
            sum += v[e];
        }
        int inc = sum;
        for (int d = 1; d < 64; d <<= 1) {
            const int u = __shfl_up(inc, d, 64);
            if (lane >= d) inc += u;
        }
        int acc = inc - sum;
        for (int e = 0; e < E; ++e) {
            const int j = lane * E + e;
            if (j < NG * W) s_off[j] = acc;
            acc += v[e];
        }
    }
    __syncthreads();
    s_idx[s_off[grp * W + wave] + rank] = tid;
    __syncthreads();
    return s_idx[tid];
}

// octant of g*p relative to its cell c (per axis: upper half or not)
__device__ __forceinline__ int octant_of(float gx, float gy, float gz, int ci, int cj, int ck)
{
    return ((gx - (float)ci >= 0.5f) << 2) | ((gy - (float)cj >= 0.5f) << 1) | (gz - (float)ck >= 0.5f);
}

// window centre of the spring pass: max(0, min(g, g*p)) per axis with the
// reference's macros (utility.h:16-17), truncated
__device__ __forceinline__ void spring_centre(int n, const float4 &p, float &cx, float &cy, float &cz, int &ci,
                                              int &cj, int &ck)
{
    const float g = (float)n;
    cx = g * p.x;
    cy = g * p.y;
    cz = g * p.z;
    float x = g > cx ? cx : g;
    x = 0.0f > x ? 0.0f : x;
    float y = g > cy ? cy : g;
    y = 0.0f > y ? 0.0f : y;
    float z = g > cz ? cz : g;
    z = 0.0f > z ? 0.0f : z;
    ci = (int)x;
    cj = (int)y;
    ck = (int)z;
}

// corrector.cpp:55-88 spring for one particle.  RAND == false: at the first
// random-jitter event stops forming the spring, counts the window's events
// and returns minus their number; RAND == true: consumes the host-drawn
// rand()%101 values in window order.
template <bool RAND>
__device__ __forceinline__ int spring_one(int64_t i, int n, const float4 *__restrict__ pos,
                                          const int32_t *__restrict__ meta, const float4 *__restrict__ nrm_id,
                                          const int32_t *__restrict__ cs, float dt, float re, const float *rv,
                                          float4 *tmp0)
{
    float4 p = pos[i];
    float cx, cy, cz;
    int ci, cj, ck;
    spring_centre(n, p, cx, cy, cz, ci, cj, ck);
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    int events = 0;
    const double rre = 0.1 * (double)re;
    const double cj_ = 0.01 * (double)re / (double)dt;
    const double cw = 0.05 * (double)re / (double)dt;
    // A neighbour at distance >= re has smooth weight exactly 0 and lies on the
    // dist > 0.1 re branch, so it adds a signed zero: skipped.  The float
    // squared distance is within a few ulp of dist*dist, so the 1e-4 margin
    // keeps the test conservative.
    const float far2 = re * re * 1.0001f;
    const WinCull wc(n, cx, cy, cz, ci, cj, ck, far2);
    bool bail = false;
    for (int si = ci - 1; si <= ci + 1; ++si) {
        if (si < 0 || si > n - 1) continue;
        for (int sj = cj - 1; sj <= cj + 1; ++sj) {
            if (sj < 0 || sj > n - 1) continue;
            float dij;
            int klo, khi;
            if (!wc.row(ci, cj, si, sj, dij)) continue;
            wc.cells(n, ck, dij, klo, khi);
            if (klo > khi) continue;
            int b = cs[cidx(n, si, sj, klo)], e = cs[cidx(n, si, sj, khi) + 1];
            for_range4(pos, b, e, [&](const float4 &q, int a) {
                if (a == i) return;
                if (length2(p.x, p.y, p.z, q.x, q.y, q.z) > far2) return;
                float dist = length3(p.x, p.y, p.z, q.x, q.y, q.z);
                if (!RAND && bail) {
                    events += !((double)dist > rre) && !is_wall_w(q.w);
                    return;
                }
                float w = (float)(50.0 * (double)mass_of(q.w) * (double)smooth_kernel(dist * dist, re));
                if ((double)dist > rre) {
                    sx += w * (p.x - q.x) / dist * re;
                    sy += w * (p.y - q.y) / dist * re;
                    sz += w * (p.z - q.z) / dist * re;
                } else if (!is_wall_w(q.w)) {
                    if (!RAND) {
                        bail = true;
                        events = 1;
                        return;
                    }
                    sx = (float)((double)sx + cj_ * (double)rv[3 * events + 0] / 100.0);
                    sy = (float)((double)sy + cj_ * (double)rv[3 * events + 1] / 100.0);
                    sz = (float)((double)sz + cj_ * (double)rv[3 * events + 2] / 100.0);
                    events++;
                } else {
                    float4 nr = nrm_id[meta_id(meta[a])];
                    sx = (float)((double)sx + cw * (double)nr.x);
                    sy = (float)((double)sy + cw * (double)nr.y);
                    sz = (float)((double)sz + cw * (double)nr.z);
                }
            });
        }
    }
    if (!RAND && bail) return -events;
    tmp0[i] = make_float4(p.x + dt * sx, p.y + dt * sy, p.z + dt * sz, 0.0f);
    return events;
}

// The spring pass in two phases per particle, so that a wave's lanes run
// the expensive near-neighbour arithmetic together: phase 1 scans the window
// with the cheap squared-distance test and lists the survivors (window
// order) in LDS; phase 2 applies spring_one's arithmetic to the list in that
// order, so every sum is formed in the same order as in spring_one.  More
// than SPRING_NEAR survivors (never seen in practice) falls back to
// spring_one itself.
constexpr int SPRING_TB = 256;
constexpr int SPRING_NEAR = 16;

__global__ __launch_bounds__(SPRING_TB) void k_spring(int64_t np, int n, const float4 *__restrict__ pos,
                                                      const int32_t *__restrict__ meta,
                                                      const float4 *__restrict__ nrm_id,
                                                      const int32_t *__restrict__ cs,
                                                      const double *__restrict__ dtv, float re, float4 *tmp0,
                                                      int32_t *list, int32_t *count)
{
    const float dt = (float)dtv[0];
    __shared__ int32_t near[SPRING_NEAR + 1][SPRING_TB];  // slot SPRING_NEAR: scratch
    __shared__ int32_t s_off[9 * SPRING_TB / 64], s_idx[SPRING_TB];
    const int64_t base = (int64_t)blockIdx.x * SPRING_TB;
    float cx, cy, cz;
    int ci, cj, ck;
    {
        const int64_t i = base + threadIdx.x;
        int grp = 8;
        if (i < np) {
            const float4 p = pos[i];
            if (!is_wall_w(p.w)) {
                spring_centre(n, p, cx, cy, cz, ci, cj, ck);
                grp = octant_of(cx, cy, cz, ci, cj, ck);
            }
        }
        const int t = octant_regroup<SPRING_TB>(grp, s_off, s_idx);
        // the thread that handles particle base + t
        if (base + t >= np) return;
    }
    const int tid = threadIdx.x;
    const int64_t i = base + s_idx[tid];
    const float4 p = pos[i];
    if (is_wall_w(p.w)) return;
    spring_centre(n, p, cx, cy, cz, ci, cj, ck);
    const float far2 = re * re * 1.0001f;  // as in spring_one
    const WinCull wc(n, cx, cy, cz, ci, cj, ck, far2);
    int cnt = 0;
    int rlo[9], rhi[9];
    window_rows(n, wc, ci, cj, ck, cs, rlo, rhi);
    // Branch-free listing: every record is written to the next free slot
    // (the scratch slot once the list is full) and counted only when it is
    // a neighbour, so a later neighbour overwrites a non-neighbour's entry.
    const int self = (int)i;
#pragma unroll
    for (int r = 0; r < 9; ++r)
        for_range8(pos, rlo[r], rhi[r], [&](const float4 &q, int a) {
            const bool nb = (a != self) & !(length2(p.x, p.y, p.z, q.x, q.y, q.z) > far2);
            near[min(cnt, SPRING_NEAR)][tid] = a;
            cnt += nb;
        });
    int r = 0;
    if (cnt > SPRING_NEAR) {
        r = spring_one<false>(i, n, pos, meta, nrm_id, cs, dt, re, nullptr, tmp0);
    } else {
        const double rre = 0.1 * (double)re;
        const double cw = 0.05 * (double)re / (double)dt;
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        for (int k = 0; k < cnt; ++k) {
            const int a = near[k][tid];
            const float4 q = pos[a];
            float dist = length3(p.x, p.y, p.z, q.x, q.y, q.z);
            if (r < 0) {  // after the first event: only count the others
                r -= !((double)dist > rre) && !is_wall_w(q.w);
                continue;
            }
            float w = (float)(50.0 * (double)mass_of(q.w) * (double)smooth_kernel(dist * dist, re));
            if ((double)dist > rre) {
                sx += w * (p.x - q.x) / dist * re;
                sy += w * (p.y - q.y) / dist * re;
                sz += w * (p.z - q.z) / dist * re;
            } else if (!is_wall_w(q.w)) {
                r = -1;  // a random-jitter event: the fix-up pass replays this particle
            } else {
                float4 nr = nrm_id[meta_id(meta[a])];
                sx = (float)((double)sx + cw * (double)nr.x);
                sy = (float)((double)sy + cw * (double)nr.y);
                sz = (float)((double)sz + cw * (double)nr.z);
            }
        }
        if (r == 0) tmp0[i] = make_float4(p.x + dt * sx, p.y + dt * sy, p.z + dt * sz, 0.0f);
    }
    if (r < 0) {  // listed with its number of jitter events, -r
        int32_t at = atomicAdd(count, 1);
        list[3 * at] = meta_id(meta[i]);
        list[3 * at + 1] = (int32_t)i;
        list[3 * at + 2] = -r;
    }
}

// ------------------------------------------- corrector jitter events
// The spring pass lists the particles with a jitter event as (creation index,
// slot, event count) triples, count in mcount[0]; the event count is every
// FLUID neighbour within 0.1 re (corrector.cpp:73-77).  The reference draws 3
// rand()%101 per event in particle order at one thread (corrector.cpp:73-77);
// here the listed particles are ranked by creation index (k_jitter_order),
// the draws made in that order from the device copy of the generator
// (k_jitter_draw), and the springs finished (k_spring_fix_wave).  No host
// round trip; every launch returns at once when nothing is listed.
constexpr int JB = RNG_CHUNK;  // block of the single-block jitter kernels (one draw per lane)

constexpr int OTB = 256;  // block of the ranking kernel (one listed particle per wave)
__global__ __launch_bounds__(OTB) void k_jitter_order(const int32_t *__restrict__ mcount,
                                                      const int32_t *__restrict__ trip, int32_t *slots,
                                                      int32_t *events)
{
    const int m = mcount[0];
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (OTB / 64);
    for (int t = blockIdx.x * (OTB / 64) + (int)(threadIdx.x >> 6); t < m; t += waves) {  // uniform over the wave
        const int32_t id = trip[3 * t];
        int below = 0;
        for (int u = lane; u < m; u += 64) below += trip[3 * u] < id;  // creation indices are unique
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) below += __shfl_xor(below, o, 64);
        if (lane == 0) {
            slots[below] = trip[3 * t + 1];
            events[below] = trip[3 * t + 2];
        }
    }
}

// offsets (3 draws per event) by a block scan, then the draws.  The glibc
// TYPE_3 generator (state s[31], indices f and r = f + 28 mod 31; a draw sets
// s[f] += s[r], advances both, and rand() is the new word >> 1) only adds
// words mod 2^32, so with the state held in canonical order z_j = s[(f + j)
// mod 31] the k-th new word is sum_j table[k][j] z_j for a table fixed by the
// recurrence alone (rng_table, built on the host for k < JB).  Each lane forms
// one draw of a chunk of JB; the chunk's last 31 words are the next state.
// mcount[1] is set when the draws would overflow rv (cap values).
__global__ __launch_bounds__(JB) void k_jitter_draw(int32_t *mcount, const int32_t *__restrict__ events,
                                                    int32_t *off, float *rv, long cap, int32_t *rng,
                                                    const uint32_t *__restrict__ table)
{
    __shared__ int32_t wsum[JB / 64];
    __shared__ uint32_t z[31], zn[31], y[JB];
    __shared__ int32_t s_total;
    const int m = mcount[0];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int carry = 0;
    for (int b = 0; b < m; b += JB) {
        const int t = b + (int)threadIdx.x;
        const int v = t < m ? events[t] : 0;
        int incl = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int x = __shfl_up(incl, o, 64);
            if (lane >= o) incl += x;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        int before = 0, tot = 0;
        for (int u = 0; u < JB / 64; ++u) {
            if (u < w) before += wsum[u];
            tot += wsum[u];
        }
        if (t < m) off[t] = 3 * (carry + before + incl - v);
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) s_total = carry;
    if (threadIdx.x < 31) z[threadIdx.x] = (uint32_t)rng[threadIdx.x];  // canonical order (f = 0)
    __syncthreads();
    const long draws = 3L * s_total;
    if (threadIdx.x == 0) mcount[3] = s_total;
    if (draws > cap) {
        if (threadIdx.x == 0) mcount[1] = 1;
        return;
    }
    for (long e0 = 0; e0 < draws; e0 += JB) {
        const int cnt = (int)min((long)JB, draws - e0);
        const int t = threadIdx.x;
        if (t < cnt) {
            const uint32_t *row = table + (size_t)t * 31;
            uint32_t acc = 0u;
#pragma unroll
            for (int j = 0; j < 31; ++j) acc += row[j] * z[j];
            y[t] = acc;
            rv[e0 + t] = (float)((int)(acc >> 1) % 101);
        }
        __syncthreads();
        if (t < 31) zn[t] = cnt >= 31 ? y[cnt - 31 + t] : (t < 31 - cnt ? z[t + cnt] : y[t - (31 - cnt)]);
        __syncthreads();
        if (t < 31) z[t] = zn[t];
        __syncthreads();
    }
    if (threadIdx.x < 31) rng[threadIdx.x] = (int32_t)z[threadIdx.x];
    if (threadIdx.x == 0) {
        rng[31] = 0;
        rng[32] = 28;
    }
}

// The window of listed particle t as one wave sees it: its rows (culled as
// spring_one culls them) and the length of their concatenation.
struct SpringWindow {
    float4 p;
    int rlo[9], rhi[9], len;
    __device__ __forceinline__ SpringWindow(int64_t i, int n, const float4 *__restrict__ pos,
                                            const int32_t *__restrict__ cs, float far2)
    {
        p = pos[i];
        float cx, cy, cz;
        int ci, cj, ck;
        spring_centre(n, p, cx, cy, cz, ci, cj, ck);
        const WinCull wc(n, cx, cy, cz, ci, cj, ck, far2);
        window_rows(n, wc, ci, cj, ck, cs, rlo, rhi);
        len = 0;
#pragma unroll
        for (int r = 0; r < 9; ++r) len += rhi[r] - rlo[r];
    }
    // the particle at window position e (< len)
    __device__ __forceinline__ int at(int e) const
    {
        int a = -1;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const int L = rhi[r] - rlo[r];
            if (a < 0 && e < L) a = rlo[r] + e;
            if (a < 0) e -= L;
        }
        return a;
    }
};

constexpr int FIX_TB = 256;  // block of the fix-up kernel (one listed particle per wave)

// Listed particles' springs with the jitter draws, one wave per particle:
// lanes form the terms of 64 window particles at a time (window order; the
// event index of a jitter term is the count of events before it) and lane 0
// adds them in that order, float adds for the distance terms and double
// rounding for the jitter and wall terms as spring_one does.
__global__ __launch_bounds__(FIX_TB) void k_spring_fix_wave(const int32_t *__restrict__ mcount,
                                                            const int32_t *__restrict__ list,
                                                            const int32_t *__restrict__ off, int n,
                                                            const float4 *__restrict__ pos,
                                                            const int32_t *__restrict__ meta,
                                                            const float4 *__restrict__ nrm_id,
                                                            const int32_t *__restrict__ cs,
                                                            const double *__restrict__ dtv, float re,
                                                            const float *__restrict__ rv, float4 *tmp0)
{
    __shared__ double tv[FIX_TB / 64][64][3];
    __shared__ int8_t tk[FIX_TB / 64][64];
    const int m = mcount[1] ? 0 : mcount[0];  // nothing when the draws overflowed
    const float dt = (float)dtv[0];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nw = gridDim.x * (FIX_TB / 64);
    const double rre = 0.1 * (double)re;
    const double cj_ = 0.01 * (double)re / (double)dt;
    const double cw = 0.05 * (double)re / (double)dt;
    const float far2 = re * re * 1.0001f;
    for (int t = blockIdx.x * (FIX_TB / 64) + wv; t < m; t += nw) {
        const int64_t i = list[t];
        const float *rvt = rv + off[t];
        const SpringWindow win(i, n, pos, cs, far2);
        const float4 p = win.p;
        const int len = win.len;
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        int events = 0;
        for (int e0 = 0; e0 < len; e0 += 64) {
            const int a = e0 + lane < len ? win.at(e0 + lane) : -1;
            int kind = 0;  // 0 none, 1 distance term (float add), 2 jitter, 3 wall (double-rounded adds)
            double v0 = 0.0, v1 = 0.0, v2 = 0.0;
            float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (a >= 0 && a != i) {
                q = pos[a];
                if (!(length2(p.x, p.y, p.z, q.x, q.y, q.z) > far2)) {
                    const float dist = length3(p.x, p.y, p.z, q.x, q.y, q.z);
                    if ((double)dist > rre) {
                        const float w = (float)(50.0 * (double)mass_of(q.w) * (double)smooth_kernel(dist * dist, re));
                        kind = 1;
                        v0 = (double)(w * (p.x - q.x) / dist * re);
                        v1 = (double)(w * (p.y - q.y) / dist * re);
                        v2 = (double)(w * (p.z - q.z) / dist * re);
                    } else {
                        kind = is_wall_w(q.w) ? 3 : 2;
                    }
                }
            }
            const unsigned long long jm = __ballot(kind == 2);
            if (kind == 2) {
                const int ev = events + __popcll(jm & ((1ull << lane) - 1ull));
                v0 = cj_ * (double)rvt[3 * ev + 0] / 100.0;
                v1 = cj_ * (double)rvt[3 * ev + 1] / 100.0;
                v2 = cj_ * (double)rvt[3 * ev + 2] / 100.0;
            } else if (kind == 3) {
                const float4 nr = nrm_id[meta_id(meta[a])];
                v0 = cw * (double)nr.x;
                v1 = cw * (double)nr.y;
                v2 = cw * (double)nr.z;
            }
            events += __popcll(jm);
            // only the lanes with a term are visited, lowest lane (window
            // position) first
            const unsigned long long nz = __ballot(kind != 0);
            tk[wv][lane] = (int8_t)kind;
            tv[wv][lane][0] = v0;
            tv[wv][lane][1] = v1;
            tv[wv][lane][2] = v2;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) {
                for (unsigned long long left = nz; left; left &= left - 1) {
                    const int u = __builtin_ctzll(left);
                    const int kd = tk[wv][u];
                    if (kd == 1) {
                        sx += (float)tv[wv][u][0];
                        sy += (float)tv[wv][u][1];
                        sz += (float)tv[wv][u][2];
                    } else if (kd >= 2) {
                        sx = (float)((double)sx + tv[wv][u][0]);
                        sy = (float)((double)sy + tv[wv][u][1]);
                        sz = (float)((double)sz + tv[wv][u][2]);
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        if (lane == 0) tmp0[i] = make_float4(p.x + dt * sx, p.y + dt * sy, p.z + dt * sz, 0.0f);
    }
}


// corrector.cpp:15-47 resample at the pseudo-moved position with the old
// positions and velocities (re = DENSITY/N); keeps u when no weight.
constexpr int RESAMPLE_TB = 256;
constexpr int RESAMPLE_NEAR = 16;

// the weighted window sum of one particle in window order (the reference's loop)
__device__ __forceinline__ void resample_window(int n, const float4 &p, const WinCull &wc, int ci, int cj, int ck,
                                                const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                                                const int32_t *__restrict__ cs, float hh, float &ax, float &ay,
                                                float &az, float &wsum)
{
    for (int si = ci - 1; si <= ci + 1; ++si) {
        if (si < 0 || si > n - 1) continue;
        for (int sj = cj - 1; sj <= cj + 1; ++sj) {
            if (sj < 0 || sj > n - 1) continue;
            float dij;
            int klo, khi;
            if (!wc.row(ci, cj, si, sj, dij)) continue;
            wc.cells(n, ck, dij, klo, khi);
            if (klo > khi) continue;
            int b = cs[cidx(n, si, sj, klo)], e = cs[cidx(n, si, sj, khi) + 1];
            for_range4(pos, b, e, [&](const float4 &q, int a) {
                if (is_wall_w(q.w)) return;
                float d2 = length2(p.x, p.y, p.z, q.x, q.y, q.z);
                if (!(d2 < hh)) return;  // weight exactly 0
                float w = q.w * sharp_kernel_hh(d2, hh);
                float4 u = vel[a];
                ax += w * u.x;
                ay += w * u.y;
                az += w * u.z;
                wsum += w;
            });
        }
    }
}

// Two phases per particle, so that a wave's lanes run the division-heavy
// weight arithmetic together: phase 1 scans the window with the squared
// distance test alone and lists the non-wall particles inside the support
// (window order) in LDS; phase 2 forms the weights and sums over the list in
// that order, exactly as resample_window would.  More than RESAMPLE_NEAR
// listed particles falls back to resample_window itself.  Phase 1 reads the
// sort's 12-byte records (WALL particles sit at +inf and fail the test).
__global__ __launch_bounds__(RESAMPLE_TB) void k_resample(int64_t np, int n, const float4 *__restrict__ pos,
                                                          const float4 *__restrict__ vel,
                                                          const int32_t *__restrict__ cs,
                                                          const float4 *__restrict__ tmp0, float4 *tmp1, float hh,
                                                          const P3 *__restrict__ xyz3)
{
    __shared__ int32_t near[RESAMPLE_NEAR + 1][RESAMPLE_TB];  // slot RESAMPLE_NEAR: scratch
    __shared__ int32_t s_off[9 * RESAMPLE_TB / 64], s_idx[RESAMPLE_TB];
    const int64_t base = (int64_t)blockIdx.x * RESAMPLE_TB;
    const float g = (float)n;
    {
        const int64_t i = base + threadIdx.x;
        int grp = 8;
        if (i < np && !is_wall_w(pos[i].w)) {
            const float4 p = tmp0[i];
            grp = octant_of(g * p.x, g * p.y, g * p.z, cell_of(n, p.x), cell_of(n, p.y), cell_of(n, p.z));
        }
        const int t = octant_regroup<RESAMPLE_TB>(grp, s_off, s_idx);
        if (base + t >= np) return;
    }
    const int tid = threadIdx.x;
    const int64_t i = base + s_idx[tid];
    if (is_wall_w(pos[i].w)) return;
    const float4 p = tmp0[i];
    const int ci = cell_of(n, p.x), cj = cell_of(n, p.y), ck = cell_of(n, p.z);
    const WinCull wc(n, g * p.x, g * p.y, g * p.z, ci, cj, ck, hh);
    int cnt = 0;
    int rlo[9], rhi[9];
    window_rows(n, wc, ci, cj, ck, cs, rlo, rhi);
    // branch-free listing, as in k_spring
#pragma unroll
    for (int r = 0; r < 9; ++r)
        for_range8(xyz3, rlo[r], rhi[r], [&](const P3 &q, int a) {
            const bool nb = length2(p.x, p.y, p.z, q.x, q.y, q.z) < hh;
            near[min(cnt, RESAMPLE_NEAR)][tid] = a;
            cnt += nb;
        });
    float ax = 0.0f, ay = 0.0f, az = 0.0f, wsum = 0.0f;
    if (cnt > RESAMPLE_NEAR) {
        resample_window(n, p, wc, ci, cj, ck, pos, vel, cs, hh, ax, ay, az, wsum);
    } else {
        for (int k = 0; k < cnt; ++k) {
            const int a = near[k][tid];
            const float4 q = pos[a];
            const float w = q.w * sharp_kernel_in(length2(p.x, p.y, p.z, q.x, q.y, q.z), hh);
            const float4 u = vel[a];
            ax += w * u.x;
            ay += w * u.y;
            az += w * u.z;
            wsum += w;
        }
    }
    if (wsum != 0.0f) {
        tmp1[i] = make_float4(ax / wsum, ay / wsum, az / wsum, 0.0f);
    } else {
        const float4 v0 = vel[i];
        tmp1[i] = make_float4(v0.x, v0.y, v0.z, 0.0f);
    }
}

// corrector.cpp:103-114 update (tmp1 null: every velocity is kept)
__global__ void k_correct_update(int64_t np, float4 *pos, float4 *vel, const float4 *__restrict__ tmp0,
                                 const float4 *__restrict__ tmp1, int32_t *nonfinite)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float4 p = pos[i];
    if (is_wall_w(p.w)) return;
    float4 a = tmp0[i];
    pos[i] = make_float4(a.x, a.y, a.z, p.w);
    bool ok = finite3(a.x, a.y, a.z);
    if (tmp1) {
        float4 b = tmp1[i];
        float4 v = vel[i];
        vel[i] = make_float4(b.x, b.y, b.z, v.w);
        ok = ok && finite3(b.x, b.y, b.z);
    }
    if (!ok) atomicOr(nonfinite, 1);
}

// dt policy 1: dt = min(cap, cfl / (N max|u|)) (no FLUID motion: the cap);
// nb == 0 writes the cap (policy 0).  fmax drops NaN speeds, as the max does.
__global__ __launch_bounds__(TB) void k_choose_dt(int nb, const double *__restrict__ part, double cap, double cfl,
                                                  int n, double *dtv)
{
    __shared__ double lds[TB / 64];
    double m = 0.0;
    for (int b = threadIdx.x; b < nb; b += TB) m = fmax(m, part[b]);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < TB / 64; ++w) r = fmax(r, lds[w]);
        const double umax = sqrt(r);
        double dt = cap;
        if (umax > 0.0) {
            const double c = cfl / (n * umax);
            if (c < dt) dt = c;
        }
        dtv[0] = dt;
    }
}

__global__ void k_selftest(int64_t m, const float *x, const float *y, float *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    float a = x[i], b = y[i];
    out[4 * i + 0] = a / b;
    out[4 * i + 1] = sqrtf(a);
    out[4 * i + 2] = hypotf_glibc(a, b);
    out[4 * i + 3] = (float)((double)a / (double)b);
    out[4 * m + i] = a < b ? sharp_kernel_in(a, b) : 0.0f;  // a = r2, b = h*h
}

// flip3D.cpp:375-402 pourWater: the disc's particles appended in disc order
// after the current ones, storage slot = creation index = np + t (the
// reference appends to its particle vector); velocity (0, uy, 0), dens =
// max_dens, m = 1, zero normal.
__global__ void k_emit(int m, int64_t np, const float4 *__restrict__ disc, float uy, float dens, float4 *pos,
                       float4 *vel, int32_t *meta, float4 *nrm_id)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    int64_t s = np + t;
    pos[s] = disc[t];
    vel[s] = make_float4(0.0f, uy, 0.0f, dens);
    meta[s] = (int32_t)((uint32_t)s << 2 | (uint32_t)FLIP_FLUID);
    nrm_id[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace

// ------------------------------------------------------------- launchers
void k_launch_emit(FlipCtx *c)
{
    if (c->emit_n <= 0) return;
    const float uy = (float)(-0.5 * c->prm.density / c->n / c->prm.dt);
    hipLaunchKernelGGL(k_emit, nblk(c->emit_n, TB), TB, 0, c->stream, c->emit_n, c->np, c->emit_pos, uy, c->max_dens,
                       c->P4(), c->V4(), c->M(), c->nrm_id);
    c->np += c->emit_n;
    c->sorted = false;
    c->bucketed = false;
}
bool k_launch_density(FlipCtx *c, bool with_force, bool p2g_records)
{
    if (c->np == 0) return false;
    const float h = (float)(4.0 * c->prm.density / c->n), hh = h * h;
    int e = 0;
    const bool pow2 = std::isnormal(hh) && std::frexp(hh, &e) == 0.5f;
    const float inv_hh = pow2 ? std::ldexp(1.0f, 1 - e) : 0.0f;  // hh = 2^(e-1)
    // the sort's packed records are current while the storage is sorted
    const bool packed = pow2 && c->unit_mass && c->sorted;
    const char *name = with_force ? "density_force" : "density";
    if (packed) {
        // the P2G records are formed here only from the velocities P2G will see
        const bool rec = p2g_records && with_force;
        launch_timed(c, name, with_force ? k_density_lds<true> : k_density_lds<false>, nblk(c->np, TB), TB, c->np,
                     c->n, c->P4(), c->V4(), c->cell_start, inv_hh, c->max_dens, c->dtv, c->prm.gravity,
                     reinterpret_cast<const P3 *>(c->xyz3), rec ? c->spos : nullptr, rec ? c->tmp0 : nullptr,
                     rec ? c->tmp1 : nullptr, rec ? c->dw : nullptr);
        return rec;
    }
    auto kern = pow2 ? (with_force ? k_density<true, true> : k_density<true, false>)
                     : (with_force ? k_density<false, true> : k_density<false, false>);
    launch_timed(c, name, kern, nblk(c->np, TB), TB, c->np, c->n, c->P4(), c->V4(), c->cell_start, h, inv_hh,
                 c->max_dens, c->dtv, c->prm.gravity);
    return false;
}

void k_launch_ext_force(FlipCtx *c)
{
    t_begin(c, "ext_force");
    if (c->np > 0)
        hipLaunchKernelGGL(k_ext_force, nblk(c->np, TB), TB, 0, c->stream, c->np, c->V4(), c->dtv, c->prm.gravity);
    t_end(c);
}

void k_launch_flags(FlipCtx *c, bool dw_ready, bool fluid_flags)
{
    float d = (float)c->prm.density;
    float n0 = (float)(1.0 / (d * d * d));
    launch_timed(c, "flags", k_flags, nblk(c->C, FLAGS_TB), FLAGS_TB, c->C, c->cell_start, c->P4(), c->V4(),
                 dw_ready ? c->dw : nullptr, c->A, c->L, 0.2 * (double)n0, fluid_flags ? c->cell_count : nullptr,
                 c->press);
}

void k_launch_p2g(FlipCtx *c, bool records_ready)
{
    const float re = (float)1.4;  // RE, mapper.cpp:14
    const float hh = re * re;
    if (c->unit_mass) {
        // packed (x, y, z, u_c) per component in spos / tmp0 / tmp1 (free until the corrector)
        t_begin(c, "p2g_prep");
        if (c->np > 0 && !records_ready)
            hipLaunchKernelGGL(k_p2g_prep_unit, nblk(c->np, TB), TB, 0, c->stream, c->np, c->n, c->P4(), c->V4(),
                               c->spos, c->tmp0, c->tmp1);
        t_end(c);
        t_begin(c, "p2g");
        int bi, bj, bk;
        p2g_bricks<0>(c->n, bi, bj, bk);
        const int nb0 = bi * bj * bk;
        p2g_bricks<1>(c->n, bi, bj, bk);
        const int nb1 = bi * bj * bk;
        p2g_bricks<2>(c->n, bi, bj, bk);
        const int nb2 = bi * bj * bk;
        hipLaunchKernelGGL(k_p2g_lds3, nb0 + nb1 + nb2, P2G_TB, 0, c->stream, c->n, c->spos, c->tmp0, c->tmp1,
                           c->cell_start, c->A, c->ug[0], c->ug[1], c->ug[2], c->us[0], c->us[1], c->us[2], hh, nb0,
                           nb1);
        t_end(c);
        return;
    }
    t_begin(c, "p2g_prep");
    if (c->np > 0) hipLaunchKernelGGL(k_p2g_prep, nblk(c->np, TB), TB, 0, c->stream, c->np, c->n, c->P4(), c->spos);
    t_end(c);
    t_begin(c, "p2g");
    unsigned nb = nblk(c->F, TB);
    hipLaunchKernelGGL(k_p2g<0>, nb, TB, 0, c->stream, c->n, c->spos, c->V4(), c->cell_start, c->A, c->ug[0],
                       c->us[0], hh);
    hipLaunchKernelGGL(k_p2g<1>, nb, TB, 0, c->stream, c->n, c->spos, c->V4(), c->cell_start, c->A, c->ug[1],
                       c->us[1], hh);
    hipLaunchKernelGGL(k_p2g<2>, nb, TB, 0, c->stream, c->n, c->spos, c->V4(), c->cell_start, c->A, c->ug[2],
                       c->us[2], hh);
    t_end(c);
}

void k_launch_g2p(FlipCtx *c, bool fresh)
{
    double a = c->prm.alpha;
    if (c->np == 0) return;
    if (fresh)  // the extrapolation has just written ud
        launch_timed(c, "g2p_flip", k_g2p_flip2, nblk(c->np, TB), TB, c->np, c->n, c->P4(), c->V4(), c->ud[0],
                     c->ud[1], c->ud[2], 1.0 - a, a);
    else
        launch_timed(c, "g2p_flip", k_g2p_flip, nblk(c->np, TB), TB, c->np, c->n, c->P4(), c->V4(), c->ug[0],
                     c->ug[1], c->ug[2], c->us[0], c->us[1], c->us[2], 1.0 - a, a);
}

void k_launch_g2p_points(FlipCtx *c, const float *g0, const float *g1, const float *g2, int64_t m, const float4 *p,
                         float4 *out)
{
    if (m > 0) hipLaunchKernelGGL(k_g2p_points, nblk(m, TB), TB, 0, c->stream, m, c->n, p, g0, g1, g2, out);
}

void k_launch_advect(FlipCtx *c, bool stuck)
{
    if (c->np > 0)
        launch_timed(c, "advect", k_advect, nblk(c->np, TB), TB, c->np, c->n, c->P4(), c->V4(), c->ug[0], c->ug[1],
                     c->ug[2], c->dtv, (float)(1.0 / c->n), 1.5 * c->prm.density / c->n, c->wall_start,
                     c->wall_near, c->wall_pos, c->wall_nrm, c->ev_count + EV_NONFINITE_ADV, c->M(), c->A,
                     stuck ? c->ev_list : nullptr, c->ev_count + EV_STUCK);
}

void k_launch_selftest(int64_t m, const float *x, const float *y, float *out, hipStream_t s)
{
    if (m > 0) hipLaunchKernelGGL(k_selftest, nblk(m, TB), TB, 0, s, m, x, y, out);
}

// The rand() stream lives on the host (scene placement, reposition) or on the
// device (the corrector's jitter draws); these move it where it is used next.
// draw k of a chunk from canonical state e_j (unit vectors): the coefficient
// of z_j in the k-th new word (k_jitter_draw), uploaded at context creation
// on the context's stream
void rng_upload_table(FlipCtx *c)
{
    static const std::vector<uint32_t> tab = [] {
        std::vector<uint32_t> t((size_t)JB * 31);
        for (int j = 0; j < 31; ++j) {
            uint32_t st[31] = {};
            st[j] = 1u;
            int f = 0, r = 28;
            for (int k = 0; k < JB; ++k) {
                st[f] += st[r];
                t[(size_t)k * 31 + j] = st[f];
                f = f == 30 ? 0 : f + 1;
                r = r == 30 ? 0 : r + 1;
            }
        }
        return t;
    }();
    FLIP_CHECK(hipMemcpyAsync(c->rng_table, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              c->stream));
    FLIP_CHECK(hipStreamSynchronize(c->stream));
}

// On the device the state is kept in canonical order (f = 0, r = 28).
void rng_to_device(FlipCtx *c)
{
    if (c->rng_on_device) return;
    if ((c->rng.f + 28) % 31 != c->rng.r) throw flip_arg_error{FLIP_E_STATE, "rand() state out of phase"};
    for (int j = 0; j < 31; ++j) c->h_rng[j] = c->rng.s[(c->rng.f + j) % 31];
    c->h_rng[31] = 0;
    c->h_rng[32] = 28;
    FLIP_CHECK(hipMemcpyAsync(c->rng_dev, c->h_rng, 33 * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    c->rng_on_device = true;
}

void rng_to_host(FlipCtx *c)
{
    if (!c->rng_on_device) return;
    FLIP_CHECK(hipMemcpyAsync(c->h_rng, c->rng_dev, 33 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    FLIP_CHECK(hipStreamSynchronize(c->stream));
    std::memcpy(c->rng.s, c->h_rng, sizeof(c->rng.s));
    c->rng.f = c->h_rng[31];
    c->rng.r = c->h_rng[32];
    c->rng_on_device = false;
}

void k_launch_clear_events(FlipCtx *c, int lo, int hi)
{
    hipLaunchKernelGGL(k_clear_events, 1, 64, 0, c->stream, c->ev_count, lo, hi);
}

__global__ void k_step_record(const int32_t *__restrict__ ev, const double *__restrict__ dtv,
                              FlipCtx::StepRecord *rec)
{
    const int t = threadIdx.x;
    if (t < EV_SLOTS) rec->ev[t] = ev[t];
    if (t == 0) rec->dt = dtv[0];
}

void k_launch_step_record(FlipCtx *c)
{
    hipLaunchKernelGGL(k_step_record, 1, 64, 0, c->stream, c->ev_count, c->dtv, c->d_rec);
}

// the step's time step into c->dtv[0] on the device (no host round trip)
void k_launch_choose_dt(FlipCtx *c)
{
    // the largest FLUID |u|^2 per block of the sort's gather (k_launch_sort
    // with c->partials), one per 256 particles
    const unsigned nb = c->prm.dt_policy == 1 && c->np > 0 ? nblk(c->np, 256) : 0;
    hipLaunchKernelGGL(k_choose_dt, 1, TB, 0, c->stream, (int)nb, c->partials, c->prm.dt, c->prm.cfl, c->n, c->dtv);
}

// a host-chosen time step (stage calls) into c->dtv[0]
void k_launch_set_dt(FlipCtx *c, double dt)
{
    hipLaunchKernelGGL(k_choose_dt, 1, TB, 0, c->stream, 0, c->partials, dt, 0.0, c->n, c->dtv);
}

// Stuck detection and repositioning (flip3D.cpp:932-968, 405-451), all on
// the device with no host round trip: the stuck list, the deep-water cells
// (k_deep, a scan, the first ones compacted), the plan (k_repo_plan: the
// list in creation order, the shuffle and positions drawn with the
// reference's rand() sequence, the resample levels), the new positions, the
// sort and the resample.  Kernels after the stuck test return at once when
// nothing is stuck.  The counts stay in ev_count[EV_STUCK..EV_NLEV]; the
// particles are left sorted on their current positions.
void k_stuck_and_reposition(FlipCtx *c)
{
    int n = c->n;
    if (c->prm.reposition && c->np > 0) {
        t_begin(c, "reposition");
        hipLaunchKernelGGL(k_deep, nblk(c->C, TB), TB, 0, c->stream, c->C, n, c->A, c->cell_count,
                           c->ev_count + EV_STUCK);
        // skipped on the device when no particle is stuck
        k_launch_scan(c, c->cell_count, c->scan_tmp, c->C, nullptr, c->ev_count + EV_STUCK);
        hipLaunchKernelGGL(k_deep_compact, nblk(c->C, TB), TB, 0, c->stream, c->C, c->cell_count, c->scan_tmp,
                           c->ev_count + EV_STUCK, c->ev_off);
        rng_to_device(c);
        hipLaunchKernelGGL(k_repo_plan, 1, RP_TB, 0, c->stream, c->ev_count, c->ev_list, c->scan_tmp + c->C,
                           c->ev_off, c->repo_ids, c->deep, c->rng_dev, c->ev_rand, c->pkey, c->repo_loff, c->pslot,
                           c->repo_order, n);
        hipLaunchKernelGGL(k_set_pos_by_id, 64, TB, 0, c->stream, c->ev_count + EV_MOVED, c->repo_ids, c->ev_rand,
                           c->inv, c->P4());
        t_end(c);
    }
    k_launch_sort(c);
    if (c->prm.reposition && c->np > 0) {
        t_begin(c, "reposition_resample");
        const float h = (float)(1.0 / n);
        hipLaunchKernelGGL(k_resample_levels, 1, 64 * RL_WAVES, 0, c->stream, c->ev_count + EV_NLEV, c->repo_loff,
                           c->repo_order, c->repo_ids, c->inv, n, c->P4(), c->V4(), c->cell_start, h * h);
        t_end(c);
    }
}

// corrector::correct (corrector.cpp:49-115); particles must be sorted.
int k_correct(FlipCtx *c)
{
    if (c->np == 0) return 0;
    // the sort's packed records (resample) are current while the storage is sorted
    if (!c->sorted) throw flip_arg_error{FLIP_E_STATE, "the corrector needs sorted particles"};
    int n = c->n;
    float re = (float)(c->prm.density / n);
    // ev_count[0]: listed particles, [1]: draw overflow, [EV_NONFINITE]: a
    // non-finite update, [3]: events (read with the step's end)
    hipLaunchKernelGGL(k_clear_events, 1, 64, 0, c->stream, c->ev_count, 0, 4);
    launch_timed(c, "spring", k_spring, nblk(c->np, SPRING_TB), SPRING_TB, c->np, n, c->P4(), c->M(), c->nrm_id,
                 c->cell_start, c->dtv, re, c->tmp0, c->ev_list, c->ev_count);
    // the jitter events in creation order with the reference's rand() draws,
    // on the device (corrector.cpp:74-76 at one thread)
    rng_to_device(c);
    launch_timed(c, "jitter_order", k_jitter_order, 256, OTB, c->ev_count, c->ev_list, c->ev_slot, c->ev_off);
    launch_timed(c, "jitter_draw", k_jitter_draw, 1, JB, c->ev_count, c->ev_off, c->ev_off, c->ev_rand,
                 (long)(3 * c->ev_cap), c->rng_dev, c->rng_table);
    launch_timed(c, "jitter_fix", k_spring_fix_wave, 256, FIX_TB, c->ev_count, c->ev_slot, c->ev_off, n, c->P4(),
                 c->M(), c->nrm_id, c->cell_start, c->dtv, re, c->ev_rand, c->tmp0);
    // With re^2 <= 1e-5 (N >= 159) every resample weight m * (re^2 /
    // fmax(r2, 1e-5) - 1) clamps to exactly 0 (r2 < re^2 puts the divisor at
    // 1e-5), so the weight sum is 0 and the reference keeps every velocity
    // (corrector.cpp:39-44): only the positions change.  Unit masses keep the
    // products exact zeros.
    const float hh = re * re;
    const bool keep = c->unit_mass && (double)hh <= 1.0e-5;
    if (!keep)
        launch_timed(c, "resample", k_resample, nblk(c->np, RESAMPLE_TB), RESAMPLE_TB, c->np, n, c->P4(), c->V4(),
                     c->cell_start, c->tmp0, c->tmp1, hh, reinterpret_cast<const P3 *>(c->xyz3));
    launch_timed(c, "correct_update", k_correct_update, nblk(c->np, TB), TB, c->np, c->P4(), c->V4(), c->tmp0,
                 keep ? nullptr : c->tmp1, c->ev_count + EV_NONFINITE);
    return 0;
}
