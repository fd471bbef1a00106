// pcg_device.hpp - device helpers shared by the pressure-solver kernels
// (k_pcg.hip: reference PCG, wavefront preconditioner, FP64 solve;
// k_learned.hip: the learned multilevel preconditioner and flexible CG).
#pragma once
#include "flip_device.hpp"
#include "flip_internal.hpp"

#include <algorithm>

namespace pcgdev {

using namespace flipdev;

constexpr int TB = 1024;   // 16-wave blocks, one per CU at the largest grid
constexpr int GMAX = 256;  // grid of the iteration kernels (partials per reduction)
inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// device scalars: a_k lives in S_A0 + (k & 1); S_RR0 is the residual the
// iteration-0 test of pcg_a compares (the initial one, or after a learned solve)
enum { S_A0 = 0, S_A1, S_ALPHA, S_BETA, S_ERR2, S_ERR2_0, S_BNORM2, S_EPS, S_RR0 };
// device int flags; I_KA is the next pcg_a iteration, I_KB the next pcg_b one;
// I_START: pcg_a(0) first tests S_RR0 against the stop threshold; I_SPEC: the
// per-cell (Jacobi-form) iterations run before the host knows whether the
// factor has P == 1.0 cells, and stop at once when it does (I_UNIT > 0)
enum { I_DONE = 0, I_ITERS, I_UNIT, I_CAPPED, I_TOLPOL, I_CAP, I_NFL, I_KA, I_KB, I_START, I_SPEC };

// neighbour kinds in the metadata word, 2 bits each in x_ref order
// (+i, -i, +j, -j, +k, -k), solver.cpp:19-26 / 161-163
enum { NB_FLUID = 0, NB_WALL = 1, NB_AIR = 2, NB_OUT = 3 };
constexpr uint32_t META_HAS_AIR = 1u << 12;

// Block-wide sum of one double per thread, fixed order, result in every thread.
__device__ __forceinline__ double block_sum_all(double v)
{
    __shared__ double lds[TB / 64];
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < TB / 64; ++i) r += lds[i];
    return r;
}

// Two block sums in one LDS round; each has the tree of block_sum_all.
__device__ __forceinline__ void block_sum2_all(double &a, double &b)
{
    __shared__ double lds[2][TB / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        lds[0][threadIdx.x >> 6] = a;
        lds[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    double ra = 0.0, rb = 0.0;
#pragma unroll
    for (int i = 0; i < TB / 64; ++i) {
        ra += lds[0][i];
        rb += lds[1][i];
    }
    a = ra;
    b = rb;
}

// Sum of cnt partials p[i*stride + off], identical in every block.
__device__ __forceinline__ double sum_partials(const double *__restrict__ p, int cnt, int stride, int off)
{
    double a = 0.0;
    for (int i = threadIdx.x; i < cnt; i += TB) a += p[i * stride + off];
    return block_sum_all(a);
}

// Sums of the cnt partial pairs (p[2i], p[2i+1]), identical in every block.
__device__ __forceinline__ void sum_partials2(const double *__restrict__ p, int cnt, double &a, double &b)
{
    a = 0.0;
    b = 0.0;
    for (int i = threadIdx.x; i < cnt; i += TB) {
        double2 v = reinterpret_cast<const double2 *>(p)[i];
        a += v.x;
        b += v.y;
    }
    block_sum2_all(a, b);
}

__device__ __forceinline__ void store_partials(double *p, int stride, double v0, double v1, double v2, int nv)
{
    double a = v0, b = v1, c = 0.0;
    if (nv > 1) block_sum2_all(a, b);
    else a = block_sum_all(v0);
    if (nv > 2) c = block_sum_all(v2);
    if (threadIdx.x == 0) {
        p[blockIdx.x * stride + 0] = a;
        if (nv > 1) p[blockIdx.x * stride + 1] = b;
        if (nv > 2) p[blockIdx.x * stride + 2] = c;
    }
}

// solver.cpp:331-337 stop test after `it` completed iterations
__device__ __forceinline__ bool stop_test(double rr, int it, const int32_t *iflags, const double *scal, bool &capped)
{
    bool stop = iflags[I_TOLPOL] == 0 ? (double)(float)rr <= scal[S_EPS] : rr <= scal[S_EPS];
    if (!(rr == rr)) stop = true;  // NaN: report instead of running to the cap
    capped = !stop && it >= iflags[I_CAP];
    return stop || capped;
}

__device__ __forceinline__ long nb_offset(int u, int n)
{
    long nn = (long)n * n;
    switch (u) {
    case 0: return nn;
    case 1: return -nn;
    case 2: return n;
    case 3: return -n;
    case 4: return 1;
    default: return -1;
    }
}

// solver.cpp:19-26 x_ref of neighbour u (kind kd) for FLUID cell c, value xs
__device__ __forceinline__ float x_ref(int kd, float xn, float xs, long c, long off, float Lc,
                                       const float *__restrict__ L, int subcell)
{
    if (kd == NB_FLUID) return xn;
    if (kd != NB_AIR) return xs;  // WALL, or clamped to the cell itself
    return subcell ? (float)((double)L[c + off] / fmin(1.0e-6, (double)Lc) * (double)xs) : 0.0f;
}

// solver.cpp:29-41 one row of compute_Ax; xv[u] holds FLUID neighbours' values
__device__ __forceinline__ float ax_row(uint32_t meta, const float xv[6], float xs, long c, int n,
                                        const float *__restrict__ L, int subcell, double h2)
{
    double acc = 6.0 * (double)xs;
    float Lc = (meta & META_HAS_AIR) ? L[c] : 0.0f;
#pragma unroll
    for (int u = 0; u < 6; ++u) {
        int kd = (meta >> (2 * u)) & 3;
        acc -= (double)x_ref(kd, xv[u], xs, c, nb_offset(u, n), Lc, L, subcell);
    }
    return (float)(acc / h2);
}

// solver.cpp:103-117 A_diag (FLOAT accumulation, neighbour order -i,+i,-j,+j,-k,+k)
__device__ __forceinline__ float a_diag(int n, long c, const int8_t *__restrict__ A, const float *__restrict__ L,
                                        int subcell)
{
    int k = (int)(c % n), j = (int)((c / n) % n), i = (int)(c / ((long)n * n));
    long nn = (long)n * n;
    float diag = 6.0f;
    float Lc = L[c];
    const int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0}, dk[6] = {0, 0, 0, 0, -1, 1};
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        int qi = i + di[m], qj = j + dj[m], qk = k + dk[m];
        if (qi < 0 || qi > n - 1 || qj < 0 || qj > n - 1 || qk < 0 || qk > n - 1) {
            diag = (float)((double)diag - 1.0);
            continue;
        }
        long q = c + di[m] * nn + dj[m] * n + dk[m];
        int8_t aq = A[q];
        if (aq == WALL) diag = (float)((double)diag - 1.0);
        else if (aq == AIR && subcell) diag = (float)((double)diag - (double)L[q] / fmin(1.0e-6, (double)Lc));
    }
    return diag;
}

// Per-cell preconditioner when no P == 1.0 exists: each coupling term
// A_ref*P_ref(P)*P_ref(.) is -0 for an in-range FLUID neighbour and +0
// otherwise (solver.cpp:147-176); the subtractions keep the signed zeros.
__device__ __forceinline__ double coupling_zero(uint32_t meta, int u)
{
    return ((meta >> (2 * u)) & 3) == NB_FLUID ? -0.0 : 0.0;
}

__device__ __forceinline__ float precond_cell(uint32_t meta, float r, double P)
{
    double t = (double)r;
    t = t - coupling_zero(meta, 1);  // -i
    t = t - coupling_zero(meta, 3);  // -j
    t = t - coupling_zero(meta, 5);  // -k
    double q = t * P;
    double u = q;
    u = u - coupling_zero(meta, 0);  // +i
    u = u - coupling_zero(meta, 2);  // +j
    u = u - coupling_zero(meta, 4);  // +k
    return (float)(u * P);
}

// One FLUID cell's operands for the stencil, all loads issued together: the
// cell's own (z, s) pair, the six neighbours' pairs (a non-FLUID neighbour's
// slot is in range and ignored; outside the grid the cell's own slot is read)
// and, for a cell next to AIR, the level set of the cell and of its
// neighbours.
struct StencilItem {
    long c;
    uint32_t m;
    float2 v;
    float2 w[6];
    float Lc;
    float Ln[6];
};

__device__ __forceinline__ void stencil_load(StencilItem &it, int t, const int32_t *__restrict__ fl,
                                             const uint32_t *__restrict__ meta, const float2 *__restrict__ rd,
                                             const float *__restrict__ L, int n)
{
    it.c = fl[t];
    it.m = meta[t];
    it.v = rd[it.c];
#pragma unroll
    for (int u = 0; u < 6; ++u) {
        const int kd = (it.m >> (2 * u)) & 3;
        it.w[u] = rd[it.c + (kd == NB_OUT ? 0 : nb_offset(u, n))];
    }
    it.Lc = 0.0f;
#pragma unroll
    for (int u = 0; u < 6; ++u) it.Ln[u] = 0.0f;
    if (it.m & META_HAS_AIR) {
        it.Lc = L[it.c];
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            const int kd = (it.m >> (2 * u)) & 3;
            it.Ln[u] = L[it.c + (kd == NB_AIR ? nb_offset(u, n) : 0)];
        }
    }
}

// solver.cpp:19-41 compute_Ax row of s = z + beta*s_old (k > 0) or s = z (k == 0)
__device__ __forceinline__ float stencil_row(const StencilItem &it, int k, float beta, int subcell, double h2,
                                             float &ss)
{
    ss = k == 0 ? it.v.x : it.v.x + beta * it.v.y;
    double acc = 6.0 * (double)ss;
#pragma unroll
    for (int u = 0; u < 6; ++u) {
        const int kd = (it.m >> (2 * u)) & 3;
        float xr;
        if (kd == NB_FLUID) xr = k == 0 ? it.w[u].x : it.w[u].x + beta * it.w[u].y;
        else if (kd != NB_AIR) xr = ss;  // WALL, or clamped to the cell itself
        else xr = subcell ? (float)((double)it.Ln[u] / fmin(1.0e-6, (double)it.Lc) * (double)ss) : 0.0f;
        acc -= (double)xr;
    }
    return (float)(acc / h2);
}

// grid of the iteration kernels for a context (fixed so graphs stay valid)
inline int pcg_grid(const FlipCtx *c)
{
    const int g = (int)std::min<unsigned>(nblk(c->C, TB), GMAX);
    return c->tune_blocks > 0 ? std::min(g, c->tune_blocks) : g;
}

}  // namespace pcgdev
