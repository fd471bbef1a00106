// k_pcg.hip - solver::solve (src/solver.cpp:183-241) on the GPU.
//
// Vectors live on the full n^3 grid but every pass runs over the compacted
// FLUID cell list: the reference's op() zeroes non-FLUID x, r, s, z and the
// ghost-fluid fetch x_ref (solver.cpp:19-26) never reads a non-FLUID entry,
// so only FLUID cells carry information.
//
// One PCG iteration k (solver.cpp:326-344) is two kernels over a grid of at
// most GMAX blocks (grid-stride loops):
//   pcg_a(k): every block sums the previous kernel's per-block partials of
//             (float)(r*r) and (float)(z*r) in one fixed order (so all blocks
//             agree bit for bit), applies the stop test, forms beta = a_k /
//             a_{k-1}; then s = z + (float)beta*s (ping-pong buffers, the six
//             neighbours recomputed), q = A s, partials of (float)(q*s).
//   pcg_b(k): every block forms alpha = a_k / (q.s); x += (float)alpha*s;
//             r += (float)(-alpha)*q; z = M r; partials of r.r and z.r.
// Reductions cross kernel boundaries only, so no block needs a device-scope
// fence (which on a multi-XCD part writes back the L2).  Products are FLOAT
// and accumulated in double as in product() (solver.cpp:44-54); the fixed
// tree order makes results deterministic run to run and differs from the
// reference's serial order only in the double's last bits, which the float
// casts of alpha and beta absorb.
//
// Preconditioner: with the reference's value test P_ref (solver.cpp:119-123)
// every coupling term is a signed zero unless some P == 1.0 exactly, so M is
// applied per cell, keeping the signed zeros so even the sign of a zero
// result matches.  If a P == 1.0 occurs (or for the flag-keyed IC(0) mode)
// the triangular solves run as i+j+k hyperplane wavefronts; each cell's
// formula and operand order are the reference's, so any topological order
// gives identical results.
#include "flip_device.hpp"
#include "flip_internal.hpp"

#include <algorithm>
#include <cmath>

using namespace flipdev;

namespace {

constexpr int TB = 256;
constexpr int GMAX = 512;
inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// device scalars: a_k lives in S_A0 + (k & 1)
enum { S_A0 = 0, S_A1, S_ALPHA, S_BETA, S_ERR2, S_ERR2_0, S_BNORM2, S_EPS };
// device int flags
enum { I_DONE = 0, I_ITERS, I_UNIT, I_CAPPED, I_TOLPOL, I_CAP };

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

// Sum of cnt partials p[i*stride + off], identical in every block.
__device__ __forceinline__ double sum_partials(const double *__restrict__ p, int cnt, int stride, int off)
{
    double a = 0.0;
    for (int i = threadIdx.x; i < cnt; i += TB) a += p[i * stride + off];
    return block_sum_all(a);
}

__device__ __forceinline__ void store_partials(double *p, int stride, double v0, double v1, double v2, int nv)
{
    double a = block_sum_all(v0);
    double b = nv > 1 ? block_sum_all(v1) : 0.0;
    double c = nv > 2 ? block_sum_all(v2) : 0.0;
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

struct Nb {
    long c[6];     // clamped neighbour cells in x_ref order: +i,-i,+j,-j,+k,-k
    int8_t a[6];   // their flags
};

__device__ __forceinline__ void neighbours(int n, long c, Nb &nb, const int8_t *__restrict__ A)
{
    int k = (int)(c % n), j = (int)((c / n) % n), i = (int)(c / ((long)n * n));
    long nn = (long)n * n;
    nb.c[0] = i + 1 <= n - 1 ? c + nn : c;
    nb.c[1] = i - 1 >= 0 ? c - nn : c;
    nb.c[2] = j + 1 <= n - 1 ? c + n : c;
    nb.c[3] = j - 1 >= 0 ? c - n : c;
    nb.c[4] = k + 1 <= n - 1 ? c + 1 : c;
    nb.c[5] = k - 1 >= 0 ? c - 1 : c;
#pragma unroll
    for (int t = 0; t < 6; ++t) nb.a[t] = A[nb.c[t]];
}

// solver.cpp:19-26 x_ref for neighbour t of FLUID cell c with value xs
__device__ __forceinline__ float x_ref(const Nb &nb, int t, float xn, float xs, float Lc, const float *__restrict__ L,
                                       int subcell)
{
    if (nb.a[t] == FLUID) return xn;
    if (nb.a[t] == WALL) return xs;
    return subcell ? (float)((double)L[nb.c[t]] / fmin(1.0e-6, (double)Lc) * (double)xs) : 0.0f;
}

// solver.cpp:29-41 one row of compute_Ax
__device__ __forceinline__ float ax_row(const Nb &nb, const float xv[6], float xs, float Lc, const float *L,
                                        int subcell, double h2)
{
    double acc = 6.0 * (double)xs;
#pragma unroll
    for (int t = 0; t < 6; ++t) acc -= (double)x_ref(nb, t, xv[t], xs, Lc, L, subcell);
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

// signed zero of A_ref(...) * 0.0f: -0 when both cells are in range and FLUID
__device__ __forceinline__ double coupling_zero(bool both_fluid) { return both_fluid ? -0.0 : 0.0; }

// per-cell preconditioner with every coupling a signed zero (no P == 1.0)
__device__ __forceinline__ float precond_cell(int n, long c, float r, double P, const int8_t *__restrict__ A)
{
    int k = (int)(c % n), j = (int)((c / n) % n), i = (int)(c / ((long)n * n));
    long nn = (long)n * n;
    double t = (double)r;
    t = t - coupling_zero(i - 1 >= 0 && A[c - nn] == FLUID);
    t = t - coupling_zero(j - 1 >= 0 && A[c - n] == FLUID);
    t = t - coupling_zero(k - 1 >= 0 && A[c - 1] == FLUID);
    double q = t * P;
    double u = q;
    u = u - coupling_zero(i + 1 <= n - 1 && A[c + nn] == FLUID);
    u = u - coupling_zero(j + 1 <= n - 1 && A[c + n] == FLUID);
    u = u - coupling_zero(k + 1 <= n - 1 && A[c + 1] == FLUID);
    return (float)(u * P);
}


// ---------------------------------------------------------------- setup
__global__ void k_negate(long C, float *b)
{
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < C) b[c] = -b[c];
}

// FLUID flags for the compaction; also zeroes x on non-FLUID cells, which the
// reference's first op(x) does (solver.cpp:205-212) and nothing reads earlier
__global__ void k_fluid_flag(long C, const int8_t *__restrict__ A, int32_t *f, float *x)
{
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    bool fl = A[c] == FLUID;
    f[c] = fl ? 1 : 0;
    if (!fl && x) x[c] = 0.0f;
}

__global__ void k_fluid_compact(long C, const int8_t *__restrict__ A, const int32_t *__restrict__ at, int32_t *fl)
{
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < C && A[c] == FLUID) fl[at[c]] = (int32_t)c;
}

// per-cell factorisation diagonal (solver.cpp:125-139 with no active coupling)
__global__ void k_build_diag(int nfl, const int32_t *__restrict__ fl, int n, const int8_t *__restrict__ A,
                             const float *__restrict__ L, int subcell, double *P, int32_t *iflags)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfl) return;
    long c = fl[t];
    double diag = (double)a_diag(n, c, A, L, subcell);
    double e = diag - 0.0 - 0.0 - 0.0;
    if (e < 0.25 * diag) e = diag;
    double p = 1.0 / (double)sqrtf((float)e);
    P[c] = p;
    if (p == 1.0) atomicAdd(&iflags[I_UNIT], 1);
}

// r = b - A x (op with a = -1), z = M r, s = z; partials r.r, z.r, b.b
template <bool JACOBI>
__global__ __launch_bounds__(TB) void k_pcg_init(int nfl, const int32_t *__restrict__ fl, int n,
                                                 const int8_t *__restrict__ A, const float *__restrict__ L,
                                                 int subcell, double h2, const float *__restrict__ x,
                                                 const float *__restrict__ b, const double *__restrict__ P, float *r,
                                                 float *z, float *s, double *part)
{
    double rr = 0.0, zr = 0.0, bb = 0.0;
    for (int t = blockIdx.x * TB + threadIdx.x; t < nfl; t += gridDim.x * TB) {
        long c = fl[t];
        Nb nb;
        neighbours(n, c, nb, A);
        float xv[6];
#pragma unroll
        for (int u = 0; u < 6; ++u) xv[u] = x[nb.c[u]];
        float ax = ax_row(nb, xv, x[c], L[c], L, subcell, h2);
        float bc = b[c];
        float rc = bc + -1.0f * ax;
        r[c] = rc;
        rr += (double)(rc * rc);
        bb += (double)(bc * bc);
        if (JACOBI) {
            float zc = precond_cell(n, c, rc, P[c], A);
            z[c] = zc;
            s[c] = zc;
            zr += (double)(zc * rc);
        }
    }
    store_partials(part, 3, rr, zr, bb, 3);
}

// single block: initial scalars from the init partials
__global__ __launch_bounds__(TB) void k_pcg_init_fin(int G, const double *__restrict__ part, double *scal)
{
    double rr = sum_partials(part, G, 3, 0);
    double zr = sum_partials(part, G, 3, 1);
    double bb = sum_partials(part, G, 3, 2);
    if (threadIdx.x == 0) {
        scal[S_ERR2_0] = rr;
        scal[S_A0] = zr;
        scal[S_BNORM2] = bb;
        scal[S_ERR2] = rr;
    }
}

// iteration part a (see the file comment)
__global__ __launch_bounds__(TB) void k_pcg_a(int nfl, const int32_t *__restrict__ fl, int n,
                                              const int8_t *__restrict__ A, const float *__restrict__ L, int subcell,
                                              double h2, const float *__restrict__ z, const float *__restrict__ s_old,
                                              float *s_new, float *q, const double *__restrict__ pb, double *pa,
                                              double *scal, int32_t *iflags, int k, int G)
{
    if (iflags[I_DONE]) return;
    float beta = 0.0f;
    if (k > 0) {
        double rr = sum_partials(pb, G, 2, 0);
        double zr = sum_partials(pb, G, 2, 1);
        bool capped;
        bool stop = stop_test(rr, k, iflags, scal, capped);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scal[S_ERR2] = iflags[I_TOLPOL] == 0 ? (double)(float)rr : rr;
            iflags[I_ITERS] = k;
            if (stop) {
                iflags[I_DONE] = 1;
                iflags[I_CAPPED] = capped ? 1 : 0;
            } else {
                scal[S_A0 + (k & 1)] = zr;
            }
        }
        if (stop) return;
        double bd = zr / scal[S_A0 + ((k - 1) & 1)];
        if (blockIdx.x == 0 && threadIdx.x == 0) scal[S_BETA] = bd;
        beta = (float)bd;
    }
    double qs = 0.0;
    for (int t = blockIdx.x * TB + threadIdx.x; t < nfl; t += gridDim.x * TB) {
        long c = fl[t];
        Nb nb;
        neighbours(n, c, nb, A);
        float sv[6];
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            long m = nb.c[u];
            sv[u] = nb.a[u] == FLUID ? (k == 0 ? z[m] : z[m] + beta * s_old[m]) : 0.0f;
        }
        float ss = k == 0 ? z[c] : z[c] + beta * s_old[c];
        float qc = ax_row(nb, sv, ss, L[c], L, subcell, h2);
        s_new[c] = ss;
        q[c] = qc;
        qs += (double)(qc * ss);
    }
    store_partials(pa, 1, qs, 0.0, 0.0, 1);
}

// iteration part b: x, r updates and (JACOBI) z = M r
template <bool JACOBI>
__global__ __launch_bounds__(TB) void k_pcg_b(int nfl, const int32_t *__restrict__ fl, int n,
                                              const int8_t *__restrict__ A, float *x, float *r,
                                              const float *__restrict__ s, const float *__restrict__ q,
                                              const double *__restrict__ P, float *z, const double *__restrict__ pa,
                                              double *pb, double *scal, const int32_t *__restrict__ iflags, int k,
                                              int G)
{
    if (iflags[I_DONE]) return;
    double qs = sum_partials(pa, G, 1, 0);
    const double alpha = scal[S_A0 + (k & 1)] / qs;
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[S_ALPHA] = alpha;
    const float a = (float)alpha, na = (float)-alpha;
    double rr = 0.0, zr = 0.0;
    for (int t = blockIdx.x * TB + threadIdx.x; t < nfl; t += gridDim.x * TB) {
        long c = fl[t];
        x[c] = x[c] + a * s[c];
        float rc = r[c] + na * q[c];
        r[c] = rc;
        rr += (double)(rc * rc);
        if (JACOBI) {
            float zc = precond_cell(n, c, rc, P[c], A);
            z[c] = zc;
            zr += (double)(zc * rc);
        }
    }
    store_partials(pb, 2, rr, zr, 0.0, 2);
}

// z.r partials after a wavefront preconditioner (into column `off` of stride `stride`)
__global__ __launch_bounds__(TB) void k_pcg_zr(int nfl, const int32_t *__restrict__ fl, const float *__restrict__ z,
                                               const float *__restrict__ r, double *part, int stride, int off,
                                               const int32_t *__restrict__ iflags)
{
    if (iflags[I_DONE]) return;
    double zr = 0.0;
    for (int t = blockIdx.x * TB + threadIdx.x; t < nfl; t += gridDim.x * TB) {
        long c = fl[t];
        zr += (double)(z[c] * r[c]);
    }
    double v = block_sum_all(zr);
    if (threadIdx.x == 0) part[blockIdx.x * stride + off] = v;
}

// single block: stop test after k completed iterations (end of a batch)
__global__ __launch_bounds__(TB) void k_pcg_check(const double *__restrict__ pb, double *scal, int32_t *iflags, int k,
                                                  int G)
{
    if (iflags[I_DONE]) return;
    double rr = sum_partials(pb, G, 2, 0);
    bool capped;
    bool stop = stop_test(rr, k, iflags, scal, capped);
    if (threadIdx.x == 0) {
        scal[S_ERR2] = iflags[I_TOLPOL] == 0 ? (double)(float)rr : rr;
        iflags[I_ITERS] = k;
        if (stop) {
            iflags[I_DONE] = 1;
            iflags[I_CAPPED] = capped ? 1 : 0;
        }
    }
}

__global__ void k_copy_fluid(int nfl, const int32_t *__restrict__ fl, const float *__restrict__ a, float *b)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nfl) b[fl[t]] = a[fl[t]];
}

// ------------------------------------------------ wavefront preconditioner
// P_ref (solver.cpp:119-123): mode 0 value test (== 1.0), mode 1 flag test
__device__ __forceinline__ float pref_d(int mode, int n, const int8_t *A, const double *P, int i, int j, int k)
{
    if (i < 0 || i > n - 1 || j < 0 || j > n - 1 || k < 0 || k > n - 1) return 0.0f;
    long c = cidx(n, i, j, k);
    if (mode == 1) return A[c] == FLUID ? (float)P[c] : 0.0f;
    if (P[c] != 1.0) return 0.0f;
    return (float)P[c];
}
__device__ __forceinline__ float pref_f(int mode, int n, const int8_t *A, const float *z, int i, int j, int k)
{
    if (i < 0 || i > n - 1 || j < 0 || j > n - 1 || k < 0 || k > n - 1) return 0.0f;
    long c = cidx(n, i, j, k);
    if (mode == 1) return A[c] == FLUID ? z[c] : 0.0f;
    if (z[c] != 1.0f) return 0.0f;
    return z[c];
}
// solver.cpp:97-101 A_ref
__device__ __forceinline__ float aref(int n, const int8_t *A, int i, int j, int k, int qi, int qj, int qk)
{
    if (i < 0 || i > n - 1 || j < 0 || j > n - 1 || k < 0 || k > n - 1 || A[cidx(n, i, j, k)] != FLUID) return 0.0f;
    if (qi < 0 || qi > n - 1 || qj < 0 || qj > n - 1 || qk < 0 || qk > n - 1 || A[cidx(n, qi, qj, qk)] != FLUID)
        return 0.0f;
    return -1.0f;
}

__device__ __forceinline__ bool plane_cell(int n, int d, int &i, int &j, int &k)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    i = t / n;
    j = t % n;
    k = d - i - j;
    return i < n && k >= 0 && k < n;
}

// buildPreconditioner on hyperplane d (solver.cpp:125-139)
__global__ void k_ic_build_plane(int n, int d, int mode, const int8_t *A, const float *L, int subcell, double *P)
{
    int i, j, k;
    if (!plane_cell(n, d, i, j, k)) return;
    long c = cidx(n, i, j, k);
    if (A[c] != FLUID) {
        P[c] = 0.0;
        return;
    }
    float l = aref(n, A, i - 1, j, k, i, j, k) * pref_d(mode, n, A, P, i - 1, j, k);
    float bo = aref(n, A, i, j - 1, k, i, j, k) * pref_d(mode, n, A, P, i, j - 1, k);
    float ba = aref(n, A, i, j, k - 1, i, j, k) * pref_d(mode, n, A, P, i, j, k - 1);
    double left = l, bottom = bo, back = ba;
    double diag = (double)a_diag(n, c, A, L, subcell);
    float sl = (float)left, sb = (float)bottom, sk = (float)back;
    double e = diag - (double)(sl * sl) - (double)(sb * sb) - (double)(sk * sk);
    if (e < 0.25 * diag) e = diag;
    P[c] = 1.0 / (double)sqrtf((float)e);
}

// forward sweep on plane d (solver.cpp:147-160)
__global__ void k_ic_fwd_plane(int n, int d, int mode, const int8_t *A, const double *P, const float *r, double *qv)
{
    int i, j, k;
    if (!plane_cell(n, d, i, j, k)) return;
    long c = cidx(n, i, j, k);
    if (A[c] != FLUID) {
        qv[c] = 0.0;
        return;
    }
    double left = (double)(aref(n, A, i - 1, j, k, i, j, k) * pref_d(mode, n, A, P, i - 1, j, k) *
                           pref_d(mode, n, A, qv, i - 1, j, k));
    double bottom = (double)(aref(n, A, i, j - 1, k, i, j, k) * pref_d(mode, n, A, P, i, j - 1, k) *
                             pref_d(mode, n, A, qv, i, j - 1, k));
    double back = (double)(aref(n, A, i, j, k - 1, i, j, k) * pref_d(mode, n, A, P, i, j, k - 1) *
                           pref_d(mode, n, A, qv, i, j, k - 1));
    double t = (double)r[c] - left - bottom - back;
    qv[c] = t * P[c];
}

// backward sweep on plane d (solver.cpp:163-176)
__global__ void k_ic_bwd_plane(int n, int d, int mode, const int8_t *A, const double *P, const double *qv, float *z)
{
    int i, j, k;
    if (!plane_cell(n, d, i, j, k)) return;
    long c = cidx(n, i, j, k);
    if (A[c] != FLUID) return;
    float ps = pref_d(mode, n, A, P, i, j, k);
    double right = (double)(aref(n, A, i, j, k, i + 1, j, k) * ps * pref_f(mode, n, A, z, i + 1, j, k));
    double top = (double)(aref(n, A, i, j, k, i, j + 1, k) * ps * pref_f(mode, n, A, z, i, j + 1, k));
    double front = (double)(aref(n, A, i, j, k, i, j, k + 1) * ps * pref_f(mode, n, A, z, i, j, k + 1));
    double t = qv[c] - right - top - front;
    z[c] = (float)(t * P[c]);
}

__global__ void k_fill_f64(long n, double *a, double v)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

}  // namespace

// ordered list of FLUID cells (zeroes non-FLUID press); returns the count (host sync)
int k_build_fluid_list(FlipCtx *c)
{
    hipLaunchKernelGGL(k_fluid_flag, nblk(c->C, TB), TB, 0, c->stream, c->C, c->A, c->cell_count, c->press);
    k_launch_scan(c, c->cell_count, c->scan_tmp, c->C);
    hipLaunchKernelGGL(k_fluid_compact, nblk(c->C, TB), TB, 0, c->stream, c->C, c->A, c->scan_tmp, c->fl);
    FLIP_CHECK(hipMemcpyAsync(c->h_pinned, c->scan_tmp + c->C, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    FLIP_CHECK(hipStreamSynchronize(c->stream));
    c->nfl = c->h_pinned[0];
    return c->nfl;
}

static void precond_wavefront(FlipCtx *c, int mode, const float *r, float *z)
{
    int n = c->n;
    unsigned nb = nblk((int64_t)n * n, TB);
    for (int d = 0; d <= 3 * n - 3; ++d)
        hipLaunchKernelGGL(k_ic_fwd_plane, nb, TB, 0, c->stream, n, d, mode, c->A, c->P, r, c->qv);
    for (int d = 3 * n - 3; d >= 0; --d)
        hipLaunchKernelGGL(k_ic_bwd_plane, nb, TB, 0, c->stream, n, d, mode, c->A, c->P, c->qv, z);
}

// PCG on the context's A, L, press (x) and div (b, negated in place).
int k_solve(FlipCtx *c, flip_step_stats *st)
{
    int n = c->n;
    long C = c->C;
    int subcell = c->prm.subcell;
    double h2 = (double)(float)(1.0 / ((double)n * n));
    int mode = c->prm.precond == FLIP_PRECOND_IC0 ? 1 : 0;

    t_begin(c, "pcg_setup");
    hipLaunchKernelGGL(k_negate, nblk(C, TB), TB, 0, c->stream, C, c->div);  // solver.cpp:231 flip(b)
    int nfl = k_build_fluid_list(c);
    int32_t hf[8] = {0, 0, 0, 0, c->prm.tol_policy, 0, 0, 0};
    long cap = c->prm.max_iters > 0 ? c->prm.max_iters : (long)n * n * n;
    hf[I_CAP] = (int32_t)std::min<long>(cap, 0x7fffffff);
    FLIP_CHECK(hipMemcpyAsync(c->iflags, hf, sizeof(hf), hipMemcpyHostToDevice, c->stream));
    double hs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    FLIP_CHECK(hipMemcpyAsync(c->scal, hs, sizeof(hs), hipMemcpyHostToDevice, c->stream));
    if (st) st->n_fluid_cells = nfl;
    if (nfl == 0) {
        t_end(c);
        if (st) {
            st->iterations = 0;
            st->error2 = st->error2_0 = st->bnorm2 = 0.0;
        }
        return 0;
    }
    unsigned nb = nblk(nfl, TB);
    int G = (int)std::min<unsigned>(nb, GMAX);
    double *pi = c->pcg_part, *pa = c->pcg_part + 3 * GMAX, *pb = c->pcg_part + 4 * GMAX;
    bool jacobi = mode == 0;
    if (mode == 0) {
        hipLaunchKernelGGL(k_build_diag, nb, TB, 0, c->stream, nfl, c->fl, n, c->A, c->L, subcell, c->P, c->iflags);
        int32_t unit = 0;
        FLIP_CHECK(hipMemcpyAsync(&unit, c->iflags + I_UNIT, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        if (st) st->n_precond_unit = unit;
        jacobi = unit == 0;
    }
    if (!jacobi) {
        hipLaunchKernelGGL(k_fill_f64, nblk(C, TB), TB, 0, c->stream, C, c->P, 0.0);
        unsigned pb2 = nblk((int64_t)n * n, TB);
        for (int d = 0; d <= 3 * n - 3; ++d)
            hipLaunchKernelGGL(k_ic_build_plane, pb2, TB, 0, c->stream, n, d, mode, c->A, c->L, subcell, c->P);
    }
    float *s_cur = c->sv[0], *s_nxt = c->sv[1];
    if (jacobi) {
        hipLaunchKernelGGL(k_pcg_init<true>, G, TB, 0, c->stream, nfl, c->fl, n, c->A, c->L, subcell, h2, c->press,
                           c->div, c->P, c->r, c->z, s_cur, pi);
    } else {
        hipLaunchKernelGGL(k_pcg_init<false>, G, TB, 0, c->stream, nfl, c->fl, n, c->A, c->L, subcell, h2, c->press,
                           c->div, c->P, c->r, c->z, s_cur, pi);
        precond_wavefront(c, mode, c->r, c->z);
        hipLaunchKernelGGL(k_copy_fluid, nb, TB, 0, c->stream, nfl, c->fl, c->z, s_cur);
        hipLaunchKernelGGL(k_pcg_zr, G, TB, 0, c->stream, nfl, c->fl, c->z, c->r, pi, 3, 1, c->iflags);
    }
    hipLaunchKernelGGL(k_pcg_init_fin, 1, TB, 0, c->stream, G, pi, c->scal);
    // stop threshold: solver.cpp:323, or the relative one
    double e0 = 0.0, b2 = 0.0;
    FLIP_CHECK(hipMemcpyAsync(&e0, c->scal + S_ERR2_0, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FLIP_CHECK(hipMemcpyAsync(&b2, c->scal + S_BNORM2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FLIP_CHECK(hipStreamSynchronize(c->stream));
    double eps = c->prm.tol_policy == 0 ? 1.0e-2 * ((double)n * n * n) : c->prm.tol * c->prm.tol * b2;
    FLIP_CHECK(hipMemcpyAsync(c->scal + S_EPS, &eps, sizeof(double), hipMemcpyHostToDevice, c->stream));
    t_end(c);

    int iters = 0;
    bool done = c->prm.tol_policy == 1 && e0 <= eps;
    // Iterations are enqueued in batches closed by a single-block stop check;
    // kernels after convergence return at once.  With profiling on, every
    // launch is timed and each batch is one iteration.
    const bool prof = c->profiling;
    int batch = prof ? 1 : std::max(1, c->last_iters - 2);
    int k = 0;
    while (!done) {
        for (int t = 0; t < batch; ++t, ++k) {
            t_begin(c, "pcg_a");
            hipLaunchKernelGGL(k_pcg_a, G, TB, 0, c->stream, nfl, c->fl, n, c->A, c->L, subcell, h2, c->z, s_cur,
                               s_nxt, c->q, pb, pa, c->scal, c->iflags, k, G);
            t_end(c);
            std::swap(s_cur, s_nxt);
            t_begin(c, "pcg_b");
            if (jacobi) {
                hipLaunchKernelGGL(k_pcg_b<true>, G, TB, 0, c->stream, nfl, c->fl, n, c->A, c->press, c->r, s_cur,
                                   c->q, c->P, c->z, pa, pb, c->scal, c->iflags, k, G);
                t_end(c);
            } else {
                hipLaunchKernelGGL(k_pcg_b<false>, G, TB, 0, c->stream, nfl, c->fl, n, c->A, c->press, c->r, s_cur,
                                   c->q, c->P, c->z, pa, pb, c->scal, c->iflags, k, G);
                t_end(c);
                t_begin(c, "precond_wavefront");
                precond_wavefront(c, mode, c->r, c->z);
                hipLaunchKernelGGL(k_pcg_zr, G, TB, 0, c->stream, nfl, c->fl, c->z, c->r, pb, 2, 1, c->iflags);
                t_end(c);
            }
        }
        hipLaunchKernelGGL(k_pcg_check, 1, TB, 0, c->stream, pb, c->scal, c->iflags, k, G);
        FLIP_CHECK(hipMemcpyAsync(c->h_pinned, c->iflags, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        done = c->h_pinned[I_DONE] != 0;
        iters = c->h_pinned[I_ITERS];
        batch = prof ? 1 : 4;
    }
    c->last_iters = iters;
    int32_t capped = 0;
    double e2 = 0.0;
    FLIP_CHECK(hipMemcpyAsync(&capped, c->iflags + I_CAPPED, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    FLIP_CHECK(hipMemcpyAsync(&e2, c->scal + S_ERR2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FLIP_CHECK(hipStreamSynchronize(c->stream));
    if (st) {
        st->iterations = iters;
        st->error2 = e2;
        st->error2_0 = e0;
        st->bnorm2 = b2;
    }
    if (!(e2 == e2)) return FLIP_E_NUMERIC;
    return capped ? 1 : 0;
}
