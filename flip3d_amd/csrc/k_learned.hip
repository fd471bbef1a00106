// k_learned.hip - FLIP_PRECOND_LEARNED: flexible CG whose preconditioner is a
// multigrid V-cycle with the U-Net as its coarsest-level solver.  No
// counterpart in the reference, which runs PCG only (src/solver.cpp:183-218);
// the reference PCG checks (and if needed finishes) every learned solve.
//
// Hierarchy.  Level 0 is the pressure system itself (the FLUID list of the
// PCG).  Level l >= 1 is a dense R_l^3 grid, R_l = n >> l, rediscretised from
// level l-1: a cell is FLUID if any of its 8 children is, WALL if all are,
// AIR otherwise; its level set is the mean over its FLUID children (FLUID
// cells) or its non-WALL children (AIR cells).  Every level uses the solver's
// ghost-fluid 7-point operator (solver.cpp:19-41) at its own spacing
// 1/R_l.  The coarsest level K has side 32 for n = 64, 128, 256 (16 for
// n = 32); the U-Net solves it.
//
// V-cycle z = M(r) (damped Jacobi, weight MG_W, two sweeps down and two up):
//   level 0:  z = S(r); b_1 = R(r - A z)
//   level l:  x_l = S(b_l); b_{l+1} = R(b_l - A_l x_l)        l = 1..K-1
//   level K:  x_K = UNet(A_K, L_K, b_K)
//   up:       x_l += P x_{l+1} / MG_S; x_l = S(x_l; b_l)       l = K-1..1
//             z += P x_1 / MG_S; z = S(z; r)
// R sums the residual over each coarse cell's FLUID children and P injects
// the coarse value into the children.  With these transfers the Galerkin
// coarse operator R A P is 16 times the rediscretised one; MG_S = 10
// over-corrects that by 1.6, as piecewise-constant transfers want.
//
// Flexible CG (Polak-Ribiere: beta_k = -alpha_{k-1} (z_k.q_{k-1}) /
// (z_{k-1}.r_{k-1})), because the network makes M nonlinear.  One iteration
// (the direction and A s, the update fused with the first smoothing sweep,
// the stop check in the level-0 residual, the rest of the V-cycle) is
// captured once as a HIP graph: the iteration index, buffer parity, FLUID
// count and the stop flag live in device memory, every kernel after
// convergence returns at once, and the host launches iteration k+1 before it
// reads iteration k's stop flag.
#include "pcg_device.hpp"
#include "unet_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace pcgdev;

namespace {

// Damped Jacobi weight.  Below 1, every mode of D^-1 A (eigenvalues in
// (0, 2]) shrinks each sweep; 0.9 lies between the classical 3-D smoothing
// optimum 6/7 and the undamped limit 1, where the checkerboard mode is no
// longer reduced (the flexible CG then stalls).  On the 128^3 dam break it
// takes 13.4 iterations per solve against 15.8 at 0.8.
constexpr float MG_W = 0.9f;
constexpr float MG_S = 10.0f;   // coarse-correction scale (see the file comment)
constexpr int FCG_MAX = 60;     // flexible iterations before the plain PCG takes over
constexpr int MG_COARSE = 32;   // target side of the coarsest (network) level
constexpr int TBL = 256;        // block of the level kernels

// fs (device doubles): z.r of the last two iterations, alpha of the last two, eps, r.r
enum { F_ZR0 = 0, F_ZR1, F_AL0, F_AL1, F_EPS, F_RR };
// fi (device ints): iteration index, stop flag, converged flag
enum { FI_K = 0, FI_DONE, FI_CONV };

inline unsigned nb(long n, int t = TBL) { return (unsigned)((n + t - 1) / t); }

struct IJK {
    int i, j, k;
};

__device__ __forceinline__ IJK ijk_of(long c, int n)
{
    const Split3 s = split3(c, n, n);
    return IJK{s.i, s.j, s.k};
}

// parent (side n/2) of the cell at p plus the offset of neighbour u (u = 6: the cell itself)
// the coarse cell (side h) over cell p's neighbour u (u = 6: p itself)
__device__ __forceinline__ long parent_nb(const IJK &p, int u, int h)
{
    const int i = p.i + (u == 0) - (u == 1), j = p.j + (u == 2) - (u == 3), k = p.k + (u == 4) - (u == 5);
    return ((long)(i >> 1) * h + (j >> 1)) * h + (k >> 1);
}

__device__ __forceinline__ float nbr_or_zero(uint32_t m, int u, const float *__restrict__ v, long c, int n)
{
    return ((m >> (2 * u)) & 3) == NB_FLUID ? v[c + nb_offset(u, n)] : 0.0f;
}

// ------------------------------------------------------------ hierarchy
// coarse flags and level set from the 8 children
__global__ void k_ml_coarsen(int Rc, int Rf, const int8_t *__restrict__ Af, const float *__restrict__ Lf, int8_t *Ac,
                             float *Lc)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long Cc = (long)Rc * Rc * Rc;
    if (g >= Cc) return;
    const Split3 s3 = split3(g, Rc, Rc);
    const int I = s3.i, J = s3.j, K = s3.k;
    int fluid = 0, wall = 0, nw = 0;
    float lf = 0.0f, ls = 0.0f;
    for (int d = 0; d < 8; ++d) {
        const int fi = 2 * I + (d >> 2), fj = 2 * J + ((d >> 1) & 1), fk = 2 * K + (d & 1);
        // children past the fine grid's end (a padded coarse level) are WALL
        const bool in = fi < Rf && fj < Rf && fk < Rf;
        const long f = in ? ((long)fi * Rf + fj) * Rf + fk : 0;
        const int a = in ? Af[f] : WALL;
        fluid += a == FLUID;
        wall += a == WALL;
        if (a == FLUID) lf += Lf[f];
        if (a != WALL) {
            ls += Lf[f];
            nw++;
        }
    }
    Ac[g] = fluid ? FLUID : (wall == 8 ? WALL : AIR);
    Lc[g] = fluid ? lf / (float)fluid : (nw ? ls / (float)nw : 1.0f);
}

// neighbour metadata and Jacobi weights of a dense level (FLUID cells)
__global__ void k_ml_meta(int R, const int8_t *__restrict__ A, const float *__restrict__ L, int subcell, double h2,
                          uint32_t *meta, float *dinv)
{
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long C = (long)R * R * R;
    if (c >= C) return;
    if (A[c] != FLUID) {
        meta[c] = 0;
        dinv[c] = 0.0f;
        return;
    }
    const Split3 s3 = split3(c, R, R);
    const int i = s3.i, j = s3.j, k = s3.k;
    const bool inr[6] = {i + 1 <= R - 1, i - 1 >= 0, j + 1 <= R - 1, j - 1 >= 0, k + 1 <= R - 1, k - 1 >= 0};
    uint32_t m = 0;
#pragma unroll
    for (int u = 0; u < 6; ++u) {
        int kd = NB_OUT;
        if (inr[u]) {
            int8_t a = A[c + nb_offset(u, R)];
            kd = a == FLUID ? NB_FLUID : (a == WALL ? NB_WALL : NB_AIR);
        }
        m |= (uint32_t)kd << (2 * u);
        if (kd == NB_AIR) m |= META_HAS_AIR;
    }
    meta[c] = m;
    float diag = a_diag(R, c, A, L, subcell);
    dinv[c] = diag > 0.0f ? (float)(h2 / (double)diag) : 0.0f;
}

// fine level: Jacobi weights h2 P^2 from the PCG diagonal on FLUID cells
__global__ void k_ml_setup0(const int32_t *__restrict__ iflags, const int32_t *__restrict__ fl,
                            const double *__restrict__ P, double h2, float *dinv)
{
    const int nfl = iflags[I_NFL];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nfl; t += gridDim.x * blockDim.x) {
        long c = fl[t];
        dinv[c] = (float)(h2 * P[c] * P[c]);
    }
}

__global__ void k_ml_zero(long n, float *a)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0f;
}

// ------------------------------------------------------------ V-cycle
// One FLUID cell's operands for the level-0 pre-smoothing, all loads issued
// together: the six FLUID neighbours' (dinv, r), the cell's own, and for a
// cell next to AIR the level set of the cell and of its AIR neighbours.
struct Pre0Item {
    long c;
    uint32_t m;
    float dq[6], rq[6];
    float rc, dc, Lc;
    float Ln[6];
};

__device__ __forceinline__ void pre0_load(Pre0Item &it, int t, const int32_t *__restrict__ fl,
                                          const uint32_t *__restrict__ meta, const float *__restrict__ r,
                                          const float *__restrict__ dinv, const float *__restrict__ L, int n)
{
    it.c = fl[t];
    it.m = meta[t];
#pragma unroll
    for (int u = 0; u < 6; ++u) {
        const int kd = (it.m >> (2 * u)) & 3;
        const long q = it.c + (kd == NB_FLUID ? nb_offset(u, n) : 0);
        it.dq[u] = dinv[q];
        it.rq[u] = r[q];
    }
    it.rc = r[it.c];
    it.dc = dinv[it.c];
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

// z1 = w D^-1 r (the neighbours' recomputed), z2 = z1 + w D^-1 (r - A z1):
// the row of ax_row with the level sets already loaded
__device__ __forceinline__ float pre0_row(const Pre0Item &it, int subcell, double h2)
{
    const float z1 = MG_W * it.dc * it.rc;
    double acc = 6.0 * (double)z1;
#pragma unroll
    for (int u = 0; u < 6; ++u) {
        const int kd = (it.m >> (2 * u)) & 3;
        float xr;
        if (kd == NB_FLUID) xr = MG_W * it.dq[u] * it.rq[u];
        else if (kd != NB_AIR) xr = z1;  // WALL, or clamped to the cell itself
        else xr = subcell ? (float)((double)it.Ln[u] / fmin(1.0e-6, (double)it.Lc) * (double)z1) : 0.0f;
        acc -= (double)xr;
    }
    const float az = (float)(acc / h2);
    return z1 + MG_W * it.dc * (it.rc - az);
}

// level 0 pre-smoothing of the first V-cycle: z1 = w D^-1 r,
// z2 = z1 + w D^-1 (r - A z1)
__global__ __launch_bounds__(TB) void k_ml_pre0(const int32_t *__restrict__ iflags, const int32_t *__restrict__ fi,
                                                const int32_t *__restrict__ fl, const uint32_t *__restrict__ meta,
                                                int n, const float *__restrict__ L, int subcell, double h2,
                                                const float *__restrict__ r, const float *__restrict__ dinv, float *z2)
{
    if (fi[FI_DONE]) return;
    const int nfl = iflags[I_NFL];
    Pre0Item it;
    for (int t = blockIdx.x * TB + threadIdx.x; t < nfl; t += gridDim.x * TB) {
        pre0_load(it, t, fl, meta, r, dinv, L, n);
        z2[it.c] = pre0_row(it, subcell, h2);
    }
}

// Flexible-CG update fused with the level-0 pre-smoothing (iteration k):
// alpha = zr_k / (q.s); x += alpha s; r' = r - alpha q into rn (the
// neighbours' r' recomputed from r and q, so r itself is only read); z2 from
// r' as k_ml_pre0 forms it; partials of r'.r'.  The values are those of the
// update in a pass of its own followed by k_ml_pre0, bit for bit.  The first
// item's loads go out before the partial sums.
struct UpdItem {
    long c;
    uint32_t m;
    float x, s, r, q, d;
    float rq[6], qq[6], dq[6];
    float Lc, Ln[6];
};

__device__ __forceinline__ void upd_load(UpdItem &it, int t, const int32_t *__restrict__ fl,
                                         const uint32_t *__restrict__ meta, const float *__restrict__ x,
                                         const float2 *__restrict__ bk, const float *__restrict__ r,
                                         const float *__restrict__ q, const float *__restrict__ dinv,
                                         const float *__restrict__ L, int n)
{
    it.c = fl[t];
    it.m = meta[t];
    it.x = x[it.c];
    it.s = bk[it.c].y;
    it.r = r[it.c];
    it.q = q[it.c];
    it.d = dinv[it.c];
#pragma unroll
    for (int u = 0; u < 6; ++u) {
        const int kd = (it.m >> (2 * u)) & 3;
        const long o = it.c + (kd == NB_FLUID ? nb_offset(u, n) : 0);
        it.rq[u] = r[o];
        it.qq[u] = q[o];
        it.dq[u] = dinv[o];
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

__global__ __launch_bounds__(TB) void k_fcg_b_pre0(const int32_t *__restrict__ iflags, const int32_t *__restrict__ fi,
                                                   const int32_t *__restrict__ fl, const uint32_t *__restrict__ meta,
                                                   int n, const float *__restrict__ L, int subcell, double h2, float *x,
                                                   const float *__restrict__ r, float *rn,
                                                   const float2 *__restrict__ B0, const float2 *__restrict__ B1,
                                                   const float *__restrict__ q, const float *__restrict__ dinv,
                                                   float *z2, const double *__restrict__ pa, double *fs, double *pb,
                                                   int G)
{
    if (fi[FI_DONE]) return;
    const int k = fi[FI_K];
    const int nfl = iflags[I_NFL];
    const float2 *__restrict__ bk = (k & 1) ? B1 : B0;
    const int t0 = blockIdx.x * TB + threadIdx.x;
    UpdItem it;
    if (t0 < nfl) upd_load(it, t0, fl, meta, x, bk, r, q, dinv, L, n);
    double qs = sum_partials(pa, G, 1, 0);
    double alpha = fs[F_ZR0 + (k & 1)] / qs;
    if (blockIdx.x == 0 && threadIdx.x == 0) fs[F_AL0 + (k & 1)] = alpha;
    const float a = (float)alpha, na = (float)-alpha;
    double rr = 0.0;
    for (int t = t0; t < nfl; t += gridDim.x * TB) {
        if (t != t0) upd_load(it, t, fl, meta, x, bk, r, q, dinv, L, n);
        x[it.c] = it.x + a * it.s;
        const float rc = it.r + na * it.q;
        rn[it.c] = rc;
        rr += (double)(rc * rc);
        // k_ml_pre0 on r'
        const float z1 = MG_W * it.d * rc;
        double acc = 6.0 * (double)z1;
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            const int kd = (it.m >> (2 * u)) & 3;
            float xr;
            if (kd == NB_FLUID) xr = MG_W * it.dq[u] * (it.rq[u] + na * it.qq[u]);
            else if (kd != NB_AIR) xr = z1;
            else xr = subcell ? (float)((double)it.Ln[u] / fmin(1.0e-6, (double)it.Lc) * (double)z1) : 0.0f;
            acc -= (double)xr;
        }
        const float az = (float)(acc / h2);
        z2[it.c] = z1 + MG_W * it.d * (rc - az);
    }
    store_partials(pb, 1, rr, 0.0, 0.0, 1);
}

// dense level pre-smoothing from zero: t = w D^-1 b, x = t + w D^-1 (b - A t)
__global__ __launch_bounds__(TBL) void k_ml_pre(const int32_t *__restrict__ fi, int R, const int8_t *__restrict__ A,
                                                const uint32_t *__restrict__ meta, const float *__restrict__ L,
                                                int subcell, double h2, const float *__restrict__ b,
                                                const float *__restrict__ dinv, float *x)
{
    if (fi[FI_DONE]) return;
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long C = (long)R * R * R;
    if (c >= C) return;
    if (A[c] != FLUID) {
        x[c] = 0.0f;
        return;
    }
    uint32_t m = meta[c];
    float tv[6];
#pragma unroll
    for (int u = 0; u < 6; ++u) {
        const int kd = (m >> (2 * u)) & 3;
        const long q = c + (kd == NB_FLUID ? nb_offset(u, R) : 0);
        tv[u] = kd == NB_FLUID ? MG_W * dinv[q] * b[q] : 0.0f;
    }
    float bc = b[c], dc = dinv[c];
    float t = MG_W * dc * bc;
    float at = ax_row(m, tv, t, c, R, L, subcell, h2);
    x[c] = t + MG_W * dc * (bc - at);
}

// level 0 residual: res = r - A z2 on FLUID cells (res is zero elsewhere).
// With check set it first applies the flexible-CG stop test of the
// iteration that has just updated r: every block sums k_fcg_b_pre0's r'.r'
// partials in the same order and stops on convergence or NaN; block 0
// records the result, advances the iteration index and also stops at kmax.
__global__ __launch_bounds__(TBL) void k_ml_resid0(const int32_t *__restrict__ iflags, int32_t *fi,
                                                   const int32_t *__restrict__ fl, const uint32_t *__restrict__ meta,
                                                   int n, const float *__restrict__ L, int subcell, double h2,
                                                   const float *__restrict__ r, const float *__restrict__ z2,
                                                   float *res, const double *__restrict__ pb, double *fs, int G,
                                                   int kmax, int check)
{
    if (fi[FI_DONE]) return;
    const int nfl = iflags[I_NFL];
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    long c0 = 0;
    uint32_t m0 = 0;
    float zv0[6], zc0 = 0.0f, rc0 = 0.0f;
    if (t0 < nfl) {
        c0 = fl[t0];
        m0 = meta[t0];
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            const int kd = (m0 >> (2 * u)) & 3;
            zv0[u] = z2[c0 + (kd == NB_FLUID ? nb_offset(u, n) : 0)];
        }
        zc0 = z2[c0];
        rc0 = r[c0];
    }
    if (check) {
        double a = 0.0;
        for (int i = threadIdx.x; i < G; i += TBL) a += pb[i];
        __shared__ double red[TBL / 64];
        a = wave_sum(a);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
        __syncthreads();
        double rr = 0.0;
#pragma unroll
        for (int w = 0; w < TBL / 64; ++w) rr += red[w];
        const bool conv = rr <= fs[F_EPS];
        const bool stop = conv || !(rr == rr);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            const int k = fi[FI_K] + 1;
            fs[F_RR] = rr;
            fi[FI_K] = k;
            if (stop || k >= kmax) {
                fi[FI_DONE] = 1;
                fi[FI_CONV] = conv ? 1 : 0;
            }
        }
        if (stop) return;
    }
    for (int t = t0; t < nfl; t += gridDim.x * blockDim.x) {
        long c = c0;
        uint32_t m = m0;
        float zv[6], zc = zc0, rc = rc0;
#pragma unroll
        for (int u = 0; u < 6; ++u) zv[u] = zv0[u];
        if (t != t0) {
            c = fl[t];
            m = meta[t];
#pragma unroll
            for (int u = 0; u < 6; ++u) {
                const int kd = (m >> (2 * u)) & 3;
                zv[u] = z2[c + (kd == NB_FLUID ? nb_offset(u, n) : 0)];
            }
            zc = z2[c];
            rc = r[c];
        }
#pragma unroll
        for (int u = 0; u < 6; ++u) zv[u] = ((m >> (2 * u)) & 3) == NB_FLUID ? zv[u] : 0.0f;
        res[c] = rc - ax_row(m, zv, zc, c, n, L, subcell, h2);
    }
}

// restriction: bc[I] = sum of the 8 children's residuals (zero off FLUID).
// With Ac given (the network's level) it also writes per-block partials
// (sum bc^2 over FLUID coarse cells, FLUID count) for the network's scaling.
__global__ __launch_bounds__(TBL) void k_ml_sum8(const int32_t *__restrict__ fi, int Rc, int Rf,
                                                 const float *__restrict__ res, float *bc,
                                                 const int8_t *__restrict__ Ac, double *part)
{
    if (fi[FI_DONE]) return;
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long Cc = (long)Rc * Rc * Rc;
    float sum = 0.0f;
    if (g < Cc) {
        const IJK p = ijk_of(g, Rc);
        const long f0 = ((long)(2 * p.i) * Rf + 2 * p.j) * Rf + 2 * p.k;
        const long di = (long)Rf * Rf, dj = Rf;
        if ((Rf & 1) == 0 && 2 * p.i + 1 < Rf && 2 * p.j + 1 < Rf && 2 * p.k + 1 < Rf) {
            const float2 a = *reinterpret_cast<const float2 *>(res + f0);
            const float2 b = *reinterpret_cast<const float2 *>(res + f0 + dj);
            const float2 c = *reinterpret_cast<const float2 *>(res + f0 + di);
            const float2 d = *reinterpret_cast<const float2 *>(res + f0 + di + dj);
            sum += a.x;
            sum += a.y;
            sum += b.x;
            sum += b.y;
            sum += c.x;
            sum += c.y;
            sum += d.x;
            sum += d.y;
        } else {
            // an odd fine side or a padded coarse level: the children inside
            // the fine grid, in the same order
            for (int d = 0; d < 8; ++d) {
                const int ci = 2 * p.i + (d >> 2), cj = 2 * p.j + ((d >> 1) & 1), ck = 2 * p.k + (d & 1);
                if (ci < Rf && cj < Rf && ck < Rf) sum += res[((long)ci * Rf + cj) * Rf + ck];
            }
        }
        bc[g] = sum;
    }
    if (!Ac) return;
    __shared__ double s0[TBL / 64], s1[TBL / 64];
    const bool fl = g < Cc && Ac[g] == FLUID;
    double a2 = fl ? (double)sum * (double)sum : 0.0, cnt = fl ? 1.0 : 0.0;
    a2 = wave_sum(a2);
    cnt = wave_sum(cnt);
    if ((threadIdx.x & 63) == 0) {
        s0[threadIdx.x >> 6] = a2;
        s1[threadIdx.x >> 6] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tc = 0.0;
        for (int w = 0; w < TBL / 64; ++w) {
            ta += s0[w];
            tc += s1[w];
        }
        part[2 * blockIdx.x] = ta;
        part[2 * blockIdx.x + 1] = tc;
    }
}

// The dense level residual b - A x on the fine level Rf (zero off FLUID)
// restricted into the coarse level Rc in one pass: eight threads per coarse
// cell form its children's residuals, and the first of them adds them in
// k_ml_sum8's order (children inside the fine grid, d = 0..7), so bc is bit
// for bit the restriction of the stored residual; the fine residual itself is
// not stored.  With Ac: per-block partials as k_ml_sum8's, one per 32 coarse
// cells (TBL / 8).
__global__ __launch_bounds__(TBL) void k_ml_resid_sum8(const int32_t *__restrict__ fi, int Rc, int Rf,
                                                       const int8_t *__restrict__ A,
                                                       const uint32_t *__restrict__ meta,
                                                       const float *__restrict__ L, int subcell, double h2,
                                                       const float *__restrict__ b, const float *__restrict__ x,
                                                       float *bc, const int8_t *__restrict__ Ac, double *part)
{
    if (fi[FI_DONE]) return;
    const long g = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 3;
    const int d = threadIdx.x & 7;
    const long Cc = (long)Rc * Rc * Rc;
    float res = 0.0f;
    int inside = 0;
    if (g < Cc) {
        const IJK p = ijk_of(g, Rc);
        const int ci = 2 * p.i + (d >> 2), cj = 2 * p.j + ((d >> 1) & 1), ck = 2 * p.k + (d & 1);
        inside = ci < Rf && cj < Rf && ck < Rf;
        if (inside) {
            const long c = ((long)ci * Rf + cj) * Rf + ck;
            if (A[c] == FLUID) {
                const uint32_t m = meta[c];
                float xv[6];
#pragma unroll
                for (int u = 0; u < 6; ++u) {
                    const int kd = (m >> (2 * u)) & 3;
                    const float w = x[c + (kd == NB_FLUID ? nb_offset(u, Rf) : 0)];
                    xv[u] = kd == NB_FLUID ? w : 0.0f;
                }
                res = b[c] - ax_row(m, xv, x[c], c, Rf, L, subcell, h2);
            }
        }
    }
    const int base = (threadIdx.x & 63) & ~7;
    float sum = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float r = __shfl(res, base + e, 64);
        if (__shfl(inside, base + e, 64)) sum += r;
    }
    const bool lead = g < Cc && d == 0;
    if (lead) bc[g] = sum;
    if (!Ac) return;
    __shared__ double s0[TBL / 64], s1[TBL / 64];
    const bool fl = lead && Ac[g] == FLUID;
    double a2 = fl ? (double)sum * (double)sum : 0.0, cnt = fl ? 1.0 : 0.0;
    a2 = wave_sum(a2);
    cnt = wave_sum(cnt);
    if ((threadIdx.x & 63) == 0) {
        s0[threadIdx.x >> 6] = a2;
        s1[threadIdx.x >> 6] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tc = 0.0;
        for (int w = 0; w < TBL / 64; ++w) {
            ta += s0[w];
            tc += s1[w];
        }
        part[2 * blockIdx.x] = ta;
        part[2 * blockIdx.x + 1] = tc;
    }
}

// dense level post-smoothing, first sweep: x3 = x + xc(parent) / MG_S,
// t = x3 + w D^-1 (b - A x3)
__global__ __launch_bounds__(TBL) void k_ml_post_a(const int32_t *__restrict__ fi, int R,
                                                   const int8_t *__restrict__ A, const uint32_t *__restrict__ meta,
                                                   const float *__restrict__ L, int subcell, double h2,
                                                   const float *__restrict__ b, const float *__restrict__ dinv,
                                                   const float *__restrict__ x, const float *__restrict__ xc,
                                                   float *t)
{
    if (fi[FI_DONE]) return;
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long C = (long)R * R * R;
    if (c >= C) return;
    if (A[c] != FLUID) {
        t[c] = 0.0f;
        return;
    }
    const float ws = 1.0f / MG_S;
    uint32_t m = meta[c];
    const IJK p = ijk_of(c, R);
    float xv[6];
#pragma unroll
    for (int u = 0; u < 6; ++u) {
        const int kd = (m >> (2 * u)) & 3;
        const bool f = kd == NB_FLUID;
        const float xn = x[c + (f ? nb_offset(u, R) : 0)];
        const float xp = xc[f ? parent_nb(p, u, R >> 1) : parent_nb(p, 6, R >> 1)];
        xv[u] = f ? xn + ws * xp : 0.0f;
    }
    float x3 = x[c] + ws * xc[parent_nb(p, 6, R >> 1)];
    float ax = ax_row(m, xv, x3, c, R, L, subcell, h2);
    t[c] = x3 + MG_W * dinv[c] * (b[c] - ax);
}

// dense level post-smoothing, second sweep: x = t + w D^-1 (b - A t)
__global__ __launch_bounds__(TBL) void k_ml_post_b(const int32_t *__restrict__ fi, int R,
                                                   const int8_t *__restrict__ A, const uint32_t *__restrict__ meta,
                                                   const float *__restrict__ L, int subcell, double h2,
                                                   const float *__restrict__ b, const float *__restrict__ dinv,
                                                   const float *__restrict__ t, float *x)
{
    if (fi[FI_DONE]) return;
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long C = (long)R * R * R;
    if (c >= C) return;
    if (A[c] != FLUID) {
        x[c] = 0.0f;
        return;
    }
    uint32_t m = meta[c];
    float tv[6];
#pragma unroll
    for (int u = 0; u < 6; ++u) tv[u] = nbr_or_zero(m, u, t, c, R);
    float ax = ax_row(m, tv, t[c], c, R, L, subcell, h2);
    x[c] = t[c] + MG_W * dinv[c] * (b[c] - ax);
}

// level 0 post-smoothing, first sweep: z3 = z2 + x1(parent) / MG_S,
// z4 = z3 + w D^-1 (r - A z3)
__global__ __launch_bounds__(TBL) void k_ml_post0_a(const int32_t *__restrict__ iflags,
                                                    const int32_t *__restrict__ fi, const int32_t *__restrict__ fl,
                                                    const uint32_t *__restrict__ meta, int n,
                                                    const float *__restrict__ L, int subcell, double h2,
                                                    const float *__restrict__ r, const float *__restrict__ dinv,
                                                    const float *__restrict__ z2, const float *__restrict__ x1,
                                                    int h1, float *z4)
{
    if (fi[FI_DONE]) return;
    const int nfl = iflags[I_NFL];
    const float ws = 1.0f / MG_S;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nfl; t += gridDim.x * blockDim.x) {
        long c = fl[t];
        uint32_t m = meta[t];
        const IJK p = ijk_of(c, n);
        float zv[6];
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            const int kd = (m >> (2 * u)) & 3;
            const bool f = kd == NB_FLUID;
            const float zn = z2[c + (f ? nb_offset(u, n) : 0)];
            const float xp = x1[f ? parent_nb(p, u, h1) : parent_nb(p, 6, h1)];
            zv[u] = f ? zn + ws * xp : 0.0f;
        }
        float z3 = z2[c] + ws * x1[parent_nb(p, 6, h1)];
        float az = ax_row(m, zv, z3, c, n, L, subcell, h2);
        z4[c] = z3 + MG_W * dinv[c] * (r[c] - az);
    }
}

// level 0 post-smoothing, second sweep: z = z4 + w D^-1 (r - A z4) into the
// .x slots of the (z, s) buffer the next flexible-CG iteration reads (parity
// (k + 1) & 1 with k = fi[FI_K]); partials z.r and, when q is given, z.q;
// with rcopy given, r is also copied there (the iteration's r' into the
// residual the next update reads)
__global__ __launch_bounds__(TB) void k_ml_post0_b(const int32_t *__restrict__ iflags, const int32_t *__restrict__ fi,
                                                   const int32_t *__restrict__ fl, const uint32_t *__restrict__ meta,
                                                   int n, const float *__restrict__ L, int subcell, double h2,
                                                   const float *__restrict__ r, const float *__restrict__ dinv,
                                                   const float *__restrict__ z4, const float *__restrict__ q,
                                                   float2 *B0, float2 *B1, double *part, int32_t *hflag,
                                                   float *rcopy)
{
    // the stop flag into mapped host memory (null outside the iteration)
    if (hflag && blockIdx.x == 0 && threadIdx.x == 0) {
        hflag[1] = fi[FI_K];
        hflag[0] = fi[FI_DONE];
    }
    if (fi[FI_DONE]) return;
    const int nfl = iflags[I_NFL];
    float2 *B = ((fi[FI_K] + 1) & 1) ? B1 : B0;
    double zr = 0.0, zq = 0.0;
    for (int t = blockIdx.x * TB + threadIdx.x; t < nfl; t += gridDim.x * TB) {
        long c = fl[t];
        uint32_t m = meta[t];
        float zv[6];
#pragma unroll
        for (int u = 0; u < 6; ++u) zv[u] = nbr_or_zero(m, u, z4, c, n);
        float rc = r[c];
        if (rcopy) rcopy[c] = rc;  // r' becomes the residual the next update reads
        float az = ax_row(m, zv, z4[c], c, n, L, subcell, h2);
        float z = z4[c] + MG_W * dinv[c] * (rc - az);
        B[c].x = z;
        zr += (double)(z * rc);
        if (q) zq += (double)(z * q[c]);
    }
    store_partials(part, 2, zr, zq, 0.0, 2);
}

// ------------------------------------------------------------ flexible CG
// the flexible CG's device state: iteration 0, the solve's stop threshold; it
// does not start (done at once) when the factor has P == 1.0 cells (the
// wavefront PCG then solves), when there is no FLUID cell, or when the start
// residual already meets a relative tolerance
__global__ void k_fcg_begin(const int32_t *__restrict__ iflags, const double *__restrict__ scal, double *fs,
                            int32_t *fi)
{
    const int t = threadIdx.x;
    if (t < 8) fs[t] = 0.0;
    if (t < 4) fi[t] = 0;
    __syncthreads();
    if (t == 0) {
        fs[F_EPS] = scal[S_EPS];
        const bool skip = iflags[I_UNIT] > 0 || iflags[I_NFL] == 0 ||
                          (iflags[I_TOLPOL] == 1 && scal[S_ERR2_0] <= scal[S_EPS]);
        fi[FI_DONE] = skip ? 1 : 0;
    }
}

// part a(k): z.r_k from the preconditioner partials, beta, s_k = z + beta
// s_{k-1} (neighbours recomputed), q = A s_k, partials of q.s.  The first
// item's loads go out before the partial sums (as in the reference PCG's
// k_pcg_a, whose stencil row this is).
__global__ __launch_bounds__(TB) void k_fcg_a(const int32_t *__restrict__ iflags, const int32_t *__restrict__ fi,
                                              const int32_t *__restrict__ fl, const uint32_t *__restrict__ meta,
                                              int n, const float *__restrict__ L, int subcell, double h2, float2 *B0,
                                              float2 *B1, float *q, const double *__restrict__ pz, double *fs,
                                              double *pa, int G)
{
    if (fi[FI_DONE]) return;
    const int k = fi[FI_K];
    const int nfl = iflags[I_NFL];
    const float2 *__restrict__ rd = (k & 1) ? B0 : B1;
    float2 *wr = (k & 1) ? B1 : B0;
    const int t0 = blockIdx.x * TB + threadIdx.x;
    StencilItem it;
    if (t0 < nfl) stencil_load(it, t0, fl, meta, rd, L, n);
    double zr, zq;
    sum_partials2(pz, G, zr, zq);
    float beta = 0.0f;
    if (k > 0) beta = (float)(-fs[F_AL0 + ((k - 1) & 1)] * zq / fs[F_ZR0 + ((k - 1) & 1)]);
    if (blockIdx.x == 0 && threadIdx.x == 0) fs[F_ZR0 + (k & 1)] = zr;
    double qs = 0.0;
    for (int t = t0; t < nfl; t += gridDim.x * TB) {
        if (t != t0) stencil_load(it, t, fl, meta, rd, L, n);
        float ss;
        const float qc = stencil_row(it, k, beta, subcell, h2, ss);
        wr[it.c] = make_float2(it.v.x, ss);
        q[it.c] = qc;
        qs += (double)(qc * ss);
    }
    store_partials(pa, 1, qs, 0.0, 0.0, 1);
}

struct Level {
    int R = 0;
    long C = 0;
    double h2 = 0.0;
    int8_t *A = nullptr;
    float *L = nullptr, *dinv = nullptr, *b = nullptr, *x = nullptr, *t = nullptr;
    uint32_t *meta = nullptr;
};

template <class T>
void dmalloc(T **p, long n)
{
    FLIP_CHECK(hipMalloc((void **)p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
}

}  // namespace

// ------------------------------------------------- learned solver state
struct LearnedSolver {
    int n = 0, K = 0;
    std::vector<Level> lv;  // lv[1..K]
    UNet *net = nullptr;
    float *dinv0 = nullptr, *z2 = nullptr, *z4 = nullptr, *res0 = nullptr, *rn = nullptr;
    float *sigma = nullptr;
    double *part = nullptr;  // [4 * GMAX]: pa, pb, pz (stride 2)
    double *fs = nullptr;
    int32_t *fi = nullptr;
    double *upart = nullptr;
    int32_t *h_flag = nullptr;  // mapped pinned host word: the stop flag (monotonic within a solve)
    int32_t *d_flag = nullptr;  // its device address
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipGraphExec_t graph = nullptr;
    long graph_key = -1;  // subcell and block count the graph was captured for
};

void learned_free(LearnedSolver *ls)
{
    if (!ls) return;
    if (ls->graph) (void)hipGraphExecDestroy(ls->graph);
    if (ls->net) unet_free(ls->net);
    for (Level &l : ls->lv) {
        void *bufs[] = {l.A, l.L, l.dinv, l.b, l.x, l.t, l.meta};
        for (void *b : bufs)
            if (b) (void)hipFree(b);
    }
    void *bufs[] = {ls->dinv0, ls->z2, ls->z4, ls->res0, ls->rn, ls->sigma, ls->part, ls->fs, ls->fi, ls->upart};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (ls->h_flag) (void)hipHostFree(ls->h_flag);
    for (hipEvent_t e : ls->ev)
        if (e) (void)hipEventDestroy(e);
    delete ls;
}

// number of coarse levels for grid side n: halve while the next level is at
// least MG_COARSE cells wide and a multiple of 8 (at least one level)
int learned_side(int n) { return (n + 15) / 16 * 16; }

int learned_levels(int n)
{
    int K = 1;
    while ((n % (1 << (K + 1))) == 0 && (n >> (K + 1)) >= MG_COARSE && ((n >> (K + 1)) % 8) == 0) ++K;
    return K;
}

static LearnedSolver *learned_get(FlipCtx *c)
{
    if (c->learned) return c->learned;
    const int n = c->n;
    if (c->learned_weights.empty())
        throw flip_arg_error{FLIP_E_STATE, "FLIP_PRECOND_LEARNED needs weights (flip_gpu_set_learned_weights)"};
    LearnedSolver *ls = new LearnedSolver();
    try {
        ls->n = n;
        // the coarse levels halve a side padded to a multiple of 16 (the
        // network's level a multiple of 8); level 1 covers the fine grid,
        // its cells past the grid's end WALL
        const int np = learned_side(n);
        ls->K = learned_levels(np);
        ls->lv.resize(ls->K + 1);
        for (int l = 1; l <= ls->K; ++l) {
            Level &L = ls->lv[l];
            L.R = np >> l;
            L.C = (long)L.R * L.R * L.R;
            L.h2 = (double)(float)(1.0 / ((double)L.R * L.R));
            dmalloc(&L.A, L.C);
            dmalloc(&L.L, L.C);
            dmalloc(&L.dinv, L.C);
            dmalloc(&L.b, L.C);
            dmalloc(&L.x, L.C);
            dmalloc(&L.t, L.C);
            dmalloc(&L.meta, L.C);
        }
        ls->net = unet_new(ls->lv[ls->K].R, 1, c->stream);
        std::vector<float> w;
        std::string err;
        if (!unet_read_weights(c->learned_weights.c_str(), ls->net->nparam, w, err))
            throw flip_arg_error{FLIP_E_ARG, err};
        FLIP_CHECK(hipMemcpyAsync(ls->net->pw, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice, c->stream));
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        unet_refresh_bf16(ls->net);
        const long C = c->C;
        dmalloc(&ls->dinv0, C);
        dmalloc(&ls->z2, C);
        dmalloc(&ls->z4, C);
        dmalloc(&ls->res0, C);
        dmalloc(&ls->rn, C);
        dmalloc(&ls->sigma, 4);
        dmalloc(&ls->part, 4 * GMAX);
        dmalloc(&ls->fs, 8);
        dmalloc(&ls->fi, 4);
        dmalloc(&ls->upart, 2 * (long)nb(8 * ls->lv[ls->K].C));
        FLIP_CHECK(hipHostMalloc((void **)&ls->h_flag, sizeof(int32_t) * 2, hipHostMallocMapped));
        FLIP_CHECK(hipHostGetDevicePointer((void **)&ls->d_flag, ls->h_flag, 0));
        FLIP_CHECK(hipEventCreateWithFlags(&ls->ev[0], hipEventDisableTiming));
        FLIP_CHECK(hipEventCreateWithFlags(&ls->ev[1], hipEventDisableTiming));
    } catch (...) {
        learned_free(ls);
        throw;
    }
    c->learned = ls;
    return ls;
}

// z = M(r) into the (z, s) buffer of parity (fi[FI_K] + 1) & 1, partials (z.r,
// z.q) in pz (q may be null).  iter == false: the first V-cycle, on c->r.
// iter == true: flexible-CG iteration k's update fused with the
// pre-smoothing (k_fcg_b_pre0, r' into ls->rn), the stop test in the level-0
// residual, and r' copied back into c->r by the last kernel.
static void vcycle(FlipCtx *c, LearnedSolver *ls, int G, double h2, int subcell, const float *q, double *pz,
                   const double *pa, double *pb, bool iter)
{
    const int n = c->n, K = ls->K;
    int32_t *fi = ls->fi;
    const unsigned G0 = (unsigned)std::min<long>(nb(c->C), 4096);  // level-0 smoothing grid (grid-stride)
    const float *r = iter ? ls->rn : c->r;
    if (iter)
        launch_timed(c, "fcg_b", k_fcg_b_pre0, G, TB, c->iflags, fi, c->fl, c->flmeta, n, c->L, subcell, h2,
                     c->press, c->r, ls->rn, c->zs[0], c->zs[1], c->q, ls->dinv0, ls->z2, pa, ls->fs, pb, G);
    else
        launch_timed(c, "ml_pre0", k_ml_pre0, G, TB, c->iflags, fi, c->fl, c->flmeta, n, c->L, subcell, h2, c->r,
                     ls->dinv0, ls->z2);
    // down: restrict level l-1's residual into b_l, smooth level l
    for (int l = 1; l <= K; ++l) {
        Level &L = ls->lv[l];
        if (l == 1) {
            launch_timed(c, "ml_resid0", k_ml_resid0, G0, TBL, c->iflags, fi, c->fl, c->flmeta, n, c->L, subcell, h2,
                         r, ls->z2, ls->res0, pb, ls->fs, G, FCG_MAX, iter ? 1 : 0);
            launch_timed(c, "ml_sum8", k_ml_sum8, nb(L.C), TBL, fi, L.R, n, ls->res0, L.b,
                               l == K ? L.A : nullptr, ls->upart);
        } else {
            Level &F = ls->lv[l - 1];
            launch_timed(c, "ml_resid", k_ml_resid_sum8, nb(8 * L.C), TBL, fi, L.R, F.R, F.A, F.meta, F.L, subcell,
                         F.h2, F.b, F.x, L.b, l == K ? L.A : nullptr, ls->upart);
        }
        if (l < K)
            launch_timed(c, "ml_pre", k_ml_pre, nb(L.C), TBL, fi, L.R, L.A, L.meta, L.L, subcell, L.h2, L.b,
                               L.dinv, L.x);
    }
    Level &KL = ls->lv[K];
    t_begin(c, "unet");
    // the network's scaling partials: k_ml_resid_sum8's (one per 32 coarse
    // cells) below two levels, else k_ml_sum8's
    const int nparts = K >= 2 ? (int)nb(8 * KL.C) : unet_partials_count(ls->net);
    unet_infer_layers(ls->net, KL.A, KL.L, KL.b, ls->upart, nparts, fi + FI_DONE, KL.x,
                      ls->sigma);
    t_end(c);
    // up: correct level l with level l+1, smooth
    for (int l = K - 1; l >= 1; --l) {
        Level &L = ls->lv[l];
        launch_timed(c, "ml_post_a", k_ml_post_a, nb(L.C), TBL, fi, L.R, L.A, L.meta, L.L, subcell, L.h2, L.b,
                           L.dinv, L.x, ls->lv[l + 1].x, L.t);
        launch_timed(c, "ml_post_b", k_ml_post_b, nb(L.C), TBL, fi, L.R, L.A, L.meta, L.L, subcell, L.h2, L.b,
                           L.dinv, L.t, L.x);
    }
    launch_timed(c, "ml_post0_a", k_ml_post0_a, G0, TBL, c->iflags, fi, c->fl, c->flmeta, n, c->L, subcell, h2, r,
                 ls->dinv0, ls->z2, ls->lv[1].x, ls->lv[1].R, ls->z4);
    launch_timed(c, "ml_post0_b", k_ml_post0_b, G, TB, c->iflags, fi, c->fl, c->flmeta, n, c->L, subcell, h2, r,
                 ls->dinv0, ls->z4, q, c->zs[0], c->zs[1], pz, iter ? ls->d_flag : nullptr, iter ? c->r : nullptr);
}

// Flexible CG with the multilevel preconditioner from the current x (press)
// and r = b - A x (c->r, from the PCG init); the FLUID list, P and iflags
// (I_NFL) are current.  Returns the iterations run; *converged when the
// recursive residual met eps.
int learned_fcg(FlipCtx *c, int G, double h2, int subcell, double eps, bool *converged)
{
    LearnedSolver *ls = learned_get(c);
    const int n = c->n, K = ls->K;
    double *pa = ls->part, *pb = ls->part + GMAX, *pz = ls->part + 2 * GMAX;
    t_begin(c, "learned_setup");
    // hierarchy of this system
    hipLaunchKernelGGL(k_ml_zero, nb(c->C), TBL, 0, c->stream, c->C, ls->res0);
    hipLaunchKernelGGL(k_ml_setup0, G, TBL, 0, c->stream, c->iflags, c->fl, c->P, h2, ls->dinv0);
    for (int l = 1; l <= K; ++l) {
        Level &L = ls->lv[l];
        const int8_t *Af = l == 1 ? c->A : ls->lv[l - 1].A;
        const float *Lf = l == 1 ? c->L : ls->lv[l - 1].L;
        const int Rf = l == 1 ? n : ls->lv[l - 1].R;
        hipLaunchKernelGGL(k_ml_coarsen, nb(L.C), TBL, 0, c->stream, L.R, Rf, Af, Lf, L.A, L.L);
        hipLaunchKernelGGL(k_ml_meta, nb(L.C), TBL, 0, c->stream, L.R, L.A, L.L, subcell, L.h2, L.meta, L.dinv);
    }
    hipLaunchKernelGGL(k_fcg_begin, 1, 64, 0, c->stream, c->iflags, c->scal, ls->fs, ls->fi);
    t_end(c);
    vcycle(c, ls, G, h2, subcell, nullptr, pz, pa, pb, false);  // z_0 into B_1.x

    const bool direct = c->profiling || c->direct_launch;
    const long key = (long)G * 2 + subcell;
    if (!ls->graph || ls->graph_key != key || direct) {
        if (ls->graph) {
            (void)hipGraphExecDestroy(ls->graph);
            ls->graph = nullptr;
        }
    }
    auto iteration = [&]() {
        launch_timed(c, "fcg_a", k_fcg_a, G, TB, c->iflags, ls->fi, c->fl, c->flmeta, n, c->L, subcell, h2, c->zs[0],
                     c->zs[1], c->q, pz, ls->fs, pa, G);
        // the update and the pre-smoothing, then the stop test (in the level-0
        // residual); the V-cycle's last kernel writes the flag
        vcycle(c, ls, G, h2, subcell, c->q, pz, pa, pb, true);
    };
    if (!direct && !ls->graph) {
        hipGraph_t g;
        FLIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        iteration();
        FLIP_CHECK(hipStreamEndCapture(c->stream, &g));
        FLIP_CHECK(hipGraphInstantiate(&ls->graph, g, nullptr, nullptr, 0));
        FLIP_CHECK(hipGraphDestroy(g));
        ls->graph_key = key;
    }
    // launch iteration k, then wait for iteration k-1 and read the stop flag
    // (a set flag stays set, so reading a later iteration's write is fine)
    volatile int32_t *flag = ls->h_flag;
    flag[0] = 0;
    for (int k = 0; k < FCG_MAX; ++k) {
        if (direct) {
            iteration();
        } else {
            FLIP_CHECK(hipGraphLaunch(ls->graph, c->stream));
        }
        FLIP_CHECK(hipEventRecord(ls->ev[k & 1], c->stream));
        if (k > 0) {
            FLIP_CHECK(hipEventSynchronize(ls->ev[(k - 1) & 1]));
            if (flag[0]) break;
        }
    }
    // the iteration count and outcome are read after the solve's next wait
    // (learned_fcg_iterations)
    *converged = false;
    (void)eps;
    return 0;
}

// flexible-CG iterations of the last learned solve, from the stop flag's
// mapped record (valid once the stream has passed its last iteration)
int learned_fcg_iterations(FlipCtx *c)
{
    return c->learned ? c->learned->h_flag[1] : 0;
}
