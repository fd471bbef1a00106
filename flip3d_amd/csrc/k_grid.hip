// k_grid.hip - MAC-grid passes of project() and solve_picflip().
#include "flip_device.hpp"
#include "flip_internal.hpp"

using namespace flipdev;

namespace {

constexpr int TB = 256;
inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// flip3D.cpp:815-821 divergence on FLUID cells, left to right as written:
// ((((u0+ - u0-) + u1+) - u1-) + u2+) - u2-, then / h.  Non-FLUID cells keep
// their value (the reference's static div array).  press is cleared on
// non-FLUID cells: the solver's first x update (solver.cpp:329, op masks with
// FLUID) zeroes them, and they are never read before that.
__global__ void k_divergence(int n, const int8_t *__restrict__ A, const float *__restrict__ u0,
                             const float *__restrict__ u1, const float *__restrict__ u2, float h, float *div,
                             float *press, int clear_press, int negate)
{
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long C = (long)n * n * n;
    if (c >= C) return;
    const Split3 s3 = split3(c, n, n);
    const int i = s3.i, j = s3.j, k = s3.k;
    if (A[c] == FLUID) {
        float d = u0[fidx0(n, i + 1, j, k)] - u0[fidx0(n, i, j, k)];
        d = d + u1[fidx1(n, i, j + 1, k)];
        d = d - u1[fidx1(n, i, j, k)];
        d = d + u2[fidx2(n, i, j, k + 1)];
        d = d - u2[fidx2(n, i, j, k)];
        div[c] = negate ? -(d / h) : d / h;
    } else if (clear_press) {
        press[c] = 0.0f;
    }
}

__device__ __forceinline__ int wall_chk(int8_t a) { return a == WALL ? 1 : -1; }

// flip3D.cpp:833-867 ghost-fluid pressure gradient, then the second
// enforce_boundary (flip3D.cpp:1067) on the same face.
template <int COMP>
__device__ __forceinline__ void gradient_face(int n, const int8_t *__restrict__ A, const float *__restrict__ L,
                                              const float *__restrict__ press, float *u, float h, int subcell, long f)
{
    const int ni = n + (COMP == 0), nj = n + (COMP == 1), nk = n + (COMP == 2);
    if (f >= (long)ni * nj * nk) return;
    const Split3 s3 = split3(f, nj, nk);
    const int i = s3.i, j = s3.j, k = s3.k;
    int a0 = COMP == 0 ? i : (COMP == 1 ? j : k);
    if (a0 == 0 || a0 == n) {
        u[f] = 0.0f;  // enforce_boundary: domain faces
        return;
    }
    long cf = cidx(n, i, j, k);
    long cb = COMP == 0 ? cidx(n, i - 1, j, k) : (COMP == 1 ? cidx(n, i, j - 1, k) : cidx(n, i, j, k - 1));
    float Lf = L[cf], Lb = L[cb];
    float pf = press[cf], pb = press[cb];
    if (subcell && Lf * Lb < 0.0f) {
        pf = Lf < 0.0f ? press[cf] : (float)((double)Lf / fmin(1.0e-3, (double)Lb) * (double)press[cb]);
        pb = Lb < 0.0f ? press[cb] : (float)((double)Lb / fmin(1.0e-6, (double)Lf) * (double)press[cf]);
    }
    float v = u[f] - (pf - pb) / h;
    if (wall_chk(A[cf]) * wall_chk(A[cb]) < 0) v = 0.0f;
    u[f] = v;
}

// the three face components in one launch, nb blocks each (the components
// are independent)
__global__ void k_gradient3(int n, const int8_t *__restrict__ A, const float *__restrict__ L,
                            const float *__restrict__ press, float *u0, float *u1, float *u2, float h, int subcell,
                            unsigned nb)
{
    const unsigned b = blockIdx.x;
    const long t = threadIdx.x;
    if (b < nb)
        gradient_face<0>(n, A, L, press, u0, h, subcell, (long)b * blockDim.x + t);
    else if (b < 2 * nb)
        gradient_face<1>(n, A, L, press, u1, h, subcell, (long)(b - nb) * blockDim.x + t);
    else
        gradient_face<2>(n, A, L, press, u2, h, subcell, (long)(b - 2 * nb) * blockDim.x + t);
}

// flip3D.cpp:1011-1054 mark / wall_mark of a face of component n
template <int COMP>
__device__ __forceinline__ void face_marks(int n, const int8_t *__restrict__ A, int i, int j, int k, bool &mark,
                                           bool &wmark)
{
    int a0 = COMP == 0 ? i : (COMP == 1 ? j : k);
    bool hasb = a0 > 0, hasf = a0 < n;
    int8_t Ab = AIR, Af = AIR;
    if (hasb) Ab = A[COMP == 0 ? cidx(n, i - 1, j, k) : (COMP == 1 ? cidx(n, i, j - 1, k) : cidx(n, i, j, k - 1))];
    if (hasf) Af = A[cidx(n, i, j, k)];
    mark = (hasb && Ab == FLUID) || (hasf && Af == FLUID);
    wmark = (!hasb || Ab == WALL) && (!hasf || Af == WALL);
}

// extrapolate_velocity (flip3D.cpp:1031-1053) fused with subtract_grid
// (flip3D.cpp:985-997): writes the extrapolated face value when the face has
// no fluid cell and walls (or the domain edge) on both sides, then
// u_save = u - u_save.  Extrapolated faces read only marked faces, which are
// never written here.
template <int COMP>
__device__ __forceinline__ void extrapolate_face(int n, const int8_t *__restrict__ A, float *u, float *us,
                                                 float2 *ud, long f)
{
    const int ni = n + (COMP == 0), nj = n + (COMP == 1), nk = n + (COMP == 2);
    if (f >= (long)ni * nj * nk) return;
    const Split3 s3 = split3(f, nj, nk);
    const int i = s3.i, j = s3.j, k = s3.k;
    bool mark, wmark;
    face_marks<COMP>(n, A, i, j, k, mark, wmark);
    float v = u[f];
    if (!mark && wmark) {
        int wsum = 0;
        float sum = 0.0f;
        const int q[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            int qi = q[t][0], qj = q[t][1], qk = q[t][2];
            if (qi >= 0 && qi < ni && qj >= 0 && qj < nj && qk >= 0 && qk < nk) {
                bool m2, w2;
                face_marks<COMP>(n, A, qi, qj, qk, m2, w2);
                if (m2) {
                    wsum++;
                    sum += u[((long)qi * nj + qj) * nk + qk];
                }
            }
        }
        if (wsum) {
            v = sum / (float)wsum;
            u[f] = v;
        }
    }
    const float d = v - us[f];
    us[f] = d;
    ud[f] = make_float2(v, d);
}

// the three face components in one launch, nb blocks each (an extrapolated
// face reads only faces of its own component)
__global__ void k_extrapolate3(int n, const int8_t *__restrict__ A, float *u0, float *u1, float *u2, float *s0,
                               float *s1, float *s2, float2 *d0, float2 *d1, float2 *d2, unsigned nb)
{
    const unsigned b = blockIdx.x;
    const long t = threadIdx.x;
    if (b < nb)
        extrapolate_face<0>(n, A, u0, s0, d0, (long)b * blockDim.x + t);
    else if (b < 2 * nb)
        extrapolate_face<1>(n, A, u1, s1, d1, (long)(b - nb) * blockDim.x + t);
    else
        extrapolate_face<2>(n, A, u2, s2, d2, (long)(b - 2 * nb) * blockDim.x + t);
}

}  // namespace

void k_launch_divergence(FlipCtx *c, bool negate)
{
    t_begin(c, "divergence");
    hipLaunchKernelGGL(k_divergence, nblk(c->C, TB), TB, 0, c->stream, c->n, c->A, c->ug[0], c->ug[1], c->ug[2],
                       (float)(1.0 / c->n), c->div, c->press, 0, negate ? 1 : 0);
    t_end(c);
}

void k_launch_gradient(FlipCtx *c)
{
    t_begin(c, "gradient");
    float h = (float)(1.0 / c->n);
    unsigned nb = nblk(c->F, TB);
    hipLaunchKernelGGL(k_gradient3, 3 * nb, TB, 0, c->stream, c->n, c->A, c->L, c->press, c->ug[0], c->ug[1],
                       c->ug[2], h, c->prm.subcell, nb);
    t_end(c);
}

void k_launch_extrapolate(FlipCtx *c)
{
    t_begin(c, "extrapolate");
    unsigned nb = nblk(c->F, TB);
    hipLaunchKernelGGL(k_extrapolate3, 3 * nb, TB, 0, c->stream, c->n, c->A, c->ug[0], c->ug[1], c->ug[2], c->us[0],
                       c->us[1], c->us[2], c->ud[0], c->ud[1], c->ud[2], nb);
    t_end(c);
}
