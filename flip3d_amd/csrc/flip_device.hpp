// flip_device.hpp - device arithmetic shared by the FLIP kernels (gfx950).
//
// Every helper reproduces the reference's rounding exactly: the library is
// built with -ffp-contract=off (no FMA fusion), float division and sqrtf are
// IEEE correctly rounded (HIP default), and double temporaries appear exactly
// where the reference's double literals put them (src/common.h:10 FLOAT=float).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flipdev {

constexpr int AIR = 0;
constexpr int FLUID = 1;
constexpr int WALL = 2;

// Particle metadata word: creation index << 2 | type.
__device__ __forceinline__ int meta_type(int32_t m) { return m & 3; }
__device__ __forceinline__ int32_t meta_id(int32_t m) { return (int32_t)((uint32_t)m >> 2); }

__device__ __forceinline__ long cidx(int n, int i, int j, int k) { return ((long)i * n + j) * n + k; }
// face grids: u0 (n+1,n,n), u1 (n,n+1,n), u2 (n,n,n+1)
__device__ __forceinline__ long fidx0(int n, int i, int j, int k) { return ((long)i * n + j) * n + k; }
__device__ __forceinline__ long fidx1(int n, int i, int j, int k) { return ((long)i * (n + 1) + j) * n + k; }
__device__ __forceinline__ long fidx2(int n, int i, int j, int k) { return ((long)i * n + j) * (n + 1) + k; }

// kernel.cpp:4-6 smooth_kernel: (float)fmax(1.0 - (double)(r2/(h*h)), 0.0).
// For float q >= 0 the double 1.0 - q is exact whenever the float result can
// differ from 1.0f, so one float subtraction rounds to the same value.
__device__ __forceinline__ float smooth_kernel(float r2, float h)
{
    float q = r2 / (h * h);
    float v = 1.0f - q;
    return v > 0.0f ? v : 0.0f;
}

// kernel.cpp:8-10 sharp_kernel: (float)fmax((double)(h*h)/fmax((double)r2,1e-5) - 1.0, 0.0).
// r2 >= h*h gives exactly 0 (quotient <= 1); NaN r2 takes the full path like fmax does.
__device__ __forceinline__ float sharp_kernel_hh(float r2, float hh)
{
    if (r2 >= hh) return 0.0f;
    double v = (double)hh / fmax((double)r2, 1.0e-5) - 1.0;
    return (float)fmax(v, 0.0);
}

// a / b correctly rounded for positive normal a and b far from the ends of
// the double range (here a is a positive normal float h*h and b >= 1e-5, so
// a, b, 1/b and a/b all lie within 2^-260..2^150).  This is the IEEE division sequence
// the compiler emits for '/' (v_rcp_f64, two Newton steps on the
// reciprocal, one remainder correction) without v_div_scale and
// v_div_fixup: those rescale operands only near overflow, underflow or
// denormals and patch zero, infinite and NaN operands, and return their
// input unchanged on this range, while v_div_fmas is then a plain fma.
// Being plain VGPR arithmetic (no VCC), several of these can be in flight
// in one wave at once.
__device__ __forceinline__ double div_pos(double a, double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    return __builtin_fma(r, y, q);
}

// sharp_kernel_hh for r2 < hh; callers discard the value for other r2
// (NaN and infinite r2 included), so no early return is needed
__device__ __forceinline__ float sharp_kernel_in(float r2, float hh)
{
    const double v = div_pos((double)hh, fmax((double)r2, 1.0e-5)) - 1.0;
    return (float)fmax(v, 0.0);
}

// smooth_kernel when h*h is a power of two and inv_hh = 1 / (h*h): the
// quotient r2 / (h*h) is then exactly r2 * inv_hh, so fma(-r2, inv_hh, 1)
// rounds 1 - r2/(h*h) once, as the float subtraction does.  The value is at
// most 1, so clamping it to [0, 1] (the fma's clamp bit) is the max with 0,
// NaN included.
__device__ __forceinline__ float smooth_kernel_pow2(float r2, float inv_hh)
{
    const float v = __builtin_fmaf(-r2, inv_hh, 1.0f);
    return fminf(fmaxf(v, 0.0f), 1.0f);
}

// utility.cpp:20-26 length2: float (a*a + b*b) + c*c of a - b
__device__ __forceinline__ float length2(float ax, float ay, float az, float bx, float by, float bz)
{
    float x = ax - bx, y = ay - by, z = az - bz;
    return x * x + y * y + z * z;
}

// glibc 2.35 hypotf: sqrt((double)x*x + (double)y*y) rounded to float, with
// INFINITY when either operand is infinite.
__device__ __forceinline__ float hypotf_glibc(float x, float y)
{
    if (__builtin_isinf(x) || __builtin_isinf(y)) return __builtin_inff();
    double s = (double)x * (double)x + (double)y * (double)y;
    return (float)__dsqrt_rn(s);
}

// utility.cpp:16-18 length: hypotf(hypotf(dx,dy),dz)
__device__ __forceinline__ float length3(float ax, float ay, float az, float bx, float by, float bz)
{
    return hypotf_glibc(hypotf_glibc(ax - bx, ay - by), az - bz);
}

// (i, j, k) of the flat index f = (i * nj + j) * nk + k.  Grids have n <= 1024,
// so every cell or face index fits in 32 bits; two 32-bit divisions replace
// the three 64-bit divisions and remainders of the plain decomposition.
struct Split3 {
    int i, j, k;
};
__device__ __forceinline__ Split3 split3(long f, int nj, int nk)
{
    const uint32_t u = (uint32_t)f, q = u / (uint32_t)nk, i = q / (uint32_t)nj;
    return Split3{(int)i, (int)(q - i * (uint32_t)nj), (int)(u - q * (uint32_t)nk)};
}

// sorter.cpp:35-37 cell key (int)fmax(0, fmin(gn-1, gn*p)); NaN maps to gn-1.
__device__ __forceinline__ int cell_of(int gn, float p)
{
    float v = (float)gn * p;
    v = fminf(v, (float)(gn - 1));
    v = fmaxf(v, 0.0f);
    return (int)v;
}

// flip3D.cpp:82,752,903: (int)fmin(N-1, fmax(0, p*N)); NaN maps to 0.
__device__ __forceinline__ int cell_of_b(int gn, float p)
{
    float v = p * (float)gn;
    v = fmaxf(v, 0.0f);
    v = fminf(v, (float)(gn - 1));
    return (int)v;
}

// Visit src[b..e) in order, four loads in flight: body(q, index) is called
// for every element strictly in index order, so order-sensitive float sums
// are unchanged; the early loads only hide memory latency.
template <class T, class F>
__device__ __forceinline__ void for_range4(const T *__restrict__ src, int b, int e, F &&body)
{
    int a = b;
    for (; a + 3 < e; a += 4) {
        T q0 = src[a], q1 = src[a + 1], q2 = src[a + 2], q3 = src[a + 3];
        body(q0, a);
        body(q1, a + 1);
        body(q2, a + 2);
        body(q3, a + 3);
    }
    for (; a < e; ++a) body(src[a], a);
}

// for_range4 with eight loads in flight (then four, then one at a time)
template <class T, class F>
__device__ __forceinline__ void for_range8(const T *__restrict__ src, int b, int e, F &&body)
{
    int a = b;
    for (; a + 7 < e; a += 8) {
        T q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) q[u] = src[a + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) body(q[u], a + u);
    }
    if (a + 3 < e) {
        T q0 = src[a], q1 = src[a + 1], q2 = src[a + 2], q3 = src[a + 3];
        body(q0, a);
        body(q1, a + 1);
        body(q2, a + 2);
        body(q3, a + 3);
        a += 4;
    }
    for (; a < e; ++a) body(src[a], a);
}

// ------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Deterministic block sum (fixed tree order) of one double per thread.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double *lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < BLOCK / 64; ++i) r += lds[i];
    }
    __syncthreads();
    return r;
}

}  // namespace flipdev
