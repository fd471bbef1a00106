// k_sort.hip - cell binning (sorter::sort, src/sorter.cpp:22-40) as a counting
// sort, plus the exclusive scan used by the sort and by ordered compactions.
//
// The reference keeps, per cell, the particles in particle-vector (creation)
// order.  Here particles are stored cell-sorted; every sort
//   1. computes each particle's cell key and claims a slot with an atomic
//      counter (arbitrary order inside the cell),
//   2. scans the per-cell counts into cell_start,
//   3. scatters particle indices to cell_start[key] + slot,
//   4. restores creation order inside each cell (cells hold ~8 particles,
//      ranked in registers; larger buckets are ranked by whole blocks), and
//   5. permutes all particle arrays into the new order (coalesced gather).
// The result is bit-identical to the reference's buckets.
#include <mutex>

#include "flip_device.hpp"
#include "flip_internal.hpp"

using namespace flipdev;

namespace {

constexpr int SB = 256;     // threads per block
constexpr int SI = 8;       // items per thread in the scan
constexpr int TILE = SB * SI;

__global__ void k_fill_i32(int32_t *a, long n, int32_t v)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of per-thread values over the block; returns the prefix and
// writes the block total to *total.
__device__ __forceinline__ int block_excl_scan(int v, int *lds, int *total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = wave_incl_scan(v);
    if (lane == 63) lds[w] = incl;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < SB / 64; ++i) {
        if (i < w) off += lds[i];
        tot += lds[i];
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

// gate (optional): the scan does nothing when *gate == 0
__global__ __launch_bounds__(SB) void k_scan_reduce(const int32_t *__restrict__ in, long n, int32_t *bsum,
                                                    const int32_t *__restrict__ gate)
{
    if (gate && *gate == 0) return;
    __shared__ int lds[SB / 64];
    long base = (long)blockIdx.x * TILE;
    int s = 0;
#pragma unroll
    for (int t = 0; t < SI; ++t) {
        long i = base + (long)t * SB + threadIdx.x;
        if (i < n) s += in[i];
    }
    int tot;
    (void)block_excl_scan(s, lds, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// single block: exclusive scan of nb block sums in place; out_total <- sum
__global__ __launch_bounds__(SB) void k_scan_top(int32_t *bsum, int nb, int32_t *out_total,
                                                 const int32_t *__restrict__ gate)
{
    if (gate && *gate == 0) return;
    __shared__ int lds[SB / 64];
    int carry = 0;
    for (int base = 0; base < nb; base += SB) {
        int i = base + threadIdx.x;
        int v = i < nb ? bsum[i] : 0;
        int tot;
        int ex = block_excl_scan(v, lds, &tot);
        if (i < nb) bsum[i] = ex + carry;
        carry += tot;
    }
    if (threadIdx.x == 0) *out_total = carry;
}

__global__ __launch_bounds__(SB) void k_scan_apply(const int32_t *__restrict__ in, long n,
                                                   const int32_t *__restrict__ bsum, int32_t *out,
                                                   const int32_t *__restrict__ gate)
{
    if (gate && *gate == 0) return;
    __shared__ int lds[SB / 64];
    long base = (long)blockIdx.x * TILE + (long)threadIdx.x * SI;
    int v[SI];
    int s = 0;
#pragma unroll
    for (int t = 0; t < SI; ++t) {
        v[t] = base + t < n ? in[base + t] : 0;
        s += v[t];
    }
    int tot;
    int ex = block_excl_scan(s, lds, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int t = 0; t < SI; ++t) {
        if (base + t < n) out[base + t] = ex;
        ex += v[t];
    }
}

// 1. cell key and slot claim.  Storage is sorted by the previous step's cells,
// so neighbouring lanes mostly share a key: each run of equal keys within a
// wave claims its slots with one atomic (slots in lane order within the run).
// Positions are float4 records (x, y, z, m) or packed xyz triples.  Thread 0
// also clears the count of large buckets for step 4.
template <bool XYZ>
__global__ void k_sort_count(const float *__restrict__ pos, int64_t np, int n, int32_t *key, int32_t *slot,
                             int32_t *count, int32_t *big)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) big[0] = 0;
    const int lane = threadIdx.x & 63;
    int32_t c = -1;
    if (i < np) {
        float x, y, z;
        if (XYZ) {
            x = pos[3 * i];
            y = pos[3 * i + 1];
            z = pos[3 * i + 2];
        } else {
            float4 p = reinterpret_cast<const float4 *>(pos)[i];
            x = p.x;
            y = p.y;
            z = p.z;
        }
        c = (int32_t)cidx(n, cell_of(n, x), cell_of(n, y), cell_of(n, z));
    }
    int32_t prev = __shfl_up(c, 1, 64);
    bool head = lane == 0 || prev != c;
    unsigned long long heads = __ballot(head);
    unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    int my_head = 63 - __clzll(heads & upto);
    unsigned long long after = my_head == 63 ? 0ull : (heads & ~((2ull << my_head) - 1));
    int next_head = after ? __ffsll((long long)after) - 1 : 64;
    int base = 0;
    if (head && c >= 0) base = atomicAdd(&count[c], next_head - my_head);
    base = __shfl(base, my_head, 64);
    if (i < np) {
        key[i] = c;
        slot[i] = base + (lane - my_head);
    }
}

// 3. scatter source indices to their bucket slots
__global__ void k_sort_scatter(int64_t np, const int32_t *__restrict__ key, const int32_t *__restrict__ slot,
                               const int32_t *__restrict__ start, const int32_t *__restrict__ meta, int32_t *perm,
                               int32_t *pid)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    int32_t d = start[key[i]] + slot[i];
    perm[d] = (int32_t)i;
    pid[d] = meta ? meta_id(meta[i]) : (int32_t)i;  // without records the input index is the creation index
}

// 4. creation order inside every bucket.  A bucket of at most SO_REG
// particles is ranked in registers (ids are unique, so every rank is
// distinct) and skipped when already in order, which the atomic slot
// claim of step 1 mostly leaves it.  A larger bucket is copied to (tpid,
// tperm) (the key/slot arrays, free after the scatter) and listed in big
// (big[0] = count, entries from big[1]); k_sort_order_big ranks it with
// whole blocks.
constexpr int SO_REG = 16;

__global__ void k_sort_order(long C, const int32_t *__restrict__ start, int32_t *perm, int32_t *pid, int32_t *tperm,
                             int32_t *tpid, int32_t *big)
{
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    int b = start[c], e = start[c + 1];
    const int m = e - b;
    if (m <= 1) return;
    if (m <= SO_REG) {
        int key[SO_REG], val[SO_REG];
#pragma unroll
        for (int t = 0; t < SO_REG; ++t) {
            key[t] = t < m ? pid[b + t] : 0x7fffffff;
            val[t] = t < m ? perm[b + t] : 0;
        }
        bool ordered = true;
#pragma unroll
        for (int t = 1; t < SO_REG; ++t) ordered = ordered && key[t - 1] < key[t];
        if (ordered) return;
#pragma unroll
        for (int t = 0; t < SO_REG; ++t) {
            if (t < m) {
                int r = 0;
#pragma unroll
                for (int u = 0; u < SO_REG; ++u) r += key[u] < key[t];
                pid[b + r] = key[t];
                perm[b + r] = val[t];
            }
        }
        return;
    }
    for (int a = b; a < e; ++a) {
        tpid[a] = pid[a];
        tperm[a] = perm[a];
    }
    int slot = atomicAdd(&big[0], 1);
    big[1 + slot] = (int32_t)c;
}

// Large buckets: chunk ch of 256 particles of listed bucket q goes to the
// block with (ch + q) % gridDim.x == blockIdx.x, so one huge bucket spreads
// over the whole grid.  A particle's rank is the number of smaller ids in
// its bucket, counted against the bucket staged through LDS 256 ids at a
// time (ids are unique, so the ranks are a permutation).
constexpr int OB = 256;

__global__ __launch_bounds__(OB) void k_sort_order_big(const int32_t *__restrict__ start,
                                                       const int32_t *__restrict__ big,
                                                       const int32_t *__restrict__ tperm,
                                                       const int32_t *__restrict__ tpid, int32_t *perm, int32_t *pid)
{
    __shared__ int32_t ids[OB];
    const int nbig = big[0];
    for (int q = 0; q < nbig; ++q) {
        const int c = big[1 + q];
        const int b = start[c], e = start[c + 1];
        const int chunks = (e - b + OB - 1) / OB;
        int ch = (int)blockIdx.x - q % (int)gridDim.x;
        if (ch < 0) ch += gridDim.x;
        for (; ch < chunks; ch += gridDim.x) {
            const int a = b + ch * OB + (int)threadIdx.x;
            const bool mine = a < e;
            const int32_t key = mine ? tpid[a] : 0;
            int rank = 0;
            for (int t0 = b; t0 < e; t0 += OB) {
                __syncthreads();
                if (t0 + (int)threadIdx.x < e) ids[threadIdx.x] = tpid[t0 + threadIdx.x];
                __syncthreads();
                const int lim = min(OB, e - t0);
                for (int u = 0; u < lim; ++u) rank += ids[u] < key;
            }
            if (mine) {
                pid[b + rank] = key;
                perm[b + rank] = tperm[a];
            }
        }
    }
}

// 5. permute particle state.  With spart, also the block's largest FLUID
// |u|^2 (double, for the CFL time step) into spart[blockIdx.x].
constexpr int GB = 256;  // block of the gather
__global__ __launch_bounds__(GB) void k_sort_gather(int64_t np, const int32_t *__restrict__ perm,
                                                    const float4 *__restrict__ pin, const float4 *__restrict__ vin,
                                                    const int32_t *__restrict__ min, float4 *pout, float4 *vout,
                                                    int32_t *mout, int32_t *inv, float *xyz3, double *spart)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = j < np;
    const int32_t s = ok ? perm[j] : 0;
    const float4 p = ok ? pin[s] : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    const float4 v = ok ? vin[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // the packed window record: a WALL particle (sign bit of m) sits at
    // +inf, so every squared distance to it is +inf
    const bool wall = __builtin_signbit(p.w);
    if (spart) {
        __shared__ double lds[GB / 64];
        double m = ok && !wall ? (double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z : 0.0;
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = 0.0;
            for (int w = 0; w < GB / 64; ++w) r = fmax(r, lds[w]);
            spart[blockIdx.x] = r;
        }
    }
    if (!ok) return;
    pout[j] = p;
    const float inf = __builtin_inff();
    xyz3[3 * j + 0] = wall ? inf : p.x;
    xyz3[3 * j + 1] = wall ? inf : p.y;
    xyz3[3 * j + 2] = wall ? inf : p.z;
    vout[j] = v;
    int32_t m = min[s];
    mout[j] = m;
    inv[meta_id(m)] = (int32_t)j;
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

namespace {
void scan_on(hipStream_t st, const int32_t *in, int32_t *out, long n, int32_t *blocks,
             const int32_t *gate = nullptr)
{
    if (n <= 0) {
        hipLaunchKernelGGL(k_fill_i32, 1, 1, 0, st, out, 1L, 0);
        return;
    }
    int nb = (int)((n + TILE - 1) / TILE);
    hipLaunchKernelGGL(k_scan_reduce, nb, SB, 0, st, in, n, blocks, gate);
    hipLaunchKernelGGL(k_scan_top, 1, SB, 0, st, blocks, nb, out + n, gate);
    hipLaunchKernelGGL(k_scan_apply, nb, SB, 0, st, in, n, blocks, out, gate);
}
}  // namespace

void k_launch_scan(FlipCtx *c, const int32_t *in, int32_t *out, long n, int32_t *blocks, const int32_t *gate)
{
    scan_on(c->stream, in, out, n, blocks ? blocks : c->scan_blocks, gate);
}

namespace {

// the launch sequence of steps 1-4 on caller-provided buffers
void launch_bucket_sort(hipStream_t st, bool xyz, const float *pos, int64_t np, int n, long C, int32_t *count,
                        int32_t *start, int32_t *scan_blocks, int32_t *key, int32_t *slot, int32_t *perm,
                        int32_t *pid, const int32_t *meta, int32_t *big)
{
    hipLaunchKernelGGL(k_fill_i32, nblk(C, 256), 256, 0, st, count, C, 0);
    if (np > 0) {
        if (xyz)
            hipLaunchKernelGGL(k_sort_count<true>, nblk(np, 256), 256, 0, st, pos, np, n, key, slot, count, big);
        else
            hipLaunchKernelGGL(k_sort_count<false>, nblk(np, 256), 256, 0, st, pos, np, n, key, slot, count, big);
    }
    scan_on(st, count, start, C, scan_blocks);
    if (np == 0) return;
    hipLaunchKernelGGL(k_sort_scatter, nblk(np, 256), 256, 0, st, np, key, slot, start, meta, perm, pid);
    hipLaunchKernelGGL(k_sort_order, nblk(C, 256), 256, 0, st, C, start, perm, pid, slot, key, big);
    hipLaunchKernelGGL(k_sort_order_big, 256, OB, 0, st, start, big, slot, key, perm, pid);
}

}  // namespace

void k_launch_sort(FlipCtx *c, double *speed_part)
{
    int64_t np = c->np;
    t_begin(c, "sort_count");
    hipLaunchKernelGGL(k_fill_i32, nblk(c->C, 256), 256, 0, c->stream, c->cell_count, c->C, 0);
    if (np > 0)
        hipLaunchKernelGGL(k_sort_count<false>, nblk(np, 256), 256, 0, c->stream, (const float *)c->P4(), np, c->n,
                           c->pkey, c->pslot, c->cell_count, c->sort_big);
    t_end(c);
    t_begin(c, "sort_scan");
    k_launch_scan(c, c->cell_count, c->cell_start, c->C);
    t_end(c);
    c->bucketed = true;
    if (np == 0) return;
    t_begin(c, "sort_scatter");
    int32_t *pid = c->deep;  // scratch sized max(C, np) + 1: creation index per bucket slot
    hipLaunchKernelGGL(k_sort_scatter, nblk(np, 256), 256, 0, c->stream, np, c->pkey, c->pslot, c->cell_start,
                       c->M(), c->perm, pid);
    hipLaunchKernelGGL(k_sort_order, nblk(c->C, 256), 256, 0, c->stream, c->C, c->cell_start, c->perm, pid, c->pslot,
                       c->pkey, c->sort_big);
    hipLaunchKernelGGL(k_sort_order_big, 256, OB, 0, c->stream, c->cell_start, c->sort_big, c->pslot, c->pkey,
                       c->perm, pid);
    t_end(c);
    t_begin(c, "sort_gather");
    int nx = c->cur ^ 1;
    hipLaunchKernelGGL(k_sort_gather, nblk(np, GB), GB, 0, c->stream, np, c->perm, c->pos[c->cur],
                       c->vel[c->cur], c->meta[c->cur], c->pos[nx], c->vel[nx], c->meta[nx], c->inv, c->xyz3,
                       speed_part);
    c->cur = nx;
    t_end(c);
}

// ------------------------------------------------ sorting with no allocation
// The scratch of flip_gpu_sort_cells_noalloc: one fixed array per GPU (each
// device holds its own copy for the life of the process), laid out as
//   key[NA_NP] | slot[NA_NP] | start[NA_C + 1] | scan blocks | big list
// Host code never passes the array's name to a kernel: it is a host-side
// stand-in, not a device address.  Each device's address comes from
// hipGetSymbolAddress on the first call on that device.
namespace {
constexpr int64_t NA_NP = int64_t(1) << 20;                // particles
constexpr long NA_C = 1L << 20;                            // cells (n^3)
constexpr long NA_BLOCKS = NA_C / TILE + 8;                // scan block sums
constexpr long NA_BIG = NA_NP / (SO_REG + 1) + 2;          // big-bucket list (count + entries)
constexpr size_t NA_WORDS = 2 * (size_t)NA_NP + (NA_C + 1) + NA_BLOCKS + NA_BIG;

__device__ __align__(256) int32_t scratch[NA_WORDS];

constexpr int NA_MAX_DEV = 64;
std::mutex g_na_mu;
int32_t *g_na_ptr[NA_MAX_DEV] = {};
}  // namespace

extern "C" int flip_gpu_sort_cells_noalloc(const float *pos, int64_t np, int n, int32_t *counts, int32_t *ids,
                                           void *stream)
{
    if (n < 1 || np < 0 || (np > 0 && (!pos || !ids)) || !counts)
        return flip_internal_fail(FLIP_E_ARG, "sort_cells_noalloc: bad argument");
    if (np > NA_NP)
        return flip_internal_fail(FLIP_E_ARG, "sort_cells_noalloc: np exceeds the scratch limit of 1048576 particles");
    const long C = (long)n * n * n;
    if (n > 1024 || C > NA_C)
        return flip_internal_fail(FLIP_E_ARG, "sort_cells_noalloc: n^3 exceeds the scratch limit of 1048576 cells");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEV)
        return flip_internal_fail(FLIP_E_HIP, "sort_cells_noalloc: no current HIP device");
    int32_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_na_mu);
        if (!g_na_ptr[dev]) {
            void *p = nullptr;
            hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return flip_internal_fail(FLIP_E_HIP, "sort_cells_noalloc: hipGetSymbolAddress of the scratch failed");
            g_na_ptr[dev] = static_cast<int32_t *>(p);
        }
        base = g_na_ptr[dev];
    }
    int32_t *key = base, *slot = key + NA_NP, *start = slot + NA_NP, *blocks = start + (NA_C + 1),
            *big = blocks + NA_BLOCKS;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // ids doubles as the scatter's source-index and creation-index arrays:
    // the input index is the creation index, so both hold the same values
    launch_bucket_sort(st, true, pos, np, n, C, counts, start, blocks, key, slot, ids, ids, nullptr, big);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return flip_internal_fail(FLIP_E_HIP, hipGetErrorString(e));
    return FLIP_OK;
}
