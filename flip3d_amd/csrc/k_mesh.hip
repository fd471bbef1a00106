// k_mesh.hip - the surface mesher on the GPU: mesher::generateMesh
// (src/mesher.cpp:137-195) with implicit::implicit_func (implicit.cpp:14-38)
// and CIsoSurface<float> (rchandra/CIsoSurface.cpp:331-777), producing the
// reference's vertices, normals and faces bit for bit.
//
// The reference builds the surface cell by cell into per-cell maps and
// vectors, merges them in x-major cell order and renumbers the vertices by
// edge id.  Here every part is a flat data-parallel pass over the lattice
// with that numbering made explicit:
//   - vertex v is the v-th cut lattice edge in edge-id order (3 * point id +
//     axis, points x fastest), found by a scan of per-edge cut flags; its
//     position is interpolated in the direction the reference's owning cell
//     uses for that edge;
//   - triangle t is the t-th triangle in the merge order (cells x-major, z
//     fastest; the case table's order inside a cell), found by a scan of
//     per-cell counts;
//   - a vertex's normal sum and its smoothing neighbour list run over the
//     triangles of the (at most four) cells around its edge, in triangle
//     order, which is the order the reference's serial loops add them in.
// The implicit field reads the particles through the current buckets
// (cell_start and the storage order of the last sort), as the reference's
// mesher reads its sorter.
#include "flip_device.hpp"
#include "flip_internal.hpp"
#include "mc_table.h"

#include <vector>

using namespace flipdev;

namespace {

constexpr int TB = 256;
inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

__constant__ char d_mc_cases[256][16] = {MC_CASE_LIST};

__device__ __forceinline__ int hexval(char ch) { return ch <= '9' ? ch - '0' : ch - 'a' + 10; }

// Bounding box of the current positions of each bucket's particles (the
// buckets may be older than the positions: after a step they are the
// corrector's, which moves particles afterwards).  Empty: lo = +inf.
__global__ void k_mesh_cell_boxes(long C, const int32_t *__restrict__ cs, const float4 *__restrict__ pos,
                                  float4 *lo, float4 *hi)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const float inf = __builtin_inff();
    float4 l = make_float4(inf, inf, inf, 0.0f), h = make_float4(-inf, -inf, -inf, 0.0f);
    for (int a = cs[c]; a < cs[c + 1]; ++a) {
        const float4 q = pos[a];
        l.x = fminf(l.x, q.x);
        l.y = fminf(l.y, q.y);
        l.z = fminf(l.z, q.z);
        h.x = fmaxf(h.x, q.x);
        h.y = fmaxf(h.y, q.y);
        h.z = fmaxf(h.z, q.z);
    }
    lo[c] = l;
    hi[c] = h;
}

// implicit.cpp:14-38 at lattice point (i, j, k) * 1/(mg-1): the distance to
// the nearest FLUID particle of the 5^3-cell window, capped at 8 DENSITY/gn,
// less DENSITY/gn; 4.5 DENSITY/gn when a WALL particle lies within
// DENSITY/gn.  The result is a minimum plus an existence test, so the
// window may be visited in any order and a cell may be skipped when the
// box of its particles is provably no nearer than both the current minimum
// and the wall radius:
//   - the 2x2x2 cells nearest the point go first, to find a small minimum
//     early;
//   - the other cells follow, pruned by their particle boxes (squared box
//     distance above the bound by a 1e-4 margin, which covers the few ulp
//     between float and exact distances);
//   - inside a cell the exact distance (glibc hypotf twice) is formed only
//     when the float squared distance is within 1e-4 of the minimum (or of
//     the wall radius).
// mesher.cpp:141-152 then clamps the boundary and stores -value, x fastest.
__global__ __launch_bounds__(TB) void k_mesh_field(int mg, int n, const float4 *__restrict__ pos,
                                                   const int32_t *__restrict__ cs, const float4 *__restrict__ blo,
                                                   const float4 *__restrict__ bhi, float density, float *f)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long m3 = (long)mg * mg * mg;
    if (t >= m3) return;
    // lanes run along k, so they share window rows
    const int k = (int)(t % mg), j = (int)((t / mg) % mg), i = (int)(t / ((long)mg * mg));
    const double h = 1.0 / (double)(mg - 1);
    const float px = (float)(i * h), py = (float)(j * h), pz = (float)(k * h);
    const int ci = cell_of_b(n, px), cj = cell_of_b(n, py), ck = cell_of_b(n, pz);
    const float gn = (float)n;
    const float dw = density / (float)n;
    const double dw2 = (double)dw * (double)dw * (1.0 + 1e-4);
    double phi = 8.0 * (double)density / n;
    double phi2 = phi * phi * (1.0 + 1e-4);
    bool wall = false;
    auto scan = [&](long c0, long c1) {
        const int b = cs[c0], e = cs[c1 + 1];
        for (int a = b; a < e; ++a) {
            const float4 q = pos[a];
            const double r2 = (double)length2(q.x, q.y, q.z, px, py, pz);
            if (__builtin_signbit(q.w)) {
                if (r2 < dw2 && length3(q.x, q.y, q.z, px, py, pz) < dw) {
                    wall = true;
                    return;
                }
                continue;
            }
            if (r2 < phi2) {
                const double d = (double)length3(q.x, q.y, q.z, px, py, pz);
                if (d < phi) {
                    phi = d;
                    phi2 = d * d * (1.0 + 1e-4);
                }
            }
        }
    };
    auto in = [&](int s) { return s >= 0 && s <= n - 1; };
    // 1. the near cells: per axis the point's cell and its neighbour on the
    // nearer side
    const int ni = gn * px - (float)ci < 0.5f ? ci - 1 : ci + 1, nj = gn * py - (float)cj < 0.5f ? cj - 1 : cj + 1,
              nk = gn * pz - (float)ck < 0.5f ? ck - 1 : ck + 1;
    const int klo1 = max(min(ck, nk), 0), khi1 = min(max(ck, nk), n - 1);
    for (int si = min(ci, ni); si <= max(ci, ni) && !wall; ++si)
        for (int sj = min(cj, nj); sj <= max(cj, nj) && !wall; ++sj)
            if (in(si) && in(sj)) scan(cidx(n, si, sj, klo1), cidx(n, si, sj, khi1));
    // 2. the rest of the window
    for (int si = max(ci - 2, 0); si <= min(ci + 2, n - 1) && !wall; ++si)
        for (int sj = max(cj - 2, 0); sj <= min(cj + 2, n - 1) && !wall; ++sj) {
            const bool near_row = (si == ci || si == ni) && (sj == cj || sj == nj);
            for (int sk = max(ck - 2, 0); sk <= min(ck + 2, n - 1) && !wall; ++sk) {
                if (near_row && sk >= klo1 && sk <= khi1) continue;  // done in 1.
                const long c = cidx(n, si, sj, sk);
                const float4 l = blo[c], u = bhi[c];
                // per-axis gaps, shortened by 1e-6 for their own rounding
                const float gx = fmaxf(fmaxf(l.x - px, px - u.x) - 1e-6f, 0.0f),
                            gy = fmaxf(fmaxf(l.y - py, py - u.y) - 1e-6f, 0.0f),
                            gz = fmaxf(fmaxf(l.z - pz, pz - u.z) - 1e-6f, 0.0f);
                const double d2 = ((double)gx * gx + (double)gy * gy + (double)gz * gz) * (1.0 - 1e-4);
                if (!(d2 < phi2) && !(d2 < dw2)) continue;  // also skips empty cells (infinite gap)
                scan(c, c);
            }
        }
    double value = wall ? 4.5 * (double)density / n : phi - (double)dw;
    if (i == 0 || i == mg - 1 || j == 0 || j == mg - 1 || k == 0 || k == mg - 1) value = fmax(value, 0.01);
    f[((long)mg * mg) * k + (long)mg * j + i] = (float)-value;
}

// cut flag of the three lattice edges leaving point V = (x, y, z) (x fastest),
// at index 3 V + axis: the ends lie on opposite sides of the iso level 0
__global__ void k_mesh_edge_flags(int mg, const float *__restrict__ f, int32_t *flag)
{
    const long V = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long m3 = (long)mg * mg * mg;
    if (V >= m3) return;
    const int x = (int)(V % mg), y = (int)((V / mg) % mg), z = (int)(V / ((long)mg * mg));
    const bool s0 = f[V] < 0.0f;
    flag[3 * V + 0] = x + 1 < mg && s0 != (f[V + 1] < 0.0f);
    flag[3 * V + 1] = y + 1 < mg && s0 != (f[V + mg] < 0.0f);
    flag[3 * V + 2] = z + 1 < mg && s0 != (f[V + (long)mg * mg] < 0.0f);
}

// CIsoSurface.cpp:587-682 for every cut edge, from the end the owning cell
// interpolates from: x edges run backwards unless they lie on the last y
// layer (edges 3 and 7 of their cell) and forwards there (edges 1 and 5);
// y edges run forwards unless they lie on the last x layer (edges 0, 4 vs 2,
// 6); z edges always run forwards (8, 9, 10, 11).
__global__ void k_mesh_vertices(int mg, const float *__restrict__ f, const int32_t *__restrict__ flag,
                                const int32_t *__restrict__ vid, float len, float4 *vpos, int32_t *vedge)
{
    const long V = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long m3 = (long)mg * mg * mg;
    if (V >= m3) return;
    const int nc = mg - 1;
    const unsigned x = (unsigned)(V % mg), y = (unsigned)((V / mg) % mg), z = (unsigned)(V / ((long)mg * mg));
    for (int a = 0; a < 3; ++a) {
        if (!flag[3 * V + a]) continue;
        unsigned ax = x, ay = y, az = z, bx = x + (a == 0), by = y + (a == 1), bz = z + (a == 2);
        const bool fwd = a == 0 ? (int)y == nc : (a == 1 ? (int)x < nc : true);
        if (!fwd) {
            unsigned tx = ax, ty = ay, tz = az;
            ax = bx; ay = by; az = bz;
            bx = tx; by = ty; bz = tz;
        }
        const float x1 = ax * len, y1 = ay * len, z1 = az * len, x2 = bx * len, y2 = by * len, z2 = bz * len;
        const float v1 = f[((long)az * mg + ay) * mg + ax], v2 = f[((long)bz * mg + by) * mg + bx];
        const float mu = (0.0f - v1) / (v2 - v1);
        const int out = vid[3 * V + a];
        vpos[out] = make_float4(x1 + mu * (x2 - x1), y1 + mu * (y2 - y1), z1 + mu * (z2 - z1), 0.0f);
        vedge[out] = (int32_t)(3 * V + a);
    }
}

// lattice point of corner b of the cube at (x, y, z): x+1 for corners 2, 3,
// 6, 7; y+1 for 1, 2, 5, 6; z+1 for 4-7 (CIsoSurface.cpp:362-377)
__device__ __forceinline__ long corner_point(int mg, int x, int y, int z, int b)
{
    const int dx = (b == 2 || b == 3 || b == 6 || b == 7), dy = (b == 1 || b == 2 || b == 5 || b == 6), dz = b >= 4;
    return ((long)(z + dz) * mg + (y + dy)) * mg + (x + dx);
}

// GetEdgeID (CIsoSurface.cpp:550-585): 3 * point id + axis of edge e
__device__ __forceinline__ long edge_id(int mg, int x, int y, int z, int e)
{
    const int ox = (e == 2 || e == 6 || e == 10 || e == 11), oy = (e == 1 || e == 5 || e == 9 || e == 10),
              oz = (e >= 4 && e <= 7);
    const int ax = (e == 1 || e == 3 || e == 5 || e == 7) ? 0 : (e >= 8 ? 2 : 1);
    return 3 * (((long)(z + oz) * mg + (y + oy)) * mg + (x + ox)) + ax;
}

// per cell in merge order L = (x nc + y) nc + z: case index and triangle count
__global__ void k_mesh_cells(int mg, const float *__restrict__ f, uint8_t *ccase, int32_t *cnt)
{
    const long L = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = mg - 1;
    if (L >= (long)nc * nc * nc) return;
    const int z = (int)(L % nc), y = (int)((L / nc) % nc), x = (int)(L / ((long)nc * nc));
    unsigned ti = 0;
    for (int b = 0; b < 8; ++b)
        if (f[corner_point(mg, x, y, z, b)] < 0.0f) ti |= 1u << b;
    int m = 0;
    while (d_mc_cases[ti][m]) m += 3;
    ccase[L] = (uint8_t)ti;
    cnt[L] = m / 3;
}

__global__ void k_mesh_tris(int mg, const uint8_t *__restrict__ ccase, const int32_t *__restrict__ toff,
                            const int32_t *__restrict__ vid, int4 *tri)
{
    const long L = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = mg - 1;
    if (L >= (long)nc * nc * nc) return;
    const int z = (int)(L % nc), y = (int)((L / nc) % nc), x = (int)(L / ((long)nc * nc));
    const char *cs = d_mc_cases[ccase[L]];
    int t = toff[L];
    for (int i = 0; cs[i]; i += 3, ++t)
        tri[t] = make_int4(vid[edge_id(mg, x, y, z, hexval(cs[i]))], vid[edge_id(mg, x, y, z, hexval(cs[i + 1]))],
                           vid[edge_id(mg, x, y, z, hexval(cs[i + 2]))], 0);
}

// CalculateNormals' triangle normal (CIsoSurface.cpp:741-756) and the
// mesher's filter (mesher.cpp:177-184): every edge shorter than 0.1
__global__ void k_mesh_tri_normals(long ntri, const int4 *__restrict__ tri, const float4 *__restrict__ vpos,
                                   float4 *tn, int32_t *keep)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
    const int4 q = tri[t];
    const float4 a = vpos[q.x], b = vpos[q.y], c = vpos[q.z];
    const float v1x = b.x - a.x, v1y = b.y - a.y, v1z = b.z - a.z;
    const float v2x = c.x - a.x, v2y = c.y - a.y, v2z = c.z - a.z;
    tn[t] = make_float4(v1z * v2y - v1y * v2z, v1x * v2z - v1z * v2x, v1y * v2x - v1x * v2y, 0.0f);
    const double m2 = 0.1 * 0.1;
    keep[t] = (double)length2(a.x, a.y, a.z, b.x, b.y, b.z) < m2 && (double)length2(b.x, b.y, b.z, c.x, c.y, c.z) < m2 &&
              (double)length2(a.x, a.y, a.z, c.x, c.y, c.z) < m2;
}

__global__ void k_mesh_faces(long ntri, const int4 *__restrict__ tri, const int32_t *__restrict__ keep,
                             const int32_t *__restrict__ fidx, int32_t *faces)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri || !keep[t]) return;
    const int4 q = tri[t];
    const long o = 3L * fidx[t];
    faces[o] = q.x;
    faces[o + 1] = q.y;
    faces[o + 2] = q.z;
}

// The cells around a vertex's edge in merge order; returns how many (<= 4).
__device__ __forceinline__ int edge_cells(int mg, int32_t eid, long *L)
{
    const int nc = mg - 1;
    const int a = eid % 3;
    const long V = eid / 3;
    const int x = (int)(V % mg), y = (int)((V / mg) % mg), z = (int)(V / ((long)mg * mg));
    // the two transverse axes (u, w) in merge-order significance: for x edges
    // (y, z), for y edges (x, z), for z edges (x, y)
    int m = 0;
    for (int du = -1; du <= 0; ++du)
        for (int dw = -1; dw <= 0; ++dw) {
            int cx = x, cy = y, cz = z;
            if (a == 0) { cy += du; cz += dw; }
            else if (a == 1) { cx += du; cz += dw; }
            else { cx += du; cy += dw; }
            if (cx < 0 || cy < 0 || cz < 0 || cx >= nc || cy >= nc || cz >= nc) continue;
            L[m++] = ((long)cx * nc + cy) * nc + cz;
        }
    return m;
}

// CalculateNormals (CIsoSurface.cpp:722-777): each vertex sums the normals
// of its triangles in triangle order, then divides by the float length;
// the mesher keeps them as doubles.
__global__ void k_mesh_vertex_normals(long nv, int mg, const int32_t *__restrict__ vedge,
                                      const int32_t *__restrict__ toff, const int4 *__restrict__ tri,
                                      const float4 *__restrict__ tn, double *Nn)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    long L[4];
    const int m = edge_cells(mg, vedge[v], L);
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int c = 0; c < m; ++c)
        for (int t = toff[L[c]]; t < toff[L[c] + 1]; ++t) {
            const int4 q = tri[t];
            if (q.x != v && q.y != v && q.z != v) continue;
            const float4 nn = tn[t];
            sx += nn.x;
            sy += nn.y;
            sz += nn.z;
        }
    const float l = sqrtf(sx * sx + sy * sy + sz * sz);
    Nn[3 * v] = (double)(sx / l);
    Nn[3 * v + 1] = (double)(sy / l);
    Nn[3 * v + 2] = (double)(sz / l);
}

__device__ __forceinline__ bool normal3(const double *p)
{
    return __builtin_isnormal(p[0]) && __builtin_isnormal(p[1]) && __builtin_isnormal(p[2]);
}

// One smoothMesh pass (mesher.cpp:13-86) with the connections of
// buildConnection (:88-101): the vertex itself, then for each kept face
// around it in face order its other two corners, each sample used when all
// six of its coordinates are normal numbers.  CV / CN keep their previous
// value where a vertex has no usable sample or a zero normal sum.
__global__ void k_mesh_smooth(long nv, int mg, const int32_t *__restrict__ vedge, const int32_t *__restrict__ toff,
                              const int4 *__restrict__ tri, const int32_t *__restrict__ keep,
                              const double *__restrict__ V, const double *__restrict__ Nn, double *CV, double *CN)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    float wsum = 0.0f;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, n0 = 0.0, n1 = 0.0, n2 = 0.0;
    auto add = [&](long u) {
        if (!normal3(V + 3 * u) || !normal3(Nn + 3 * u)) return;
        const float w = 1.0f;
        s0 += (double)w * V[3 * u];
        s1 += (double)w * V[3 * u + 1];
        s2 += (double)w * V[3 * u + 2];
        n0 += (double)w * Nn[3 * u];
        n1 += (double)w * Nn[3 * u + 1];
        n2 += (double)w * Nn[3 * u + 2];
        wsum += w;
    };
    add(v);
    long L[4];
    const int m = edge_cells(mg, vedge[v], L);
    for (int c = 0; c < m; ++c)
        for (int t = toff[L[c]]; t < toff[L[c] + 1]; ++t) {
            if (!keep[t]) continue;
            const int4 q = tri[t];
            if (q.x == v) {
                add(q.y);
                add(q.z);
            } else if (q.y == v) {
                add(q.x);
                add(q.z);
            } else if (q.z == v) {
                add(q.x);
                add(q.y);
            }
        }
    const float l = hypotf_glibc((float)n0, hypotf_glibc((float)n1, (float)n2));
    if (wsum != 0.0f) {
        CV[3 * v] = s0 / (double)wsum;
        CV[3 * v + 1] = s1 / (double)wsum;
        CV[3 * v + 2] = s2 / (double)wsum;
    }
    if (l != 0.0f) {
        CN[3 * v] = n0 / (double)l;
        CN[3 * v + 1] = n1 / (double)l;
        CN[3 * v + 2] = n2 / (double)l;
    }
}

__global__ void k_mesh_zero(double *a, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

__global__ void k_mesh_positions(long nv, const float4 *__restrict__ vpos, double *V)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const float4 p = vpos[v];
    V[3 * v] = p.x;
    V[3 * v + 1] = p.y;
    V[3 * v + 2] = p.z;
}

template <class T> void grow(T **p, size_t &cap, size_t need)
{
    if (need <= cap && *p) return;
    if (*p) FLIP_CHECK(hipFree(*p));
    *p = nullptr;
    FLIP_CHECK(hipMalloc((void **)p, std::max<size_t>(need, 1) * sizeof(T)));
    cap = need;
}

}  // namespace

// Device buffers of the last mesh (kept for flip_gpu_mesh_get and reuse).
struct MeshState {
    int mg = 0;
    long nv = 0, ntri = 0, nf = 0;
    float *f = nullptr;
    int32_t *flag = nullptr, *vid = nullptr, *vedge = nullptr, *cnt = nullptr, *toff = nullptr, *keep = nullptr,
            *fidx = nullptr, *faces = nullptr, *blocks = nullptr;
    uint8_t *ccase = nullptr;
    float4 *vpos = nullptr, *tn = nullptr, *blo = nullptr, *bhi = nullptr;
    int4 *tri = nullptr;
    double *V = nullptr, *N = nullptr, *CV = nullptr, *CN = nullptr;
    size_t c_f = 0, c_flag = 0, c_vid = 0, c_vedge = 0, c_cnt = 0, c_toff = 0, c_keep = 0, c_fidx = 0, c_faces = 0,
           c_blocks = 0, c_ccase = 0, c_vpos = 0, c_tn = 0, c_tri = 0, c_V = 0, c_N = 0, c_CV = 0, c_CN = 0, c_blo = 0,
           c_bhi = 0;
};

void mesh_free(MeshState *m)
{
    if (!m) return;
    void *ps[] = {m->f, m->flag, m->vid, m->vedge, m->cnt, m->toff, m->keep, m->fidx, m->faces, m->blocks,
                  m->ccase, m->vpos, m->tn, m->tri, m->V, m->N, m->CV, m->CN, m->blo, m->bhi};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    delete m;
}

// mesher::generateMesh on an mg^3 lattice; returns (vertex, face) counts
void mesh_build(FlipCtx *c, int mg, long *nv_out, long *nf_out)
{
    if (!c->mesh) c->mesh = new MeshState();
    MeshState &m = *c->mesh;
    hipStream_t s = c->stream;
    const int n = c->n, nc = mg - 1;
    const long m3 = (long)mg * mg * mg, ncell = (long)nc * nc * nc;
    if (!c->bucketed) {
        k_launch_sort(c);
        c->sorted = true;
    }
    m.mg = mg;
    grow(&m.f, m.c_f, (size_t)m3);
    grow(&m.flag, m.c_flag, (size_t)(3 * m3));
    grow(&m.vid, m.c_vid, (size_t)(3 * m3 + 1));
    grow(&m.blocks, m.c_blocks, (size_t)(std::max(3 * m3, 5 * ncell) / 2048 + 16));  // every scan below
    grow(&m.ccase, m.c_ccase, (size_t)ncell);
    grow(&m.cnt, m.c_cnt, (size_t)ncell);
    grow(&m.toff, m.c_toff, (size_t)ncell + 1);
    grow(&m.blo, m.c_blo, (size_t)c->C);
    grow(&m.bhi, m.c_bhi, (size_t)c->C);
    launch_timed(c, "mesh_boxes", k_mesh_cell_boxes, nblk(c->C, TB), TB, c->C, c->cell_start, c->P4(), m.blo, m.bhi);
    launch_timed(c, "mesh_field", k_mesh_field, nblk(m3, TB), TB, mg, n, c->P4(), c->cell_start, m.blo, m.bhi,
                 (float)c->prm.density, m.f);
    launch_timed(c, "mesh_edges", k_mesh_edge_flags, nblk(m3, TB), TB, mg, m.f, m.flag);
    k_launch_scan(c, m.flag, m.vid, 3 * m3, m.blocks);
    launch_timed(c, "mesh_cells", k_mesh_cells, nblk(ncell, TB), TB, mg, m.f, m.ccase, m.cnt);
    k_launch_scan(c, m.cnt, m.toff, ncell, m.blocks);
    int32_t tot[2];
    FLIP_CHECK(hipMemcpyAsync(&tot[0], m.vid + 3 * m3, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FLIP_CHECK(hipMemcpyAsync(&tot[1], m.toff + ncell, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FLIP_CHECK(hipStreamSynchronize(s));
    const long nv = tot[0], ntri = tot[1];
    grow(&m.vpos, m.c_vpos, (size_t)nv);
    grow(&m.vedge, m.c_vedge, (size_t)nv);
    grow(&m.tri, m.c_tri, (size_t)ntri);
    grow(&m.tn, m.c_tn, (size_t)ntri);
    grow(&m.keep, m.c_keep, (size_t)ntri);
    grow(&m.fidx, m.c_fidx, (size_t)ntri + 1);
    grow(&m.V, m.c_V, (size_t)(3 * nv));
    grow(&m.N, m.c_N, (size_t)(3 * nv));
    grow(&m.CV, m.c_CV, (size_t)(3 * nv));
    grow(&m.CN, m.c_CN, (size_t)(3 * nv));
    launch_timed(c, "mesh_vertices", k_mesh_vertices, nblk(m3, TB), TB, mg, m.f, m.flag, m.vid,
                 (float)(1.0 / (mg - 1)), m.vpos, m.vedge);
    if (ntri > 0) {
        launch_timed(c, "mesh_tris", k_mesh_tris, nblk(ncell, TB), TB, mg, m.ccase, m.toff, m.vid, m.tri);
        launch_timed(c, "mesh_tri_normals", k_mesh_tri_normals, nblk(ntri, TB), TB, ntri, m.tri, m.vpos, m.tn,
                     m.keep);
    }
    k_launch_scan(c, m.keep, m.fidx, ntri, m.blocks);
    int32_t nf = 0;
    FLIP_CHECK(hipMemcpyAsync(&nf, m.fidx + ntri, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FLIP_CHECK(hipStreamSynchronize(s));
    grow(&m.faces, m.c_faces, (size_t)(3 * (long)nf));
    if (ntri > 0)
        launch_timed(c, "mesh_faces", k_mesh_faces, nblk(ntri, TB), TB, ntri, m.tri, m.keep, m.fidx, m.faces);
    if (nv > 0) {
        launch_timed(c, "mesh_positions", k_mesh_positions, nblk(nv, TB), TB, nv, m.vpos, m.V);
        launch_timed(c, "mesh_normals", k_mesh_vertex_normals, nblk(nv, TB), TB, nv, mg, m.vedge, m.toff, m.tri,
                     m.tn, m.N);
        // smoothMesh(.., 2, 1, 1): the copies start at zero (vector::resize)
        hipLaunchKernelGGL(k_mesh_zero, nblk(3 * nv, TB), TB, 0, s, m.CV, 3 * nv);
        hipLaunchKernelGGL(k_mesh_zero, nblk(3 * nv, TB), TB, 0, s, m.CN, 3 * nv);
        for (int it = 0; it < 2; ++it) {
            launch_timed(c, "mesh_smooth", k_mesh_smooth, nblk(nv, TB), TB, nv, mg, m.vedge, m.toff, m.tri, m.keep,
                         m.V, m.N, m.CV, m.CN);
            FLIP_CHECK(hipMemcpyAsync(m.V, m.CV, 3 * nv * sizeof(double), hipMemcpyDeviceToDevice, s));
            FLIP_CHECK(hipMemcpyAsync(m.N, m.CN, 3 * nv * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
    }
    FLIP_CHECK(hipStreamSynchronize(s));
    FLIP_CHECK(hipGetLastError());
    m.nv = nv;
    m.ntri = ntri;
    m.nf = nf;
    *nv_out = nv;
    *nf_out = nf;
}

void mesh_counts(const FlipCtx *c, long *nv, long *nf)
{
    *nv = c->mesh ? c->mesh->nv : 0;
    *nf = c->mesh ? c->mesh->nf : 0;
}

void mesh_copy_out(FlipCtx *c, double *v, double *nrm, int32_t *faces)
{
    MeshState &m = *c->mesh;
    hipStream_t s = c->stream;
    if (v && m.nv) FLIP_CHECK(hipMemcpyAsync(v, m.V, 3 * m.nv * sizeof(double), hipMemcpyDeviceToHost, s));
    if (nrm && m.nv) FLIP_CHECK(hipMemcpyAsync(nrm, m.N, 3 * m.nv * sizeof(double), hipMemcpyDeviceToHost, s));
    if (faces && m.nf)
        FLIP_CHECK(hipMemcpyAsync(faces, m.faces, 3 * m.nf * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FLIP_CHECK(hipStreamSynchronize(s));
}
