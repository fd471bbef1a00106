// scene.hip - flip3D::init placement (src/flip3D.cpp:580-764) on the host, and
// the max_dens calibration (flip3D.cpp:689-709) run through the GPU density
// kernel.  Scene setup runs once per simulation, outside the timed step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "flip_internal.hpp"

namespace {

struct Obj {  // common.h:24-32
    int type, shape;  // shape 0 box, 1 sphere
    float r, c[3], lo[3], hi[3];
};

float length3(const float *a, const float *b)  // utility.cpp:16-18
{
    return hypotf(hypotf(a[0] - b[0], a[1] - b[1]), a[2] - b[2]);
}

// placeWalls (flip3D.cpp:300-362): six one-cell boxes
void add_walls(std::vector<Obj> &o, int N)
{
    float t = (float)(1.0 / N);
    float lo[6][3] = {{0, 0, 0}, {(float)(1.0 - t), 0, 0}, {0, 0, 0}, {0, (float)(1.0 - t), 0}, {0, 0, 0},
                      {0, 0, (float)(1.0 - t)}};
    float hi[6][3] = {{t, 1, 1}, {1, 1, 1}, {1, t, 1}, {1, 1, 1}, {1, 1, t}, {1, 1, 1}};
    for (int w = 0; w < 6; ++w) {
        Obj b{};
        b.type = FLIP_WALL;
        b.shape = 0;
        for (int d = 0; d < 3; ++d) {
            b.lo[d] = lo[w][d];
            b.hi[d] = hi[w][d];
        }
        o.push_back(b);
    }
}

// pushParticle's object lookup (flip3D.cpp:580-627): index of the object of
// `type` containing the point, -1 if none or if it lies deep inside a wall
int lookup(const std::vector<Obj> &o, int N, double x, double y, double z, int type)
{
    for (size_t n = 0; n < o.size(); ++n) {
        bool found = false;
        float thick = (float)(3.0 / N);
        const Obj &b = o[n];
        if (b.shape == 0) {
            if (x > b.lo[0] && x < b.hi[0] && y > b.lo[1] && y < b.hi[1] && z > b.lo[2] && z < b.hi[2]) {
                if (b.type == FLIP_WALL && x > b.lo[0] + thick && x < b.hi[0] - thick && y > b.lo[1] + thick &&
                    y < b.hi[1] - thick && z > b.lo[2] + thick && z < b.hi[2] - thick)
                    return -1;
                found = true;
            }
        } else {
            float p[3] = {(float)x, (float)y, (float)z};
            float len = length3(p, b.c);
            if (len < b.r) {
                found = true;
                if (b.type == FLIP_WALL && len < b.r - thick) return -1;
            }
        }
        if (found && b.type == type) return (int)n;
    }
    return -1;
}

uint64_t splitmix_at(uint64_t seed, uint64_t k)
{
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Scene {
    std::vector<float> p, nrm;
    std::vector<int8_t> type;
    void push(float x, float y, float z, int t)
    {
        p.push_back(x);
        p.push_back(y);
        p.push_back(z);
        type.push_back((int8_t)t);
    }
    size_t size() const { return type.size(); }
};

// fluid lattice at spacing DENSITY/N (flip3D.cpp:712-727); jit(x,y,z,idx)
// returns the placed position
template <class J> void place_fluid(Scene &s, const std::vector<Obj> &o, int N, double dens, J jit)
{
    double w = dens * (1.0 / N);
    double wt = 1.0 / N;
    int lim = (int)(N / dens);
    for (int i = 0; i < lim; ++i)
        for (int j = 0; j < lim; ++j)
            for (int k = 0; k < lim; ++k) {
                double x = i * w + w / 2.0, y = j * w + w / 2.0, z = k * w + w / 2.0;
                if (x > wt && x < 1.0 - wt && y > wt && y < 1.0 - wt && z > wt && z < 1.0 - wt) {
                    int ob = lookup(o, N, x, y, z, FLIP_FLUID);
                    if (ob >= 0) jit(s, x, y, z);
                }
            }
}

// wall particles at wall-cell centres (flip3D.cpp:730-740)
template <class R> void place_walls(Scene &s, const std::vector<Obj> &o, int N, R on_push)
{
    double w = 1.0 / N;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                double x = i * w + w / 2.0, y = j * w + w / 2.0, z = k * w + w / 2.0;
                if (lookup(o, N, x, y, z, FLIP_WALL) >= 0) {
                    on_push();
                    s.push((float)x, (float)y, (float)z, FLIP_WALL);
                }
            }
}

int cell_b(int N, float p) { return (int)fmin((double)(N - 1), fmax(0.0, (double)(p * (float)N))); }

// flip3D.cpp:742-761 drop fluid particles in wall cells, then
// compute_wall_normal (flip3D.cpp:75-130) particle normals
void finish(Scene &s, int N)
{
    std::vector<int8_t> wallcell((size_t)N * N * N, 0);
    for (size_t i = 0; i < s.size(); ++i)
        if (s.type[i] == FLIP_WALL) {
            int ci = (int)fmax(0.0, fmin((double)(N - 1), (double)((float)N * s.p[3 * i])));
            int cj = (int)fmax(0.0, fmin((double)(N - 1), (double)((float)N * s.p[3 * i + 1])));
            int ck = (int)fmax(0.0, fmin((double)(N - 1), (double)((float)N * s.p[3 * i + 2])));
            wallcell[((size_t)ci * N + cj) * N + ck] = 1;
        }
    Scene t;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s.type[i] != FLIP_WALL) {
            int ci = cell_b(N, s.p[3 * i]), cj = cell_b(N, s.p[3 * i + 1]), ck = cell_b(N, s.p[3 * i + 2]);
            if (wallcell[((size_t)ci * N + cj) * N + ck]) continue;
        }
        t.push(s.p[3 * i], s.p[3 * i + 1], s.p[3 * i + 2], s.type[i]);
    }
    s = t;
    s.nrm.assign(3 * s.size(), 0.0f);
    double wt = 1.0 / N;
    // wall particles by (cell, index): a cell's bucket in bucket order
    auto cell_key = [&](const float *q) {
        int qi = (int)fmax(0.0, fmin((double)(N - 1), (double)((float)N * q[0])));
        int qj = (int)fmax(0.0, fmin((double)(N - 1), (double)((float)N * q[1])));
        int qk = (int)fmax(0.0, fmin((double)(N - 1), (double)((float)N * q[2])));
        return ((long)qi * N + qj) * N + qk;
    };
    std::vector<std::pair<long, size_t>> walls;
    for (size_t i = 0; i < s.size(); ++i)
        if (s.type[i] == FLIP_WALL) walls.push_back({cell_key(&s.p[3 * i]), i});
    std::sort(walls.begin(), walls.end());
    for (size_t i = 0; i < s.size(); ++i) {
        if (s.type[i] != FLIP_WALL) continue;
        float *p = &s.p[3 * i], *nr = &s.nrm[3 * i];
        if (p[0] <= 1.1 * wt) nr[0] = 1.0f;
        if (p[0] >= 1.0 - 1.1 * wt) nr[0] = -1.0f;
        if (p[1] <= 1.1 * wt) nr[1] = 1.0f;
        if (p[1] >= 1.0 - 1.1 * wt) nr[1] = -1.0f;
        if (p[2] <= 1.1 * wt) nr[2] = 1.0f;
        if (p[2] >= 1.0 - 1.1 * wt) nr[2] = -1.0f;
        if (nr[0] == 0.0f && nr[1] == 0.0f && nr[2] == 0.0f) {
            // interior wall: inverse-distance sum over wall particles in the
            // 7x7x7 cell window, visited in bucket order (flip3D.cpp:107-118)
            int ci = cell_b(N, p[0]), cj = cell_b(N, p[1]), ck = cell_b(N, p[2]);
            std::vector<size_t> cand;
            for (int si = std::max(ci - 3, 0); si <= std::min(ci + 3, N - 1); ++si)
                for (int sj = std::max(cj - 3, 0); sj <= std::min(cj + 3, N - 1); ++sj) {
                    long k0 = ((long)si * N + sj) * N + std::max(ck - 3, 0);
                    long k1 = ((long)si * N + sj) * N + std::min(ck + 3, N - 1);
                    auto a = std::lower_bound(walls.begin(), walls.end(), std::make_pair(k0, (size_t)0));
                    for (; a != walls.end() && a->first <= k1; ++a)
                        if (a->second != i) cand.push_back(a->second);
                }
            for (size_t t2 : cand) {
                const float *q = &s.p[3 * t2];
                float d = length3(p, q);
                float w = (float)(1.0 / d);
                nr[0] += w * (p[0] - q[0]) / d;
                nr[1] += w * (p[1] - q[1]) / d;
                nr[2] += w * (p[2] - q[2]) / d;
            }
        }
        float d = hypotf(hypotf(nr[0], nr[1]), nr[2]);
        if (d != 0.0f) {
            nr[0] /= d;
            nr[1] /= d;
            nr[2] /= d;
        }
    }
}

}  // namespace

extern "C" int flip_gpu_upload_particles(flip_ctx *ctx, int64_t np, const float *pos, const float *vel,
                                         const float *nrm, const int8_t *type, const float *mass, const float *dens);

static Obj box_obj(int type, double x0, double x1, double y0, double y1, double z0, double z1)
{
    Obj b{};
    b.type = type;
    b.shape = 0;
    b.lo[0] = (float)x0; b.hi[0] = (float)x1;
    b.lo[1] = (float)y0; b.hi[1] = (float)y1;
    b.lo[2] = (float)z0; b.hi[2] = (float)z1;
    return b;
}

static Obj sphere_obj(int type, double cx, double cy, double cz, double r)
{
    Obj b{};
    b.type = type;
    b.shape = 1;
    b.c[0] = (float)cx; b.c[1] = (float)cy; b.c[2] = (float)cz;
    b.r = (float)r;
    return b;
}

// The reference's scene functions, appended after placeWalls' six walls:
// damBreakTest (flip3D.cpp:277-297), spherePourTest (:203-228, a wall
// sphere over a thin pool, pouring for 200 steps), cliffPourTest (:249-275,
// a wall box cliff beside a thin pool, pouring for 200 steps) and
// waterDropTest (:230-247).  The pour scenes set the emitter as they do.
static void test_scene_objects(FlipCtx *c, int scene, std::vector<Obj> &o)
{
    const double wt = 1.0 / c->n;
    switch (scene) {
    case FLIP_SCENE_REFERENCE:
        o.push_back(box_obj(FLIP_FLUID, 0.2, 0.4, wt, 0.4, 0.2, 0.8));
        o.push_back(box_obj(FLIP_FLUID, wt, 1.0 - wt, wt, 0.06, wt, 1.0 - wt));
        break;
    case FLIP_SCENE_SPHERE_POUR:
        o.push_back(sphere_obj(FLIP_WALL, 0.5, 0.4, 0.47, 0.15));
        o.push_back(box_obj(FLIP_FLUID, wt, 1.0 - wt, wt, 0.05, wt, 1.0 - wt));
        set_emitter(c, 0.5f, 0.5f, 0.1f, 200);
        break;
    case FLIP_SCENE_CLIFF_POUR:
        o.push_back(box_obj(FLIP_WALL, 0.0, 0.35, 0.0, 0.4, -0.002, 1.002));
        o.push_back(box_obj(FLIP_FLUID, 0.35, 1.0 - wt, wt, 0.05, wt, 1.0 - wt));
        set_emitter(c, 0.2f, 0.2f, 0.1f, 200);
        break;
    case FLIP_SCENE_WATER_DROP:
        o.push_back(box_obj(FLIP_FLUID, wt, 1.0 - wt, wt, 0.2, wt, 1.0 - wt));
        o.push_back(sphere_obj(FLIP_FLUID, 0.5, 0.6, 0.5, 0.12));
        break;
    }
}

// max_dens (flip3D.cpp:689-709): densities of a 10^3 lattice at
// (i+0.5)*DENSITY/N with max_dens = 1, maximum over the lattice
static float calibrate_max_dens(flip_ctx *ctx, FlipCtx *c)
{
    int N = c->n;
    float h = (float)(c->prm.density / N);
    std::vector<float> p;
    std::vector<int8_t> t;
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j)
            for (int k = 0; k < 10; ++k) {
                p.push_back((float)((i + 0.5) * h));
                p.push_back((float)((j + 0.5) * h));
                p.push_back((float)((k + 0.5) * h));
                t.push_back(FLIP_FLUID);
            }
    int rc = flip_gpu_upload_particles(ctx, 1000, p.data(), nullptr, nullptr, t.data(), nullptr, nullptr);
    if (rc) throw flip_arg_error{rc, flip_gpu_last_error()};
    float save = c->max_dens;
    c->max_dens = 1.0f;
    k_launch_sort(c);
    k_launch_density(c);
    std::vector<float4> v(1000);
    FLIP_CHECK(hipStreamSynchronize(c->stream));
    FLIP_CHECK(hipMemcpy(v.data(), c->V4(), 1000 * sizeof(float4), hipMemcpyDeviceToHost));
    c->max_dens = save;
    float md = 0.0f;
    for (int i = 0; i < 1000; ++i) md = (float)fmax((double)md, (double)v[i].w);
    return md;
}

// objects.dat records (flip3D.cpp:498-508) of a scene.  Each reference
// scene function fills one reused stack Object, so a field it set for an
// earlier object carries over to the next; fields it never sets are
// indeterminate in the reference and written as 0 here.
namespace {
using SceneObject = FlipCtx::SceneObject;

SceneObject rec_of(const Obj &b, int material, int visible)
{
    SceneObject r{};
    r.type = b.type;
    r.shape = b.shape;
    r.material = material;
    r.visible = visible;
    r.f[0] = b.r;
    for (int d = 0; d < 3; ++d) {
        r.f[1 + d] = b.c[d];
        r.f[4 + d] = b.lo[d];
        r.f[7 + d] = b.hi[d];
    }
    return r;
}

// the second object of a scene function inherits the first's fields that it does not set
SceneObject carry(SceneObject second, const SceneObject &first, bool sets_r_c, bool sets_p, bool sets_material,
                  bool sets_visible)
{
    if (!sets_r_c)
        for (int t = 0; t < 4; ++t) second.f[t] = first.f[t];
    if (!sets_p)
        for (int t = 4; t < 10; ++t) second.f[t] = first.f[t];
    if (!sets_material) second.material = first.material;
    if (!sets_visible) second.visible = first.visible;
    return second;
}

// placeWalls (flip3D.cpp:300-362): GLASS walls, a GRAY floor, none visible
void wall_records(const std::vector<Obj> &o, std::vector<SceneObject> &out)
{
    for (int w = 0; w < 6; ++w) out.push_back(rec_of(o[w], w == 2 ? 2 : 1, 0));
}

void scene_records(int scene, const std::vector<Obj> &o, std::vector<SceneObject> &out)
{
    wall_records(o, out);
    if (o.size() < 8) {
        for (size_t t = 6; t < o.size(); ++t) out.push_back(rec_of(o[t], 0, 1));
        return;
    }
    const Obj &a = o[6], &b = o[7];
    switch (scene) {
    case FLIP_SCENE_REFERENCE:  // damBreakTest: both boxes set type, shape, visible, p
        out.push_back(rec_of(a, 0, 1));
        out.push_back(rec_of(b, 0, 1));
        break;
    case FLIP_SCENE_SPHERE_POUR: {  // a RED wall sphere, then a box (r, c, material carried)
        SceneObject s0 = rec_of(a, 3, 1);
        out.push_back(s0);
        out.push_back(carry(rec_of(b, 0, 1), s0, false, true, false, true));
        break;
    }
    case FLIP_SCENE_CLIFF_POUR: {  // a RED wall box, then a box (material carried)
        SceneObject s0 = rec_of(a, 3, 1);
        out.push_back(s0);
        out.push_back(carry(rec_of(b, 0, 1), s0, true, true, false, true));
        break;
    }
    case FLIP_SCENE_WATER_DROP: {  // a box, then a sphere (p carried); visible never set
        SceneObject s0 = rec_of(a, 0, 0);
        out.push_back(s0);
        out.push_back(carry(rec_of(b, 0, 0), s0, true, false, false, false));
        break;
    }
    case FLIP_SCENE_BENCHMARK: {  // a box, then a sphere (p carried)
        SceneObject s0 = rec_of(a, 0, 1);
        out.push_back(s0);
        out.push_back(carry(rec_of(b, 0, 1), s0, true, false, false, true));
        break;
    }
    default:
        for (size_t t = 6; t < o.size(); ++t) out.push_back(rec_of(o[t], 0, 1));
    }
}
}  // namespace

extern "C" int flip_gpu_build_scene(flip_ctx *ctx, int scene, uint64_t seed, int64_t *np_out)
{
    if (!ctx) return -1;
    FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
    try {
        rng_to_host(c);  // the reference scenes draw rand() on the host
        int N = c->n;
        double dens = c->prm.density;
        std::vector<Obj> o;
        add_walls(o, N);
        Scene s;
        float wt = (float)(1.0 / N);
        c->step = 0;
        set_emitter(c, 0.0f, 0.0f, 0.12f, -1);  // flip3D.cpp:41-43 defaults: no pour
        if (scene == FLIP_SCENE_REFERENCE || scene == FLIP_SCENE_SPHERE_POUR || scene == FLIP_SCENE_CLIFF_POUR ||
            scene == FLIP_SCENE_WATER_DROP) {
            // the reference's test scenes after placeWalls, placed through
            // pushParticle: 3 rand() per placed particle
            test_scene_objects(c, scene, o);
            GlibcRand &rng = c->rng;
            place_fluid(s, o, N, dens, [&](Scene &sc, double x, double y, double z) {
                float px = (float)(x + 0.01 * 1 * 0.2 * ((rng.next() % 101) / 50.0 - 1.0) / N);
                float py = (float)(y + 0.01 * 1 * 0.2 * ((rng.next() % 101) / 50.0 - 1.0) / N);
                float pz = (float)(z + 0.01 * 1 * 0.2 * ((rng.next() % 101) / 50.0 - 1.0) / N);
                sc.push(px, py, pz, FLIP_FLUID);
            });
            place_walls(s, o, N, [&]() {
                rng.next();
                rng.next();
                rng.next();
            });
        } else if (scene == FLIP_SCENE_RANDOM || scene == FLIP_SCENE_RANDOM_BOXES) {
            // FLIP_SCENE_RANDOM: 1-4 liquid boxes (edges U(0.1, 0.6)) or spheres
            // (radius U(0.05, 0.2)); FLIP_SCENE_RANDOM_BOXES: 1-3 boxes.  Centres
            // uniform inside the walls; draws from SplitMix64(seed) at indices
            // 1e9 + i so the jitter stream below stays independent
            const bool boxes = scene == FLIP_SCENE_RANDOM_BOXES;
            const int maxobj = boxes ? 3 : 4;
            uint64_t draw = 1000000000ull;
            auto uni = [&](double a, double b) {
                double uu = (double)(splitmix_at(seed, draw++) >> 11) * (1.0 / 9007199254740992.0);
                return a + (b - a) * uu;
            };
            int nobj = 1 + (int)(uni(0.0, 1.0) * (double)maxobj);
            if (nobj > maxobj) nobj = maxobj;
            for (int m = 0; m < nobj; ++m) {
                Obj b{};
                b.type = FLIP_FLUID;
                double cx = uni(wt, 1.0 - wt), cy = uni(wt, 1.0 - wt), cz = uni(wt, 1.0 - wt);
                if (boxes || uni(0.0, 1.0) < 0.5) {
                    b.shape = 0;
                    double e[3] = {uni(0.1, 0.6), uni(0.1, 0.6), uni(0.1, 0.6)};
                    double c3[3] = {cx, cy, cz};
                    for (int d = 0; d < 3; ++d) {
                        b.lo[d] = (float)(c3[d] - 0.5 * e[d]);
                        b.hi[d] = (float)(c3[d] + 0.5 * e[d]);
                    }
                } else {
                    b.shape = 1;
                    b.r = (float)uni(0.05, 0.2);
                    b.c[0] = (float)cx;
                    b.c[1] = (float)cy;
                    b.c[2] = (float)cz;
                }
                o.push_back(b);
            }
            double dx = 1.0 / N;
            uint64_t fi = 0;
            place_fluid(s, o, N, dens, [&](Scene &sc, double x, double y, double z) {
                double j3[3];
                for (int ax = 0; ax < 3; ++ax) {
                    uint64_t v = splitmix_at(seed, 3 * fi + (uint64_t)ax);
                    double uu = (double)(v >> 11) * (1.0 / 9007199254740992.0);
                    j3[ax] = (uu * 0.5 - 0.25) * dx;
                }
                sc.push((float)(x + j3[0]), (float)(y + j3[1]), (float)(z + j3[2]), FLIP_FLUID);
                fi++;
            });
            place_walls(s, o, N, []() {});
        } else if (scene == FLIP_SCENE_BENCHMARK) {
            Obj a{}, b{};
            a.type = FLIP_FLUID;
            a.lo[0] = 0; a.hi[0] = 0.4f; a.lo[1] = 0; a.hi[1] = 0.6f; a.lo[2] = 0; a.hi[2] = 1;
            b.type = FLIP_FLUID;
            b.shape = 1;
            b.r = 0.12f;
            b.c[0] = 0.6f; b.c[1] = 0.7f; b.c[2] = 0.5f;
            o.push_back(a);
            o.push_back(b);
            double dx = 1.0 / N;
            uint64_t fi = 0;
            place_fluid(s, o, N, dens, [&](Scene &sc, double x, double y, double z) {
                double j3[3];
                for (int ax = 0; ax < 3; ++ax) {
                    uint64_t v = splitmix_at(seed, 3 * fi + (uint64_t)ax);
                    double uu = (double)(v >> 11) * (1.0 / 9007199254740992.0);
                    j3[ax] = (uu * 0.5 - 0.25) * dx;
                }
                sc.push((float)(x + j3[0]), (float)(y + j3[1]), (float)(z + j3[2]), FLIP_FLUID);
                fi++;
            });
            place_walls(s, o, N, []() {});
        } else {
            return flip_internal_fail(FLIP_E_ARG, "unknown scene id");
        }
        finish(s, N);
        std::vector<SceneObject> recs;
        scene_records(scene, o, recs);
        float md = calibrate_max_dens(ctx, c);
        int rc = flip_gpu_upload_particles(ctx, (int64_t)s.size(), s.p.data(), nullptr, s.nrm.data(), s.type.data(),
                                           nullptr, nullptr);
        if (rc) return rc;
        c->max_dens = md;
        c->objects = recs;
        c->objects_saved = false;
        if (np_out) *np_out = (int64_t)s.size();
        return FLIP_OK;
    } catch (const flip_hip_error &e) {
        return flip_internal_fail(FLIP_E_HIP, e.what.c_str());
    } catch (const flip_arg_error &e) {
        return flip_internal_fail(e.code, e.what.c_str());
    }
}
