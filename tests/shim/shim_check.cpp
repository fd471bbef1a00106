// shim_check.cpp - the reference-signature shim against the flat C ABI.
//
// Every check runs the same operation twice, once through the reference's
// C++ entry points (jagged grids, vectors of particle records) and once
// through the C ABI on flat arrays, and requires identical bits.  Run from an
// empty working directory (checkpoints go to render/save/).
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "flip3d_shim.h"

static int failures = 0;

#define EXPECT(cond, what)                                     \
    do {                                                       \
        if (!(cond)) {                                         \
            std::printf("FAIL %s (line %d)\n", what, __LINE__); \
            failures++;                                        \
        }                                                      \
    } while (0)

static void ok(int rc, const char *what)
{
    if (rc != FLIP_OK) {
        std::printf("FAIL %s: %s\n", what, flip_gpu_last_error());
        std::exit(2);
    }
}

template <class T>
static T ***alloc3(int w, int h, int d)
{
    T ***f = new T **[w + 1];
    for (int i = 0; i < w; ++i) {
        f[i] = new T *[h + 1];
        for (int j = 0; j < h; ++j) f[i][j] = new T[d]();
        f[i][h] = nullptr;
    }
    f[w] = nullptr;
    return f;
}

template <class T>
static void to_flat(T ***g, int w, int h, int d, std::vector<T> &o)
{
    o.resize((size_t)w * h * d);
    for (int i = 0; i < w; ++i)
        for (int j = 0; j < h; ++j)
            for (int k = 0; k < d; ++k) o[((size_t)i * h + j) * d + k] = g[i][j][k];
}

template <class T>
static void from_flat(const std::vector<T> &o, int w, int h, int d, T ***g)
{
    for (int i = 0; i < w; ++i)
        for (int j = 0; j < h; ++j)
            for (int k = 0; k < d; ++k) g[i][j][k] = o[((size_t)i * h + j) * d + k];
}

static bool same(const std::vector<float> &a, const std::vector<float> &b)
{
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0;
}

static unsigned rng_state = 12345u;
static float frand()
{
    rng_state = rng_state * 1664525u + 1013904223u;
    return (float)((rng_state >> 8) * (1.0 / 16777216.0));
}

static void check_solver()
{
    const int n = 16;
    char ***A = alloc3<char>(n, n, n);
    float ***L = alloc3<float>(n, n, n), ***x = alloc3<float>(n, n, n), ***b = alloc3<float>(n, n, n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                bool wall = i == 0 || j == 0 || k == 0 || i == n - 1 || j == n - 1 || k == n - 1;
                bool air = !wall && j > 10;
                A[i][j][k] = wall ? FLIP_WALL : air ? FLIP_AIR : FLIP_FLUID;
                L[i][j][k] = wall ? 1.0f : air ? 0.3f + 0.4f * frand() : -0.5f;
                x[i][j][k] = 0.1f * frand();
                b[i][j][k] = A[i][j][k] == FLIP_FLUID ? frand() - 0.5f : 0.0f;
            }
    std::vector<char> fa;
    std::vector<float> fl, fx, fb;
    to_flat(A, n, n, n, fa);
    to_flat(L, n, n, n, fl);
    to_flat(x, n, n, n, fx);
    to_flat(b, n, n, n, fb);
    solver::setSubcell(1);
    solver::solve(A, L, x, b, n);
    flip_params p;
    ok(flip_gpu_default_params(&p, n), "default_params");
    flip_ctx *c = nullptr;
    ok(flip_gpu_create(&p, &c), "create");
    ok(flip_gpu_solve(c, reinterpret_cast<const int8_t *>(fa.data()), fl.data(), fx.data(), fb.data(), 1, nullptr),
       "solve");
    std::vector<float> sx, sb;
    to_flat(x, n, n, n, sx);
    to_flat(b, n, n, n, sb);
    EXPECT(same(sx, fx), "solver::solve x matches flip_gpu_solve");
    EXPECT(same(sb, fb), "solver::solve negates b like flip_gpu_solve");
    bool moved = false;
    for (float v : sx) moved = moved || v != 0.0f;
    EXPECT(moved, "solver::solve produced a pressure");
    flip_gpu_destroy(c);
    std::printf("solver ok\n");
}

static void check_mapper()
{
    const int gn = 12;
    const int np = 4000;
    std::vector<particle> store(np);
    std::vector<particle *> ps(np);
    std::vector<float> pos(3 * np), vel(3 * np);
    std::vector<int8_t> type(np);
    for (int i = 0; i < np; ++i) {
        particle &q = store[i];
        std::memset(&q, 0, sizeof(q));
        for (int a = 0; a < 3; ++a) {
            q.p[a] = 0.05f + 0.9f * frand();
            q.u[a] = frand() - 0.5f;
            pos[3 * i + a] = q.p[a];
            vel[3 * i + a] = q.u[a];
        }
        q.type = (i % 17 == 0) ? FLIP_WALL : FLIP_FLUID;
        q.m = 1.0f;
        q.dens = 10.0f;
        type[i] = q.type;
        ps[i] = &q;
    }
    float ***grid[3];
    for (int k = 0; k < 3; ++k) grid[k] = alloc3<float>(gn + (k == 0), gn + (k == 1), gn + (k == 2));
    mapper::mapP2G(nullptr, ps, grid, gn);
    flip_params p;
    ok(flip_gpu_default_params(&p, gn), "default_params");
    flip_ctx *c = nullptr;
    ok(flip_gpu_create(&p, &c), "create");
    ok(flip_gpu_upload_particles(c, np, pos.data(), vel.data(), nullptr, type.data(), nullptr, nullptr), "upload");
    size_t F = (size_t)(gn + 1) * gn * gn;
    std::vector<float> u[3];
    for (auto &v : u) v.resize(F);
    ok(flip_gpu_map_p2g(c, u[0].data(), u[1].data(), u[2].data()), "map_p2g");
    for (int k = 0; k < 3; ++k) {
        std::vector<float> s;
        to_flat(grid[k], gn + (k == 0), gn + (k == 1), gn + (k == 2), s);
        EXPECT(same(s, u[k]), "mapper::mapP2G matches flip_gpu_map_p2g");
    }
    // G2P of a random field
    for (int k = 0; k < 3; ++k) {
        for (auto &v : u[k]) v = frand() - 0.5f;
        from_flat(u[k], gn + (k == 0), gn + (k == 1), gn + (k == 2), grid[k]);
    }
    mapper::mapG2P(ps, grid, gn);
    std::vector<float> out(3 * np);
    ok(flip_gpu_map_g2p(c, u[0].data(), u[1].data(), u[2].data(), np, pos.data(), out.data()), "map_g2p");
    bool eq = true;
    for (int i = 0; i < np; ++i)
        for (int a = 0; a < 3; ++a) eq = eq && std::memcmp(&store[i].u[a], &out[3 * i + a], 4) == 0;
    EXPECT(eq, "mapper::mapG2P matches flip_gpu_map_g2p");
    float pt[3] = {store[5].p[0], store[5].p[1], store[5].p[2]}, uv[3];
    mapper::fetchVelocity(pt, uv, grid, gn);
    EXPECT(std::memcmp(uv, &out[15], 12) == 0, "mapper::fetchVelocity matches mapG2P");
    flip_gpu_destroy(c);
    std::printf("mapper ok\n");
}

static void particles_of(flip_ctx *c, std::vector<float> &pos, std::vector<float> &vel)
{
    int64_t np = 0;
    ok(flip_gpu_num_particles(c, &np), "num_particles");
    pos.resize(3 * np);
    vel.resize(3 * np);
    ok(flip_gpu_download_particles(c, pos.data(), vel.data(), nullptr, nullptr, nullptr), "download");
}

static void check_step_and_checkpoint()
{
    const int n = 16;
    flip3d_shim::set_grid_size(n);
    flip3D::init();
    flip_params p;
    ok(flip_gpu_default_params(&p, n), "default_params");
    flip_ctx *c = nullptr;
    ok(flip_gpu_create(&p, &c), "create");
    int64_t np = 0;
    ok(flip_gpu_build_scene(c, FLIP_SCENE_REFERENCE, 0, &np), "build_scene");
    for (int s = 0; s < 3; ++s) {
        flip3D::simulateStep();
        ok(flip_gpu_step(c, nullptr), "step");
    }
    std::vector<float> p1, v1, p2, v2;
    particles_of(flip3d_shim::context(), p1, v1);
    particles_of(c, p2, v2);
    EXPECT(same(p1, p2) && same(v1, v2), "flip3D::simulateStep matches flip_gpu_step");

    // checkpoint at step 3, reload through flip3D::init(3), continue both
    mkdir("render", 0755);
    mkdir("render/save", 0755);
    ok(flip_gpu_save_state(flip3d_shim::context(), "render/save/3_scene.dat"), "save_state");
    flip3D::init(3);
    std::vector<particle> rec;
    flip3d_shim::download(rec);
    EXPECT((int64_t)rec.size() == np, "reloaded particle count");
    // a checkpoint carries no rand() state (neither does the reference's):
    // restart both streams alike
    ok(flip_gpu_srand(flip3d_shim::context(), 77), "srand");
    ok(flip_gpu_srand(c, 77), "srand");
    for (int s = 0; s < 2; ++s) {
        flip3D::simulateStep();
        ok(flip_gpu_step(c, nullptr), "step");
    }
    particles_of(flip3d_shim::context(), p1, v1);
    particles_of(c, p2, v2);
    EXPECT(same(p1, p2) && same(v1, v2), "steps after a checkpoint reload match an uninterrupted run");
    flip_gpu_destroy(c);
    std::printf("step and checkpoint ok\n");
}

static std::string slurp(const char *path)
{
    std::string out;
    FILE *f = std::fopen(path, "rb");
    if (!f) return out;
    char buf[65536];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, k);
    std::fclose(f);
    return out;
}

// flip3d_shim::set_outputs: simulateStep writes the reference's per-step files
// (preview at the start of the step, after the pour; mesh files at the end),
// equal to what the C ABI writes for the same steps
static void check_outputs()
{
    const int n = 16;
    flip3d_shim::set_grid_size(n);
    flip3d_shim::set_scene(FLIP_SCENE_SPHERE_POUR);
    flip3d_shim::set_outputs(true, true);
    flip3D::init();
    flip_params p;
    ok(flip_gpu_default_params(&p, n), "default_params");
    flip_ctx *c = nullptr;
    ok(flip_gpu_create(&p, &c), "create");
    int64_t np = 0;
    ok(flip_gpu_build_scene(c, FLIP_SCENE_SPHERE_POUR, 0, &np), "build_scene");
    mkdir("abi", 0755);
    for (int s = 1; s <= 2; ++s) {
        flip3D::simulateStep();
        ok(flip_gpu_step_begin(c), "step_begin");
        char path[64];
        std::snprintf(path, sizeof(path), "abi/frame_%d.bmp", s);
        ok(flip_gpu_write_preview(c, path), "write_preview");
        ok(flip_gpu_step(c, nullptr), "step");
        int64_t nv = 0, nf = 0;
        ok(flip_gpu_mesh(c, 0, &nv, &nf), "mesh");
        std::snprintf(path, sizeof(path), "abi/%d_scene.obj", s);
        ok(flip_gpu_write_obj(c, path), "write_obj");
        std::snprintf(path, sizeof(path), "abi/%d_scene.sc", s);
        ok(flip_gpu_write_sunflow(c, path, "abi/template.sc"), "write_sunflow");
    }
    int64_t np2 = 0;
    ok(flip_gpu_num_particles(flip3d_shim::context(), &np2), "num_particles");
    EXPECT(np2 > np, "the pour scene poured through simulateStep");
    const char *pairs[][2] = {{"render/preview/frame_1.bmp", "abi/frame_1.bmp"},
                              {"render/preview/frame_2.bmp", "abi/frame_2.bmp"},
                              {"render/obj/2_scene.obj", "abi/2_scene.obj"},
                              {"render/sunflow/2_scene.sc", "abi/2_scene.sc"},
                              {"render/sunflow/template.sc", "abi/template.sc"}};
    for (auto &pr : pairs) {
        std::string a = slurp(pr[0]), b = slurp(pr[1]);
        EXPECT(!a.empty() && a == b, pr[0]);
    }
    flip3d_shim::set_outputs(false, false);
    flip3d_shim::set_scene(FLIP_SCENE_REFERENCE);
    flip_gpu_destroy(c);
    std::printf("outputs ok\n");
}

int main()
{
    try {
        check_solver();
        check_mapper();
        check_step_and_checkpoint();
        check_outputs();
    } catch (const std::exception &e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 2;
    }
    flip3d_shim::shutdown();
    std::printf(failures ? "FAILED %d\n" : "ALL OK\n", failures);
    return failures ? 1 : 0;
}
