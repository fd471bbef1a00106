// flip3d_shim.cpp - the reference's C++ entry points over the C ABI.
//
// Each function keeps the reference's signature and in/out meaning and moves
// the data between the reference's host structures (vectors of particle
// records, jagged [i][j][k] grids) and a device context of the GPU library.
// Contexts are kept per grid size, the way the reference keeps static scratch
// sized at first call (solver.cpp:226).
#include "../../include/flip3d_shim.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

void check(int rc, const char *what)
{
    if (rc != FLIP_OK) throw std::runtime_error(std::string(what) + ": " + flip_gpu_last_error());
}

struct Shim {
    std::map<long, flip_ctx *> ctx;  // (grid size, density key) -> context
    flip_ctx *sim = nullptr;         // flip3D::init / simulateStep state
    int sim_n = 32;                  // the reference's #define N (flip3D.cpp:25)
    int sim_scene = FLIP_SCENE_REFERENCE;  // placeObjects' damBreakTest (flip3D.cpp:365-373)
    bool out_preview = false, out_mesh = false;  // write_frame / write_mesh files per step
    char subcell = 0;                // solver.cpp:13 default
    ~Shim() { release(); }
    void release()
    {
        for (auto &kv : ctx) flip_gpu_destroy(kv.second);
        ctx.clear();
        if (sim) flip_gpu_destroy(sim);
        sim = nullptr;
    }
};

Shim &shim()
{
    static Shim s;
    return s;
}

// A context for grid size n whose corrector radius DENSITY/n equals re.
flip_ctx *context_for(int n, double density = 0.5)
{
    long key = (long)n * 1000003L + (long)std::lround(density * 1e6);
    auto &m = shim().ctx;
    auto it = m.find(key);
    if (it != m.end()) return it->second;
    flip_params p;
    check(flip_gpu_default_params(&p, n), "flip_gpu_default_params");
    p.density = density;
    flip_ctx *c = nullptr;
    check(flip_gpu_create(&p, &c), "flip_gpu_create");
    m[key] = c;
    return c;
}

void upload(flip_ctx *c, const std::vector<particle *> &ps)
{
    size_t np = ps.size();
    std::vector<float> pos(3 * np), vel(3 * np), nrm(3 * np), mass(np), dens(np);
    std::vector<int8_t> type(np);
    for (size_t i = 0; i < np; ++i) {
        const particle *p = ps[i];
        for (int a = 0; a < 3; ++a) {
            pos[3 * i + a] = p->p[a];
            vel[3 * i + a] = p->u[a];
            nrm[3 * i + a] = p->n[a];
        }
        type[i] = (int8_t)p->type;
        mass[i] = p->m;
        dens[i] = p->dens;
    }
    check(flip_gpu_upload_particles(c, (int64_t)np, pos.data(), vel.data(), nrm.data(), type.data(), mass.data(),
                                    dens.data()),
          "flip_gpu_upload_particles");
}

// jagged alloc3D grid <-> flat [i][j][k]
template <class T>
void flatten(T ***g, int w, int h, int d, std::vector<T> &out)
{
    out.resize((size_t)w * h * d);
    for (int i = 0; i < w; ++i)
        for (int j = 0; j < h; ++j) std::memcpy(&out[((size_t)i * h + j) * d], g[i][j], sizeof(T) * d);
}

template <class T>
void unflatten(const std::vector<T> &in, int w, int h, int d, T ***g)
{
    for (int i = 0; i < w; ++i)
        for (int j = 0; j < h; ++j) std::memcpy(g[i][j], &in[((size_t)i * h + j) * d], sizeof(T) * d);
}

void face_dims(int c, int gn, int &w, int &h, int &d)
{
    w = gn + (c == 0);
    h = gn + (c == 1);
    d = gn + (c == 2);
}

}  // namespace

// ------------------------------------------------------------------ mapper
// mapper.cpp:17-88: the caller has sorted `sort`; the GPU bins the same
// particles itself in the same bucket order, so `sort` is not read.
void mapper::mapP2G(sorter *, std::vector<particle *> &particles, FLOAT ****grid, int gn)
{
    flip_ctx *c = context_for(gn);
    upload(c, particles);
    std::vector<float> u[3];
    size_t F = (size_t)(gn + 1) * gn * gn;
    for (auto &v : u) v.resize(F);
    check(flip_gpu_map_p2g(c, u[0].data(), u[1].data(), u[2].data()), "flip_gpu_map_p2g");
    for (int k = 0; k < 3; ++k) {
        int w, h, d;
        face_dims(k, gn, w, h, d);
        unflatten(u[k], w, h, d, grid[k]);
    }
}

// mapper.cpp:90-94
void mapper::mapG2P(std::vector<particle *> &particles, FLOAT ****grid, int gn)
{
    flip_ctx *c = context_for(gn);
    std::vector<float> u[3];
    for (int k = 0; k < 3; ++k) {
        int w, h, d;
        face_dims(k, gn, w, h, d);
        flatten(grid[k], w, h, d, u[k]);
    }
    size_t np = particles.size();
    std::vector<float> pos(3 * np), vel(3 * np);
    for (size_t i = 0; i < np; ++i)
        for (int a = 0; a < 3; ++a) pos[3 * i + a] = particles[i]->p[a];
    check(flip_gpu_map_g2p(c, u[0].data(), u[1].data(), u[2].data(), (int64_t)np, pos.data(), vel.data()),
          "flip_gpu_map_g2p");
    for (size_t i = 0; i < np; ++i)
        for (int a = 0; a < 3; ++a) particles[i]->u[a] = vel[3 * i + a];
}

// mapper.cpp:108-112 (one point: the same trilinear sampling as mapG2P)
void mapper::fetchVelocity(FLOAT p[3], FLOAT u[3], FLOAT ****grid, int gn)
{
    particle q;
    std::memset(&q, 0, sizeof(q));
    for (int a = 0; a < 3; ++a) q.p[a] = p[a];
    std::vector<particle *> one{&q};
    mapG2P(one, grid, gn);
    for (int a = 0; a < 3; ++a) u[a] = q.u[a];
}

// ------------------------------------------------------------------ solver
void solver::setSubcell(char value) { shim().subcell = value; }

// solver.cpp:224-241: negates b in place, builds the preconditioner and runs
// PCG from the warm start x; x receives the pressure.
void solver::solve(char ***A, FLOAT ***L, FLOAT ***x, FLOAT ***b, int n)
{
    flip_ctx *c = context_for(n);
    std::vector<char> a;
    std::vector<float> l, xv, bv;
    flatten(A, n, n, n, a);
    flatten(L, n, n, n, l);
    flatten(x, n, n, n, xv);
    flatten(b, n, n, n, bv);
    check(flip_gpu_solve(c, reinterpret_cast<const int8_t *>(a.data()), l.data(), xv.data(), bv.data(),
                         shim().subcell, nullptr),
          "flip_gpu_solve");
    unflatten(xv, n, n, n, x);
    unflatten(bv, n, n, n, b);
}

// --------------------------------------------------------------- corrector
// corrector.cpp:49-115 with re = DENSITY / gn (the only radius flip3D passes,
// flip3D.cpp:1242); dt is the caller's.
void corrector::correct(sorter *sort, std::vector<particle *> &particles, FLOAT dt, FLOAT re)
{
    if (particles.empty()) return;
#ifdef FLIP3D_SHIM_REFERENCE_TYPES
    const int gn = sort->getCellSize();
#else
    // without the reference's sorter class, the grid size follows from
    // re = DENSITY / gn with DENSITY = 0.5, the only radius flip3D passes
    (void)sort;
    const int gn = std::max(2, (int)std::lround(0.5 / (double)re));
#endif
    // context density with DENSITY / gn == re exactly ((double)re * gn is exact)
    const double density = (double)re * gn;
    flip_ctx *c = context_for(gn, density);
    upload(c, particles);
    check(flip_gpu_set_dt(c, (double)dt), "flip_gpu_set_dt");
    check(flip_gpu_stage(c, FLIP_STAGE_CORRECT, nullptr), "flip_gpu_stage(CORRECT)");
    size_t np = particles.size();
    std::vector<float> pos(3 * np), vel(3 * np), nrm(3 * np), dens(np);
    std::vector<int8_t> type(np);
    check(flip_gpu_download_particles(c, pos.data(), vel.data(), nrm.data(), type.data(), dens.data()),
          "flip_gpu_download_particles");
    for (size_t i = 0; i < np; ++i)
        for (int a = 0; a < 3; ++a) {
            particles[i]->p[a] = pos[3 * i + a];
            particles[i]->u[a] = vel[3 * i + a];
        }
}

// ------------------------------------------------------------------ flip3D
void flip3D::init(int load_step)
{
    Shim &s = shim();
    if (s.sim) flip_gpu_destroy(s.sim);
    s.sim = nullptr;
    flip_params p;
    check(flip_gpu_default_params(&p, s.sim_n), "flip_gpu_default_params");
    check(flip_gpu_create(&p, &s.sim), "flip_gpu_create");
    if (load_step > 0) {
        // flip3D.cpp:515-578 loadState
        char path[256];
        std::snprintf(path, sizeof(path), "render/save/%d_scene.dat", load_step);
        check(flip_gpu_load_state(s.sim, path), "flip_gpu_load_state");
        check(flip_gpu_set_step(s.sim, load_step), "flip_gpu_set_step");  // flip3D.cpp:577
        return;
    }
    int64_t np = 0;
    check(flip_gpu_build_scene(s.sim, s.sim_scene, 0, &np), "flip_gpu_build_scene");
}

void flip3D::simulateStep()
{
    Shim &s = shim();
    if (!s.sim) throw std::runtime_error("flip3D::simulateStep before flip3D::init");
    if (!s.out_preview && !s.out_mesh) {
        check(flip_gpu_step(s.sim, nullptr), "flip_gpu_step");
        return;
    }
    // flip3D.cpp:1194-1257: counter and pour, the preview frame, the step,
    // then the mesh files, under render/ as the reference names them
    check(flip_gpu_step_begin(s.sim), "flip_gpu_step_begin");
    int step = 0;
    check(flip_gpu_get_step(s.sim, &step), "flip_gpu_get_step");
    char path[256], tmpl[256];
    if (s.out_preview) {
        std::filesystem::create_directories("render/preview");
        std::snprintf(path, sizeof(path), "render/preview/frame_%d.bmp", step);
        check(flip_gpu_write_preview(s.sim, path), "flip_gpu_write_preview");
    }
    check(flip_gpu_step(s.sim, nullptr), "flip_gpu_step");
    if (s.out_mesh) {
        int64_t nv = 0, nf = 0;
        check(flip_gpu_mesh(s.sim, 0, &nv, &nf), "flip_gpu_mesh");
        std::filesystem::create_directories("render/obj");
        std::filesystem::create_directories("render/sunflow");
        std::snprintf(path, sizeof(path), "render/obj/%d_scene.obj", step);
        check(flip_gpu_write_obj(s.sim, path), "flip_gpu_write_obj");
        std::snprintf(path, sizeof(path), "render/sunflow/%d_scene.sc", step);
        std::snprintf(tmpl, sizeof(tmpl), "render/sunflow/template.sc");
        check(flip_gpu_write_sunflow(s.sim, path, tmpl), "flip_gpu_write_sunflow");
    }
}

// ------------------------------------------------------------- shim extras
void flip3d_shim::set_grid_size(int n) { shim().sim_n = n; }

void flip3d_shim::set_scene(int scene) { shim().sim_scene = scene; }

void flip3d_shim::set_outputs(bool preview, bool mesh)
{
    shim().out_preview = preview;
    shim().out_mesh = mesh;
}

flip_ctx *flip3d_shim::context() { return shim().sim; }

void flip3d_shim::download(std::vector<particle> &out)
{
    flip_ctx *c = shim().sim;
    if (!c) throw std::runtime_error("flip3d_shim::download before flip3D::init");
    int64_t np = 0;
    check(flip_gpu_num_particles(c, &np), "flip_gpu_num_particles");
    std::vector<float> pos(3 * np), vel(3 * np), nrm(3 * np), dens(np);
    std::vector<int8_t> type(np);
    check(flip_gpu_download_particles(c, pos.data(), vel.data(), nrm.data(), type.data(), dens.data()),
          "flip_gpu_download_particles");
    out.assign((size_t)np, particle());
    for (int64_t i = 0; i < np; ++i) {
        particle &p = out[i];
        std::memset(&p, 0, sizeof(p));
        for (int a = 0; a < 3; ++a) {
            p.p[a] = pos[3 * i + a];
            p.u[a] = vel[3 * i + a];
            p.n[a] = nrm[3 * i + a];
        }
        p.type = (char)type[i];
        p.visible = type[i] == FLIP_FLUID;
        p.m = 1.0f;
        p.dens = dens[i];
    }
}

void flip3d_shim::shutdown() { shim().release(); }
