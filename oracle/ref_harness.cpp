// ref_harness.cpp - drives the reference flip3d (mmmovania/flip3d src/) stage
// by stage to produce golden vectors.  TEST INFRASTRUCTURE ONLY; built by
// oracle/Makefile into oracle/_ref/libflipref.so from the sources where they
// lie (nothing is copied).  flip3D.cpp is included in place so its file-static
// step functions and state are reachable; it fixes N at 32 (flip3D.cpp:25).
// The module functions (sorter, mapper, solver, corrector, kernel) take their
// grid size at run time and are driven at any n through the ref_kat_* calls.
//
// The process must run with a working directory that it may write render/
// into (flip3D::init creates render/log, and solver.cpp:209 logs there).
#include <cmath>
#include <limits>
#include <vector>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <cstdint>

#include "flip3D.cpp"

namespace {
template <class T> void fill_grid(T ***g, int w, int h, int d, T *out)
{
    for (int i = 0; i < w; i++)
        for (int j = 0; j < h; j++)
            for (int k = 0; k < d; k++) *out++ = g[i][j][k];
}
template <class T> void load_grid(T ***g, int w, int h, int d, const T *in)
{
    for (int i = 0; i < w; i++)
        for (int j = 0; j < h; j++)
            for (int k = 0; k < d; k++) g[i][j][k] = *in++;
}
void clear_particles()
{
    for (size_t n = 0; n < particles.size(); n++) delete particles[n];
    particles.clear();
}
}  // namespace

extern "C" {

int ref_grid_n() { return N; }

// flip3D::init(0): allocation, max_dens calibration, damBreak scene, walls,
// wall normals.  Returns the particle count.
int ref_init()
{
    flip3D::init(0);
    return (int)particles.size();
}

void ref_srand(unsigned s) { srand(s); }
int ref_rand() { return rand(); }
float ref_get_max_dens() { return max_dens; }
void ref_set_max_dens(float v) { max_dens = v; }
int ref_num_particles() { return (int)particles.size(); }

void ref_get_particles(float *p, float *u, float *nrm, int8_t *type, float *dens)
{
    for (size_t n = 0; n < particles.size(); n++) {
        particle *q = particles[n];
        for (int c = 0; c < 3; c++) {
            if (p) p[3 * n + c] = q->p[c];
            if (u) u[3 * n + c] = q->u[c];
            if (nrm) nrm[3 * n + c] = q->n[c];
        }
        if (type) type[n] = q->type;
        if (dens) dens[n] = q->dens;
    }
}

// Replace the particle vector (the loadState injection route, flip3D.cpp:540-547).
void ref_set_particles(int np, const float *p, const float *u, const float *nrm, const int8_t *type,
                       const float *dens)
{
    clear_particles();
    for (int n = 0; n < np; n++) {
        particle *q = new particle;
        memset(q, 0, sizeof(particle));
        for (int c = 0; c < 3; c++) {
            q->p[c] = p[3 * n + c];
            q->u[c] = u ? u[3 * n + c] : 0.0f;
            q->n[c] = nrm ? nrm[3 * n + c] : 0.0f;
        }
        q->type = type[n];
        q->visible = 1;
        q->m = 1.0f;
        q->dens = dens ? dens[n] : 10.0f;
        particles.push_back(q);
    }
}

void ref_compute_wall_normal() { compute_wall_normal(); }

// flip3D::init(0) with one of the reference's test scenes in place of the
// damBreakTest that placeObjects (flip3D.cpp:365-373) hard-wires:
// 0 damBreakTest, 1 spherePourTest, 2 cliffPourTest, 3 waterDropTest.
// init(0) runs first for its allocation and max_dens calibration (no rand()
// draws); its particles are then discarded and the placement of
// flip3D.cpp:711-764 is driven again through the reference's own
// pushParticle, sorter, markWater and compute_wall_normal, with the rand()
// stream reseeded to `seed` before the placement.
int ref_init_test(int which, unsigned seed)
{
    // init's calibration runs over whatever the particle vector holds
    clear_particles();
    objects.clear();
    flip3D::init(0);
    clear_particles();
    objects.clear();
    pourTime = -1;
    step = 0;
    placeWalls();
    switch (which) {
    case 0: damBreakTest(); break;
    case 1: spherePourTest(); break;
    case 2: cliffPourTest(); break;
    case 3: waterDropTest(); break;
    case 4: {
        // the benchmark scene's liquid (box x<0.4, y<0.6, all z, plus a
        // sphere r=0.12 at (0.6,0.7,0.5)) as reference Objects, so that the
        // reference's own object test (pushParticle, flip3D.cpp:580-633)
        // decides which sub-lattice points hold liquid
        Object obj;
        obj.type = FLUID;
        obj.shape = BOX;
        obj.visible = true;
        obj.p[0][0] = 0.0; obj.p[1][0] = 0.4;
        obj.p[0][1] = 0.0; obj.p[1][1] = 0.6;
        obj.p[0][2] = 0.0; obj.p[1][2] = 1.0;
        objects.push_back(obj);
        obj.shape = SPHERE;
        obj.c[0] = 0.6; obj.c[1] = 0.7; obj.c[2] = 0.5;
        obj.r = 0.12;
        objects.push_back(obj);
        break;
    }
    default: return -1;
    }
    srand(seed);
    const double lw = DENSITY * WALL_THICK;
    const int lim = (int)(N / DENSITY);
    for (int a = 0; a < lim; a++)
        for (int b = 0; b < lim; b++)
            for (int c = 0; c < lim; c++) {
                double x = a * lw + lw / 2.0, y = b * lw + lw / 2.0, z = c * lw + lw / 2.0;
                bool in = x > WALL_THICK && x < 1.0 - WALL_THICK && y > WALL_THICK && y < 1.0 - WALL_THICK &&
                          z > WALL_THICK && z < 1.0 - WALL_THICK;
                if (in) pushParticle(x, y, z, FLUID);
            }
    const double cw = 1.0 / N;
    for (int a = 0; a < N; a++)
        for (int b = 0; b < N; b++)
            for (int c = 0; c < N; c++) pushParticle(a * cw + cw / 2.0, b * cw + cw / 2.0, c * cw + cw / 2.0, WALL);
    sort->sort(particles);
    sort->markWater(A, wall_normal[3], DENSITY);
    std::vector<particle *> kept;
    for (size_t n = 0; n < particles.size(); n++) {
        particle *q = particles[n];
        int ci = fmin(N - 1, fmax(0, q->p[0] * N)), cj = fmin(N - 1, fmax(0, q->p[1] * N)),
            ck = fmin(N - 1, fmax(0, q->p[2] * N));
        if (q->type != WALL && A[ci][cj][ck] == WALL) {
            delete q;
            continue;
        }
        kept.push_back(q);
    }
    particles = kept;
    compute_wall_normal();
    return (int)particles.size();
}

// mesher::generateMesh as write_mesh calls it (flip3D.cpp:1002), through the
// sorter's current buckets; mg stays fixed per process (the mesher and
// CIsoSurface keep static arrays sized at their first call)
int ref_mesh(int mg)
{
    mesher::generateMesh(A, sort, particles, DENSITY, mg, vertices, normals, faces);
    return (int)(vertices.size() / 3);
}
int ref_mesh_faces() { return (int)(faces.size() / 3); }
void ref_mesh_get(double *v, double *nr, int32_t *f)
{
    for (size_t i = 0; i < vertices.size(); i++) v[i] = vertices[i];
    for (size_t i = 0; i < normals.size(); i++) nr[i] = normals[i];
    for (size_t i = 0; i < faces.size(); i++) f[i] = faces[i];
}

// The reference's per-step output files as simulateStep writes them, for
// frame `at`: the preview BMP (write_frame, flip3D.cpp:1133-1192, into
// render/preview/frame_<at>.bmp) and the mesh files (write_mesh, :999-1009:
// render/obj/<at>_scene.obj, render/sunflow/<at>_scene.sc and template.sc)
void ref_write_outputs(int at)
{
    step = at;
    write_frame();
    write_mesh();
}

// simulateStep's step counter and pourWater (flip3D.cpp:1196-1199)
void ref_set_step(int s) { step = s; }
int ref_get_step() { return step; }
int ref_pour()
{
    size_t before = particles.size();
    pourWater(pourTime);
    return (int)(particles.size() - before);
}

// grids: 0..2 u, 3..5 u_save, 6 A, 7 L, 8 press
void ref_get_grid(int which, void *out)
{
    switch (which) {
    case 0: fill_grid(u[0], N + 1, N, N, (float *)out); break;
    case 1: fill_grid(u[1], N, N + 1, N, (float *)out); break;
    case 2: fill_grid(u[2], N, N, N + 1, (float *)out); break;
    case 3: fill_grid(u_save[0], N + 1, N, N, (float *)out); break;
    case 4: fill_grid(u_save[1], N, N + 1, N, (float *)out); break;
    case 5: fill_grid(u_save[2], N, N, N + 1, (float *)out); break;
    case 6: fill_grid(A, N, N, N, (char *)out); break;
    case 7: fill_grid(L, N, N, N, (float *)out); break;
    case 8: fill_grid(press, N, N, N, (float *)out); break;
    }
}

void ref_set_press(const float *in) { load_grid(press, N, N, N, in); }

// Stages of flip3D::simulateStep (flip3D.cpp:1216-1247), called in order.
void ref_stage_density()
{
    sort->sort(particles);  // :1221
    computeDensity();       // :1222
}
void ref_stage_ext_force() { add_ExtForce(); }  // :1230
// P2G probe: the first two calls of solve_picflip (:1059-1060) plus
// markWater (:1061); solve_picflip recomputes all three from the same state.
void ref_probe_p2g()
{
    sort->sort(particles);
    mapper::mapP2G(sort, particles, u, N);
    sort->markWater(A, wall_normal[3], DENSITY);
}
void ref_stage_solve_picflip() { solve_picflip(); }  // :1231
void ref_stage_advect() { advect_particle(); }        // :1236
int ref_num_stuck() { return gNumStuck; }
void ref_stage_correct() { corrector::correct(sort, particles, DT, DENSITY / N); }  // :1242
void ref_step()
{
    ref_stage_density();
    ref_stage_ext_force();
    ref_stage_solve_picflip();
    ref_stage_advect();
    ref_stage_correct();
}
// saveState (flip3D.cpp:453-513) as step `at`: writes render/save/<at>_scene.dat
// under the working directory
void ref_save_state(int at)
{
    step = at;
    saveState();
}

// ---- run-time-n module KATs (sorter / mapper / corrector / solver) ----
static std::vector<particle *> make_particles(int np, const float *p, const float *uu, const float *nrm,
                                              const int8_t *type)
{
    std::vector<particle *> v;
    for (int n = 0; n < np; n++) {
        particle *q = new particle;
        memset(q, 0, sizeof(particle));
        for (int c = 0; c < 3; c++) {
            q->p[c] = p[3 * n + c];
            q->u[c] = uu ? uu[3 * n + c] : 0.0f;
            q->n[c] = nrm ? nrm[3 * n + c] : 0.0f;
        }
        q->type = type[n];
        q->m = 1.0f;
        q->dens = 1.0f;
        v.push_back(q);
    }
    return v;
}
static void free_particles(std::vector<particle *> &v)
{
    for (size_t n = 0; n < v.size(); n++) delete v[n];
    v.clear();
}

// sorter::sort bucket contents: for every cell in [i][j][k] order, the
// particle indices in bucket order.  counts: gn^3, ids: np.
void ref_kat_sort(int gn, int np, const float *p, const int8_t *type, int32_t *counts, int32_t *ids)
{
    std::vector<particle *> v = make_particles(np, p, NULL, NULL, type);
    sorter s(gn);
    s.sort(v);
    int at = 0;
    for (int i = 0; i < gn; i++)
        for (int j = 0; j < gn; j++)
            for (int k = 0; k < gn; k++) {
                std::vector<particle *> c = s.getNeigboringParticles_cell(i, j, k, 0, 0, 0);
                counts[(i * gn + j) * gn + k] = (int32_t)c.size();
                for (size_t a = 0; a < c.size(); a++) {
                    for (int n = 0; n < np; n++)
                        if (v[n] == c[a]) { ids[at++] = n; break; }
                }
            }
    free_particles(v);
}

// mapper::mapP2G at grid size gn
void ref_kat_p2g(int gn, int np, const float *p, const float *uu, const int8_t *type, float *g0, float *g1,
                 float *g2)
{
    std::vector<particle *> v = make_particles(np, p, uu, NULL, type);
    sorter s(gn);
    s.sort(v);
    FLOAT ***g[3] = {alloc3D<FLOAT>(gn + 1, gn, gn), alloc3D<FLOAT>(gn, gn + 1, gn), alloc3D<FLOAT>(gn, gn, gn + 1)};
    mapper::mapP2G(&s, v, g, gn);
    fill_grid(g[0], gn + 1, gn, gn, g0);
    fill_grid(g[1], gn, gn + 1, gn, g1);
    fill_grid(g[2], gn, gn, gn + 1, g2);
    for (int c = 0; c < 3; c++) free3D(g[c]);
    free_particles(v);
}

// mapper::fetchVelocity for np points on caller grids (fluid-range points only:
// the reference reads past grid[1]'s rows for z in the top half cell)
void ref_kat_g2p(int gn, const float *g0, const float *g1, const float *g2, int np, const float *p, float *uo)
{
    FLOAT ***g[3] = {alloc3D<FLOAT>(gn + 1, gn, gn), alloc3D<FLOAT>(gn, gn + 1, gn), alloc3D<FLOAT>(gn, gn, gn + 1)};
    load_grid(g[0], gn + 1, gn, gn, g0);
    load_grid(g[1], gn, gn + 1, gn, g1);
    load_grid(g[2], gn, gn, gn + 1, g2);
    for (int n = 0; n < np; n++) {
        FLOAT pp[3] = {p[3 * n], p[3 * n + 1], p[3 * n + 2]};
        mapper::fetchVelocity(pp, uo + 3 * n, g, gn);
    }
    for (int c = 0; c < 3; c++) free3D(g[c]);
}

// corrector::correct at grid size gn (rand() stream consumed as in the reference)
void ref_kat_correct(int gn, int np, float *p, float *uu, const float *nrm, const int8_t *type, float dt, float re)
{
    std::vector<particle *> v = make_particles(np, p, uu, nrm, type);
    sorter s(gn);
    corrector::correct(&s, v, dt, re);
    for (int n = 0; n < np; n++)
        for (int c = 0; c < 3; c++) {
            p[3 * n + c] = v[n]->p[c];
            uu[3 * n + c] = v[n]->u[c];
        }
    free_particles(v);
}

// solver::solve at size n (static scratch: one n per process, solver.cpp:226)
void ref_kat_solve(int n, const int8_t *Af, const float *Lf, float *x, float *b, int subcell)
{
    char ***AA = alloc3D<char>(n, n, n);
    FLOAT ***LL = alloc3D<FLOAT>(n, n, n), ***XX = alloc3D<FLOAT>(n, n, n), ***BB = alloc3D<FLOAT>(n, n, n);
    load_grid(AA, n, n, n, (const char *)Af);
    load_grid(LL, n, n, n, Lf);
    load_grid(XX, n, n, n, x);
    load_grid(BB, n, n, n, b);
    solver::setSubcell((char)subcell);
    solver::solve(AA, LL, XX, BB, n);
    fill_grid(XX, n, n, n, x);
    fill_grid(BB, n, n, n, b);
    free3D(AA); free3D(LL); free3D(XX); free3D(BB);
}

void ref_kat_kernels(int m, const float *r2, const float *h, float *smooth, float *sharp)
{
    for (int i = 0; i < m; i++) {
        smooth[i] = kernel::smooth_kernel(r2[i], h[i]);
        sharp[i] = kernel::sharp_kernel(r2[i], h[i]);
    }
}

void ref_kat_length(int m, const float *a, const float *b, float *len, float *len2)
{
    for (int i = 0; i < m; i++) {
        FLOAT pa[3] = {a[3 * i], a[3 * i + 1], a[3 * i + 2]}, pb[3] = {b[3 * i], b[3 * i + 1], b[3 * i + 2]};
        len[i] = length(pa, pb);
        len2[i] = length2(pa, pb);
    }
}

}  // extern "C"
