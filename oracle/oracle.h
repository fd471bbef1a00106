/*
 * oracle.h - CPU restatement of the flip3d per-step FLIP path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path and the timed CPU baseline ("port") beside it.  Only tests/, the
 * smoke() entry of __graft_entry__.py and the cpu_baseline leg of bench.py may
 * load it; the product library never links or calls it.
 *
 * Every function restates the reference's arithmetic in the reference's order
 * (float/double mixing, loop order, serial sections), citing src/<file>:<line>
 * of mmmovania/flip3d.  Built twice: liboracle_f32.so (FLOAT = float, the
 * reference type) and liboracle_f64.so (FLOAT = double, the FP64 oracle used
 * for error statements).
 *
 * Layout: particles are structure-of-arrays in creation ("particle vector")
 * order; grids are flat [i][j][k] with k fastest, matching the jagged
 * alloc3D<T>(w,h,d) indexing of src/utility.h:40-51.
 */
#ifndef FLIP_ORACLE_H
#define FLIP_ORACLE_H

#include <stdint.h>

#ifdef ORACLE_F64
typedef double real;
#else
typedef float real;
#endif

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_AIR = 0, ORC_FLUID = 1, ORC_WALL = 2 };

/* Step policies.  "faithful" reproduces the reference's compile-time defaults;
 * "config" is the benchmark policy (CFL-capped dt, relative tolerance). */
typedef struct orc_params {
    int n;               /* grid resolution N (flip3D.cpp:25) */
    double dt;           /* DT (flip3D.cpp:31) or the CFL cap when dt_policy=1 */
    int dt_policy;       /* 0 fixed dt, 1 dt = min(dt, cfl*dx/max|u|) */
    double cfl;          /* 0.6 */
    double alpha;        /* ALPHA FLIP/PIC blend (flip3D.cpp:30) */
    double gravity;      /* GRAVITY (flip3D.cpp:33) */
    double density;      /* DENSITY (flip3D.cpp:32) */
    int subcell;         /* SUBCELL (flip3D.cpp:37) */
    int tol_policy;      /* 0 reference absolute stop (solver.cpp:323), 1 relative */
    double tol;          /* relative tolerance for tol_policy=1 */
    int precond;         /* 0 reference value-test MIC(0) form, 1 IC(0) keyed on flags */
    int max_iters;       /* <=0: n^3 cap as in solver.cpp:326 */
    int correction;      /* !DISABLE_CORRECTION (flip3D.cpp:39) */
    int reposition;      /* stuck-particle repositioning (flip3D.cpp:932-968) */
} orc_params;

typedef struct orc_stats {
    int iterations;      /* PCG iterations taken */
    double error2;       /* final sum r^2 over fluid cells */
    double bnorm2;       /* sum b^2 over fluid cells (b = -div) */
    double error2_0;     /* initial residual r.r */
    double dt;           /* dt used this step */
    int n_fluid_cells;
    int n_reposition;    /* particles repositioned */
    int n_jitter;        /* corrector random-jitter events */
    int n_precond_unit;  /* cells with P == 1.0 (value-test coupling) */
} orc_stats;

typedef struct orc_sim orc_sim;

void orc_default_params(orc_params *p, int n);
orc_sim *orc_create(const orc_params *prm);
void orc_destroy(orc_sim *s);

/* particles: p,u,nrm are 3*np interleaved (x,y,z per particle) */
int orc_set_particles(orc_sim *s, int np, const real *p, const real *u,
                      const real *nrm, const int8_t *type, const real *m,
                      const real *dens);
int orc_num_particles(const orc_sim *s);
void orc_get_particles(const orc_sim *s, real *p, real *u, real *nrm,
                       int8_t *type, real *dens);
void orc_set_max_dens(orc_sim *s, real v);
real orc_get_max_dens(const orc_sim *s);

/* grid ids for orc_get_grid/orc_set_grid */
enum { ORC_G_U0 = 0, ORC_G_U1, ORC_G_U2, ORC_G_US0, ORC_G_US1, ORC_G_US2,
       ORC_G_A, ORC_G_L, ORC_G_PRESS, ORC_G_DIV };
long orc_grid_size(const orc_sim *s, int which);
void orc_get_grid(const orc_sim *s, int which, void *out);
void orc_set_grid(orc_sim *s, int which, const void *in);
void orc_get_sorted(const orc_sim *s, int32_t *cell_start, int32_t *cell_pid);

/* rand(): the glibc random_r TYPE_3 stream, private to each sim */
void orc_srand(orc_sim *s, unsigned seed);
int orc_rand(orc_sim *s);

/* reference stages (file:line in oracle.c) */
void orc_sort(orc_sim *s);
void orc_compute_density(orc_sim *s);
void orc_add_ext_force(orc_sim *s);
void orc_map_p2g(orc_sim *s);
void orc_mark_water(orc_sim *s);
void orc_save_grid(orc_sim *s);
void orc_enforce_boundary(orc_sim *s);
void orc_project(orc_sim *s, orc_stats *st);
void orc_extrapolate_velocity(orc_sim *s);
void orc_subtract_grid(orc_sim *s);
void orc_flip_g2p(orc_sim *s);
void orc_solve_picflip(orc_sim *s, orc_stats *st);
void orc_advect(orc_sim *s, orc_stats *st);
void orc_correct(orc_sim *s, orc_stats *st);
void orc_step(orc_sim *s, orc_stats *st);
double orc_current_dt(const orc_sim *s);
void orc_set_dt(orc_sim *s, double dt);

/* stand-alone pressure solve on caller grids (solver.cpp:224-241):
 * x in/out warm start, b negated in place as in the reference. */
int orc_solve(const orc_params *prm, const int8_t *A, const real *L, real *x,
              real *b, orc_stats *st);
/* single kernels for known-answer checks */
real orc_smooth_kernel(real r2, real h);
real orc_sharp_kernel(real r2, real h);
real orc_length(const real *a, const real *b);
real orc_length2(const real *a, const real *b);
void orc_fetch_velocity(const orc_sim *s, int grid_set, const real *p, real *u);
void orc_compute_ax(const orc_params *prm, const int8_t *A, const real *L,
                    const real *x, real *ans);
void orc_build_precond(const orc_params *prm, const int8_t *A, const real *L,
                       double *P);
void orc_apply_precond(const orc_params *prm, const int8_t *A, const double *P,
                       const real *r, real *z);

/* scenes: reference damBreak placement (flip3D.cpp:277-297,580-647,649-764)
 * and the seeded benchmark scene.  Returns particle count; arrays sized by
 * orc_scene_capacity().  rand_state: sim whose rand() stream is consumed
 * (reference scene) or NULL. */
long orc_scene_capacity(int n);
real orc_calibrate_max_dens(int n, double density);
int orc_scene_reference(orc_sim *s);
/* the reference's test scenes: 0 damBreakTest, 1 spherePourTest, 2
 * cliffPourTest, 3 waterDropTest (pour scenes also set the emitter) */
int orc_scene_test(orc_sim *s, int which);
/* pourWater emitter (flip3D.cpp:375-402): disc centre (x, z), radius, last
 * step; orc_step advances the step counter and pours first */
void orc_set_emitter(orc_sim *s, real x, real z, real r, int pour_time);
void orc_set_step(orc_sim *s, int step);
int orc_get_step(const orc_sim *s);
int orc_pour(orc_sim *s);
/* mesher::generateMesh on the mg^3 lattice through the current buckets;
 * returns the vertex count, keeps the mesh for orc_mesh_get (one mesh per
 * process: the reference's mesher keeps static state too) */
int orc_mesh(orc_sim *s, int mg);
int orc_mesh_faces(void);
void orc_mesh_get(double *v, double *n, int32_t *f);
int orc_scene_benchmark(orc_sim *s, uint64_t seed);
void orc_compute_wall_normals(orc_sim *s);

#ifdef __cplusplus
}
#endif
#endif
