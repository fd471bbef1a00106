/*
 * flip3d_gpu.h - C ABI of the MI355X (gfx950) FLIP step library.
 *
 * Drop-in boundary for the per-step path of mmmovania/flip3d.  The reference
 * exposes plain C++ namespace functions (src/flip3D.h:9-12, src/mapper.h:11-14,
 * src/solver.h:9-12, src/sorter.h:11-29, src/corrector.h:10-13); each entry
 * below names the reference interface it replaces.  All state is
 * device-resident behind an opaque context; calls are synchronous at return,
 * like the reference's.  Every function returns 0 on success or a negative
 * FLIP_E* code, with a message from flip_gpu_last_error(); nothing throws.
 *
 * Layout conventions (identical to the reference's alloc3D<T>(w,h,d) jagged
 * arrays flattened in [i][j][k] order, k fastest, src/utility.h:40-51):
 *   cell grids  A (int8), L, press, div : n*n*n
 *   face grids  u0 : (n+1)*n*n, u1 : n*(n+1)*n, u2 : n*n*(n+1)
 * Particles cross the boundary as structure-of-arrays in the reference's
 * particle-vector order (src/flip3D.cpp:51); positions, velocities and
 * normals are interleaved xyz triples.
 */
#ifndef FLIP3D_GPU_H
#define FLIP3D_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLIP_OK 0
#define FLIP_E_ARG (-1)      /* bad argument */
#define FLIP_E_HIP (-2)      /* HIP runtime error (no device, out of memory, launch) */
#define FLIP_E_STATE (-3)    /* call out of order (e.g. step before upload) */
#define FLIP_E_NUMERIC (-4)  /* solver produced NaN/Inf or hit its iteration cap */

/* cell flags, src/common.h:11-13 */
#define FLIP_AIR 0
#define FLIP_FLUID 1
#define FLIP_WALL 2

/* grid ids for flip_gpu_get_grid / flip_gpu_set_grid */
#define FLIP_GRID_U0 0
#define FLIP_GRID_U1 1
#define FLIP_GRID_U2 2
#define FLIP_GRID_US0 3 /* u_save (flip3D.cpp:46): raw P2G, then the FLIP delta */
#define FLIP_GRID_US1 4
#define FLIP_GRID_US2 5
#define FLIP_GRID_A 6
#define FLIP_GRID_L 7
#define FLIP_GRID_PRESS 8
#define FLIP_GRID_DIV 9

/* preconditioners */
#define FLIP_PRECOND_REFERENCE 0 /* solver.cpp:119-180 as written (value-test P_ref) */
#define FLIP_PRECOND_IC0 1       /* same factorisation with couplings keyed on flags */
#define FLIP_PRECOND_LEARNED 2   /* flexible CG, two-grid preconditioner with the U-Net as coarse solver;
                                    the reference PCG then checks (and if needed finishes) the result */

/* Step parameters.  flip_gpu_default_params fills the reference's compile-time
 * defaults (src/flip3D.cpp:25-39, mapper.cpp:14, corrector.cpp:13,
 * solver.cpp:16). */
typedef struct flip_params {
    int n;              /* grid resolution N */
    double dt;          /* DT, or the cap when dt_policy == 1 */
    int dt_policy;      /* 0 fixed dt; 1 dt = min(dt, cfl / (N max|u_fluid|)) */
    double cfl;         /* 0.6 */
    double alpha;       /* ALPHA, FLIP/PIC blend 0.95 */
    double gravity;     /* GRAVITY 9.8 */
    double density;     /* DENSITY 0.5 (2x2x2 particles per cell) */
    int subcell;        /* SUBCELL ghost-fluid boundary, 1 */
    int tol_policy;     /* 0 reference stop sum r^2 <= 1e-2 n^3; 1 ||r|| <= tol ||b|| */
    double tol;         /* relative tolerance for tol_policy 1 */
    int precond;        /* FLIP_PRECOND_* */
    int max_iters;      /* <= 0: n^3 as in solver.cpp:326 */
    int correction;     /* !DISABLE_CORRECTION */
    int reposition;     /* stuck-particle repositioning, flip3D.cpp:932-968 */
    unsigned rand_seed; /* glibc-compatible rand() stream seed; the reference's is 1 */
    int refine;         /* > 0: at most this many FP64-residual iterative-refinement sweeps: the
                           residual b - A x is formed in FP64 on an FP64 copy of x, each sweep
                           solves A d = r with the FP32 PCG to relative tol (tol_policy 1
                           required) and adds d to the FP64 x */
    double refine_tol;  /* outer relative tolerance ||b - A x|| <= refine_tol ||b|| for refine > 0 */
    int solve_fp64;     /* 1: the pressure solve runs entirely in FP64 (Jacobi-preconditioned CG on
                           the same operator, stop at relative tol; tol_policy 1 required); the
                           error reference for the FP32 and refined solves.  No counterpart in
                           the reference, whose PCG is FP32 with FP64 reductions. */
} flip_params;

typedef struct flip_step_stats {
    int iterations;      /* PCG iterations of this step's solve */
    double error2;       /* final sum r^2 over fluid cells */
    double bnorm2;       /* sum b^2 over fluid cells */
    double error2_0;     /* initial sum r^2 */
    double dt;           /* dt used */
    int n_fluid_cells;
    int n_reposition;    /* particles repositioned (flip3D.cpp:405-451) */
    int n_jitter;        /* corrector random-jitter events (corrector.cpp:73-77) */
    int n_precond_unit;  /* cells with P == 1.0, which enable value-test couplings */
    int refine_sweeps;   /* outer refinement sweeps used */
    double ms_step;      /* device time of the step, HIP events */
    double ms_solve;     /* device time of the pressure solve */
    int learned_iterations; /* FLIP_PRECOND_LEARNED: flexible-CG iterations before the plain-PCG check */
} flip_step_stats;

typedef struct flip_ctx flip_ctx;

const char *flip_gpu_last_error(void);
int flip_gpu_default_params(flip_params *p, int n);

/* Context lifetime.  Replaces flip3D::init's allocation half
 * (src/flip3D.cpp:649-683) and the static solver scratch (solver.cpp:226). */
int flip_gpu_create(const flip_params *p, flip_ctx **out);
int flip_gpu_destroy(flip_ctx *ctx);
int flip_gpu_synchronize(flip_ctx *ctx);

/* Particle state in reference order.  pos/vel/nrm: 3*np floats; type: FLIP_FLUID
 * or FLIP_WALL; mass, dens may be NULL (1.0, 10.0 as pushParticle sets,
 * flip3D.cpp:641-644).  Replaces the loadState particle read (flip3D.cpp:540-547). */
int flip_gpu_upload_particles(flip_ctx *ctx, int64_t np, const float *pos, const float *vel,
                              const float *nrm, const int8_t *type, const float *mass,
                              const float *dens);
int flip_gpu_download_particles(flip_ctx *ctx, float *pos, float *vel, float *nrm, int8_t *type,
                                float *dens);
int flip_gpu_num_particles(flip_ctx *ctx, int64_t *np);

/* pourWater (flip3D.cpp:375-402) and the step counter it reads (flip3D.cpp:59,
 * 1196).  flip_gpu_step first advances the counter; while it is <= pour_time
 * the step then appends a disc of FLUID particles (centre (x, z), radius r, on
 * the DENSITY/N lattice just under the ceiling wall, velocity -0.5 DENSITY/N/dt
 * along y, dens = max_dens) before anything else, as simulateStep does.
 * pour_time < 0 switches the emitter off.  Particle storage grows as needed.
 * set_step is loadState's step = load_step (flip3D.cpp:577). */
int flip_gpu_set_emitter(flip_ctx *ctx, float x, float z, float r, int pour_time);

/* Surface mesh.  Replaces mesher::generateMesh (src/mesher.cpp:137-195, with
 * implicit.cpp:14-38 and rchandra/CIsoSurface.cpp:331-777) as write_mesh calls
 * it (flip3D.cpp:1002): implicit distance field on an mg^3 lattice over the unit
 * cube (mg <= 0: N + 1), marching cubes at iso level 0, faces with an edge of
 * 0.1 or longer dropped, two smoothing passes.  The field reads the particles
 * through the buckets of the last sort, as the reference's mesher reads its
 * sorter (after a step: the corrector's sort, with the corrected positions).
 * flip_gpu_mesh builds the mesh on the device and returns its counts;
 * flip_gpu_mesh_get copies it out: vertices and normals as 3*nvert doubles
 * (the reference's vector<double>), faces as 3*nface vertex indices (any
 * pointer may be NULL). */
int flip_gpu_mesh(flip_ctx *ctx, int mg, int64_t *nvert, int64_t *nface);
int flip_gpu_mesh_get(flip_ctx *ctx, double *vertices, double *normals, int32_t *faces);

/* The reference's per-frame output files, byte for byte.
 * flip_gpu_render_preview: write_frame's point-splat preview
 *   (src/flip3D.cpp:1105-1192) of the current FLUID particles, rendered on the
 *   GPU, as 1024 x 700 RGBA bytes (rows in write_frame's order);
 * flip_gpu_write_preview: the same written by write_bmp (write_bmp.cpp:52-118);
 * flip_gpu_write_obj / flip_gpu_write_sunflow: the last flip_gpu_mesh as
 *   write_obj (exporter.cpp:127-164) / writeSunflow (exporter.cpp:72-125) write
 *   it, with wall thickness 1/N; template_path (may be NULL) also receives the
 *   template.sc the scene file includes.
 * flip_gpu_step_begin: the first part of simulateStep (flip3D.cpp:1196-1199:
 *   the step counter and pourWater) on its own, so that a caller can write the
 *   preview where the reference does, before the rest of the step; the next
 *   flip_gpu_step then skips that part. */
int flip_gpu_render_preview(flip_ctx *ctx, uint8_t *rgba);
int flip_gpu_write_preview(flip_ctx *ctx, const char *path);
int flip_gpu_write_obj(flip_ctx *ctx, const char *path);
int flip_gpu_write_sunflow(flip_ctx *ctx, const char *path, const char *template_path);
int flip_gpu_step_begin(flip_ctx *ctx);
int flip_gpu_set_step(flip_ctx *ctx, int step);
int flip_gpu_get_step(flip_ctx *ctx, int *step);
int flip_gpu_set_max_dens(flip_ctx *ctx, float max_dens); /* flip3D.cpp:58 */
int flip_gpu_get_max_dens(flip_ctx *ctx, float *max_dens);
int flip_gpu_srand(flip_ctx *ctx, unsigned seed);          /* stdlib srand */
int flip_gpu_set_dt(flip_ctx *ctx, double dt);             /* dt for stage calls */

/* Grids (see FLIP_GRID_*): A is int8, all others float. */
int flip_gpu_get_grid(flip_ctx *ctx, int which, void *out);
int flip_gpu_set_grid(flip_ctx *ctx, int which, const void *in);

/* One simulation step.  Replaces flip3D::simulateStep (flip3D.cpp:1194-1264),
 * its timed region :1216-:1247 (no preview, mesh or checkpoint I/O). */
int flip_gpu_step(flip_ctx *ctx, flip_step_stats *stats);

/* Stages of the step, in order, for parity harnesses:
 *   DENSITY    sorter::sort + computeDensity       flip3D.cpp:1221-1222
 *   EXT_FORCE  add_ExtForce                        flip3D.cpp:1230
 *   P2G        sort + mapper::mapP2G + markWater + save_grid + enforce_boundary
 *                                                  flip3D.cpp:1059-1065
 *   PROJECT    project + enforce_boundary          flip3D.cpp:1066-1067
 *   EXTRAP     extrapolate_velocity + subtract_grid flip3D.cpp:1068-1069
 *   G2P        FLIP/PIC blend via mapper::mapG2P   flip3D.cpp:1071-1097
 *   ADVECT     advect_particle (push-out, stuck, reposition) flip3D.cpp:877-969
 *   CORRECT    corrector::correct                  corrector.cpp:49-115 */
#define FLIP_STAGE_DENSITY 0
#define FLIP_STAGE_EXT_FORCE 1
#define FLIP_STAGE_P2G 2
#define FLIP_STAGE_PROJECT 3
#define FLIP_STAGE_EXTRAP 4
#define FLIP_STAGE_G2P 5
#define FLIP_STAGE_ADVECT 6
#define FLIP_STAGE_CORRECT 7
int flip_gpu_stage(flip_ctx *ctx, int stage, flip_step_stats *stats);

/* mapper::mapP2G (mapper.cpp:17-88) on the context's particles: sorts, then
 * writes the three face grids (u0,u1,u2 layouts above); any pointer may be NULL. */
int flip_gpu_map_p2g(flip_ctx *ctx, float *u0, float *u1, float *u2);
/* mapper::mapG2P / fetchVelocity (mapper.cpp:90-112): trilinear samples of the
 * given face grids at np positions. */
int flip_gpu_map_g2p(flip_ctx *ctx, const float *u0, const float *u1, const float *u2, int64_t np,
                     const float *pos, float *vel);
/* solver::setSubcell + solver::solve (solver.cpp:220-241) on caller grids of the
 * context's size n: x is the warm start and the result; b is negated in place as
 * the reference does (solver.cpp:231). */
int flip_gpu_solve(flip_ctx *ctx, const int8_t *A, const float *L, float *x, float *b, int subcell,
                   flip_step_stats *stats);
/* sorter::sort (sorter.cpp:22-40) result in reference bucket order: counts per
 * cell (n^3) and particle indices (reference order) cell by cell (np). */
int flip_gpu_sorted_cells(flip_ctx *ctx, int32_t *counts, int32_t *ids);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the buckets of flip_gpu_sorted_cells, computed by the same
 * kernels with no context and no allocation.  All pointers are device memory
 * owned by the caller: pos holds 3*np floats (xyz per particle, the particle
 * order is the creation order), counts receives n^3 per-cell counts and ids
 * the np particle indices cell by cell in creation order.  Scratch is a fixed
 * array of 12.3 MiB per GPU inside the library (held until the process ends),
 * so np <= 1048576 and n^3 <= 1048576; larger sizes fail with FLIP_E_ARG.  The
 * kernels are queued on `stream` (a hipStream_t; NULL = the null stream) and
 * the call returns without waiting.  Every call on a GPU shares that GPU's
 * scratch: calls that overlap on one GPU race. */
int flip_gpu_sort_cells_noalloc(const float *pos, int64_t np, int n, int32_t *counts, int32_t *ids, void *stream);

/* Per-kernel device timings of the last step (ms), in launch-name order;
 * names is filled with a ';'-separated list.  Returns the count. */
int flip_gpu_kernel_times(flip_ctx *ctx, char *names, int names_cap, double *ms, int ms_cap);
/* Enable/disable per-kernel timing of the next steps: single-kernel stages record
 * their dispatch begin/end timestamps, multi-kernel stages an event pair. */
int flip_gpu_set_profiling(flip_ctx *ctx, int on);

/* Scene generation (host, once; flip3D::init placement, flip3D.cpp:685-764).
 * scene 0: the reference damBreakTest with rand() jitter (consumes the
 * context's rand stream); scene 1: the benchmark scene (box x<0.4,y<0.6 plus
 * sphere r=0.12 at (0.6,0.7,0.5), SplitMix64 jitter U(+-0.25dx)).  Uploads the
 * particles and max_dens into the context; returns the particle count via np. */
#define FLIP_SCENE_REFERENCE 0
#define FLIP_SCENE_BENCHMARK 1
#define FLIP_SCENE_RANDOM 2 /* 1-4 liquid boxes (edges U(0.1,0.6)) / spheres (r U(0.05,0.2)), seeded */
#define FLIP_SCENE_RANDOM_BOXES 3 /* 1-3 liquid boxes (edges U(0.1,0.6)), seeded (ensemble scenes) */
/* The reference's other test scenes, placed like scene 0 (rand() jitter):
 * spherePourTest (flip3D.cpp:203-228): a wall sphere over a thin pool, pouring;
 * cliffPourTest (flip3D.cpp:249-275): a wall box cliff beside a thin pool, pouring;
 * waterDropTest (flip3D.cpp:230-247): a pool and a liquid sphere above it.
 * Every scene resets the step counter; the pour scenes set the emitter as
 * their reference functions do (200 steps), the others switch it off. */
#define FLIP_SCENE_SPHERE_POUR 4
#define FLIP_SCENE_CLIFF_POUR 5
#define FLIP_SCENE_WATER_DROP 6
int flip_gpu_build_scene(flip_ctx *ctx, int scene, uint64_t seed, int64_t *np);

/* Checkpoints in the reference's render/save/<step>_scene.dat layout
 * (saveState flip3D.cpp:453-513, loadState :515-578): int count, float
 * max_dens, int pourTime, float pourPos[2], float pourRad, count 72-byte
 * particle records (src/common.h:34-45) in creation order, the pressure grid
 * (n^3 floats, [i][j][k]), four wall-normal floats per cell.  Saving writes
 * the context's emitter (pour time, disc position and radius) and the
 * wall-normal grid as compute_wall_normal (flip3D.cpp:75-164) builds it; on
 * the context's first save it also writes objects.dat beside the file, the
 * scene's object list in saveState's text format (flip3D.cpp:498-508; fields
 * the reference leaves indeterminate are written as 0).  Loading takes the
 * particles, max_dens, pressure and emitter, checks the file size against n,
 * takes the wall normals from the particle records (the file's wall-normal
 * grid is derived from them), and on the context's first load reads
 * objects.dat from beside the file when it is present. */
int flip_gpu_save_state(flip_ctx *ctx, const char *path);
int flip_gpu_load_state(flip_ctx *ctx, const char *path);

/* Weights file (flip_unet_save layout) of the coarse-level network used by
 * FLIP_PRECOND_LEARNED; the network solves the coarsest level of the
 * multigrid hierarchy (32^3 for n = 64, 128 and 256; 16^3 for n = 32).  The
 * levels halve n rounded up to a multiple of 16, the coarse cells past the
 * grid's end being WALL, so any n works.  No counterpart in the reference. */
int flip_gpu_set_learned_weights(flip_ctx *ctx, const char *path);

/* Launch shape of the pressure solver's FLUID-list kernels: pcg_blocks caps
 * their block count (1..256; 0 = the built-in min(ceil(n^3 / 1024), 256)),
 * pcg_batch sets the PCG iterations per captured graph replay (1..64; 0 =
 * 8).  The block count fixes how the dot products' per-block partials are
 * grouped, so results can differ in the last bits from the default; the
 * batch size changes nothing but the launch pattern.  The values the
 * autotuner (flip3d_amd/autotune.py) chose per grid size are in
 * flip3d_amd/tuning.json.  No counterpart in the reference. */
int flip_gpu_set_launch(flip_ctx *ctx, int pcg_blocks, int pcg_batch);

/* Step graphs (on by default): flip_gpu_step replays the kernels before the
 * pressure solve (time step through divergence) and after it (gradient
 * through the corrector) from HIP graphs captured on first use, and
 * captures them again when the particle storage, the wall bins or a kernel
 * parameter change; the solve between them keeps its own captured
 * iterations.  Steps of the pourWater emitter, profiled steps and contexts
 * created with FLIP_DIRECT_LAUNCH=1 launch directly.  Results do not depend
 * on the setting.  No counterpart in the reference. */
int flip_gpu_set_step_graphs(flip_ctx *ctx, int on);

/* Ensembles of independent scenes.  Steps every context of ctxs once (each
 * exactly as flip_gpu_step would), with `threads` host threads driving the
 * contexts concurrently, each context on its own stream; returns when all
 * are done.  stats and ms (the wall-clock of each context's step) may be
 * NULL.  On failure returns the first failing context's code and message.
 * No counterpart in the reference, which simulates one scene per process. */
int flip_gpu_step_many(flip_ctx **ctxs, int count, int threads, flip_step_stats *stats, double *ms);

/* Ensembles in one launch per part of the step: steps every context of ctxs
 * once, as flip_gpu_step would, with the kernels before the pressure solve
 * of all contexts replayed as ONE graph launch (each context's captured part
 * a child graph, the contexts independent of each other), then the solves
 * (each on its context's stream, `threads` host threads driving them), then
 * the kernels after the solves of all contexts as one more graph launch.
 * Every context must be able to use step graphs (flip_gpu_set_step_graphs;
 * no emitter step, profiling, direct launch, refinement or FP64 solve).
 * ms[i] (optional) and stats[i].ms_step are the batch's wall time.  Calls
 * are serialised (one batch's graphs are kept and rebuilt when the contexts
 * or their captures change).  No counterpart in the reference. */
int flip_gpu_step_batch(flip_ctx **ctxs, int count, int threads, flip_step_stats *stats, double *ms);

/* Device math self-check for the ops whose rounding the parity argument
 * relies on: out[4*i+0] = x/y, +1 = sqrtf(x), +2 = (float)sqrt((double)x*x+(double)y*y),
 * +3 = (float)((double)x / (double)y); and out[4*m+i] = kernel.cpp:8-10's
 * sharp_kernel(x, sqrt(y)) as the device forms it for x < y (with y = h*h;
 * 0 otherwise).  out holds 5*m floats. */
int flip_gpu_selftest_math(int64_t m, const float *x, const float *y, float *out);

#ifdef __cplusplus
}
#endif
#endif
