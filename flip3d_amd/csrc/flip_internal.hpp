// flip_internal.hpp - host-side context and kernel launch entry points.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/flip3d_gpu.h"

#define FLIP_CHECK(x)                                                          \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess) throw flip_hip_error(e_, #x, __FILE__, __LINE__); \
    } while (0)

struct flip_hip_error {
    hipError_t err;
    std::string what;
    flip_hip_error(hipError_t e, const char *expr, const char *file, int line);
};

struct flip_arg_error {
    int code;
    std::string what;
};

// glibc random_r TYPE_3 generator (additive lagged, taps 31/3), seeded like
// srandom_r; the reference's rand() stream with seed 1.
struct GlibcRand {
    int32_t s[31];
    int f = 3, r = 0;
    void seed(unsigned seed);
    int next();
};

struct LearnedSolver;  // learned multilevel flexible-CG state (k_learned.hip)
struct MeshState;      // buffers of the last surface mesh (k_mesh.hip)
void learned_free(LearnedSolver *ls);

// Device buffers of one context.  Particles live in cell-sorted order and are
// re-permuted after every sort; meta carries creation index << 2 | type.
constexpr int RNG_CHUNK = 1024;  // rand() draws per chunk of the device's generator (its table's rows)
// slots of ev_count (8 ints, zeroed where a step begins): the corrector's
// listed particles [0], draw overflow [1], non-finite update [EV_NONFINITE]
// and events [3]; the advection's stuck particles [EV_STUCK], the
// repositioned ones [EV_MOVED], their resample levels [EV_NLEV] and its
// non-finite flag [EV_NONFINITE_ADV]
constexpr int EV_NONFINITE = 2;
constexpr int EV_STUCK = 4, EV_MOVED = 5, EV_NLEV = 6, EV_NONFINITE_ADV = 7, EV_SLOTS = 8;

struct FlipCtx {
    flip_params prm;
    int n = 0;
    long C = 0, F = 0;  // cells n^3, face-grid size (n+1) n^2
    hipStream_t stream = nullptr;
    bool profiling = false;
    // FLIP_DIRECT_LAUNCH=1 at creation: solver iterations are launched kernel
    // by kernel instead of replayed from captured HIP graphs (for tracing
    // tools that cannot follow graph launches); results are unchanged
    bool direct_launch = false;
    // launch shape of the solver's FLUID-list kernels (flip_gpu_set_launch;
    // 0 = built in): their block count (at most GMAX) and the PCG iterations
    // per captured batch
    int tune_blocks = 0, tune_batch = 0;

    // particles (double-buffered for the permutation gather)
    int64_t np = 0, cap = 0;
    int cur = 0;
    float4 *pos[2] = {nullptr, nullptr};  // x, y, z, m
    float4 *vel[2] = {nullptr, nullptr};  // ux, uy, uz, dens
    int32_t *meta[2] = {nullptr, nullptr};
    float4 *nrm_id = nullptr;   // wall normals by creation index
    int32_t *inv = nullptr;     // creation index -> storage slot
    float4 *spos = nullptr;     // P2G positions: clamp(n p, 0, n), m (far away for non-fluid)
    // (x, y, z) of each stored particle, +inf for WALL particles, written by
    // every sort: the 12-byte window records of the density and resample
    // scans, valid while `sorted` holds (positions unchanged since the sort)
    float *xyz3 = nullptr;
    float *dw = nullptr;  // per particle: density, or -1 for WALL (the step's density pass writes it for flags)
    float4 *tmp0 = nullptr, *tmp1 = nullptr;  // corrector pseudo-position / velocity
    int32_t *pkey = nullptr, *pslot = nullptr, *perm = nullptr;
    int32_t *sort_big = nullptr;  // buckets over 16 particles: count, then cell ids

    // cells
    int32_t *cell_count = nullptr, *cell_start = nullptr;
    int32_t *wall_start = nullptr;  // static wall bins
    int8_t *wall_near = nullptr;    // per cell: a wall particle lies in its 3x3x3 window
    float4 *wall_pos = nullptr, *wall_nrm = nullptr;
    int64_t nwall = 0;

    // grids
    float *ug[3] = {nullptr, nullptr, nullptr}, *us[3] = {nullptr, nullptr, nullptr};
    // (u, u - u_save) per face as written by the extrapolation, for the
    // step's G2P: one 8-byte load serves both interpolations
    float2 *ud[3] = {nullptr, nullptr, nullptr};
    int8_t *A = nullptr;
    float *L = nullptr, *press = nullptr, *div = nullptr;

    // solver
    int32_t *fl = nullptr;  // fluid cell list (ascending cell index)
    int nfl = 0;
    double *P = nullptr;
    float *r = nullptr, *z = nullptr, *q = nullptr;
    float2 *zs[2] = {nullptr, nullptr};  // (z, s) pairs, ping-pong by iteration parity
    double *partials = nullptr;  // per-block partial sums
    double *pcg_part = nullptr;  // PCG per-block partials (8 * 1024)
    uint32_t *flmeta = nullptr;  // per FLUID cell: neighbour kinds (2 bits x 6) + has-AIR bit
    hipGraphExec_t pcg_graph = nullptr;  // captured batch of PCG iterations
    long pcg_graph_key = -1;             // subcell, block count and batch it was captured for
    // the solver's status after each batch of iterations, written by the
    // device into mapped host memory (no copy, read after the batch's event)
    struct SolveStatus {
        int32_t done, iters, capped, unit, nfl, pad[3];
        double err2, err2_0, bnorm2;
    };
    SolveStatus *h_status = nullptr, *d_status = nullptr;
    hipEvent_t ev_batch[2] = {nullptr, nullptr};
    bool solve_timed = false;  // ev_s0/ev_s1 bracket a solve whose time is not read yet
    bool corrected = false;    // a corrector ran; its jitter-event count is not read yet
    bool advected = false;     // an advection ran; its reposition count and checks are not read yet
    // the step's divergence pass wrote -div (solver.cpp:231's flip) and its
    // flags pass the FLUID flags and the non-FLUID press zeroes: the next
    // PCG setup skips both
    bool solve_prepared = false;
    hipGraphExec_t ic_graph = nullptr;  // captured batch of wavefront-preconditioned iterations
    long ic_graph_key = -1;             // subcell, preconditioner mode and block count it was captured for
    double *qv = nullptr;        // forward-sweep intermediate of the wavefront preconditioner
    // FP64 solve / iterative refinement vectors (allocated on first use)
    double *x64 = nullptr, *b64 = nullptr, *r64 = nullptr, *s64 = nullptr, *q64 = nullptr;
    double *part64 = nullptr;    // per-block partials of the FP64 kernels
    double *h_part64 = nullptr;  // pinned host copy of part64
    double *scal = nullptr;      // device scalars (see pcg kernels)
    int32_t *iflags = nullptr;   // device int flags / counters
    unsigned *tickets = nullptr;
    int last_iters = 0;

    // scan scratch
    int32_t *scan_tmp = nullptr, *scan_blocks = nullptr;
    int32_t *deep = nullptr;

    // event buffers (host-drawn rand() consumers)
    int32_t *ev_list = nullptr;  // (slot) of particles with events
    double *dtv = nullptr;        // [0] the time step the kernels use (chosen on the device)
    int32_t *ev_count = nullptr;  // EV_SLOTS event counts and flags (see EV_STUCK)
    int32_t *repo_ids = nullptr, *repo_order = nullptr, *repo_loff = nullptr;  // the reposition plan
    float *ev_rand = nullptr;    // rand()%101 values, 3 per event
    int32_t *ev_off = nullptr;   // per listed particle: offset into ev_rand
    int32_t *ev_slot = nullptr;  // listed particles' slots in creation order
    // the glibc rand() state (s[31], f, r) on the device, and whether it (not
    // rng) is the current one
    int32_t *rng_dev = nullptr, *h_rng = nullptr;
    uint32_t *rng_table = nullptr;  // draw coefficients of the device's chunked rand() (k_jitter_draw)
    bool rng_on_device = false;
    int64_t ev_cap = 0;
    int32_t *h_pinned = nullptr;  // pinned host scalars
    // Step graphs: the step's kernels before the pressure solve (time step
    // through divergence) and after it (gradient through the corrector and
    // the step record) captured once per particle count and buffer parity
    // and replayed; the solve between them keeps its own captured iterations.
    bool step_graphs = true;
    hipGraphExec_t step_gx[2] = {nullptr, nullptr};
    hipGraph_t step_g[2] = {nullptr, nullptr};  // the captured graphs (children of a batch's graphs)
    uint64_t step_gen = 0;                       // process-wide number of the context's last capture
    std::vector<uint64_t> step_graph_sig;  // the pointers and scalars the captured launches bake in
    // what the step's end reads, written by the device into mapped host
    // memory (k_step_record): the event counts and the time step
    struct StepRecord {
        int32_t ev[EV_SLOTS];
        double dt;
    };
    StepRecord *h_rec = nullptr, *d_rec = nullptr;

    std::string learned_weights;  // U-Net weights file for FLIP_PRECOND_LEARNED
    // the scene's object list as saveState writes it to objects.dat
    // (flip3D.cpp:498-508): type, shape, material, visible, then r, c[3],
    // p[0][3], p[1][3]; written on the context's first save
    struct SceneObject {
        int type, shape, material, visible;
        float f[10];
    };
    std::vector<SceneObject> objects;
    bool objects_saved = false, objects_loaded = false;
    LearnedSolver *learned = nullptr;

    float max_dens = 1.0f;
    double dt = 6e-3;
    // pourWater emitter (flip3D.cpp:41-43,375-402) and the step counter it
    // reads (flip3D.cpp:59,1196); emit_pos holds the disc's particles
    int step = 0, pour_time = -1;
    float pour_pos[2] = {0.0f, 0.0f}, pour_rad = 0.12f;
    float4 *emit_pos = nullptr;
    int emit_n = 0;
    GlibcRand rng;
    bool uploaded = false;
    bool sorted = false;  // storage order matches the current positions
    bool bucketed = false;  // cell_start describes the storage (the last sort), maybe of older positions
    bool begun = false;     // flip_gpu_step_begin ran for the coming step
    MeshState *mesh = nullptr;
    bool unit_mass = false;  // every particle has m == 1 (the reference's only mass)

    // timing
    std::vector<std::string> tnames;
    std::vector<hipEvent_t> tev;
    std::vector<double> tms;
    hipEvent_t ev_a = nullptr, ev_b = nullptr, ev_s0 = nullptr, ev_s1 = nullptr;

    float4 *P4() const { return pos[cur]; }
    float4 *V4() const { return vel[cur]; }
    int32_t *M() const { return meta[cur]; }
};

// ---- kernels (launch wrappers; each enqueues on ctx->stream) ----
// counting sort + per-cell id order + permute; speed_part: also the gather's
// per-block largest FLUID |u|^2 (one per 256 particles)
void k_launch_sort(FlipCtx *c, double *speed_part = nullptr);
// with_force: add_ExtForce fused in; p2g_records: also write P2G's packed
// records and the flags pass's dw (returns whether it did)
bool k_launch_density(FlipCtx *c, bool with_force = false, bool p2g_records = false);
void k_launch_ext_force(FlipCtx *c);
// A and L from the sorted buckets; dw_ready: read the density pass's dw;
// fluid_flags: also the PCG setup's FLUID flags (cell_count) and press zeroes
void k_launch_flags(FlipCtx *c, bool dw_ready = false, bool fluid_flags = false);
void k_launch_p2g(FlipCtx *c, bool records_ready = false);  // P2G + save_grid + enforce_boundary
void k_launch_divergence(FlipCtx *c, bool negate = false);  // negate: write -div
void k_launch_gradient(FlipCtx *c);             // ghost-fluid gradient + enforce_boundary
void k_launch_extrapolate(FlipCtx *c);          // extrapolate_velocity + subtract_grid
void k_launch_g2p(FlipCtx *c, bool fresh = false);  // FLIP/PIC blend; fresh: right after k_launch_extrapolate
// advect + clamp + wall push-out; with stuck, the stuck particles' creation
// indices are listed in ev_list (count in ev_count[EV_STUCK], cleared by the caller)
void k_launch_advect(FlipCtx *c, bool stuck);
void k_stuck_and_reposition(FlipCtx *c);  // the listed stuck particles; counts in ev_count[EV_STUCK..]
void k_launch_clear_events(FlipCtx *c, int lo, int hi);  // ev_count[lo, hi) = 0 on the stream
void k_launch_step_record(FlipCtx *c);  // the event counts and the time step into c->d_rec (mapped host memory)
int k_correct(FlipCtx *c);                      // corrector; returns jitter events
void k_launch_emit(FlipCtx *c);                 // append the emitter's disc (capacity already ensured)
void ensure_particle_capacity(FlipCtx *c, int64_t need);  // grow, keeping the particles
void set_emitter(FlipCtx *c, float x, float z, float r, int pour_time);
// exclusive scan, out[n] = total; `blocks` holds n / 2048 + 1 block sums
// (default: the context's scratch, sized for max(particles, cells))
// gate (optional, device): the scan does nothing while *gate == 0
void k_launch_scan(FlipCtx *c, const int32_t *in, int32_t *out, long n, int32_t *blocks = nullptr,
                   const int32_t *gate = nullptr);
// surface mesher (k_mesh.hip)
void mesh_build(FlipCtx *c, int mg, long *nv, long *nf);
void mesh_copy_out(FlipCtx *c, double *v, double *nrm, int32_t *faces);
void mesh_free(MeshState *m);
void mesh_counts(const FlipCtx *c, long *nv, long *nf);
// output files (export.hip)
void preview_render(FlipCtx *c, uint8_t *rgba_host);  // write_frame's RGBA image
int preview_width();
int preview_height();
bool write_bmp_file(const char *path, const uint8_t *rgba, int width, int height);
bool write_obj_file(const char *path, const std::vector<double> &v, const std::vector<double> &n,
                    const std::vector<int32_t> &f, float wall_thick);
bool write_sunflow_file(const char *path, const std::vector<double> &v, const std::vector<double> &n,
                        const std::vector<int32_t> &f, float wall_thick);
bool write_sunflow_template(const char *path);
int k_build_fluid_list(FlipCtx *c);             // returns nfl (host sync)
int k_solve(FlipCtx *c, flip_step_stats *st);   // PCG on (A, L, press, div)
// FLIP_PRECOND_LEARNED: flexible CG from the current press and r (after the
// PCG init); returns its iterations, *converged when ||r||^2 <= eps
int learned_fcg(FlipCtx *c, int G, double h2, int subcell, double eps, bool *converged);
int learned_fcg_iterations(FlipCtx *c);
int learned_side(int n);    // the side the learned preconditioner's levels halve (n padded to 16)
int learned_levels(int n);  // coarse levels of the learned preconditioner for side n
void k_launch_g2p_points(FlipCtx *c, const float *g0, const float *g1, const float *g2, int64_t m,
                         const float4 *p, float4 *out);
void k_launch_selftest(int64_t m, const float *x, const float *y, float *out, hipStream_t s);
void rng_upload_table(FlipCtx *c);  // the chunked-draw coefficients (context creation)
void rng_to_device(FlipCtx *c);  // the rand() state to the device, if it is on the host
void rng_to_host(FlipCtx *c);    // and back (waits for the stream)
// the step's time step on the device, from the speeds of the sort's gather
// (k_launch_sort with c->partials, just before)
void k_launch_choose_dt(FlipCtx *c);
void k_launch_set_dt(FlipCtx *c, double dt);    // a given dt into dtv[0]

// records the message for flip_gpu_last_error and returns code
int flip_internal_fail(int code, const char *msg);

// timing helpers
void t_begin(FlipCtx *c, const char *name);
void t_end(FlipCtx *c);

// Launch one kernel on the context stream.  When profiling, the runtime writes
// the dispatch's own begin/end timestamps into a named event pair, so the
// recorded duration is the kernel alone, as a kernel trace reports it.
template <typename... Formals, typename... Actuals>
void launch_timed(FlipCtx *c, const char *name, void (*kernel)(Formals...), dim3 grid, dim3 block,
                  Actuals... args)
{
    if (!c->profiling) {
        hipLaunchKernelGGL(kernel, grid, block, 0, c->stream, args...);
        return;
    }
    hipEvent_t e0, e1;
    FLIP_CHECK(hipEventCreate(&e0));
    FLIP_CHECK(hipEventCreate(&e1));
    c->tnames.push_back(name);
    c->tev.push_back(e0);
    c->tev.push_back(e1);
    hipExtLaunchKernelGGL(kernel, grid, block, 0, c->stream, e0, e1, 0, args...);
    FLIP_CHECK(hipGetLastError());
}
