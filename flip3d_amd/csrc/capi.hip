// capi.hip - C ABI (include/flip3d_gpu.h): context, particle I/O, step
// orchestration in the order of flip3D::simulateStep (src/flip3D.cpp:1194-1264).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "flip_internal.hpp"

namespace {
thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

template <class F> int guarded(F &&f)
{
    try {
        return f();
    } catch (const flip_hip_error &e) {
        return fail(FLIP_E_HIP, e.what);
    } catch (const flip_arg_error &e) {
        return fail(e.code, e.what);
    } catch (const std::bad_alloc &) {
        return fail(FLIP_E_HIP, "host allocation failed");
    } catch (...) {
        return fail(FLIP_E_STATE, "unexpected exception");
    }
}

template <class T> void dalloc(T **p, size_t count)
{
    if (*p) FLIP_CHECK(hipFree(*p));
    *p = nullptr;
    if (count == 0) count = 1;
    FLIP_CHECK(hipMalloc((void **)p, count * sizeof(T)));
}

template <class T> void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

void require(bool ok, const char *msg, int code = FLIP_E_ARG)
{
    if (!ok) throw flip_arg_error{code, msg};
}

void free_particles(FlipCtx *c)
{
    for (int b = 0; b < 2; ++b) {
        dfree(c->pos[b]);
        dfree(c->vel[b]);
        dfree(c->meta[b]);
    }
    dfree(c->nrm_id);
    dfree(c->inv);
    dfree(c->spos);
    dfree(c->xyz3);
    dfree(c->dw);
    dfree(c->tmp0);
    dfree(c->tmp1);
    dfree(c->pkey);
    dfree(c->pslot);
    dfree(c->sort_big);
    dfree(c->perm);
    dfree(c->deep);
    dfree(c->ev_list);
    dfree(c->ev_off);
    dfree(c->repo_ids);
    dfree(c->repo_order);
    dfree(c->repo_loff);
    dfree(c->ev_slot);
    dfree(c->ev_rand);
    dfree(c->partials);
    dfree(c->scan_blocks);
    c->cap = 0;
}

void alloc_particles(FlipCtx *c, int64_t cap)
{
    free_particles(c);
    size_t n = (size_t)std::max<int64_t>(cap, 1);
    for (int b = 0; b < 2; ++b) {
        dalloc(&c->pos[b], n);
        dalloc(&c->vel[b], n);
        dalloc(&c->meta[b], n);
    }
    dalloc(&c->nrm_id, n);
    dalloc(&c->inv, n);
    dalloc(&c->spos, n);
    dalloc(&c->xyz3, 3 * n);
    dalloc(&c->dw, n);
    dalloc(&c->tmp0, n);
    dalloc(&c->tmp1, n);
    dalloc(&c->pkey, n);
    dalloc(&c->pslot, n);
    dalloc(&c->sort_big, n / 17 + 2);  // at most np / 17 buckets hold more than 16
    dalloc(&c->perm, n);
    size_t big = std::max<size_t>(n, (size_t)c->C) + 1;
    dalloc(&c->deep, big);
    dalloc(&c->ev_list, 3 * n + 3);  // the corrector's (creation index, slot, events) triples
    dalloc(&c->ev_off, big);
    dalloc(&c->ev_slot, n);
    dalloc(&c->repo_ids, n);
    dalloc(&c->repo_order, n);
    dalloc(&c->repo_loff, n + 1);
    c->ev_cap = 2 * (int64_t)n + 1024;
    dalloc(&c->ev_rand, 3 * (size_t)c->ev_cap);
    dalloc(&c->partials, 4 * (big / 256 + 8));
    dalloc(&c->scan_blocks, big / 2048 + 8);
    c->cap = cap;
}

}  // namespace

// Room for `need` particles.  Growing reallocates every per-particle buffer;
// the persistent state (current positions, velocities, ids and the normals
// by creation index) is carried over through the host.  The next sort
// rebuilds everything else.
void ensure_particle_capacity(FlipCtx *c, int64_t need)
{
    if (need <= c->cap) return;
    require(need < (int64_t(1) << 29), "too many particles (limit 2^29)");
    const int64_t np = c->np;
    std::vector<float4> p4(np), v4(np), n4(np);
    std::vector<int32_t> mt(np);
    hipStream_t st = c->stream;
    if (np > 0) {
        FLIP_CHECK(hipMemcpyAsync(p4.data(), c->P4(), np * sizeof(float4), hipMemcpyDeviceToHost, st));
        FLIP_CHECK(hipMemcpyAsync(v4.data(), c->V4(), np * sizeof(float4), hipMemcpyDeviceToHost, st));
        FLIP_CHECK(hipMemcpyAsync(mt.data(), c->M(), np * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        FLIP_CHECK(hipMemcpyAsync(n4.data(), c->nrm_id, np * sizeof(float4), hipMemcpyDeviceToHost, st));
    }
    FLIP_CHECK(hipStreamSynchronize(st));
    alloc_particles(c, std::max<int64_t>(need, 2 * c->cap));
    c->cur = 0;
    if (np > 0) {
        FLIP_CHECK(hipMemcpyAsync(c->pos[0], p4.data(), np * sizeof(float4), hipMemcpyHostToDevice, st));
        FLIP_CHECK(hipMemcpyAsync(c->vel[0], v4.data(), np * sizeof(float4), hipMemcpyHostToDevice, st));
        FLIP_CHECK(hipMemcpyAsync(c->meta[0], mt.data(), np * sizeof(int32_t), hipMemcpyHostToDevice, st));
        FLIP_CHECK(hipMemcpyAsync(c->nrm_id, n4.data(), np * sizeof(float4), hipMemcpyHostToDevice, st));
    }
    FLIP_CHECK(hipStreamSynchronize(st));  // the host vectors go out of scope
}

// pourWater's disc (flip3D.cpp:375-402): lattice points at spacing DENSITY/N
// in x and z, the float coordinate stepped by the double spacing, kept when
// hypotf(x - px, z - pz) < r; y just under the ceiling wall.
void set_emitter(FlipCtx *c, float x, float z, float r, int pour_time)
{
    c->pour_pos[0] = x;
    c->pour_pos[1] = z;
    c->pour_rad = r;
    c->pour_time = pour_time;
    const int N = c->n;
    const double w = c->prm.density / N;
    const float y = (float)(1.0 - 1.0 / N - 2.5 * c->prm.density / N);
    std::vector<float4> disc;
    if (pour_time >= 0)
        for (float px = (float)(w + w / 2.0); px < 1.0 - w / 2.0; px = (float)((double)px + w))
            for (float pz = (float)(w + w / 2.0); pz < 1.0 - w / 2.0; pz = (float)((double)pz + w))
                if (hypotf(px - x, pz - z) < r) disc.push_back(make_float4(px, y, pz, 1.0f));
    dalloc(&c->emit_pos, std::max<size_t>(disc.size(), 1));
    if (!disc.empty()) {
        FLIP_CHECK(hipMemcpyAsync(c->emit_pos, disc.data(), disc.size() * sizeof(float4), hipMemcpyHostToDevice,
                                  c->stream));
        FLIP_CHECK(hipStreamSynchronize(c->stream));
    }
    c->emit_n = (int)disc.size();
}

namespace {

void alloc_grids(FlipCtx *c)
{
    for (int k = 0; k < 3; ++k) {
        dalloc(&c->ug[k], c->F);
        dalloc(&c->us[k], c->F);
        dalloc(&c->ud[k], c->F);
        FLIP_CHECK(hipMemset(c->ug[k], 0, c->F * sizeof(float)));
        FLIP_CHECK(hipMemset(c->us[k], 0, c->F * sizeof(float)));
    }
    dalloc(&c->A, c->C);
    dalloc(&c->L, c->C);
    dalloc(&c->press, c->C);
    dalloc(&c->div, c->C);
    FLIP_CHECK(hipMemset(c->A, 0, c->C));
    FLIP_CHECK(hipMemset(c->L, 0, c->C * sizeof(float)));
    FLIP_CHECK(hipMemset(c->press, 0, c->C * sizeof(float)));
    FLIP_CHECK(hipMemset(c->div, 0, c->C * sizeof(float)));
    dalloc(&c->cell_count, c->C + 1);
    dalloc(&c->cell_start, c->C + 1);
    dalloc(&c->wall_start, c->C + 1);
    dalloc(&c->wall_near, c->C);
    FLIP_CHECK(hipMemset(c->wall_near, 0, c->C * sizeof(int8_t)));
    dalloc(&c->scan_tmp, c->C + 1);
    dalloc(&c->fl, c->C);
    dalloc(&c->P, c->C);
    dalloc(&c->r, c->C);
    dalloc(&c->z, c->C);
    dalloc(&c->q, c->C);
    dalloc(&c->zs[0], c->C);
    dalloc(&c->zs[1], c->C);
    FLIP_CHECK(hipMemset(c->zs[0], 0, c->C * sizeof(float2)));
    FLIP_CHECK(hipMemset(c->zs[1], 0, c->C * sizeof(float2)));
    FLIP_CHECK(hipMemset(c->z, 0, c->C * sizeof(float)));
    dalloc(&c->pcg_part, 8 * 1024);
    dalloc(&c->flmeta, c->C);
    dalloc(&c->qv, c->C);
    dalloc(&c->scal, 32);
    dalloc(&c->iflags, 16);
    dalloc(&c->tickets, 8);
    FLIP_CHECK(hipMemset(c->tickets, 0, 8 * sizeof(unsigned)));
    dalloc(&c->ev_count, EV_SLOTS);
    dalloc(&c->dtv, 4);
    dalloc(&c->rng_dev, 64);
    dalloc(&c->rng_table, (size_t)RNG_CHUNK * 31);
    rng_upload_table(c);
    FLIP_CHECK(hipHostMalloc((void **)&c->h_rng, 64 * sizeof(int32_t)));
    FLIP_CHECK(hipHostMalloc((void **)&c->h_pinned, 64 * sizeof(int32_t)));
    FLIP_CHECK(hipHostMalloc((void **)&c->h_rec, sizeof(FlipCtx::StepRecord), hipHostMallocMapped));
    FLIP_CHECK(hipHostGetDevicePointer((void **)&c->d_rec, c->h_rec, 0));
    memset(c->h_rec, 0, sizeof(FlipCtx::StepRecord));
    FLIP_CHECK(hipHostMalloc((void **)&c->h_status, sizeof(FlipCtx::SolveStatus), hipHostMallocMapped));
    FLIP_CHECK(hipHostGetDevicePointer((void **)&c->d_status, c->h_status, 0));
    memset(c->h_status, 0, sizeof(FlipCtx::SolveStatus));
    for (auto &e : c->ev_batch) FLIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

void free_all(FlipCtx *c)
{
    free_particles(c);
    dfree(c->wall_pos);
    dfree(c->wall_nrm);
    dfree(c->emit_pos);
    for (int k = 0; k < 3; ++k) {
        dfree(c->ug[k]);
        dfree(c->us[k]);
        dfree(c->ud[k]);
    }
    dfree(c->A);
    dfree(c->L);
    dfree(c->press);
    dfree(c->div);
    dfree(c->cell_count);
    dfree(c->cell_start);
    dfree(c->wall_start);
    dfree(c->wall_near);
    dfree(c->scan_tmp);
    dfree(c->fl);
    dfree(c->P);
    dfree(c->r);
    dfree(c->z);
    dfree(c->q);
    dfree(c->zs[0]);
    dfree(c->zs[1]);
    dfree(c->pcg_part);
    dfree(c->flmeta);
    dfree(c->qv);
    dfree(c->x64);
    dfree(c->b64);
    dfree(c->r64);
    dfree(c->s64);
    dfree(c->q64);
    dfree(c->part64);
    if (c->h_part64) (void)hipHostFree(c->h_part64);
    c->h_part64 = nullptr;
    if (c->pcg_graph) (void)hipGraphExecDestroy(c->pcg_graph);
    c->pcg_graph = nullptr;
    if (c->ic_graph) (void)hipGraphExecDestroy(c->ic_graph);
    c->ic_graph = nullptr;
    learned_free(c->learned);
    c->learned = nullptr;
    mesh_free(c->mesh);
    c->mesh = nullptr;
    dfree(c->scal);
    dfree(c->iflags);
    dfree(c->tickets);
    dfree(c->ev_count);
    dfree(c->dtv);
    dfree(c->rng_dev);
    dfree(c->rng_table);
    if (c->h_rng) (void)hipHostFree(c->h_rng);
    c->h_rng = nullptr;
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    c->h_pinned = nullptr;
    if (c->h_status) (void)hipHostFree(c->h_status);
    c->h_status = c->d_status = nullptr;
    for (int w = 0; w < 2; ++w) {
        if (c->step_gx[w]) (void)hipGraphExecDestroy(c->step_gx[w]);
        if (c->step_g[w]) (void)hipGraphDestroy(c->step_g[w]);
        c->step_gx[w] = nullptr;
        c->step_g[w] = nullptr;
    }
    if (c->h_rec) (void)hipHostFree(c->h_rec);
    c->h_rec = c->d_rec = nullptr;
    for (auto &e : c->ev_batch)
        if (e) (void)hipEventDestroy(e);
    for (auto e : c->tev) (void)hipEventDestroy(e);
    c->tev.clear();
    hipEvent_t evs[4] = {c->ev_a, c->ev_b, c->ev_s0, c->ev_s1};
    for (auto e : evs)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

float grid_cell_of(int gn, float p)  // sorter.cpp:35-37 on the host
{
    return (float)(int)fmax(0.0, fmin((double)(gn - 1), (double)((float)gn * p)));
}

void build_wall_bins(FlipCtx *c, const std::vector<float4> &pos, const std::vector<int32_t> &meta,
                     const std::vector<float4> &nrm_id)
{
    int n = c->n;
    std::vector<std::pair<int64_t, int32_t>> walls;  // (cell, id)
    for (size_t i = 0; i < pos.size(); ++i) {
        if ((meta[i] & 3) != FLIP_WALL) continue;
        int ci = (int)grid_cell_of(n, pos[i].x), cj = (int)grid_cell_of(n, pos[i].y), ck = (int)grid_cell_of(n, pos[i].z);
        walls.push_back({((int64_t)ci * n + cj) * n + ck, (int32_t)i});
    }
    std::sort(walls.begin(), walls.end());
    std::vector<int32_t> start(c->C + 1, 0);
    std::vector<float4> wp(walls.size()), wn(walls.size());
    for (size_t t = 0; t < walls.size(); ++t) {
        start[walls[t].first + 1]++;
        wp[t] = pos[walls[t].second];
        wn[t] = nrm_id[(uint32_t)meta[walls[t].second] >> 2];
    }
    // cells whose 3x3x3 window holds a wall particle (the advection's push-out
    // visits only these windows)
    std::vector<int8_t> near(c->C, 0);
    for (long k = 0; k < c->C; ++k) {
        if (start[k + 1] == 0) continue;
        const int ci = (int)(k / ((long)n * n)), cj = (int)((k / n) % n), ck = (int)(k % n);
        for (int a = std::max(ci - 1, 0); a <= std::min(ci + 1, n - 1); ++a)
            for (int b = std::max(cj - 1, 0); b <= std::min(cj + 1, n - 1); ++b)
                for (int d = std::max(ck - 1, 0); d <= std::min(ck + 1, n - 1); ++d) near[((long)a * n + b) * n + d] = 1;
    }
    for (long k = 0; k < c->C; ++k) start[k + 1] += start[k];
    c->nwall = (int64_t)walls.size();
    dalloc(&c->wall_pos, walls.size());
    dalloc(&c->wall_nrm, walls.size());
    FLIP_CHECK(hipMemcpy(c->wall_start, start.data(), (c->C + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    FLIP_CHECK(hipMemcpy(c->wall_near, near.data(), c->C * sizeof(int8_t), hipMemcpyHostToDevice));
    if (!walls.empty()) {
        FLIP_CHECK(hipMemcpy(c->wall_pos, wp.data(), wp.size() * sizeof(float4), hipMemcpyHostToDevice));
        FLIP_CHECK(hipMemcpy(c->wall_nrm, wn.data(), wn.size() * sizeof(float4), hipMemcpyHostToDevice));
    }
}

void ensure_sorted(FlipCtx *c)
{
    if (!c->sorted) {
        k_launch_sort(c);
        c->sorted = true;
    }
}

double elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.0f;
    FLIP_CHECK(hipEventElapsedTime(&ms, a, b));
    return ms;
}

}  // namespace

int flip_internal_fail(int code, const char *msg) { return fail(code, msg); }

flip_hip_error::flip_hip_error(hipError_t e, const char *expr, const char *file, int line) : err(e)
{
    char buf[512];
    snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", expr, hipGetErrorString(e), file, line);
    what = buf;
}

void GlibcRand::seed(unsigned sd)
{
    if (sd == 0) sd = 1;
    int32_t w = (int32_t)sd;
    s[0] = w;
    for (int i = 1; i < 31; ++i) {
        long hi = w / 127773, lo = w % 127773;
        long t = 16807 * lo - 2836 * hi;
        if (t < 0) t += 2147483647;
        w = (int32_t)t;
        s[i] = w;
    }
    f = 3;
    r = 0;
    for (int k = 0; k < 310; ++k) (void)next();
}

int GlibcRand::next()
{
    uint32_t v = (uint32_t)s[f] + (uint32_t)s[r];
    s[f] = (int32_t)v;
    f = (f + 1) % 31;
    r = (r + 1) % 31;
    return (int)(v >> 1);
}

void t_begin(FlipCtx *c, const char *name)
{
    if (!c->profiling) return;
    hipEvent_t e;
    FLIP_CHECK(hipEventCreate(&e));
    FLIP_CHECK(hipEventRecord(e, c->stream));
    c->tnames.push_back(name);
    c->tev.push_back(e);
}

void t_end(FlipCtx *c)
{
    if (!c->profiling) return;
    hipEvent_t e;
    FLIP_CHECK(hipEventCreate(&e));
    FLIP_CHECK(hipEventRecord(e, c->stream));
    c->tev.push_back(e);
}

static void timing_reset(FlipCtx *c)
{
    for (auto e : c->tev) (void)hipEventDestroy(e);
    c->tev.clear();
    c->tnames.clear();
    c->tms.clear();
}

static void timing_collect(FlipCtx *c)
{
    if (!c->profiling) return;
    FLIP_CHECK(hipStreamSynchronize(c->stream));
    c->tms.clear();
    for (size_t t = 0; t + 1 < c->tev.size(); t += 2) c->tms.push_back(elapsed(c->tev[t], c->tev[t + 1]));
}

static void run_stage(FlipCtx *c, int stage, flip_step_stats *st)
{
    switch (stage) {
    case FLIP_STAGE_DENSITY:
        k_launch_sort(c);
        c->sorted = true;
        k_launch_density(c);
        break;
    case FLIP_STAGE_EXT_FORCE:
        k_launch_ext_force(c);
        break;
    case FLIP_STAGE_P2G:
        ensure_sorted(c);
        k_launch_flags(c);
        k_launch_p2g(c);
        break;
    case FLIP_STAGE_PROJECT: {
        k_launch_divergence(c);
        if (c->profiling || st) {
            if (!c->ev_s0) {
                FLIP_CHECK(hipEventCreate(&c->ev_s0));
                FLIP_CHECK(hipEventCreate(&c->ev_s1));
            }
            FLIP_CHECK(hipEventRecord(c->ev_s0, c->stream));
        }
        int rc = k_solve(c, st);
        if (c->ev_s0 && (c->profiling || st)) {
            // read once the step (or stage call) has finished
            FLIP_CHECK(hipEventRecord(c->ev_s1, c->stream));
            c->solve_timed = true;
        }
        if (rc == FLIP_E_NUMERIC) throw flip_arg_error{FLIP_E_NUMERIC, "pressure solve produced NaN"};
        k_launch_gradient(c);
        break;
    }
    case FLIP_STAGE_EXTRAP:
        k_launch_extrapolate(c);
        break;
    case FLIP_STAGE_G2P:
        k_launch_g2p(c);
        break;
    case FLIP_STAGE_ADVECT: {
        // the advection also lists the stuck particles (flip3D.cpp:933-955)
        k_launch_clear_events(c, EV_STUCK, EV_NONFINITE_ADV + 1);
        k_launch_advect(c, c->prm.reposition != 0);
        c->sorted = false;
        k_stuck_and_reposition(c);  // leaves the particles sorted; counts read at the next wait
        c->sorted = true;
        c->advected = true;
        break;
    }
    case FLIP_STAGE_CORRECT: {
        ensure_sorted(c);
        k_correct(c);  // its event count (ev_count[3]) is read when the call ends
        c->sorted = false;
        c->corrected = true;
        break;
    }
    default:
        throw flip_arg_error{FLIP_E_ARG, "unknown stage"};
    }
}

// ------------------------------------------------------------------ C ABI
extern "C" {

const char *flip_gpu_last_error(void) { return g_last_error.c_str(); }

int flip_gpu_default_params(flip_params *p, int n)
{
    if (!p) return fail(FLIP_E_ARG, "null params");
    memset(p, 0, sizeof(*p));
    p->n = n;
    p->dt = 0.6e-2;
    p->dt_policy = 0;
    p->cfl = 0.6;
    p->alpha = 0.95;
    p->gravity = 9.8;
    p->density = 0.5;
    p->subcell = 1;
    p->tol_policy = 0;
    p->tol = 1e-5;
    p->precond = FLIP_PRECOND_REFERENCE;
    p->max_iters = 0;
    p->correction = 1;
    p->reposition = 1;
    p->rand_seed = 1;
    p->refine = 0;
    p->refine_tol = 1e-7;
    p->solve_fp64 = 0;
    return FLIP_OK;
}

int flip_gpu_create(const flip_params *p, flip_ctx **out)
{
    return guarded([&]() -> int {
        require(p && out, "null argument");
        require(p->n >= 2 && p->n <= 1024, "n out of range [2, 1024]");
        require(p->precond >= 0 && p->precond <= 2, "unknown preconditioner");
        require(p->refine >= 0, "refine must be >= 0");
        require(p->refine == 0 || (p->tol_policy == 1 && p->refine_tol > 0.0),
                "iterative refinement needs tol_policy 1 and refine_tol > 0");
        require(p->solve_fp64 == 0 || p->solve_fp64 == 1, "solve_fp64 must be 0 or 1");
        require(p->solve_fp64 == 0 || p->tol_policy == 1, "the FP64 solve needs tol_policy 1");
        int ndev = 0;
        FLIP_CHECK(hipGetDeviceCount(&ndev));
        require(ndev > 0, "no HIP device", FLIP_E_HIP);
        FlipCtx *c = new FlipCtx();
        try {
            c->prm = *p;
            c->n = p->n;
            c->C = (long)p->n * p->n * p->n;
            c->F = (long)(p->n + 1) * p->n * p->n;
            c->dt = p->dt;
            c->rng.seed(p->rand_seed);
            const char *dl = std::getenv("FLIP_DIRECT_LAUNCH");
            c->direct_launch = dl && dl[0] == '1';
            FLIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            alloc_grids(c);
            alloc_particles(c, 1);
        } catch (...) {
            free_all(c);
            delete c;
            throw;
        }
        *out = reinterpret_cast<flip_ctx *>(c);
        return FLIP_OK;
    });
}

int flip_gpu_destroy(flip_ctx *ctx)
{
    if (!ctx) return FLIP_OK;
    FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return FLIP_OK;
}

int flip_gpu_synchronize(flip_ctx *ctx)
{
    return guarded([&]() -> int {
        require(ctx, "null context");
        FLIP_CHECK(hipStreamSynchronize(reinterpret_cast<FlipCtx *>(ctx)->stream));
        return FLIP_OK;
    });
}

int flip_gpu_upload_particles(flip_ctx *ctx, int64_t np, const float *pos, const float *vel, const float *nrm,
                              const int8_t *type, const float *mass, const float *dens)
{
    return guarded([&]() -> int {
        require(ctx && np >= 0, "bad argument");
        require(np == 0 || (pos && type), "pos and type are required");
        require(np < (int64_t(1) << 29), "too many particles (limit 2^29)");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        std::vector<float4> p4(np), v4(np), n4(np);
        std::vector<int32_t> mt(np);
        bool unit = true;
        for (int64_t i = 0; i < np; ++i) {
            int t = type[i];
            require(t == FLIP_FLUID || t == FLIP_WALL, "particle type must be FLUID or WALL");
            float m = mass ? mass[i] : 1.0f;
            require(m >= 0.0f, "negative mass");
            unit = unit && m == 1.0f;
            float w = t == FLIP_WALL ? -m : m;
            if (t == FLIP_WALL && m == 0.0f) w = -0.0f;
            p4[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], w);
            v4[i] = make_float4(vel ? vel[3 * i] : 0.0f, vel ? vel[3 * i + 1] : 0.0f, vel ? vel[3 * i + 2] : 0.0f,
                                dens ? dens[i] : 10.0f);
            n4[i] = make_float4(nrm ? nrm[3 * i] : 0.0f, nrm ? nrm[3 * i + 1] : 0.0f, nrm ? nrm[3 * i + 2] : 0.0f,
                                0.0f);
            mt[i] = (int32_t)((uint32_t)i << 2 | (uint32_t)t);
        }
        if (np > c->cap || np == 0) alloc_particles(c, std::max<int64_t>(np, 1));
        c->np = np;
        c->cur = 0;
        if (np > 0) {
            FLIP_CHECK(hipMemcpy(c->pos[0], p4.data(), np * sizeof(float4), hipMemcpyHostToDevice));
            FLIP_CHECK(hipMemcpy(c->vel[0], v4.data(), np * sizeof(float4), hipMemcpyHostToDevice));
            FLIP_CHECK(hipMemcpy(c->meta[0], mt.data(), np * sizeof(int32_t), hipMemcpyHostToDevice));
            FLIP_CHECK(hipMemcpy(c->nrm_id, n4.data(), np * sizeof(float4), hipMemcpyHostToDevice));
        }
        build_wall_bins(c, p4, mt, n4);
        c->unit_mass = unit;
        c->sorted = false;
        c->bucketed = false;
        c->uploaded = true;
        return FLIP_OK;
    });
}

int flip_gpu_download_particles(flip_ctx *ctx, float *pos, float *vel, float *nrm, int8_t *type, float *dens)
{
    return guarded([&]() -> int {
        require(ctx, "null context");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        int64_t np = c->np;
        std::vector<float4> p4(np), v4(np), n4(np);
        std::vector<int32_t> mt(np);
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        if (np > 0) {
            FLIP_CHECK(hipMemcpy(p4.data(), c->P4(), np * sizeof(float4), hipMemcpyDeviceToHost));
            FLIP_CHECK(hipMemcpy(v4.data(), c->V4(), np * sizeof(float4), hipMemcpyDeviceToHost));
            FLIP_CHECK(hipMemcpy(mt.data(), c->M(), np * sizeof(int32_t), hipMemcpyDeviceToHost));
            FLIP_CHECK(hipMemcpy(n4.data(), c->nrm_id, np * sizeof(float4), hipMemcpyDeviceToHost));
        }
        for (int64_t s = 0; s < np; ++s) {
            int64_t id = (uint32_t)mt[s] >> 2;
            require(id < np, "corrupt particle index", FLIP_E_STATE);
            if (pos) {
                pos[3 * id] = p4[s].x;
                pos[3 * id + 1] = p4[s].y;
                pos[3 * id + 2] = p4[s].z;
            }
            if (vel) {
                vel[3 * id] = v4[s].x;
                vel[3 * id + 1] = v4[s].y;
                vel[3 * id + 2] = v4[s].z;
            }
            if (dens) dens[id] = v4[s].w;
            if (type) type[id] = (int8_t)(mt[s] & 3);
        }
        if (nrm)
            for (int64_t i = 0; i < np; ++i) {
                nrm[3 * i] = n4[i].x;
                nrm[3 * i + 1] = n4[i].y;
                nrm[3 * i + 2] = n4[i].z;
            }
        return FLIP_OK;
    });
}

int flip_gpu_set_emitter(flip_ctx *ctx, float x, float z, float r, int pour_time)
{
    return guarded([&]() -> int {
        require(ctx, "null context");
        require(std::isfinite(x) && std::isfinite(z) && std::isfinite(r) && r >= 0.0f, "bad emitter disc");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        set_emitter(c, x, z, r, pour_time);
        return FLIP_OK;
    });
}

int flip_gpu_mesh(flip_ctx *ctx, int mg, int64_t *nvert, int64_t *nface)
{
    return guarded([&]() -> int {
        require(ctx, "null context");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        require(c->uploaded, "no particles uploaded", FLIP_E_STATE);
        if (mg <= 0) mg = c->n + 1;
        require(mg >= 2 && (int64_t)mg * mg * mg * 3 < (int64_t(1) << 31), "lattice size out of range");
        long nv = 0, nf = 0;
        timing_reset(c);
        mesh_build(c, mg, &nv, &nf);
        timing_collect(c);
        if (nvert) *nvert = nv;
        if (nface) *nface = nf;
        return FLIP_OK;
    });
}

int flip_gpu_mesh_get(flip_ctx *ctx, double *vertices, double *normals, int32_t *faces)
{
    return guarded([&]() -> int {
        require(ctx, "null context");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        require(c->mesh != nullptr, "no mesh built", FLIP_E_STATE);
        mesh_copy_out(c, vertices, normals, faces);
        return FLIP_OK;
    });
}

int flip_gpu_set_step(flip_ctx *ctx, int step)
{
    if (!ctx) return fail(FLIP_E_ARG, "null context");
    if (step < 0) return fail(FLIP_E_ARG, "negative step");
    reinterpret_cast<FlipCtx *>(ctx)->step = step;
    return FLIP_OK;
}

int flip_gpu_get_step(flip_ctx *ctx, int *step)
{
    if (!ctx || !step) return fail(FLIP_E_ARG, "null argument");
    *step = reinterpret_cast<FlipCtx *>(ctx)->step;
    return FLIP_OK;
}

int flip_gpu_num_particles(flip_ctx *ctx, int64_t *np)
{
    if (!ctx || !np) return fail(FLIP_E_ARG, "null argument");
    *np = reinterpret_cast<FlipCtx *>(ctx)->np;
    return FLIP_OK;
}

int flip_gpu_set_max_dens(flip_ctx *ctx, float v)
{
    if (!ctx) return fail(FLIP_E_ARG, "null context");
    reinterpret_cast<FlipCtx *>(ctx)->max_dens = v;
    return FLIP_OK;
}

int flip_gpu_get_max_dens(flip_ctx *ctx, float *v)
{
    if (!ctx || !v) return fail(FLIP_E_ARG, "null argument");
    *v = reinterpret_cast<FlipCtx *>(ctx)->max_dens;
    return FLIP_OK;
}

int flip_gpu_srand(flip_ctx *ctx, unsigned seed)
{
    if (!ctx) return fail(FLIP_E_ARG, "null context");
    FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
    c->rng.seed(seed);
    c->rng_on_device = false;
    return FLIP_OK;
}

int flip_gpu_set_dt(flip_ctx *ctx, double dt)
{
    if (!ctx) return fail(FLIP_E_ARG, "null context");
    reinterpret_cast<FlipCtx *>(ctx)->dt = dt;
    return FLIP_OK;
}

static void *grid_ptr(FlipCtx *c, int which, size_t &bytes)
{
    switch (which) {
    case FLIP_GRID_U0: case FLIP_GRID_U1: case FLIP_GRID_U2:
        bytes = c->F * sizeof(float);
        return c->ug[which];
    case FLIP_GRID_US0: case FLIP_GRID_US1: case FLIP_GRID_US2:
        bytes = c->F * sizeof(float);
        return c->us[which - FLIP_GRID_US0];
    case FLIP_GRID_A:
        bytes = c->C;
        return c->A;
    case FLIP_GRID_L:
        bytes = c->C * sizeof(float);
        return c->L;
    case FLIP_GRID_PRESS:
        bytes = c->C * sizeof(float);
        return c->press;
    case FLIP_GRID_DIV:
        bytes = c->C * sizeof(float);
        return c->div;
    }
    throw flip_arg_error{FLIP_E_ARG, "unknown grid id"};
}

int flip_gpu_get_grid(flip_ctx *ctx, int which, void *out)
{
    return guarded([&]() -> int {
        require(ctx && out, "null argument");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        size_t bytes = 0;
        void *p = grid_ptr(c, which, bytes);
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        FLIP_CHECK(hipMemcpy(out, p, bytes, hipMemcpyDeviceToHost));
        return FLIP_OK;
    });
}

int flip_gpu_set_grid(flip_ctx *ctx, int which, const void *in)
{
    return guarded([&]() -> int {
        require(ctx && in, "null argument");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        size_t bytes = 0;
        void *p = grid_ptr(c, which, bytes);
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        FLIP_CHECK(hipMemcpy(p, in, bytes, hipMemcpyHostToDevice));
        return FLIP_OK;
    });
}

int flip_gpu_stage(flip_ctx *ctx, int stage, flip_step_stats *st)
{
    return guarded([&]() -> int {
        require(ctx, "null context");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        require(c->uploaded, "no particles uploaded", FLIP_E_STATE);
        if (st) memset(st, 0, sizeof(*st));
        k_launch_set_dt(c, c->dt);
        c->solve_timed = false;
        run_stage(c, stage, st);
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        FLIP_CHECK(hipGetLastError());
        if (st && c->solve_timed) st->ms_solve = elapsed(c->ev_s0, c->ev_s1);
        c->solve_timed = false;
        int32_t ev[EV_SLOTS];
        if (c->corrected || c->advected) FLIP_CHECK(hipMemcpy(ev, c->ev_count, sizeof(ev), hipMemcpyDeviceToHost));
        if (c->advected) {
            c->advected = false;
            if (st) st->n_reposition = ev[EV_MOVED];
            if (ev[EV_NONFINITE_ADV])
                return fail(FLIP_E_NUMERIC, "non-finite particle position or velocity in the advection");
        }
        if (c->corrected) {
            c->corrected = false;
            if (st) st->n_jitter = ev[3];
            if (ev[1]) return fail(FLIP_E_STATE, "too many corrector jitter events");
        }
        return FLIP_OK;
    });
}

// flip3D.cpp:1196-1199: the step counter, then pourWater
static void begin_step(FlipCtx *c)
{
    c->step++;
    if (c->step <= c->pour_time && c->emit_n > 0) {
        ensure_particle_capacity(c, c->np + c->emit_n);
        k_launch_emit(c);
    }
    c->begun = true;
}

int flip_gpu_step_begin(flip_ctx *ctx)
{
    return guarded([&]() -> int {
        require(ctx, "null context");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        require(c->uploaded, "no particles uploaded", FLIP_E_STATE);
        require(!c->begun, "flip_gpu_step_begin twice before a step", FLIP_E_STATE);
        begin_step(c);
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        return FLIP_OK;
    });
}

int flip_gpu_render_preview(flip_ctx *ctx, uint8_t *rgba)
{
    return guarded([&]() -> int {
        require(ctx && rgba, "null argument");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        require(c->uploaded, "no particles uploaded", FLIP_E_STATE);
        preview_render(c, rgba);
        return FLIP_OK;
    });
}

int flip_gpu_write_preview(flip_ctx *ctx, const char *path)
{
    return guarded([&]() -> int {
        require(ctx && path, "null argument");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        require(c->uploaded, "no particles uploaded", FLIP_E_STATE);
        std::vector<uint8_t> img(4 * (size_t)preview_width() * preview_height());
        preview_render(c, img.data());
        require(write_bmp_file(path, img.data(), preview_width(), preview_height()), "cannot write the preview file");
        return FLIP_OK;
    });
}

static void mesh_host(FlipCtx *c, std::vector<double> &v, std::vector<double> &n, std::vector<int32_t> &f)
{
    require(c->mesh != nullptr, "no mesh built (flip_gpu_mesh)", FLIP_E_STATE);
    long nv = 0, nf = 0;
    mesh_counts(c, &nv, &nf);
    v.resize(3 * nv);
    n.resize(3 * nv);
    f.resize(3 * nf);
    mesh_copy_out(c, v.data(), n.data(), f.data());
}

int flip_gpu_write_obj(flip_ctx *ctx, const char *path)
{
    return guarded([&]() -> int {
        require(ctx && path, "null argument");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        std::vector<double> v, n;
        std::vector<int32_t> f;
        mesh_host(c, v, n, f);
        require(write_obj_file(path, v, n, f, (float)(1.0 / c->n)), "cannot write the OBJ file");
        return FLIP_OK;
    });
}

int flip_gpu_write_sunflow(flip_ctx *ctx, const char *path, const char *template_path)
{
    return guarded([&]() -> int {
        require(ctx && path, "null argument");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        std::vector<double> v, n;
        std::vector<int32_t> f;
        mesh_host(c, v, n, f);
        if (template_path) require(write_sunflow_template(template_path), "cannot write the Sunflow template");
        require(write_sunflow_file(path, v, n, f, (float)(1.0 / c->n)), "cannot write the Sunflow scene file");
        return FLIP_OK;
    });
}

// What the captured step parts bake into their launches: the particle count,
// the parity of the double-buffered storage, the host scalars and parameters
// the kernels take, and every device buffer that can be reallocated (particle
// storage, its scratch, the wall bins).  A change means a new capture.
static std::vector<uint64_t> step_graph_sig(const FlipCtx *c)
{
    std::vector<uint64_t> s;
    auto ptr = [&](const void *p) { s.push_back((uint64_t)(uintptr_t)p); };
    s.push_back((uint64_t)c->np);
    s.push_back((uint64_t)c->cur);
    s.push_back(c->unit_mass ? 1u : 0u);
    uint32_t md;
    memcpy(&md, &c->max_dens, sizeof(md));
    s.push_back(md);
    s.push_back((uint64_t)c->ev_cap);
    const unsigned char *pb = reinterpret_cast<const unsigned char *>(&c->prm);
    for (size_t i = 0; i < sizeof(c->prm); i += 8) {
        uint64_t w = 0;
        memcpy(&w, pb + i, std::min<size_t>(8, sizeof(c->prm) - i));
        s.push_back(w);
    }
    for (int b = 0; b < 2; ++b) {
        ptr(c->pos[b]);
        ptr(c->vel[b]);
        ptr(c->meta[b]);
    }
    const void *bufs[] = {c->nrm_id,   c->inv,       c->spos,      c->xyz3,     c->dw,       c->tmp0,     c->tmp1,
                          c->pkey,     c->pslot,     c->perm,      c->sort_big, c->deep,     c->ev_list,
                          c->ev_off,   c->ev_slot,   c->repo_ids,  c->repo_order, c->repo_loff, c->ev_rand,
                          c->wall_start, c->wall_near, c->wall_pos, c->wall_nrm};
    for (const void *b : bufs) ptr(b);
    return s;
}

// ---- the step in parts (flip_gpu_step, flip_gpu_step_batch)
// flip3D.cpp:1194-1264 before the pressure solve
// whether the step's passes prepare the FP32 PCG's setup (the FP64 and
// refinement solves prepare their own)
static bool step_solve_prepared(const FlipCtx *c) { return !c->prm.solve_fp64 && c->prm.refine <= 0; }

static void step_launch_pre(FlipCtx *c)
{
    // flip3D.cpp:1221-1222 sort + computeDensity; the time step (kept on the
    // device, read back with the step's end) from the speeds the sort's
    // gather forms
    k_launch_sort(c, c->prm.dt_policy == 1 ? c->partials : nullptr);
    c->sorted = true;
    k_launch_choose_dt(c);
    // with add_ExtForce (:1230) fused in, and the inputs of flags and P2G
    // written (densities with the wall mark, P2G's packed records)
    const bool records = k_launch_density(c, true, true);
    // the FP32 PCG's setup (negated div, FLUID flags) formed by the passes
    // that produce its inputs (step_solve tells the solve)
    const bool prep = step_solve_prepared(c);
    // solve_picflip (:1057-1103); its sort (:1059) sees unchanged positions
    k_launch_flags(c, records, prep);
    k_launch_p2g(c, records);
    k_launch_divergence(c, prep);  // project (:807-868): div, solve, gradient
}

// and after it, ending with the step record in mapped host memory
static void step_launch_post(FlipCtx *c, flip_step_stats *st)
{
    k_launch_gradient(c);
    k_launch_extrapolate(c);
    k_launch_g2p(c, true);  // reads the extrapolation's (u, u - u_save) pairs
    run_stage(c, FLIP_STAGE_ADVECT, st);  // :1236
    if (c->prm.correction) run_stage(c, FLIP_STAGE_CORRECT, st);  // :1242
    k_launch_step_record(c);
}

static std::atomic<uint64_t> g_step_capture_gen{0};

// the start of a step: statistics, timing, step counter and emitter
static void step_open(FlipCtx *c, flip_step_stats *st)
{
    require(c->uploaded, "no particles uploaded", FLIP_E_STATE);
    memset(st, 0, sizeof(*st));
    timing_reset(c);
    hipEvent_t *evs[4] = {&c->ev_a, &c->ev_b, &c->ev_s0, &c->ev_s1};
    for (hipEvent_t *e : evs)
        if (!*e) FLIP_CHECK(hipEventCreate(e));
    if (!c->begun) begin_step(c);
    c->begun = false;
}

// whether this step replays its captured parts; drops them when what they
// bake in has changed
static bool step_graphs_usable(FlipCtx *c)
{
    const bool emitting = c->step <= c->pour_time && c->emit_n > 0;
    if (!c->step_graphs || c->profiling || c->direct_launch || emitting || c->prm.refine != 0 || c->prm.solve_fp64)
        return false;
    std::vector<uint64_t> sig = step_graph_sig(c);
    if (sig != c->step_graph_sig) {
        for (int w = 0; w < 2; ++w) {
            if (c->step_gx[w]) FLIP_CHECK(hipGraphExecDestroy(c->step_gx[w]));
            if (c->step_g[w]) FLIP_CHECK(hipGraphDestroy(c->step_g[w]));
            c->step_gx[w] = nullptr;
            c->step_g[w] = nullptr;
        }
        c->step_graph_sig = std::move(sig);
    }
    rng_to_device(c);  // no host copy inside a captured part
    return true;
}

// capture part `which` (0: before the solve, 1: after it) unless it is
// captured; the graph is kept for flip_gpu_step_batch
static void step_capture(FlipCtx *c, int which, flip_step_stats *st)
{
    if (c->step_gx[which]) return;
    hipGraph_t g;
    FLIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    if (which == 0) step_launch_pre(c);
    else step_launch_post(c, st);
    FLIP_CHECK(hipStreamEndCapture(c->stream, &g));
    FLIP_CHECK(hipGraphInstantiate(&c->step_gx[which], g, nullptr, nullptr, 0));
    c->step_g[which] = g;
    c->step_gen = ++g_step_capture_gen;
}

// the pressure solve between the parts, on the context's stream
static void step_solve(FlipCtx *c, flip_step_stats *st)
{
    c->solve_prepared = step_solve_prepared(c);
    FLIP_CHECK(hipEventRecord(c->ev_s0, c->stream));
    const int rc = k_solve(c, st);
    FLIP_CHECK(hipEventRecord(c->ev_s1, c->stream));
    c->solve_timed = true;
    if (rc == FLIP_E_NUMERIC) throw flip_arg_error{FLIP_E_NUMERIC, "pressure solve produced NaN"};
}

// the host's view of the storage after the step (what the launches set when
// they are captured; a replay does not run them)
static void step_host_state(FlipCtx *c)
{
    c->bucketed = true;
    c->sorted = !c->prm.correction;
    c->advected = true;
    c->corrected = c->prm.correction != 0;
}

// the step's record, once the device has passed it
static int step_close(FlipCtx *c, flip_step_stats *st)
{
    const FlipCtx::StepRecord &rec = *c->h_rec;
    c->dt = rec.dt;
    st->dt = c->dt;
    if (c->solve_timed) st->ms_solve = elapsed(c->ev_s0, c->ev_s1);
    c->solve_timed = false;
    c->advected = false;
    st->n_reposition = rec.ev[EV_MOVED];
    if (rec.ev[EV_NONFINITE_ADV])
        return fail(FLIP_E_NUMERIC, "non-finite particle position or velocity in the advection");
    if (c->corrected) {
        st->n_jitter = rec.ev[3];
        c->corrected = false;
        if (rec.ev[1]) return fail(FLIP_E_STATE, "too many corrector jitter events");
        if (rec.ev[EV_NONFINITE])
            return fail(FLIP_E_NUMERIC, "non-finite particle position or velocity after the corrector");
    }
    timing_collect(c);
    return FLIP_OK;
}

int flip_gpu_step(flip_ctx *ctx, flip_step_stats *st)
{
    return guarded([&]() -> int {
        require(ctx, "null context");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        flip_step_stats local;
        if (!st) st = &local;
        step_open(c, st);
        const bool graphs = step_graphs_usable(c);
        FLIP_CHECK(hipEventRecord(c->ev_a, c->stream));
        if (graphs) {
            step_capture(c, 0, st);
            FLIP_CHECK(hipGraphLaunch(c->step_gx[0], c->stream));
        } else {
            step_launch_pre(c);
        }
        step_solve(c, st);
        if (graphs) {
            step_capture(c, 1, st);
            FLIP_CHECK(hipGraphLaunch(c->step_gx[1], c->stream));
        } else {
            step_launch_post(c, st);
        }
        step_host_state(c);
        FLIP_CHECK(hipEventRecord(c->ev_b, c->stream));
        FLIP_CHECK(hipEventSynchronize(c->ev_b));
        FLIP_CHECK(hipGetLastError());
        st->ms_step = elapsed(c->ev_a, c->ev_b);
        return step_close(c, st);
    });
}

// flip_gpu_step_batch's graphs: one child graph per context for each part,
// built again when the context list or any context's capture changes
namespace {
struct BatchGraphs {
    std::vector<std::pair<const FlipCtx *, uint64_t>> key;  // (context, capture number)
    hipGraphExec_t gx[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> done;  // one per context: its solve has been issued
    hipEvent_t pre = nullptr, end = nullptr;
    void clear()
    {
        for (auto &g : gx)
            if (g) (void)hipGraphExecDestroy(g);
        gx[0] = gx[1] = nullptr;
        for (auto e : done) (void)hipEventDestroy(e);
        done.clear();
        if (pre) (void)hipEventDestroy(pre);
        if (end) (void)hipEventDestroy(end);
        pre = end = nullptr;
        key.clear();
    }
};
std::mutex g_batch_mu;
BatchGraphs g_batch;
}  // namespace

int flip_gpu_step_batch(flip_ctx **ctxs, int count, int threads, flip_step_stats *stats, double *ms)
{
    if (!ctxs || count <= 0) return fail(FLIP_E_ARG, "no contexts");
    if (threads <= 0) return fail(FLIP_E_ARG, "threads must be >= 1");
    for (int i = 0; i < count; ++i)
        if (!ctxs[i]) return fail(FLIP_E_ARG, "null context");
    for (int i = 0; i < count; ++i)
        for (int j = i + 1; j < count; ++j)
            if (ctxs[i] == ctxs[j]) return fail(FLIP_E_ARG, "a context appears twice");
    std::lock_guard<std::mutex> lk(g_batch_mu);
    std::vector<flip_step_stats> local(stats ? 0 : count);
    flip_step_stats *st = stats ? stats : local.data();
    return guarded([&]() -> int {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<FlipCtx *> cs(count);
        for (int i = 0; i < count; ++i) {
            FlipCtx *c = reinterpret_cast<FlipCtx *>(ctxs[i]);
            cs[i] = c;
            step_open(c, &st[i]);
            require(step_graphs_usable(c),
                    "flip_gpu_step_batch needs step graphs on every context (no emitter step, profiling, direct "
                    "launch, refinement or FP64 solve)",
                    FLIP_E_STATE);
            // both parts captured before any runs (their host bookkeeping
            // ends where it began: two sorts)
            step_capture(c, 0, &st[i]);
            step_capture(c, 1, &st[i]);
        }
        BatchGraphs &B = g_batch;
        std::vector<std::pair<const FlipCtx *, uint64_t>> key(count);
        for (int i = 0; i < count; ++i) key[i] = {cs[i], cs[i]->step_gen};
        if (key != B.key) {
            B.clear();
            for (int w = 0; w < 2; ++w) {
                hipGraph_t g;
                FLIP_CHECK(hipGraphCreate(&g, 0));
                for (FlipCtx *c : cs) {
                    hipGraphNode_t node;
                    FLIP_CHECK(hipGraphAddChildGraphNode(&node, g, nullptr, 0, c->step_g[w]));
                }
                FLIP_CHECK(hipGraphInstantiate(&B.gx[w], g, nullptr, nullptr, 0));
                FLIP_CHECK(hipGraphDestroy(g));
            }
            B.done.resize(count, nullptr);
            for (auto &e : B.done) FLIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            FLIP_CHECK(hipEventCreateWithFlags(&B.pre, hipEventDisableTiming));
            FLIP_CHECK(hipEventCreateWithFlags(&B.end, hipEventDisableTiming));
            B.key = key;
        }
        hipStream_t master = cs[0]->stream;
        // every context's kernels before its solve: one launch
        FLIP_CHECK(hipGraphLaunch(B.gx[0], master));
        FLIP_CHECK(hipEventRecord(B.pre, master));
        for (int i = 1; i < count; ++i) FLIP_CHECK(hipStreamWaitEvent(cs[i]->stream, B.pre, 0));
        // the solves, each on its context's stream, driven by host threads
        threads = std::min(threads, count);
        std::atomic<int> next{0};
        std::mutex mu;
        int first_rc = FLIP_OK, first_at = count;
        std::string first_msg;
        auto worker = [&]() {
            for (;;) {
                const int i = next.fetch_add(1);
                if (i >= count) return;
                const int rc = guarded([&]() -> int {
                    step_solve(cs[i], &st[i]);
                    FLIP_CHECK(hipEventRecord(B.done[i], cs[i]->stream));
                    return FLIP_OK;
                });
                if (rc < 0) {
                    std::lock_guard<std::mutex> l2(mu);
                    if (i < first_at) {
                        first_at = i;
                        first_rc = rc;
                        first_msg = flip_gpu_last_error();
                    }
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < threads; ++t) pool.emplace_back(worker);
        worker();
        for (auto &th : pool) th.join();
        if (first_rc < 0) throw flip_arg_error{first_rc, "context " + std::to_string(first_at) + ": " + first_msg};
        // every context's kernels after its solve: one launch
        for (int i = 0; i < count; ++i) FLIP_CHECK(hipStreamWaitEvent(master, B.done[i], 0));
        FLIP_CHECK(hipGraphLaunch(B.gx[1], master));
        FLIP_CHECK(hipEventRecord(B.end, master));
        FLIP_CHECK(hipEventSynchronize(B.end));
        FLIP_CHECK(hipGetLastError());
        const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        int rc = FLIP_OK;
        for (int i = 0; i < count; ++i) {
            FlipCtx *c = cs[i];
            step_host_state(c);
            st[i].ms_step = wall;
            if (ms) ms[i] = wall;
            const int r = step_close(c, &st[i]);
            if (r < 0 && rc == FLIP_OK) rc = r;
        }
        return rc;
    });
}

int flip_gpu_step_many(flip_ctx **ctxs, int count, int threads, flip_step_stats *stats, double *ms)
{
    if (!ctxs || count <= 0) return fail(FLIP_E_ARG, "no contexts");
    if (threads <= 0) return fail(FLIP_E_ARG, "threads must be >= 1");
    for (int i = 0; i < count; ++i)
        if (!ctxs[i]) return fail(FLIP_E_ARG, "null context");
    for (int i = 0; i < count; ++i)
        for (int j = i + 1; j < count; ++j)
            if (ctxs[i] == ctxs[j]) return fail(FLIP_E_ARG, "a context appears twice");
    threads = std::min(threads, count);
    std::atomic<int> next{0};
    std::mutex mu;
    int first_rc = FLIP_OK, first_at = count;
    std::string first_msg;
    auto worker = [&]() {
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= count) return;
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = flip_gpu_step(ctxs[i], stats ? &stats[i] : nullptr);
            const auto t1 = std::chrono::steady_clock::now();
            if (ms) ms[i] = std::chrono::duration<double, std::milli>(t1 - t0).count();
            if (rc < 0) {
                std::lock_guard<std::mutex> lk(mu);
                if (i < first_at) {
                    first_at = i;
                    first_rc = rc;
                    first_msg = flip_gpu_last_error();
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
    if (first_rc < 0) return fail(first_rc, ("context " + std::to_string(first_at) + ": " + first_msg).c_str());
    return FLIP_OK;
}

int flip_gpu_map_p2g(flip_ctx *ctx, float *u0, float *u1, float *u2)
{
    return guarded([&]() -> int {
        require(ctx, "null context");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        require(c->uploaded, "no particles uploaded", FLIP_E_STATE);
        run_stage(c, FLIP_STAGE_P2G, nullptr);
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        float *o[3] = {u0, u1, u2};
        for (int k = 0; k < 3; ++k)
            if (o[k]) FLIP_CHECK(hipMemcpy(o[k], c->us[k], c->F * sizeof(float), hipMemcpyDeviceToHost));
        return FLIP_OK;
    });
}

int flip_gpu_map_g2p(flip_ctx *ctx, const float *u0, const float *u1, const float *u2, int64_t m, const float *pos,
                     float *vel)
{
    return guarded([&]() -> int {
        require(ctx && u0 && u1 && u2 && (m == 0 || (pos && vel)) && m >= 0, "bad argument");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        float *g[3] = {nullptr, nullptr, nullptr};
        float4 *p = nullptr, *o = nullptr;
        try {
            const float *src[3] = {u0, u1, u2};
            for (int k = 0; k < 3; ++k) {
                dalloc(&g[k], c->F);
                FLIP_CHECK(hipMemcpy(g[k], src[k], c->F * sizeof(float), hipMemcpyHostToDevice));
            }
            std::vector<float4> hp(m), ho(m);
            for (int64_t i = 0; i < m; ++i) hp[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.0f);
            dalloc(&p, m);
            dalloc(&o, m);
            if (m > 0) FLIP_CHECK(hipMemcpy(p, hp.data(), m * sizeof(float4), hipMemcpyHostToDevice));
            k_launch_g2p_points(c, g[0], g[1], g[2], m, p, o);
            FLIP_CHECK(hipStreamSynchronize(c->stream));
            if (m > 0) FLIP_CHECK(hipMemcpy(ho.data(), o, m * sizeof(float4), hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < m; ++i) {
                vel[3 * i] = ho[i].x;
                vel[3 * i + 1] = ho[i].y;
                vel[3 * i + 2] = ho[i].z;
            }
        } catch (...) {
            for (int k = 0; k < 3; ++k) dfree(g[k]);
            dfree(p);
            dfree(o);
            throw;
        }
        for (int k = 0; k < 3; ++k) dfree(g[k]);
        dfree(p);
        dfree(o);
        return FLIP_OK;
    });
}

int flip_gpu_solve(flip_ctx *ctx, const int8_t *A, const float *L, float *x, float *b, int subcell,
                   flip_step_stats *st)
{
    return guarded([&]() -> int {
        require(ctx && A && L && x && b, "null argument");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        flip_step_stats local;
        if (!st) st = &local;
        memset(st, 0, sizeof(*st));
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        FLIP_CHECK(hipMemcpy(c->A, A, c->C, hipMemcpyHostToDevice));
        FLIP_CHECK(hipMemcpy(c->L, L, c->C * sizeof(float), hipMemcpyHostToDevice));
        FLIP_CHECK(hipMemcpy(c->press, x, c->C * sizeof(float), hipMemcpyHostToDevice));
        FLIP_CHECK(hipMemcpy(c->div, b, c->C * sizeof(float), hipMemcpyHostToDevice));
        int save = c->prm.subcell;
        c->prm.subcell = subcell;
        int rc = 0;
        try {
            rc = k_solve(c, st);
        } catch (...) {
            c->prm.subcell = save;
            throw;
        }
        c->prm.subcell = save;
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        // the reference's op() leaves non-FLUID x at 0 after the first iteration
        FLIP_CHECK(hipMemcpy(x, c->press, c->C * sizeof(float), hipMemcpyDeviceToHost));
        FLIP_CHECK(hipMemcpy(b, c->div, c->C * sizeof(float), hipMemcpyDeviceToHost));
        if (rc == FLIP_E_NUMERIC) return fail(FLIP_E_NUMERIC, "pressure solve produced NaN");
        return FLIP_OK;
    });
}

int flip_gpu_sorted_cells(flip_ctx *ctx, int32_t *counts, int32_t *ids)
{
    return guarded([&]() -> int {
        require(ctx, "null context");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        require(c->uploaded, "no particles uploaded", FLIP_E_STATE);
        k_launch_sort(c);
        c->sorted = true;
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        if (counts) {
            std::vector<int32_t> st(c->C + 1);
            FLIP_CHECK(hipMemcpy(st.data(), c->cell_start, (c->C + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
            for (long k = 0; k < c->C; ++k) counts[k] = st[k + 1] - st[k];
        }
        if (ids && c->np > 0) {
            std::vector<int32_t> mt(c->np);
            FLIP_CHECK(hipMemcpy(mt.data(), c->M(), c->np * sizeof(int32_t), hipMemcpyDeviceToHost));
            for (int64_t s = 0; s < c->np; ++s) ids[s] = (int32_t)((uint32_t)mt[s] >> 2);
        }
        return FLIP_OK;
    });
}

int flip_gpu_kernel_times(flip_ctx *ctx, char *names, int names_cap, double *ms, int ms_cap)
{
    if (!ctx) return fail(FLIP_E_ARG, "null context");
    FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
    std::string all;
    int cnt = (int)c->tms.size();
    for (int i = 0; i < cnt; ++i) {
        if (i) all += ";";
        all += c->tnames[i];
        if (ms && i < ms_cap) ms[i] = c->tms[i];
    }
    if (names && names_cap > 0) {
        strncpy(names, all.c_str(), names_cap - 1);
        names[names_cap - 1] = 0;
    }
    return cnt;
}

int flip_gpu_set_launch(flip_ctx *ctx, int pcg_blocks, int pcg_batch)
{
    return guarded([&]() -> int {
        require(ctx, "null context");
        // 256: the length of the solver's per-block partials arrays
        require(pcg_blocks >= 0 && pcg_blocks <= 256, "pcg_blocks must lie in [0, 256]");
        require(pcg_batch >= 0 && pcg_batch <= 64, "pcg_batch must lie in [0, 64]");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        c->tune_blocks = pcg_blocks;
        c->tune_batch = pcg_batch;
        return FLIP_OK;
    });
}

int flip_gpu_set_step_graphs(flip_ctx *ctx, int on)
{
    return guarded([&]() -> int {
        require(ctx, "null context");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        c->step_graphs = on != 0;
        return FLIP_OK;
    });
}

int flip_gpu_set_learned_weights(flip_ctx *ctx, const char *path)
{
    return guarded([&]() -> int {
        require(ctx && path, "null argument");
        FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
        FILE *f = fopen(path, "rb");
        require(f != nullptr, "cannot open U-Net weights file");
        fclose(f);
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        learned_free(c->learned);  // reloaded on the next learned solve
        c->learned = nullptr;
        c->learned_weights = path;
        return FLIP_OK;
    });
}

int flip_gpu_set_profiling(flip_ctx *ctx, int on)
{
    if (!ctx) return fail(FLIP_E_ARG, "null context");
    reinterpret_cast<FlipCtx *>(ctx)->profiling = on != 0;
    return FLIP_OK;
}

int flip_gpu_selftest_math(int64_t m, const float *x, const float *y, float *out)
{
    return guarded([&]() -> int {
        require(m >= 0 && (m == 0 || (x && y && out)), "bad argument");
        float *dx = nullptr, *dy = nullptr, *dout = nullptr;
        dalloc(&dx, m);
        dalloc(&dy, m);
        dalloc(&dout, 5 * m);
        FLIP_CHECK(hipMemcpy(dx, x, m * sizeof(float), hipMemcpyHostToDevice));
        FLIP_CHECK(hipMemcpy(dy, y, m * sizeof(float), hipMemcpyHostToDevice));
        k_launch_selftest(m, dx, dy, dout, nullptr);
        FLIP_CHECK(hipDeviceSynchronize());
        FLIP_CHECK(hipMemcpy(out, dout, 5 * m * sizeof(float), hipMemcpyDeviceToHost));
        dfree(dx);
        dfree(dy);
        dfree(dout);
        return FLIP_OK;
    });
}

}  // extern "C"
