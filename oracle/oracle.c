/*
 * oracle.c - CPU restatement of the flip3d per-step FLIP path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * timed CPU baseline ("port").  Never linked into the product library.
 *
 * Each function states the reference's arithmetic in the reference's order;
 * the file:line citations are into mmmovania/flip3d src/.  Conventions that
 * matter for bit-exactness:
 *   - FLOAT ops round to FLOAT; expressions with a double literal evaluate in
 *     double and round on assignment (common.h:10 + C promotion rules).
 *   - No FMA contraction: built with -ffp-contract=off.
 *   - Loop orders, serial sections and window orders follow the reference.
 * Where the reference reads out of bounds (mapper.cpp:102 for grid[1] with
 * k+1 == d), this restatement reads 0.0; that read only ever carries weight
 * (z-k) == 0 for wall particles, whose velocities are never used.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_F64
#define R_SQRT sqrt
#define R_HYPOT hypot
#else
#define R_SQRT sqrtf
#define R_HYPOT hypotf
#endif

#define AIR ORC_AIR
#define FLUID ORC_FLUID
#define WALL ORC_WALL

struct orc_sim {
    orc_params prm;
    int n;
    long C;
    int np, cap;
    real *p, *u, *nrm, *m, *dens, *tmp0, *tmp1;
    int8_t *type, *remove;
    int32_t *cell_start, *cell_pid, *cell_fill;
    real *ug[3], *us[3];
    int8_t *A;
    real *L, *press, *div;
    double *P, *q;
    real *r, *z, *s, *tmp;
    real max_dens;
    double cur_dt;
    /* step counter and pourWater emitter (flip3D.cpp:41-43,59,375-402) */
    int step, pour_time;
    real pour_pos[2], pour_rad;
    /* glibc random_r TYPE_3 state: 31 words, front/rear taps 3 apart */
    int32_t rs[31];
    int rf, rr;
};

/* ------------------------------------------------------------------ rand */
/* glibc random_r TYPE_3 (additive lagged generator, taps 31 and 3) seeded as
 * srandom_r does: 16807 LCG fill (Schrage form), then 310 discarded outputs.
 * The reference never calls srand, so its stream is seed 1. */
static int32_t orc_rand_next(orc_sim *s)
{
    uint32_t v = (uint32_t)s->rs[s->rf] + (uint32_t)s->rs[s->rr];
    s->rs[s->rf] = (int32_t)v;
    s->rf = (s->rf + 1) % 31;
    s->rr = (s->rr + 1) % 31;
    return (int32_t)(v >> 1);
}

void orc_srand(orc_sim *s, unsigned seed)
{
    if (seed == 0) seed = 1;
    int32_t w = (int32_t)seed;
    s->rs[0] = w;
    for (int i = 1; i < 31; ++i) {
        long hi = w / 127773, lo = w % 127773;
        long t = 16807 * lo - 2836 * hi;
        if (t < 0) t += 2147483647;
        w = (int32_t)t;
        s->rs[i] = w;
    }
    s->rf = 3;
    s->rr = 0;
    for (int k = 0; k < 310; ++k) (void)orc_rand_next(s);
}

int orc_rand(orc_sim *s) { return orc_rand_next(s); }

/* -------------------------------------------------------------- helpers */
static inline long cidx(int n, int i, int j, int k) { return ((long)i * n + j) * n + k; }
/* face grids: u0 (n+1,n,n), u1 (n,n+1,n), u2 (n,n,n+1) */
static inline long fidx(int n, int c, int i, int j, int k)
{
    switch (c) {
    case 0: return ((long)i * n + j) * n + k;
    case 1: return ((long)i * (n + 1) + j) * n + k;
    default: return ((long)i * n + j) * (n + 1) + k;
    }
}

/* kernel.cpp:4-6 */
real orc_smooth_kernel(real r2, real h) { return (real)fmax(1.0 - (double)(r2 / (h * h)), 0.0); }
/* kernel.cpp:8-10 */
real orc_sharp_kernel(real r2, real h) { return (real)fmax((double)(h * h) / fmax((double)r2, 1.0e-5) - 1.0, 0.0); }
/* utility.cpp:16-18 */
real orc_length(const real *a, const real *b) { return R_HYPOT(R_HYPOT(a[0] - b[0], a[1] - b[1]), a[2] - b[2]); }
/* utility.cpp:20-26: a*a + b*b + c*c in FLOAT */
real orc_length2(const real *a, const real *b)
{
    real x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return x * x + y * y + z * z;
}

/* sorter.cpp:35-37 cell key: (int) fmax(0, fmin(gn-1, gn*p)) */
static inline int cell_of(int gn, real p) { return (int)fmax(0.0, fmin((double)(gn - 1), (double)((real)gn * p))); }
/* flip3D.cpp:82,752,903: (int) fmin(N-1, fmax(0, p*N)) */
static inline int cell_of_b(int gn, real p) { return (int)fmin((double)(gn - 1), fmax(0.0, (double)(p * (real)gn))); }

/* ------------------------------------------------------------- lifecycle */
void orc_default_params(orc_params *p, int n)
{
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
    p->precond = 0;
    p->max_iters = 0;
    p->correction = 1;
    p->reposition = 1;
}

orc_sim *orc_create(const orc_params *prm)
{
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    s->prm = *prm;
    int n = prm->n;
    s->n = n;
    s->C = (long)n * n * n;
    long F = (long)(n + 1) * n * n;
    for (int c = 0; c < 3; ++c) {
        s->ug[c] = (real *)calloc(F, sizeof(real));
        s->us[c] = (real *)calloc(F, sizeof(real));
    }
    s->A = (int8_t *)calloc(s->C, 1);
    s->L = (real *)calloc(s->C, sizeof(real));
    s->press = (real *)calloc(s->C, sizeof(real));
    s->div = (real *)calloc(s->C, sizeof(real));
    s->P = (double *)calloc(s->C, sizeof(double));
    s->q = (double *)calloc(s->C, sizeof(double));
    s->r = (real *)calloc(s->C, sizeof(real));
    s->z = (real *)calloc(s->C, sizeof(real));
    s->s = (real *)calloc(s->C, sizeof(real));
    s->tmp = (real *)calloc(s->C, sizeof(real));
    s->cell_start = (int32_t *)calloc(s->C + 1, sizeof(int32_t));
    s->cell_fill = (int32_t *)calloc(s->C, sizeof(int32_t));
    s->cur_dt = prm->dt;
    s->pour_time = -1;
    s->pour_rad = (real)0.12;
    orc_srand(s, 1);
    return s;
}

static void free_particles(orc_sim *s)
{
    free(s->p); free(s->u); free(s->nrm); free(s->m); free(s->dens);
    free(s->tmp0); free(s->tmp1); free(s->type); free(s->remove); free(s->cell_pid);
    s->p = s->u = s->nrm = s->m = s->dens = s->tmp0 = s->tmp1 = NULL;
    s->type = s->remove = NULL;
    s->cell_pid = NULL;
    s->np = s->cap = 0;
}

static void reserve_particles(orc_sim *s, int cap)
{
    free_particles(s);
    s->cap = cap;
    s->p = (real *)calloc(3 * (size_t)cap + 3, sizeof(real));
    s->u = (real *)calloc(3 * (size_t)cap + 3, sizeof(real));
    s->nrm = (real *)calloc(3 * (size_t)cap + 3, sizeof(real));
    s->tmp0 = (real *)calloc(3 * (size_t)cap + 3, sizeof(real));
    s->tmp1 = (real *)calloc(3 * (size_t)cap + 3, sizeof(real));
    s->m = (real *)calloc((size_t)cap + 1, sizeof(real));
    s->dens = (real *)calloc((size_t)cap + 1, sizeof(real));
    s->type = (int8_t *)calloc((size_t)cap + 1, 1);
    s->remove = (int8_t *)calloc((size_t)cap + 1, 1);
    s->cell_pid = (int32_t *)calloc((size_t)cap + 1, sizeof(int32_t));
}

/* capacity for at least `need` particles, keeping the first np */
static void grow_particles(orc_sim *s, int need)
{
    if (need <= s->cap) return;
    int cap = need > 2 * s->cap ? need : 2 * s->cap;
    size_t v3 = (3 * (size_t)cap + 3) * sizeof(real), v1 = ((size_t)cap + 1) * sizeof(real);
    s->p = (real *)realloc(s->p, v3);
    s->u = (real *)realloc(s->u, v3);
    s->nrm = (real *)realloc(s->nrm, v3);
    s->tmp0 = (real *)realloc(s->tmp0, v3);
    s->tmp1 = (real *)realloc(s->tmp1, v3);
    s->m = (real *)realloc(s->m, v1);
    s->dens = (real *)realloc(s->dens, v1);
    s->type = (int8_t *)realloc(s->type, (size_t)cap + 1);
    s->remove = (int8_t *)realloc(s->remove, (size_t)cap + 1);
    s->cell_pid = (int32_t *)realloc(s->cell_pid, ((size_t)cap + 1) * sizeof(int32_t));
    s->cap = cap;
}

void orc_destroy(orc_sim *s)
{
    if (!s) return;
    free_particles(s);
    for (int c = 0; c < 3; ++c) { free(s->ug[c]); free(s->us[c]); }
    free(s->A); free(s->L); free(s->press); free(s->div); free(s->P); free(s->q);
    free(s->r); free(s->z); free(s->s); free(s->tmp); free(s->cell_start); free(s->cell_fill);
    free(s);
}

int orc_set_particles(orc_sim *s, int np, const real *p, const real *u, const real *nrm,
                      const int8_t *type, const real *m, const real *dens)
{
    reserve_particles(s, np);
    s->np = np;
    memcpy(s->p, p, 3 * (size_t)np * sizeof(real));
    if (u) memcpy(s->u, u, 3 * (size_t)np * sizeof(real));
    if (nrm) memcpy(s->nrm, nrm, 3 * (size_t)np * sizeof(real));
    memcpy(s->type, type, (size_t)np);
    for (int i = 0; i < np; ++i) {
        s->m[i] = m ? m[i] : (real)1.0;
        s->dens[i] = dens ? dens[i] : (real)10.0;
    }
    return 0;
}

int orc_num_particles(const orc_sim *s) { return s->np; }

void orc_get_particles(const orc_sim *s, real *p, real *u, real *nrm, int8_t *type, real *dens)
{
    size_t n3 = 3 * (size_t)s->np;
    if (p) memcpy(p, s->p, n3 * sizeof(real));
    if (u) memcpy(u, s->u, n3 * sizeof(real));
    if (nrm) memcpy(nrm, s->nrm, n3 * sizeof(real));
    if (type) memcpy(type, s->type, (size_t)s->np);
    if (dens) memcpy(dens, s->dens, (size_t)s->np * sizeof(real));
}

void orc_set_max_dens(orc_sim *s, real v) { s->max_dens = v; }
real orc_get_max_dens(const orc_sim *s) { return s->max_dens; }
double orc_current_dt(const orc_sim *s) { return s->cur_dt; }
void orc_set_dt(orc_sim *s, double dt) { s->cur_dt = dt; }

long orc_grid_size(const orc_sim *s, int which)
{
    if (which <= ORC_G_US2) return (long)(s->n + 1) * s->n * s->n;
    return s->C;
}

static void *grid_ptr(const orc_sim *s, int which)
{
    switch (which) {
    case ORC_G_U0: return s->ug[0];
    case ORC_G_U1: return s->ug[1];
    case ORC_G_U2: return s->ug[2];
    case ORC_G_US0: return s->us[0];
    case ORC_G_US1: return s->us[1];
    case ORC_G_US2: return s->us[2];
    case ORC_G_A: return s->A;
    case ORC_G_L: return s->L;
    case ORC_G_PRESS: return s->press;
    default: return s->div;
    }
}

void orc_get_grid(const orc_sim *s, int which, void *out)
{
    size_t es = which == ORC_G_A ? 1 : sizeof(real);
    memcpy(out, grid_ptr(s, which), (size_t)orc_grid_size(s, which) * es);
}

void orc_set_grid(orc_sim *s, int which, const void *in)
{
    size_t es = which == ORC_G_A ? 1 : sizeof(real);
    memcpy(grid_ptr(s, which), in, (size_t)orc_grid_size(s, which) * es);
}

void orc_get_sorted(const orc_sim *s, int32_t *cell_start, int32_t *cell_pid)
{
    if (cell_start) memcpy(cell_start, s->cell_start, (s->C + 1) * sizeof(int32_t));
    if (cell_pid) memcpy(cell_pid, s->cell_pid, (size_t)s->np * sizeof(int32_t));
}

/* ---------------------------------------------------------------- sorter */
/* sorter.cpp:22-40: bucket particles by cell in particle-vector order.  The
 * buckets are a stable counting sort: cell_pid[cell_start[c] ..] lists the
 * particles of cell c in ascending particle index. */
void orc_sort(orc_sim *s)
{
    int gn = s->n;
    memset(s->cell_start, 0, (s->C + 1) * sizeof(int32_t));
    for (int n = 0; n < s->np; ++n) {
        const real *p = s->p + 3 * (size_t)n;
        long c = cidx(gn, cell_of(gn, p[0]), cell_of(gn, p[1]), cell_of(gn, p[2]));
        s->cell_start[c + 1]++;
    }
    for (long c = 0; c < s->C; ++c) s->cell_start[c + 1] += s->cell_start[c];
    memcpy(s->cell_fill, s->cell_start, s->C * sizeof(int32_t));
    for (int n = 0; n < s->np; ++n) {
        const real *p = s->p + 3 * (size_t)n;
        long c = cidx(gn, cell_of(gn, p[0]), cell_of(gn, p[1]), cell_of(gn, p[2]));
        s->cell_pid[s->cell_fill[c]++] = n;
    }
}

/* Window loops.  getNeigboringParticles_cell (sorter.cpp:54-64) visits
 * [i-w, i+w]; getNeigboringParticles_wall (:42-52) visits [i-w, i+w-1]; both
 * in si -> sj -> sk order with out-of-range cells skipped, then bucket order. */
#define WINDOW_BEGIN(S, ilo, ihi, jlo, jhi, klo, khi, PID)                          \
    for (int si_ = (ilo); si_ <= (ihi); si_++)                                      \
        for (int sj_ = (jlo); sj_ <= (jhi); sj_++)                                  \
            for (int sk_ = (klo); sk_ <= (khi); sk_++) {                            \
                if (si_ < 0 || si_ > (S)->n - 1 || sj_ < 0 || sj_ > (S)->n - 1 ||   \
                    sk_ < 0 || sk_ > (S)->n - 1)                                    \
                    continue;                                                       \
                long c_ = cidx((S)->n, si_, sj_, sk_);                              \
                for (int a_ = (S)->cell_start[c_]; a_ < (S)->cell_start[c_ + 1]; ++a_) { \
                    int PID = (S)->cell_pid[a_];
#define WINDOW_END \
    }              \
    }

/* -------------------------------------------------------------- density */
/* flip3D.cpp:166-190 computeDensity */
static void compute_density_into(orc_sim *s, real max_dens)
{
    int gn = s->n;
    real h = (real)(4.0 * s->prm.density / gn);
#pragma omp parallel for schedule(static)
    for (int n = 0; n < s->np; ++n) {
        if (s->type[n] == WALL) {
            s->dens[n] = (real)1.0;
            continue;
        }
        const real *p = s->p + 3 * (size_t)n;
        int ci = cell_of(gn, p[0]), cj = cell_of(gn, p[1]), ck = cell_of(gn, p[2]);
        real wsum = 0.0;
        WINDOW_BEGIN(s, ci - 1, ci + 1, cj - 1, cj + 1, ck - 1, ck + 1, q)
            if (s->type[q] == WALL) continue;
            real d2 = orc_length2(s->p + 3 * (size_t)q, p);
            real w = s->m[q] * orc_smooth_kernel(d2, h);
            wsum += w;
        WINDOW_END
        s->dens[n] = wsum / max_dens;
    }
}

void orc_compute_density(orc_sim *s) { compute_density_into(s, s->max_dens); }

/* flip3D.cpp:870-875 add_ExtForce: u_y += -DT*GRAVITY for every particle */
void orc_add_ext_force(orc_sim *s)
{
    double g = -s->cur_dt * s->prm.gravity;
    for (int n = 0; n < s->np; ++n) s->u[3 * (size_t)n + 1] = (real)((double)s->u[3 * (size_t)n + 1] + g);
}

/* ------------------------------------------------------------------ P2G */
/* mapper.cpp:17-88 mapP2G: per face, normalised sharp-kernel average over the
 * _wall window (2x4x4 cells for X faces etc.), FLUID particles only. */
static real p2g_face(const orc_sim *s, int c, int i, int j, int k)
{
    int gn = s->n;
    real fp[3];
    int wi = 2, wj = 2, wk = 2;
    if (c == 0) { fp[0] = (real)i; fp[1] = (real)(j + 0.5); fp[2] = (real)(k + 0.5); wi = 1; }
    else if (c == 1) { fp[0] = (real)(i + 0.5); fp[1] = (real)j; fp[2] = (real)(k + 0.5); wj = 1; }
    else { fp[0] = (real)(i + 0.5); fp[1] = (real)(j + 0.5); fp[2] = (real)k; wk = 1; }
    real sumw = 0.0, sumx = 0.0;
    real re = (real)1.4; /* RE, mapper.cpp:14 */
    WINDOW_BEGIN(s, i - wi, i + wi - 1, j - wj, j + wj - 1, k - wk, k + wk - 1, q)
        if (s->type[q] != FLUID) continue;
        const real *pp = s->p + 3 * (size_t)q;
        real pos[3];
        for (int d = 0; d < 3; ++d) pos[d] = (real)fmax(0.0, fmin((double)gn, (double)((real)gn * pp[d])));
        real w = s->m[q] * orc_sharp_kernel(orc_length2(pos, fp), re);
        sumx += w * s->u[3 * (size_t)q + c];
        sumw += w;
    WINDOW_END
    return sumw != 0 ? sumx / sumw : (real)0.0;
}

void orc_map_p2g(orc_sim *s)
{
    int gn = s->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < gn + 1; ++i)
        for (int j = 0; j < gn + 1; ++j)
            for (int k = 0; k < gn + 1; ++k) {
                if (j < gn && k < gn) s->ug[0][fidx(gn, 0, i, j, k)] = p2g_face(s, 0, i, j, k);
                if (i < gn && k < gn) s->ug[1][fidx(gn, 1, i, j, k)] = p2g_face(s, 1, i, j, k);
                if (i < gn && j < gn) s->ug[2][fidx(gn, 2, i, j, k)] = p2g_face(s, 2, i, j, k);
            }
}

/* ----------------------------------------------------------- flags / L */
/* sorter.cpp:70-81 levelset */
static real levelset(const orc_sim *s, long c)
{
    real accm = 0.0;
    for (int a = s->cell_start[c]; a < s->cell_start[c + 1]; ++a) {
        int q = s->cell_pid[a];
        if (s->type[q] == FLUID) accm += s->dens[q];
        else return (real)1.0;
    }
    real dd = (real)s->prm.density;
    real n0 = (real)(1.0 / (dd * dd * dd));
    return (real)(0.2 * n0 - accm);
}

/* sorter.cpp:83-93 markWater */
void orc_mark_water(orc_sim *s)
{
    for (long c = 0; c < s->C; ++c) {
        int8_t a = AIR;
        for (int t = s->cell_start[c]; t < s->cell_start[c + 1]; ++t)
            if (s->type[s->cell_pid[t]] == WALL) a = WALL;
        if (a != WALL) a = levelset(s, c) < 0.0 ? FLUID : AIR;
        s->A[c] = a;
    }
}

/* -------------------------------------------------------- grid passes */
/* flip3D.cpp:971-983 save_grid */
void orc_save_grid(orc_sim *s)
{
    long F = (long)(s->n + 1) * s->n * s->n;
    for (int c = 0; c < 3; ++c) memcpy(s->us[c], s->ug[c], F * sizeof(real));
}

/* flip3D.cpp:985-997 subtract_grid */
void orc_subtract_grid(orc_sim *s)
{
    long F = (long)(s->n + 1) * s->n * s->n;
    for (int c = 0; c < 3; ++c)
        for (long f = 0; f < F; ++f) s->us[c][f] = s->ug[c][f] - s->us[c][f];
}

/* flip3D.cpp:779-805 enforce_boundary */
static inline int wall_chk(int8_t a) { return a == WALL ? 1 : -1; }
void orc_enforce_boundary(orc_sim *s)
{
    int N = s->n;
    const int8_t *A = s->A;
    for (int i = 0; i < N + 1; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                real *u = &s->ug[0][fidx(N, 0, i, j, k)];
                if (i == 0 || i == N) *u = 0.0;
                if (i < N && i > 0 && wall_chk(A[cidx(N, i, j, k)]) * wall_chk(A[cidx(N, i - 1, j, k)]) < 0) *u = 0.0;
            }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N + 1; ++j)
            for (int k = 0; k < N; ++k) {
                real *u = &s->ug[1][fidx(N, 1, i, j, k)];
                if (j == 0 || j == N) *u = 0.0;
                if (j < N && j > 0 && wall_chk(A[cidx(N, i, j, k)]) * wall_chk(A[cidx(N, i, j - 1, k)]) < 0) *u = 0.0;
            }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N + 1; ++k) {
                real *u = &s->ug[2][fidx(N, 2, i, j, k)];
                if (k == 0 || k == N) *u = 0.0;
                if (k < N && k > 0 && wall_chk(A[cidx(N, i, j, k)]) * wall_chk(A[cidx(N, i, j, k - 1)]) < 0) *u = 0.0;
            }
}

/* --------------------------------------------------------------- solver */
/* solver.cpp:19-26 x_ref (clamped fetch, ghost-fluid ratio for AIR) */
static inline real x_ref(const orc_params *prm, const int8_t *A, const real *L, const real *x,
                         long f, int i, int j, int k)
{
    int n = prm->n;
    i = i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    j = j < 0 ? 0 : (j > n - 1 ? n - 1 : j);
    k = k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
    long c = cidx(n, i, j, k);
    if (A[c] == FLUID) return x[c];
    if (A[c] == WALL) return x[f];
    return prm->subcell ? (real)((double)L[c] / fmin(1.0e-6, (double)L[f]) * (double)x[f]) : (real)0.0;
}

/* solver.cpp:29-41 compute_Ax; FOR_EVERY_COMP visits i fastest, k slowest */
static void compute_ax(const orc_params *prm, const int8_t *A, const real *L, const real *x, real *ans)
{
    int n = prm->n;
    real h2 = (real)(1.0 / (n * n));
    long C = (long)n * n * n;
#pragma omp parallel for schedule(static)
    for (long g = 0; g < C; ++g) {
        int i = (int)((g % ((long)n * n)) % n), j = (int)((g % ((long)n * n)) / n), k = (int)(g / ((long)n * n));
        long c = cidx(n, i, j, k);
        if (A[c] == FLUID) {
            double acc = 6.0 * (double)x[c];
            acc -= x_ref(prm, A, L, x, c, i + 1, j, k);
            acc -= x_ref(prm, A, L, x, c, i - 1, j, k);
            acc -= x_ref(prm, A, L, x, c, i, j + 1, k);
            acc -= x_ref(prm, A, L, x, c, i, j - 1, k);
            acc -= x_ref(prm, A, L, x, c, i, j, k + 1);
            acc -= x_ref(prm, A, L, x, c, i, j, k - 1);
            ans[c] = (real)(acc / (double)h2);
        } else {
            ans[c] = 0.0;
        }
    }
}

void orc_compute_ax(const orc_params *prm, const int8_t *A, const real *L, const real *x, real *ans)
{
    compute_ax(prm, A, L, x, ans);
}

/* solver.cpp:44-54 product: serial (the omp for is orphaned), FLOAT product
 * accumulated in double, FOR_EVERY_COMP order */
static double product(int n, const int8_t *A, const real *x, const real *y)
{
    double ans = 0.0;
    long nn = (long)n * n;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                long c = ((long)i * n + j) * n + k;
                (void)nn;
                if (A[c] == FLUID) ans += x[c] * y[c];
            }
    return ans;
}

/* solver.cpp:78-85 op: ans = FLUID ? x + a*y : 0 (via tmp + copy) */
static void op(int n, const int8_t *A, const real *x, const real *y, real *ans, real a, real *tmp)
{
    long C = (long)n * n * n;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < C; ++c) tmp[c] = A[c] == FLUID ? x[c] + a * y[c] : (real)0.0;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < C; ++c) ans[c] = tmp[c];
}

/* solver.cpp:97-101 A_ref: -1 between two in-range FLUID cells */
static inline real A_ref(int n, const int8_t *A, int i, int j, int k, int qi, int qj, int qk)
{
    if (i < 0 || i > n - 1 || j < 0 || j > n - 1 || k < 0 || k > n - 1 || A[cidx(n, i, j, k)] != FLUID) return 0.0;
    if (qi < 0 || qi > n - 1 || qj < 0 || qj > n - 1 || qk < 0 || qk > n - 1 || A[cidx(n, qi, qj, qk)] != FLUID) return 0.0;
    return -1.0;
}

/* solver.cpp:103-117 A_diag (ghost-fluid diagonal, FLOAT accumulation) */
static real A_diag(const orc_params *prm, const int8_t *A, const real *L, int i, int j, int k)
{
    int n = prm->n;
    real diag = 6.0;
    long c = cidx(n, i, j, k);
    if (A[c] != FLUID) return diag;
    int q[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
    for (int m = 0; m < 6; ++m) {
        int qi = q[m][0], qj = q[m][1], qk = q[m][2];
        if (qi < 0 || qi > n - 1 || qj < 0 || qj > n - 1 || qk < 0 || qk > n - 1 || A[cidx(n, qi, qj, qk)] == WALL)
            diag = (real)((double)diag - 1.0);
        else if (A[cidx(n, qi, qj, qk)] == AIR && prm->subcell)
            diag = (real)((double)diag - (double)L[cidx(n, qi, qj, qk)] / fmin(1.0e-6, (double)L[c]));
    }
    return diag;
}

/* solver.cpp:119-123 P_ref.  precond == 0 keeps the reference's value test
 * (P != FLUID compares the stored value with 1); precond == 1 keys the
 * coupling on the cell flag, i.e. true IC(0). */
static inline real P_ref_d(const orc_params *prm, const int8_t *A, const double *P, int i, int j, int k)
{
    int n = prm->n;
    if (i < 0 || i > n - 1 || j < 0 || j > n - 1 || k < 0 || k > n - 1) return 0.0;
    long c = cidx(n, i, j, k);
    if (prm->precond == 1) return A[c] == FLUID ? (real)P[c] : (real)0.0;
    if (P[c] != FLUID) return 0.0;
    return (real)P[c];
}
static inline real P_ref_f(const orc_params *prm, const int8_t *A, const real *z, int i, int j, int k)
{
    int n = prm->n;
    if (i < 0 || i > n - 1 || j < 0 || j > n - 1 || k < 0 || k > n - 1) return 0.0;
    long c = cidx(n, i, j, k);
    if (prm->precond == 1) return A[c] == FLUID ? z[c] : (real)0.0;
    if (z[c] != FLUID) return 0.0;
    return z[c];
}

static inline real square_r(real a) { return a * a; }

/* solver.cpp:125-139 buildPreconditioner (serial, FOR_EVERY_COMP order) */
static void build_precond(const orc_params *prm, const int8_t *A, const real *L, double *P)
{
    int n = prm->n;
    long C = (long)n * n * n;
    memset(P, 0, C * sizeof(double));
    double a = 0.25;
    for (long g = 0; g < C; ++g) {
        int i = (int)((g % ((long)n * n)) % n), j = (int)((g % ((long)n * n)) / n), k = (int)(g / ((long)n * n));
        long c = cidx(n, i, j, k);
        if (A[c] != FLUID) continue;
        double left = A_ref(n, A, i - 1, j, k, i, j, k) * P_ref_d(prm, A, P, i - 1, j, k);
        double bottom = A_ref(n, A, i, j - 1, k, i, j, k) * P_ref_d(prm, A, P, i, j - 1, k);
        double back = A_ref(n, A, i, j, k - 1, i, j, k) * P_ref_d(prm, A, P, i, j, k - 1);
        double diag = A_diag(prm, A, L, i, j, k);
        double e = diag - square_r((real)left) - square_r((real)bottom) - square_r((real)back);
        if (e < a * diag) e = diag;
        P[c] = 1.0 / R_SQRT((real)e);
    }
}

void orc_build_precond(const orc_params *prm, const int8_t *A, const real *L, double *P)
{
    build_precond(prm, A, L, P);
}

/* solver.cpp:141-180 applyPreconditioner: forward sweep i->j->k, backward in
 * reverse; q is double scratch cleared each call. */
static void apply_precond(const orc_params *prm, const int8_t *A, const double *P, const real *r,
                          real *z, double *q)
{
    int n = prm->n;
    long C = (long)n * n * n;
    memset(q, 0, C * sizeof(double));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                long c = cidx(n, i, j, k);
                if (A[c] != FLUID) continue;
                /* P_ref on the double q array uses the same value test */
                double left = A_ref(n, A, i - 1, j, k, i, j, k) * P_ref_d(prm, A, P, i - 1, j, k) * P_ref_d(prm, A, q, i - 1, j, k);
                double bottom = A_ref(n, A, i, j - 1, k, i, j, k) * P_ref_d(prm, A, P, i, j - 1, k) * P_ref_d(prm, A, q, i, j - 1, k);
                double back = A_ref(n, A, i, j, k - 1, i, j, k) * P_ref_d(prm, A, P, i, j, k - 1) * P_ref_d(prm, A, q, i, j, k - 1);
                double t = (double)r[c] - left - bottom - back;
                q[c] = t * P[c];
            }
    for (int i = n - 1; i >= 0; --i)
        for (int j = n - 1; j >= 0; --j)
            for (int k = n - 1; k >= 0; --k) {
                long c = cidx(n, i, j, k);
                if (A[c] != FLUID) continue;
                double right = A_ref(n, A, i, j, k, i + 1, j, k) * P_ref_d(prm, A, P, i, j, k) * P_ref_f(prm, A, z, i + 1, j, k);
                double top = A_ref(n, A, i, j, k, i, j + 1, k) * P_ref_d(prm, A, P, i, j, k) * P_ref_f(prm, A, z, i, j + 1, k);
                double front = A_ref(n, A, i, j, k, i, j, k + 1) * P_ref_d(prm, A, P, i, j, k) * P_ref_f(prm, A, z, i, j, k + 1);
                double t = q[c] - right - top - front;
                z[c] = (real)(t * P[c]);
            }
}

void orc_apply_precond(const orc_params *prm, const int8_t *A, const double *P, const real *r, real *z)
{
    long C = (long)prm->n * prm->n * prm->n;
    double *q = (double *)calloc(C, sizeof(double));
    apply_precond(prm, A, P, r, z, q);
    free(q);
}

typedef struct {
    double *P, *q;
    real *r, *z, *s, *tmp;
} solver_scratch;

/* solver.cpp:183-218 conjGrad.  tol_policy 0 is the reference's absolute
 * stop (float error2 <= 1e-2 n^3); tol_policy 1 stops at ||r|| <= tol ||b||
 * and skips the loop when the warm start already meets it. */
static void conj_grad(const orc_params *prm, const int8_t *A, const real *L, real *x, const real *b,
                      solver_scratch *w, orc_stats *st)
{
    int n = prm->n;
    long C = (long)n * n * n;
    real *r = w->r, *z = w->z, *s = w->s;
    compute_ax(prm, A, L, x, z);
    op(n, A, b, z, r, (real)-1.0, w->tmp);
    double error2_0 = product(n, A, r, r);
    double bnorm2 = product(n, A, b, b);
    apply_precond(prm, A, w->P, r, z, w->q);
#pragma omp parallel for schedule(static)
    for (long c = 0; c < C; ++c) s[c] = z[c];
    double eps = prm->tol_policy == 0 ? 1.0e-2 * (n * n * n) : prm->tol * prm->tol * bnorm2;
    double a = product(n, A, z, r);
    long cap = prm->max_iters > 0 ? prm->max_iters : (long)n * n * n;
    int it = 0;
    double err = error2_0;
    if (st) { st->error2_0 = error2_0; st->bnorm2 = bnorm2; }
    if (prm->tol_policy == 1 && error2_0 <= eps) {
        /* no iteration: apply the masking of the loop's first op(x) anyway */
        cap = 0;
        for (long c = 0; c < C; ++c)
            if (A[c] != FLUID) x[c] = 0.0;
    }
    for (long k = 0; k < cap; ++k) {
        compute_ax(prm, A, L, s, z);
        double alpha = a / product(n, A, z, s);
        op(n, A, x, s, x, (real)alpha, w->tmp);
        op(n, A, r, z, r, (real)-alpha, w->tmp);
        it++;
        if (prm->tol_policy == 0) {
            real error2 = (real)product(n, A, r, r);
            err = error2;
            if (error2 <= eps) break;
        } else {
            err = product(n, A, r, r);
            if (err <= eps) break;
        }
        apply_precond(prm, A, w->P, r, z, w->q);
        double a2 = product(n, A, z, r);
        double beta = a2 / a;
        op(n, A, z, s, s, (real)beta, w->tmp);
        a = a2;
    }
    if (st) { st->iterations = it; st->error2 = err; }
}

static int count_unit_p(int n, const int8_t *A, const double *P)
{
    long C = (long)n * n * n;
    int cnt = 0;
    for (long c = 0; c < C; ++c)
        if (A[c] == FLUID && P[c] == 1.0) cnt++;
    return cnt;
}

/* solver.cpp:224-241 solve: b = -b; build P; PCG */
static void solve_with(const orc_params *prm, const int8_t *A, const real *L, real *x, real *b,
                       solver_scratch *w, orc_stats *st)
{
    int n = prm->n;
    long C = (long)n * n * n;
    for (long c = 0; c < C; ++c) b[c] = -b[c];
    build_precond(prm, A, L, w->P);
    if (st) st->n_precond_unit = count_unit_p(n, A, w->P);
    conj_grad(prm, A, L, x, b, w, st);
    if (st) {
        int nf = 0;
        for (long c = 0; c < C; ++c) nf += A[c] == FLUID;
        st->n_fluid_cells = nf;
    }
}

int orc_solve(const orc_params *prm, const int8_t *A, const real *L, real *x, real *b, orc_stats *st)
{
    long C = (long)prm->n * prm->n * prm->n;
    solver_scratch w;
    w.P = (double *)calloc(C, sizeof(double));
    w.q = (double *)calloc(C, sizeof(double));
    w.r = (real *)calloc(C, sizeof(real));
    w.z = (real *)calloc(C, sizeof(real));
    w.s = (real *)calloc(C, sizeof(real));
    w.tmp = (real *)calloc(C, sizeof(real));
    solve_with(prm, A, L, x, b, &w, st);
    free(w.P); free(w.q); free(w.r); free(w.z); free(w.s); free(w.tmp);
    return 0;
}

/* -------------------------------------------------------------- project */
/* flip3D.cpp:807-868 project: divergence on FLUID cells, L from levelset,
 * solve, ghost-fluid pressure gradient subtraction */
void orc_project(orc_sim *s, orc_stats *st)
{
    int N = s->n;
    real h = (real)(1.0 / N);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                long c = cidx(N, i, j, k);
                if (s->A[c] == FLUID) {
                    real d = s->ug[0][fidx(N, 0, i + 1, j, k)] - s->ug[0][fidx(N, 0, i, j, k)];
                    d = d + s->ug[1][fidx(N, 1, i, j + 1, k)];
                    d = d - s->ug[1][fidx(N, 1, i, j, k)];
                    d = d + s->ug[2][fidx(N, 2, i, j, k + 1)];
                    d = d - s->ug[2][fidx(N, 2, i, j, k)];
                    s->div[c] = d / h;
                }
            }
    for (long c = 0; c < s->C; ++c) s->L[c] = levelset(s, c);

    solver_scratch w = {s->P, s->q, s->r, s->z, s->s, s->tmp};
    orc_params prm = s->prm;
    prm.n = N;
    solve_with(&prm, s->A, s->L, s->press, s->div, &w, st);

    const real *L = s->L, *press = s->press;
    int sub = s->prm.subcell;
    for (int i = 0; i < N + 1; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                if (!(i > 0 && i < N)) continue;
                long f = cidx(N, i, j, k), b = cidx(N, i - 1, j, k);
                real pf = press[f], pb = press[b];
                if (sub && L[f] * L[b] < 0.0) {
                    pf = L[f] < 0.0 ? press[f] : (real)((double)L[f] / fmin(1.0e-3, (double)L[b]) * (double)press[b]);
                    pb = L[b] < 0.0 ? press[b] : (real)((double)L[b] / fmin(1.0e-6, (double)L[f]) * (double)press[f]);
                }
                real *u = &s->ug[0][fidx(N, 0, i, j, k)];
                *u = *u - (pf - pb) / h;
            }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N + 1; ++j)
            for (int k = 0; k < N; ++k) {
                if (!(j > 0 && j < N)) continue;
                long f = cidx(N, i, j, k), b = cidx(N, i, j - 1, k);
                real pf = press[f], pb = press[b];
                if (sub && L[f] * L[b] < 0.0) {
                    pf = L[f] < 0.0 ? press[f] : (real)((double)L[f] / fmin(1.0e-3, (double)L[b]) * (double)press[b]);
                    pb = L[b] < 0.0 ? press[b] : (real)((double)L[b] / fmin(1.0e-6, (double)L[f]) * (double)press[f]);
                }
                real *u = &s->ug[1][fidx(N, 1, i, j, k)];
                *u = *u - (pf - pb) / h;
            }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N + 1; ++k) {
                if (!(k > 0 && k < N)) continue;
                long f = cidx(N, i, j, k), b = cidx(N, i, j, k - 1);
                real pf = press[f], pb = press[b];
                if (sub && L[f] * L[b] < 0.0) {
                    pf = L[f] < 0.0 ? press[f] : (real)((double)L[f] / fmin(1.0e-3, (double)L[b]) * (double)press[b]);
                    pb = L[b] < 0.0 ? press[b] : (real)((double)L[b] / fmin(1.0e-6, (double)L[f]) * (double)press[f]);
                }
                real *u = &s->ug[2][fidx(N, 2, i, j, k)];
                *u = *u - (pf - pb) / h;
            }
}

/* flip3D.cpp:1011-1054 extrapolate_velocity */
void orc_extrapolate_velocity(orc_sim *s)
{
    int N = s->n;
    long F = (long)(N + 1) * N * N;
    int8_t *mark[3], *wmark[3];
    for (int c = 0; c < 3; ++c) {
        mark[c] = (int8_t *)calloc(F, 1);
        wmark[c] = (int8_t *)calloc(F, 1);
    }
    const int8_t *A = s->A;
    for (int i = 0; i < N + 1; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                long f = fidx(N, 0, i, j, k);
                mark[0][f] = (i > 0 && A[cidx(N, i - 1, j, k)] == FLUID) || (i < N && A[cidx(N, i, j, k)] == FLUID);
                wmark[0][f] = (i <= 0 || A[cidx(N, i - 1, j, k)] == WALL) && (i >= N || A[cidx(N, i, j, k)] == WALL);
            }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N + 1; ++j)
            for (int k = 0; k < N; ++k) {
                long f = fidx(N, 1, i, j, k);
                mark[1][f] = (j > 0 && A[cidx(N, i, j - 1, k)] == FLUID) || (j < N && A[cidx(N, i, j, k)] == FLUID);
                wmark[1][f] = (j <= 0 || A[cidx(N, i, j - 1, k)] == WALL) && (j >= N || A[cidx(N, i, j, k)] == WALL);
            }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N + 1; ++k) {
                long f = fidx(N, 2, i, j, k);
                mark[2][f] = (k > 0 && A[cidx(N, i, j, k - 1)] == FLUID) || (k < N && A[cidx(N, i, j, k)] == FLUID);
                wmark[2][f] = (k <= 0 || A[cidx(N, i, j, k - 1)] == WALL) && (k >= N || A[cidx(N, i, j, k)] == WALL);
            }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N + 1; ++i)
        for (int j = 0; j < N + 1; ++j)
            for (int k = 0; k < N + 1; ++k)
                for (int n = 0; n < 3; ++n) {
                    if (n != 0 && i > N - 1) continue;
                    if (n != 1 && j > N - 1) continue;
                    if (n != 2 && k > N - 1) continue;
                    long f = fidx(N, n, i, j, k);
                    if (!mark[n][f] && wmark[n][f]) {
                        int wsum = 0;
                        real sum = 0.0;
                        int q[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
                        for (int t = 0; t < 6; ++t) {
                            if (q[t][0] >= 0 && q[t][0] < N + (n == 0) && q[t][1] >= 0 && q[t][1] < N + (n == 1) &&
                                q[t][2] >= 0 && q[t][2] < N + (n == 2)) {
                                long g = fidx(N, n, q[t][0], q[t][1], q[t][2]);
                                if (mark[n][g]) {
                                    wsum++;
                                    sum += s->ug[n][g];
                                }
                            }
                        }
                        if (wsum) s->ug[n][f] = sum / wsum;
                    }
                }
    for (int c = 0; c < 3; ++c) { free(mark[c]); free(wmark[c]); }
}

/* ------------------------------------------------------------------ G2P */
/* mapper.cpp:96-106 linear (note k clamps with h-2, as the reference does) */
static real linear(const real *q, int c, int gn, real x, real y, real z, int w, int h, int d)
{
    x = (real)fmax(0.0, fmin((double)w, (double)x));
    y = (real)fmax(0.0, fmin((double)h, (double)y));
    z = (real)fmax(0.0, fmin((double)d, (double)z));
    int i = (int)(x > (real)(w - 2) ? (real)(w - 2) : x);
    int j = (int)(y > (real)(h - 2) ? (real)(h - 2) : y);
    int k = (int)(z > (real)(h - 2) ? (real)(h - 2) : z);
#define Q(a, b, e) (((e) < d) ? q[fidx(gn, c, (a), (b), (e))] : (real)0.0)
    real ix = (real)(i + 1) - x, xi = x - (real)i;
    real jy = (real)(j + 1) - y, yj = y - (real)j;
    real kz = (real)(k + 1) - z, zk = z - (real)k;
    real lo = (ix * Q(i, j, k) + xi * Q(i + 1, j, k)) * jy + (ix * Q(i, j + 1, k) + xi * Q(i + 1, j + 1, k)) * yj;
    real hi = (ix * Q(i, j, k + 1) + xi * Q(i + 1, j, k + 1)) * jy + (ix * Q(i, j + 1, k + 1) + xi * Q(i + 1, j + 1, k + 1)) * yj;
#undef Q
    return kz * lo + zk * hi;
}

/* mapper.cpp:108-112 fetchVelocity */
static void fetch_velocity(real *const g[3], int gn, const real *p, real *u)
{
    real gx = (real)gn * p[0], gy = (real)gn * p[1], gz = (real)gn * p[2];
    u[0] = linear(g[0], 0, gn, gx, (real)((double)gy - 0.5), (real)((double)gz - 0.5), gn + 1, gn, gn);
    u[1] = linear(g[1], 1, gn, (real)((double)gx - 0.5), gy, (real)((double)gz - 0.5), gn, gn + 1, gn);
    u[2] = linear(g[2], 2, gn, (real)((double)gx - 0.5), (real)((double)gy - 0.5), gz, gn, gn, gn + 1);
}

void orc_fetch_velocity(const orc_sim *s, int grid_set, const real *p, real *u)
{
    fetch_velocity(grid_set ? s->us : s->ug, s->n, p, u);
}

/* flip3D.cpp:1071-1097 FLIP: tmp = u; u = G2P(du); tmp = u + tmp;
 * u = G2P(u_grid); u = (1-ALPHA)*u + ALPHA*tmp (double) */
void orc_flip_g2p(orc_sim *s)
{
    double al = s->prm.alpha;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < s->np; ++n) {
        real *u = s->u + 3 * (size_t)n;
        const real *p = s->p + 3 * (size_t)n;
        real t[3] = {u[0], u[1], u[2]};
        real du[3], pic[3];
        fetch_velocity(s->us, s->n, p, du);
        for (int k = 0; k < 3; ++k) t[k] = du[k] + t[k];
        fetch_velocity(s->ug, s->n, p, pic);
        for (int k = 0; k < 3; ++k) u[k] = (real)((1.0 - al) * (double)pic[k] + al * (double)t[k]);
        s->tmp0[3 * (size_t)n + 0] = t[0];
        s->tmp0[3 * (size_t)n + 1] = t[1];
        s->tmp0[3 * (size_t)n + 2] = t[2];
    }
}

/* flip3D.cpp:1057-1103 solve_picflip (sort already current) */
void orc_solve_picflip(orc_sim *s, orc_stats *st)
{
    orc_sort(s);
    orc_map_p2g(s);
    orc_mark_water(s);
    orc_save_grid(s);
    orc_enforce_boundary(s);
    orc_project(s, st);
    orc_enforce_boundary(s);
    orc_extrapolate_velocity(s);
    orc_subtract_grid(s);
    orc_flip_g2p(s);
}

/* ------------------------------------------------------------ corrector */
/* corrector.cpp:15-47 resample */
static void resample(const orc_sim *s, const real *p, real *u, real re)
{
    int cs = s->n;
    real wsum = 0.0;
    real save[3] = {u[0], u[1], u[2]};
    u[0] = u[1] = u[2] = 0.0;
    int ci = cell_of(cs, p[0]), cj = cell_of(cs, p[1]), ck = cell_of(cs, p[2]);
    WINDOW_BEGIN(s, ci - 1, ci + 1, cj - 1, cj + 1, ck - 1, ck + 1, q)
        if (s->type[q] != FLUID) continue;
        real dist2 = orc_length2(p, s->p + 3 * (size_t)q);
        real w = s->m[q] * orc_sharp_kernel(dist2, re);
        u[0] += w * s->u[3 * (size_t)q + 0];
        u[1] += w * s->u[3 * (size_t)q + 1];
        u[2] += w * s->u[3 * (size_t)q + 2];
        wsum += w;
    WINDOW_END
    if (wsum != 0) {
        u[0] /= wsum;
        u[1] /= wsum;
        u[2] /= wsum;
    } else {
        u[0] = save[0];
        u[1] = save[1];
        u[2] = save[2];
    }
}

/* corrector.cpp:55-88 spring for particle n.  With draw == 0 the function
 * only reports whether a random-jitter event occurs (returns its count);
 * with draw != 0 it consumes rand() in window order. */
static int spring_particle(orc_sim *s, int n, real dt, real re, int draw)
{
    int cs = s->n;
    const real *p = s->p + 3 * (size_t)n;
    real spring[3] = {0.0, 0.0, 0.0};
    real cx = (real)cs * p[0], cy = (real)cs * p[1], cz = (real)cs * p[2];
    /* max(0,min(cs,cs*p)) with the reference's macros (utility.h:16-17) */
    real x = (real)cs > cx ? cx : (real)cs;
    x = 0 > x ? (real)0 : x;
    real y = (real)cs > cy ? cy : (real)cs;
    y = 0 > y ? (real)0 : y;
    real z = (real)cs > cz ? cz : (real)cs;
    z = 0 > z ? (real)0 : z;
    int ci = (int)x, cj = (int)y, ck = (int)z;
    int events = 0;
    WINDOW_BEGIN(s, ci - 1, ci + 1, cj - 1, cj + 1, ck - 1, ck + 1, q)
        if (q == n) continue;
        const real *np = s->p + 3 * (size_t)q;
        real dist = orc_length(p, np);
        real w = (real)(50.0 * (double)s->m[q] * (double)orc_smooth_kernel(dist * dist, re));
        if ((double)dist > 0.1 * (double)re) {
            spring[0] += w * (p[0] - np[0]) / dist * re;
            spring[1] += w * (p[1] - np[1]) / dist * re;
            spring[2] += w * (p[2] - np[2]) / dist * re;
        } else if (s->type[q] == FLUID) {
            events++;
            if (draw) {
                for (int c = 0; c < 3; ++c) {
                    int rv = orc_rand_next(s) % 101;
                    spring[c] = (real)((double)spring[c] + 0.01 * (double)re / (double)dt * rv / 100.0);
                }
            }
        } else {
            for (int c = 0; c < 3; ++c)
                spring[c] = (real)((double)spring[c] + 0.05 * (double)re / (double)dt * (double)s->nrm[3 * (size_t)q + c]);
        }
    WINDOW_END
    if (events && !draw) return events;
    real *t = s->tmp0 + 3 * (size_t)n;
    t[0] = p[0] + dt * spring[0];
    t[1] = p[1] + dt * spring[1];
    t[2] = p[2] + dt * spring[2];
    return events;
}

/* corrector.cpp:49-115 correct.  The spring pass runs in parallel; particles
 * with random-jitter events are redone serially in particle order, which is
 * the order a single-threaded reference draws rand() in. */
void orc_correct(orc_sim *s, orc_stats *st)
{
    real dt = (real)s->cur_dt;
    real re = (real)(s->prm.density / s->n);
    orc_sort(s);
    int *ev = (int *)calloc((size_t)s->np + 1, sizeof(int));
#pragma omp parallel for schedule(static)
    for (int n = 0; n < s->np; ++n)
        if (s->type[n] == FLUID) ev[n] = spring_particle(s, n, dt, re, 0);
    int nj = 0;
    for (int n = 0; n < s->np; ++n)
        if (s->type[n] == FLUID && ev[n]) nj += spring_particle(s, n, dt, re, 1);
    free(ev);
    if (st) st->n_jitter = nj;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < s->np; ++n) {
        if (s->type[n] != FLUID) continue;
        real *t1 = s->tmp1 + 3 * (size_t)n;
        t1[0] = s->u[3 * (size_t)n + 0];
        t1[1] = s->u[3 * (size_t)n + 1];
        t1[2] = s->u[3 * (size_t)n + 2];
        resample(s, s->tmp0 + 3 * (size_t)n, t1, re);
    }
    for (int n = 0; n < s->np; ++n) {
        if (s->type[n] != FLUID) continue;
        for (int c = 0; c < 3; ++c) {
            s->p[3 * (size_t)n + c] = s->tmp0[3 * (size_t)n + c];
            s->u[3 * (size_t)n + c] = s->tmp1[3 * (size_t)n + c];
        }
    }
}

/* --------------------------------------------------------------- advect */
/* libstdc++ std::random_shuffle(first,last) without a generator: for i in
 * [1, n): swap(a[i], a[rand() % (i+1)]) */
static void random_shuffle_cells(orc_sim *s, int32_t *a, int n)
{
    for (int i = 1; i < n; ++i) {
        int j = orc_rand_next(s) % (i + 1);
        if (i != j) {
            int32_t t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
}

/* flip3D.cpp:405-451 reposition */
static int reposition(orc_sim *s, const int32_t *idx, int m)
{
    if (m == 0) return 0;
    int gn = s->n, N = s->n;
    int32_t *waters = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
    int nw = 0;
    while (nw < m) {
        int before = nw;
        for (int i = 0; i < gn && nw < m; ++i)
            for (int j = 0; j < gn && nw < m; ++j)
                for (int k = 0; k < gn && nw < m; ++k) {
                    const int8_t *A = s->A;
                    if (i > 0 && A[cidx(N, i - 1, j, k)] != FLUID) continue;
                    if (i < N - 1 && A[cidx(N, i + 1, j, k)] != FLUID) continue;
                    if (j > 0 && A[cidx(N, i, j - 1, k)] != FLUID) continue;
                    if (j < N - 1 && A[cidx(N, i, j + 1, k)] != FLUID) continue;
                    if (k > 0 && A[cidx(N, i, j, k - 1)] != FLUID) continue;
                    if (k < N - 1 && A[cidx(N, i, j, k + 1)] != FLUID) continue;
                    if (A[cidx(N, i, j, k)] != FLUID) continue;
                    waters[nw++] = (int32_t)cidx(N, i, j, k);
                }
        if (nw == before) {
            free(waters);
            return 0;
        }
    }
    random_shuffle_cells(s, waters, nw);
    real h = (real)(1.0 / gn);
    for (int t = 0; t < m; ++t) {
        real *p = s->p + 3 * (size_t)idx[t];
        int wi = waters[t] / (N * N), wj = (waters[t] / N) % N, wk = waters[t] % N;
        int r0 = orc_rand_next(s) % 101;
        p[0] = (real)((double)h * (wi + 0.25 + 0.5 * r0 / 100));
        int r1 = orc_rand_next(s) % 101;
        p[1] = (real)((double)h * (wj + 0.25 + 0.5 * r1 / 100));
        int r2 = orc_rand_next(s) % 101;
        p[2] = (real)((double)h * (wk + 0.25 + 0.5 * r2 / 100));
    }
    orc_sort(s);
    for (int t = 0; t < m; ++t) {
        real *p = s->p + 3 * (size_t)idx[t];
        real u[3] = {0.0, 0.0, 0.0};
        resample(s, p, u, h);
        for (int c = 0; c < 3; ++c) s->u[3 * (size_t)idx[t] + c] = u[c];
    }
    free(waters);
    return m;
}

/* flip3D.cpp:877-969 advect_particle */
void orc_advect(orc_sim *s, orc_stats *st)
{
    int N = s->n;
    double dt = s->cur_dt;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < s->np; ++n) {
        if (s->type[n] != FLUID) continue;
        real vel[3];
        real *p = s->p + 3 * (size_t)n;
        fetch_velocity(s->ug, N, p, vel);
        for (int k = 0; k < 3; ++k) p[k] = (real)((double)p[k] + dt * (double)vel[k]);
    }
    orc_sort(s);
    real r = (real)(1.0 / N);
    double re = 1.5 * s->prm.density / N;
    for (int n = 0; n < s->np; ++n) {
        if (s->type[n] != FLUID) continue;
        real *p = s->p + 3 * (size_t)n;
        real *u = s->u + 3 * (size_t)n;
        for (int k = 0; k < 3; ++k) p[k] = (real)fmax((double)r, fmin(1.0 - (double)r, (double)p[k]));
        int i = cell_of_b(N, p[0]), j = cell_of_b(N, p[1]), k = cell_of_b(N, p[2]);
        WINDOW_BEGIN(s, i - 1, i + 1, j - 1, j + 1, k - 1, k + 1, q)
            if (s->type[q] != WALL) continue;
            const real *np = s->p + 3 * (size_t)q;
            real dist = orc_length(p, np);
            if ((double)dist < re) {
                real nrm[3] = {s->nrm[3 * (size_t)q], s->nrm[3 * (size_t)q + 1], s->nrm[3 * (size_t)q + 2]};
                if (nrm[0] == 0.0 && nrm[1] == 0.0 && nrm[2] == 0.0 && dist != 0)
                    for (int c = 0; c < 3; ++c) nrm[c] = (p[c] - np[c]) / dist;
                p[0] = (real)((double)p[0] + (re - (double)dist) * (double)nrm[0]);
                p[1] = (real)((double)p[1] + (re - (double)dist) * (double)nrm[1]);
                p[2] = (real)((double)p[2] + (re - (double)dist) * (double)nrm[2]);
                real dot = u[0] * nrm[0] + u[1] * nrm[1] + u[2] * nrm[2];
                u[0] -= dot * nrm[0];
                u[1] -= dot * nrm[1];
                u[2] -= dot * nrm[2];
            }
        WINDOW_END
    }
    int nrep = 0;
    if (s->prm.reposition) {
        int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * ((size_t)s->np + 1));
        int m = 0;
        for (int n = 0; n < s->np; ++n) {
            s->remove[n] = 0;
            if (s->type[n] != FLUID) continue;
            const real *p = s->p + 3 * (size_t)n;
            if (s->A[cidx(N, cell_of_b(N, p[0]), cell_of_b(N, p[1]), cell_of_b(N, p[2]))] == WALL) s->remove[n] = 1;
            int i = (int)fmin((double)(N - 3), fmax(2.0, (double)(p[0] * (real)N)));
            int j = (int)fmin((double)(N - 3), fmax(2.0, (double)(p[1] * (real)N)));
            int k = (int)fmin((double)(N - 3), fmax(2.0, (double)(p[2] * (real)N)));
            int jm = j - 1 > 0 ? j - 1 : 0, jp = j + 1 < N - 1 ? j + 1 : N - 1;
            if ((double)s->dens[n] < 0.04 && (s->A[cidx(N, i, jm, k)] == WALL || s->A[cidx(N, i, jp, k)] == WALL))
                s->remove[n] = 1;
        }
        for (int n = 0; n < s->np; ++n)
            if (s->remove[n]) {
                s->remove[n] = 0;
                idx[m++] = n;
            }
        nrep = reposition(s, idx, m);
        free(idx);
    }
    if (st) st->n_reposition = nrep;
}

/* ------------------------------------------------------------------ step */
/* dt policy 1: dt = min(dt_cap, cfl*dx/max|u|) over FLUID particles */
static void choose_dt(orc_sim *s)
{
    if (s->prm.dt_policy == 0) {
        s->cur_dt = s->prm.dt;
        return;
    }
    double umax = 0.0;
    for (int n = 0; n < s->np; ++n) {
        if (s->type[n] != FLUID) continue;
        const real *u = s->u + 3 * (size_t)n;
        double m2 = (double)u[0] * u[0] + (double)u[1] * u[1] + (double)u[2] * u[2];
        if (m2 > umax) umax = m2;
    }
    umax = sqrt(umax);
    double dt = s->prm.dt;
    if (umax > 0.0) {
        double c = s->prm.cfl / (s->n * umax);
        if (c < dt) dt = c;
    }
    s->cur_dt = dt;
}

/* flip3D.cpp:1194-1264 simulateStep, timed region :1216 -> :1247 */
void orc_set_emitter(orc_sim *s, real x, real z, real r, int pour_time)
{
    s->pour_pos[0] = x;
    s->pour_pos[1] = z;
    s->pour_rad = r;
    s->pour_time = pour_time;
}

void orc_set_step(orc_sim *s, int step) { s->step = step; }
int orc_get_step(const orc_sim *s) { return s->step; }

/* flip3D.cpp:375-402 pourWater(limit): while step <= limit, a disc of FLUID
 * particles on the DENSITY/N lattice in x and z (float coordinates stepped by
 * the double spacing), just under the ceiling wall, moving down half a
 * lattice spacing per DT.  Returns the number added. */
int orc_pour(orc_sim *s)
{
    if (s->step > s->pour_time) return 0;
    int N = s->n;
    double w = s->prm.density / N;
    int cnt = 0;
    for (real x = (real)(w + w / 2.0); x < 1.0 - w / 2.0; x += w) {
        for (real z = (real)(w + w / 2.0); z < 1.0 - w / 2.0; z += w) {
            if (R_HYPOT(x - s->pour_pos[0], z - s->pour_pos[1]) < s->pour_rad) {
                grow_particles(s, s->np + 1);
                int q = s->np++;
                s->p[3 * q + 0] = x;
                s->p[3 * q + 1] = (real)(1.0 - 1.0 / N - 2.5 * s->prm.density / N);
                s->p[3 * q + 2] = z;
                s->u[3 * q + 0] = 0.0;
                s->u[3 * q + 1] = (real)(-0.5 * s->prm.density / N / s->prm.dt);
                s->u[3 * q + 2] = 0.0;
                s->nrm[3 * q] = s->nrm[3 * q + 1] = s->nrm[3 * q + 2] = 0.0;
                s->type[q] = FLUID;
                s->dens[q] = s->max_dens;
                s->m[q] = 1.0;
                cnt++;
            }
        }
    }
    return cnt;
}

void orc_step(orc_sim *s, orc_stats *st)
{
    orc_stats tmp;
    if (!st) st = &tmp;
    memset(st, 0, sizeof(*st));
    s->step++; /* flip3D.cpp:1196 */
    orc_pour(s); /* :1199 */
    choose_dt(s);
    st->dt = s->cur_dt;
    orc_sort(s);
    orc_compute_density(s);
    orc_add_ext_force(s);
    orc_solve_picflip(s, st);
    orc_advect(s, st);
    if (s->prm.correction) orc_correct(s, st);
}

/* ---------------------------------------------------------------- scenes */
long orc_scene_capacity(int n)
{
    long f = (long)(2 * n) * (2 * n) * (2 * n);
    return f + (long)n * n * n;
}

/* flip3D.cpp:689-709 max_dens calibration: 10^3 lattice at (i+0.5)*DENSITY/N */
real orc_calibrate_max_dens(int n, double density)
{
    orc_params prm;
    orc_default_params(&prm, n);
    prm.density = density;
    orc_sim *t = orc_create(&prm);
    reserve_particles(t, 1000);
    t->np = 1000;
    real h = (real)(density / n);
    int q = 0;
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j)
            for (int k = 0; k < 10; ++k, ++q) {
                t->p[3 * q + 0] = (real)((i + 0.5) * h);
                t->p[3 * q + 1] = (real)((j + 0.5) * h);
                t->p[3 * q + 2] = (real)((k + 0.5) * h);
                t->type[q] = FLUID;
                t->m[q] = 1.0;
            }
    orc_sort(t);
    compute_density_into(t, (real)1.0);
    real md = 0.0;
    for (int i = 0; i < 1000; ++i) md = (real)fmax((double)md, (double)t->dens[i]);
    orc_destroy(t);
    return md;
}

/* scene object as in common.h:24-32 (BOX/SPHERE, FLUID/WALL) */
typedef struct {
    int type, shape;
    real r, c[3], lo[3], hi[3];
} scene_obj;

static int scene_walls(scene_obj *o, int N)
{
    real t = (real)(1.0 / N);
    real lo[6][3] = {{0, 0, 0}, {(real)(1.0 - t), 0, 0}, {0, 0, 0}, {0, (real)(1.0 - t), 0}, {0, 0, 0}, {0, 0, (real)(1.0 - t)}};
    real hi[6][3] = {{t, 1, 1}, {1, 1, 1}, {1, t, 1}, {1, 1, 1}, {1, 1, t}, {1, 1, 1}};
    for (int w = 0; w < 6; ++w) {
        o[w].type = WALL;
        o[w].shape = 0;
        for (int d = 0; d < 3; ++d) { o[w].lo[d] = lo[w][d]; o[w].hi[d] = hi[w][d]; }
    }
    return 6;
}

/* flip3D.cpp:580-627 pushParticle object lookup; returns the matching object
 * index or -1 */
static int scene_lookup(const scene_obj *o, int no, int N, double x, double y, double z, int type)
{
    for (int n = 0; n < no; ++n) {
        int found = 0;
        real thick = (real)(3.0 / N);
        if (o[n].shape == 0) {
            if (x > o[n].lo[0] && x < o[n].hi[0] && y > o[n].lo[1] && y < o[n].hi[1] && z > o[n].lo[2] && z < o[n].hi[2]) {
                if (o[n].type == WALL && x > o[n].lo[0] + thick && x < o[n].hi[0] - thick && y > o[n].lo[1] + thick &&
                    y < o[n].hi[1] - thick && z > o[n].lo[2] + thick && z < o[n].hi[2] - thick)
                    return -1;
                found = 1;
            }
        } else {
            real p[3] = {(real)x, (real)y, (real)z};
            real len = orc_length(p, o[n].c);
            if (len < o[n].r) {
                if (o[n].type == WALL) {
                    found = 1;
                    if (len < o[n].r - thick) return -1;
                } else {
                    found = 1;
                }
            }
        }
        if (found && o[n].type == type) return n;
    }
    return -1;
}

static void scene_push(orc_sim *s, real x, real y, real z, int type)
{
    int q = s->np++;
    s->p[3 * q + 0] = x;
    s->p[3 * q + 1] = y;
    s->p[3 * q + 2] = z;
    s->u[3 * q] = s->u[3 * q + 1] = s->u[3 * q + 2] = 0.0;
    s->nrm[3 * q] = s->nrm[3 * q + 1] = s->nrm[3 * q + 2] = 0.0;
    s->type[q] = (int8_t)type;
    s->m[q] = 1.0;
    s->dens[q] = 10.0;
}

/* flip3D.cpp:75-164 compute_wall_normal (particle normals only; the per-cell
 * wall_normal grids are not read by the step) */
void orc_compute_wall_normals(orc_sim *s)
{
    int N = s->n;
    orc_sort(s);
    double wt = 1.0 / N;
    for (int n = 0; n < s->np; ++n) {
        real *p = s->p + 3 * (size_t)n, *nr = s->nrm + 3 * (size_t)n;
        int i = cell_of_b(N, p[0]), j = cell_of_b(N, p[1]), k = cell_of_b(N, p[2]);
        nr[0] = nr[1] = nr[2] = 0.0;
        if (s->type[n] == WALL) {
            if (p[0] <= 1.1 * wt) nr[0] = 1.0;
            if (p[0] >= 1.0 - 1.1 * wt) nr[0] = -1.0;
            if (p[1] <= 1.1 * wt) nr[1] = 1.0;
            if (p[1] >= 1.0 - 1.1 * wt) nr[1] = -1.0;
            if (p[2] <= 1.1 * wt) nr[2] = 1.0;
            if (p[2] >= 1.0 - 1.1 * wt) nr[2] = -1.0;
            if (nr[0] == 0.0 && nr[1] == 0.0 && nr[2] == 0.0) {
                WINDOW_BEGIN(s, i - 3, i + 3, j - 3, j + 3, k - 3, k + 3, q)
                    if (q == n || s->type[q] != WALL) continue;
                    const real *np = s->p + 3 * (size_t)q;
                    real d = orc_length(p, np);
                    real w = (real)(1.0 / d);
                    nr[0] += w * (p[0] - np[0]) / d;
                    nr[1] += w * (p[1] - np[1]) / d;
                    nr[2] += w * (p[2] - np[2]) / d;
                WINDOW_END
            }
        }
        real d = R_HYPOT(R_HYPOT(nr[0], nr[1]), nr[2]);
        if (d != 0) {
            nr[0] /= d;
            nr[1] /= d;
            nr[2] /= d;
        }
    }
}

/* remove fluid particles that sit in WALL cells (flip3D.cpp:742-761) */
static void scene_finish(orc_sim *s)
{
    int N = s->n;
    orc_sort(s);
    orc_mark_water(s);
    int w = 0;
    for (int n = 0; n < s->np; ++n) {
        const real *p = s->p + 3 * (size_t)n;
        if (s->type[n] != WALL && s->A[cidx(N, cell_of_b(N, p[0]), cell_of_b(N, p[1]), cell_of_b(N, p[2]))] == WALL)
            continue;
        if (w != n) {
            memcpy(s->p + 3 * (size_t)w, s->p + 3 * (size_t)n, 3 * sizeof(real));
            memcpy(s->u + 3 * (size_t)w, s->u + 3 * (size_t)n, 3 * sizeof(real));
            s->type[w] = s->type[n];
            s->m[w] = s->m[n];
            s->dens[w] = s->dens[n];
        }
        w++;
    }
    s->np = w;
    orc_compute_wall_normals(s);
    memset(s->A, AIR, s->C);
}

static void box_obj(scene_obj *o, int type, double x0, double x1, double y0, double y1, double z0, double z1)
{
    memset(o, 0, sizeof(*o));
    o->type = type;
    o->shape = 0;
    o->lo[0] = (real)x0; o->hi[0] = (real)x1;
    o->lo[1] = (real)y0; o->hi[1] = (real)y1;
    o->lo[2] = (real)z0; o->hi[2] = (real)z1;
}

static void sphere_obj(scene_obj *o, int type, double cx, double cy, double cz, double r)
{
    memset(o, 0, sizeof(*o));
    o->type = type;
    o->shape = 1;
    o->c[0] = (real)cx; o->c[1] = (real)cy; o->c[2] = (real)cz;
    o->r = (real)r;
}

/* The reference's scene functions after placeWalls (flip3D.cpp:300-362):
 * 0 damBreakTest (:277-297), 1 spherePourTest (:203-228), 2 cliffPourTest
 * (:249-275), 3 waterDropTest (:230-247).  Sets the emitter of the pour
 * scenes.  Returns the object count, -1 for an unknown scene. */
static int scene_objects(orc_sim *s, int which, scene_obj *o)
{
    int N = s->n;
    double wt = 1.0 / N;
    int no = scene_walls(o, N);
    switch (which) {
    case 0:
        box_obj(&o[no++], FLUID, 0.2, 0.4, wt, 0.4, 0.2, 0.8);
        box_obj(&o[no++], FLUID, wt, 1.0 - wt, wt, 0.06, wt, 1.0 - wt);
        break;
    case 1:
        sphere_obj(&o[no++], WALL, 0.5, 0.4, 0.47, 0.15);
        box_obj(&o[no++], FLUID, wt, 1.0 - wt, wt, 0.05, wt, 1.0 - wt);
        orc_set_emitter(s, (real)0.5, (real)0.5, (real)0.1, 200);
        break;
    case 2:
        box_obj(&o[no++], WALL, 0.0, 0.35, 0.0, 0.4, -0.002, 1.002);
        box_obj(&o[no++], FLUID, 0.35, 1.0 - wt, wt, 0.05, wt, 1.0 - wt);
        orc_set_emitter(s, (real)0.2, (real)0.2, (real)0.1, 200);
        break;
    case 3:
        box_obj(&o[no++], FLUID, wt, 1.0 - wt, wt, 0.2, wt, 1.0 - wt);
        sphere_obj(&o[no++], FLUID, 0.5, 0.6, 0.5, 0.12);
        break;
    default:
        return -1;
    }
    return no;
}

/* flip3D.cpp:649-764 init(0) with one of the reference's test scenes: fluid
 * lattice with rand() jitter (pushParticle, :580-647), then wall particles
 * at wall-cell centres.  Consumes 3 rand() per placed particle, as the
 * reference does. */
int orc_scene_test(orc_sim *s, int which)
{
    int N = s->n;
    double dens = s->prm.density;
    scene_obj o[10];
    s->pour_time = -1;
    int no = scene_objects(s, which, o);
    if (no < 0) return -1;
    reserve_particles(s, (int)orc_scene_capacity(N));
    s->np = 0;
    s->step = 0;
    s->max_dens = orc_calibrate_max_dens(N, dens);
    double w = dens * (1.0 / N);
    double wthick = 1.0 / N;
    int lim = (int)(N / dens);
    for (int i = 0; i < lim; ++i)
        for (int j = 0; j < lim; ++j)
            for (int k = 0; k < lim; ++k) {
                double x = i * w + w / 2.0, y = j * w + w / 2.0, z = k * w + w / 2.0;
                if (x > wthick && x < 1.0 - wthick && y > wthick && y < 1.0 - wthick && z > wthick && z < 1.0 - wthick) {
                    int ob = scene_lookup(o, no, N, x, y, z, FLUID);
                    if (ob >= 0) {
                        int fl = o[ob].type == FLUID;
                        real px = (real)(x + 0.01 * fl * 0.2 * ((orc_rand_next(s) % 101) / 50.0 - 1.0) / N);
                        real py = (real)(y + 0.01 * fl * 0.2 * ((orc_rand_next(s) % 101) / 50.0 - 1.0) / N);
                        real pz = (real)(z + 0.01 * fl * 0.2 * ((orc_rand_next(s) % 101) / 50.0 - 1.0) / N);
                        scene_push(s, px, py, pz, FLUID);
                    }
                }
            }
    w = 1.0 / N;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                double x = i * w + w / 2.0, y = j * w + w / 2.0, z = k * w + w / 2.0;
                int ob = scene_lookup(o, no, N, x, y, z, WALL);
                if (ob >= 0) {
                    (void)orc_rand_next(s);
                    (void)orc_rand_next(s);
                    (void)orc_rand_next(s);
                    scene_push(s, (real)x, (real)y, (real)z, WALL);
                }
            }
    scene_finish(s);
    return s->np;
}

int orc_scene_reference(orc_sim *s) { return orc_scene_test(s, 0); }

/* SplitMix64 finaliser; value k of the stream seeded with `seed` is
 * mix(seed + (k+1)*golden). */
static uint64_t splitmix_at(uint64_t seed, uint64_t k)
{
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Benchmark scene: liquid = box {x<0.4, y<0.6, all z} U sphere(c=(0.6,0.7,
 * 0.5), r=0.12) on the reference's 2x2x2 sub-lattice with the reference's
 * object test, then jitter U(-0.25,0.25)*dx per axis from SplitMix64 keyed
 * by (3*fluid_index + axis); walls as in the reference. */
int orc_scene_benchmark(orc_sim *s, uint64_t seed)
{
    int N = s->n;
    double dens = s->prm.density;
    reserve_particles(s, (int)orc_scene_capacity(N));
    s->np = 0;
    scene_obj o[8];
    int no = scene_walls(o, N);
    o[no].type = FLUID; o[no].shape = 0;
    o[no].lo[0] = 0; o[no].hi[0] = (real)0.4;
    o[no].lo[1] = 0; o[no].hi[1] = (real)0.6;
    o[no].lo[2] = 0; o[no].hi[2] = 1;
    no++;
    o[no].type = FLUID; o[no].shape = 1;
    o[no].r = (real)0.12;
    o[no].c[0] = (real)0.6; o[no].c[1] = (real)0.7; o[no].c[2] = (real)0.5;
    no++;
    s->max_dens = orc_calibrate_max_dens(N, dens);
    double w = dens * (1.0 / N);
    double wthick = 1.0 / N;
    double dx = 1.0 / N;
    int lim = (int)(N / dens);
    uint64_t fi = 0;
    for (int i = 0; i < lim; ++i)
        for (int j = 0; j < lim; ++j)
            for (int k = 0; k < lim; ++k) {
                double x = i * w + w / 2.0, y = j * w + w / 2.0, z = k * w + w / 2.0;
                if (x > wthick && x < 1.0 - wthick && y > wthick && y < 1.0 - wthick && z > wthick && z < 1.0 - wthick) {
                    if (scene_lookup(o, no, N, x, y, z, FLUID) >= 0) {
                        double jit[3];
                        for (int a = 0; a < 3; ++a) {
                            uint64_t v = splitmix_at(seed, 3 * fi + (uint64_t)a);
                            double uu = (double)(v >> 11) * (1.0 / 9007199254740992.0);
                            jit[a] = (uu * 0.5 - 0.25) * dx;
                        }
                        scene_push(s, (real)(x + jit[0]), (real)(y + jit[1]), (real)(z + jit[2]), FLUID);
                        fi++;
                    }
                }
            }
    w = 1.0 / N;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                double x = i * w + w / 2.0, y = j * w + w / 2.0, z = k * w + w / 2.0;
                if (scene_lookup(o, no, N, x, y, z, WALL) >= 0) scene_push(s, (real)x, (real)y, (real)z, WALL);
            }
    scene_finish(s);
    return s->np;
}

/* ---------------------------------------------------------------- mesher */
#include "../flip3d_amd/csrc/mc_table.h"

static const char mc_cases[256][16] = {MC_CASE_LIST};
static const unsigned char mc_edge_corners[12] = {MC_EDGE_CORNERS};

/* implicit.cpp:14-38: distance to the nearest FLUID particle in the 5^3-cell
 * window (capped at 8 DENSITY/gn) less DENSITY/gn; 4.5 DENSITY/gn when a WALL
 * particle lies within DENSITY/gn.  The window is read through the sorter's
 * current buckets. */
static double implicit_at(const orc_sim *s, const real p[3], real density)
{
    int gn = s->n;
    int ci = cell_of_b(gn, p[0]), cj = cell_of_b(gn, p[1]), ck = cell_of_b(gn, p[2]);
    double phi = 8.0 * density / gn;
    WINDOW_BEGIN(s, ci - 2, ci + 2, cj - 2, cj + 2, ck - 2, ck + 2, q)
        const real *np = s->p + 3 * (size_t)q;
        if (s->type[q] == WALL) {
            if (orc_length(np, p) < density / gn) return 4.5 * density / gn;
            continue;
        }
        double d = orc_length(np, p);
        if (d < phi) phi = d;
    WINDOW_END
    return phi - density / gn;
}

typedef struct {
    int set;
    float x, y, z;
} mc_vert;

/* CIsoSurface.cpp:587-682: the point on edge e of cell (x,y,z), interpolated
 * from the edge's first to its second end as the class orders them */
static mc_vert mc_intersect(const float *f, int nc, float len, int x, int y, int z, int e)
{
    static const signed char d1[12][3] = {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 1}, {0, 1, 1},
                                          {1, 1, 1}, {1, 0, 1}, {0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}};
    static const signed char d2[12][3] = {{0, 1, 0}, {1, 1, 0}, {1, 0, 0}, {0, 0, 0}, {0, 1, 1}, {1, 1, 1},
                                          {1, 0, 1}, {0, 0, 1}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}};
    unsigned v1x = x + d1[e][0], v1y = y + d1[e][1], v1z = z + d1[e][2];
    unsigned v2x = x + d2[e][0], v2y = y + d2[e][1], v2z = z + d2[e][2];
    float x1 = v1x * len, y1 = v1y * len, z1 = v1z * len, x2 = v2x * len, y2 = v2y * len, z2 = v2z * len;
    unsigned np1 = nc + 1, slice = np1 * np1;
    float a = f[v1z * slice + v1y * np1 + v1x], b = f[v2z * slice + v2y * np1 + v2x];
    float mu = (0.0f - a) / (b - a);
    mc_vert r = {1, x1 + mu * (x2 - x1), y1 + mu * (y2 - y1), z1 + mu * (z2 - z1)};
    return r;
}

/* CIsoSurface.cpp:550-585 GetEdgeID */
static unsigned mc_edge_id(int nc, int x, int y, int z, int e)
{
    static const signed char o[12][4] = {{0, 0, 0, 1}, {0, 1, 0, 0}, {1, 0, 0, 1}, {0, 0, 0, 0},
                                         {0, 0, 1, 1}, {0, 1, 1, 0}, {1, 0, 1, 1}, {0, 0, 1, 0},
                                         {0, 0, 0, 2}, {0, 1, 0, 2}, {1, 1, 0, 2}, {1, 0, 0, 2}};
    unsigned np1 = nc + 1;
    unsigned vx = x + o[e][0], vy = y + o[e][1], vz = z + o[e][2];
    return 3 * (vz * np1 * np1 + vy * np1 + vx) + o[e][3];
}

static int hexval(char c) { return c <= '9' ? c - '0' : c - 'a' + 10; }

static int orc_isnormal(double v) { return isnormal(v) != 0; }

struct orc_mesh_out {
    int nv, nf;
    double *v, *n;
    int32_t *f;
};
static struct orc_mesh_out g_mesh;

/* mesher.cpp:137-195 generateMesh(A, sort, particles, DENSITY, mg): implicit
 * field on the mg^3 lattice (x fastest, negated), CIsoSurface<float> at iso
 * 0 with cell length 1/(mg-1) (CIsoSurface.cpp:331-503: per-cell vertex
 * inserts on the owned edges, triangles cell by cell in x-major order,
 * vertices renumbered by edge id, CalculateNormals :722-777), the 0.1 edge
 * length filter, buildConnection and two smoothMesh passes (mesher.cpp:13-101).
 * Returns the vertex count. */
int orc_mesh(orc_sim *s, int mg)
{
    int nc = mg - 1;
    real density = (real)s->prm.density;
    size_t npts = (size_t)mg * mg * mg;
    float *f = (float *)malloc(npts * sizeof(float));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < mg; ++i)
        for (int j = 0; j < mg; ++j)
            for (int k = 0; k < mg; ++k) {
                double h = 1.0 / (double)(mg - 1);
                double x = i * h, y = j * h, z = k * h;
                real p[3] = {(real)x, (real)y, (real)z};
                double value = implicit_at(s, p, density);
                if (i == 0 || i == mg - 1 || j == 0 || j == mg - 1 || k == 0 || k == mg - 1) value = fmax(value, 0.01);
                f[(size_t)(mg * mg) * k + mg * j + i] = (float)-value;
            }
    float len = (float)(1.0 / (mg - 1));
    size_t nid = 3 * npts;
    mc_vert *verts = (mc_vert *)calloc(nid, sizeof(mc_vert));
    size_t ncell = (size_t)nc * nc * nc;
    unsigned char *tcount = (unsigned char *)calloc(ncell, 1);
    unsigned *tri = (unsigned *)malloc(ncell * 15 * sizeof(unsigned));
    unsigned np1 = nc + 1, slice = np1 * np1;
    for (int n = 0; n < nc * nc * nc; n++) {
        int z = n / (nc * nc), y = (n % (nc * nc)) / nc, x = (n % (nc * nc)) % nc;
        static const signed char co[8][3] = {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0},
                                             {0, 0, 1}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}};
        unsigned ti = 0;
        for (int b = 0; b < 8; ++b)
            if (f[(z + co[b][2]) * slice + (y + co[b][1]) * np1 + (x + co[b][0])] < 0.0f) ti |= 1u << b;
        unsigned mask = 0;
        for (int e = 0; e < 12; ++e) {
            int a = mc_edge_corners[e] >> 4, b = mc_edge_corners[e] & 15;
            if (((ti >> a) ^ (ti >> b)) & 1) mask |= 1u << e;
        }
        if (!mask) continue;
        int own[12] = {0};
        own[3] = own[0] = own[8] = 1;
        if (x == nc - 1) own[2] = own[11] = 1;
        if (y == nc - 1) own[1] = own[9] = 1;
        if (z == nc - 1) own[4] = own[7] = 1;
        if (x == nc - 1 && y == nc - 1) own[10] = 1;
        if (x == nc - 1 && z == nc - 1) own[6] = 1;
        if (y == nc - 1 && z == nc - 1) own[5] = 1;
        for (int e = 0; e < 12; ++e)
            if (own[e] && (mask >> e & 1)) {
                unsigned id = mc_edge_id(nc, x, y, z, e);
                if (!verts[id].set) verts[id] = mc_intersect(f, nc, len, x, y, z, e);
            }
        size_t L = ((size_t)x * nc + y) * nc + z; /* the merge order of :486-505 */
        const char *cs = mc_cases[ti];
        int t = 0;
        for (int i = 0; cs[i]; i += 3, ++t)
            for (int c = 0; c < 3; ++c) tri[L * 15 + 3 * t + c] = mc_edge_id(nc, x, y, z, hexval(cs[i + c]));
        tcount[L] = (unsigned char)t;
    }
    /* renumber by edge id (RenameVerticesAndTriangles, :684-720) */
    int32_t *rid = (int32_t *)malloc(nid * sizeof(int32_t));
    int nv = 0;
    for (size_t id = 0; id < nid; ++id) rid[id] = verts[id].set ? nv++ : -1;
    float *vx = (float *)malloc(3 * (size_t)(nv + 1) * sizeof(float));
    for (size_t id = 0; id < nid; ++id)
        if (verts[id].set) {
            vx[3 * rid[id]] = verts[id].x;
            vx[3 * rid[id] + 1] = verts[id].y;
            vx[3 * rid[id] + 2] = verts[id].z;
        }
    size_t ntri = 0;
    for (size_t L = 0; L < ncell; ++L) ntri += tcount[L];
    int32_t *tv = (int32_t *)malloc(3 * (ntri + 1) * sizeof(int32_t));
    size_t t = 0;
    for (size_t L = 0; L < ncell; ++L)
        for (int k = 0; k < tcount[L]; ++k, ++t)
            for (int c = 0; c < 3; ++c) tv[3 * t + c] = rid[tri[L * 15 + 3 * k + c]];
    /* CalculateNormals: triangle normals summed in triangle order, then
     * divided by their float length */
    float *nm = (float *)calloc(3 * (size_t)(nv + 1), sizeof(float));
    for (size_t i = 0; i < ntri; ++i) {
        int id0 = tv[3 * i], id1 = tv[3 * i + 1], id2 = tv[3 * i + 2];
        float v1[3], v2[3], no[3];
        for (int c = 0; c < 3; ++c) {
            v1[c] = vx[3 * id1 + c] - vx[3 * id0 + c];
            v2[c] = vx[3 * id2 + c] - vx[3 * id0 + c];
        }
        no[0] = v1[2] * v2[1] - v1[1] * v2[2];
        no[1] = v1[0] * v2[2] - v1[2] * v2[0];
        no[2] = v1[1] * v2[0] - v1[0] * v2[1];
        for (int c = 0; c < 3; ++c) {
            nm[3 * id0 + c] += no[c];
            nm[3 * id1 + c] += no[c];
            nm[3 * id2 + c] += no[c];
        }
    }
    for (int i = 0; i < nv; ++i) {
        float l = sqrtf(nm[3 * i] * nm[3 * i] + nm[3 * i + 1] * nm[3 * i + 1] + nm[3 * i + 2] * nm[3 * i + 2]);
        nm[3 * i] /= l;
        nm[3 * i + 1] /= l;
        nm[3 * i + 2] /= l;
    }
    /* mesher.cpp:167-185 faces with every edge shorter than 0.1 */
    int32_t *faces = (int32_t *)malloc(3 * (ntri + 1) * sizeof(int32_t));
    int nf = 0;
    double maxlen2 = 0.1 * 0.1;
    for (size_t i = 0; i < ntri; ++i) {
        const int32_t *q = tv + 3 * i;
        real a[3], b[3], c[3];
        for (int d = 0; d < 3; ++d) {
            a[d] = (real)vx[3 * q[0] + d];
            b[d] = (real)vx[3 * q[1] + d];
            c[d] = (real)vx[3 * q[2] + d];
        }
        if (orc_length2(a, b) < maxlen2 && orc_length2(b, c) < maxlen2 && orc_length2(a, c) < maxlen2) {
            faces[3 * nf] = q[0];
            faces[3 * nf + 1] = q[1];
            faces[3 * nf + 2] = q[2];
            nf++;
        }
    }
    double *V = (double *)malloc(3 * (size_t)(nv + 1) * sizeof(double));
    double *Nn = (double *)malloc(3 * (size_t)(nv + 1) * sizeof(double));
    for (size_t i = 0; i < 3 * (size_t)nv; ++i) {
        V[i] = vx[i];
        Nn[i] = nm[i];
    }
    /* buildConnection (mesher.cpp:88-101): per vertex, the other two corners
     * of each face in face order */
    int *cnt = (int *)calloc((size_t)nv + 1, sizeof(int));
    for (int i = 0; i < nf; ++i)
        for (int c = 0; c < 3; ++c) cnt[faces[3 * i + c]] += 2;
    size_t *off = (size_t *)malloc(((size_t)nv + 1) * sizeof(size_t));
    size_t acc = 0;
    for (int i = 0; i < nv; ++i) {
        off[i] = acc;
        acc += cnt[i];
    }
    off[nv] = acc;
    int32_t *nb = (int32_t *)malloc((acc + 1) * sizeof(int32_t));
    memset(cnt, 0, ((size_t)nv + 1) * sizeof(int));
    for (int i = 0; i < nf; ++i) {
        int a = faces[3 * i], b = faces[3 * i + 1], c = faces[3 * i + 2];
        nb[off[a] + cnt[a]++] = b;
        nb[off[a] + cnt[a]++] = c;
        nb[off[b] + cnt[b]++] = a;
        nb[off[b] + cnt[b]++] = c;
        nb[off[c] + cnt[c]++] = a;
        nb[off[c] + cnt[c]++] = b;
    }
    /* smoothMesh(.., 2, 1, 1), mesher.cpp:13-86: the copies start at zero and
     * keep their last value where a vertex has no usable sample */
    double *CV = (double *)calloc(3 * (size_t)(nv + 1), sizeof(double));
    double *CN = (double *)calloc(3 * (size_t)(nv + 1), sizeof(double));
    for (int it = 0; it < 2; ++it) {
        for (int vi = 0; vi < nv; ++vi) {
            float wsum = 0.0f;
            double v[3] = {0.0, 0.0, 0.0}, vn[3] = {0.0, 0.0, 0.0};
            for (int k = -1; k < (int)(off[vi + 1] - off[vi]); ++k) {
                int ni = k < 0 ? vi : nb[off[vi] + k];
                int ok = 1;
                for (int d = 0; d < 3; ++d)
                    if (!orc_isnormal(V[3 * ni + d]) || !orc_isnormal(Nn[3 * ni + d])) ok = 0;
                if (!ok) continue;
                float w = 1.0f;
                for (int d = 0; d < 3; ++d) {
                    v[d] += w * V[3 * ni + d];
                    vn[d] += w * Nn[3 * ni + d];
                }
                wsum += w;
            }
            float l = hypotf((float)vn[0], hypotf((float)vn[1], (float)vn[2]));
            for (int d = 0; d < 3; ++d) {
                if (wsum != 0.0f) CV[3 * vi + d] = v[d] / (double)wsum;
                if (l != 0.0f) CN[3 * vi + d] = vn[d] / l;
            }
        }
        memcpy(V, CV, 3 * (size_t)nv * sizeof(double));
        memcpy(Nn, CN, 3 * (size_t)nv * sizeof(double));
    }
    free(g_mesh.v);
    free(g_mesh.n);
    free(g_mesh.f);
    g_mesh.nv = nv;
    g_mesh.nf = nf;
    g_mesh.v = V;
    g_mesh.n = Nn;
    g_mesh.f = faces;
    free(f); free(verts); free(tcount); free(tri); free(rid); free(vx); free(tv); free(nm);
    free(cnt); free(off); free(nb); free(CV); free(CN);
    return nv;
}

int orc_mesh_faces(void) { return g_mesh.nf; }

void orc_mesh_get(double *v, double *n, int32_t *f)
{
    if (v) memcpy(v, g_mesh.v, 3 * (size_t)g_mesh.nv * sizeof(double));
    if (n) memcpy(n, g_mesh.n, 3 * (size_t)g_mesh.nv * sizeof(double));
    if (f) memcpy(f, g_mesh.f, 3 * (size_t)g_mesh.nf * sizeof(int32_t));
}
