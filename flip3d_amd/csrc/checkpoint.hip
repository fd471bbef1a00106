// checkpoint.hip - state files in the reference's render/save/<step>_scene.dat
// layout (flip3D.cpp:453-513 saveState, :515-578 loadState).
//
// File: int count, float max_dens, int pourTime, float pourPos[2], float
// pourRad, then count 72-byte particle records in creation order, then the
// pressure grid (n^3 floats, i-j-k order, k fastest), then four wall-normal
// floats per cell in the same order.  Host code only.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "flip_internal.hpp"

namespace {

// src/common.h:34-45, 72 bytes
struct RefParticle {
    float p[3];
    float u[3];
    float n[3];
    char type;
    char visible;
    char remove;
    char thinparticle;
    float tmp[2][3];
    float m;
    float dens;
};
static_assert(sizeof(RefParticle) == 72, "reference particle record is 72 bytes");

constexpr long HEADER = 4 + 4 + 4 + 8 + 4;

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
};

// compute_wall_normal (flip3D.cpp:75-164) grid: the normal of the wall
// particle in each wall cell, and for non-wall cells next to walls the
// normalised sum of the adjacent wall-cell normals with flag 1.
void wall_normal_grid(int N, const std::vector<RefParticle> &ps, std::vector<float> &wn)
{
    const size_t C = (size_t)N * N * N;
    wn.assign(4 * C, 0.0f);
    std::vector<int8_t> wall(C, 0);
    auto cell = [N](const float *p) {
        int i = (int)fmin((double)(N - 1), fmax(0.0, (double)(p[0] * (float)N)));
        int j = (int)fmin((double)(N - 1), fmax(0.0, (double)(p[1] * (float)N)));
        int k = (int)fmin((double)(N - 1), fmax(0.0, (double)(p[2] * (float)N)));
        return ((size_t)i * N + j) * N + k;
    };
    for (const RefParticle &p : ps) {
        size_t c = cell(p.p);
        wn[4 * c] = wn[4 * c + 1] = wn[4 * c + 2] = 0.0f;
        if (p.n[0] != 0.0f || p.n[1] != 0.0f || p.n[2] != 0.0f)
            for (int a = 0; a < 3; ++a) wn[4 * c + a] = p.n[a];
    }
    // markWater flags a cell WALL when it holds a wall particle (sorter.cpp:70-93)
    for (const RefParticle &p : ps)
        if (p.type == FLIP_WALL) {
            int i = (int)fmax(0.0, fmin((double)(N - 1), (double)((float)N * p.p[0])));
            int j = (int)fmax(0.0, fmin((double)(N - 1), (double)((float)N * p.p[1])));
            int k = (int)fmax(0.0, fmin((double)(N - 1), (double)((float)N * p.p[2])));
            wall[((size_t)i * N + j) * N + k] = 1;
        }
    std::vector<float> out(wn);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                size_t c = ((size_t)i * N + j) * N + k;
                out[4 * c + 3] = 0.0f;
                if (wall[c]) continue;
                int sum = 0;
                float nr[3] = {0.0f, 0.0f, 0.0f};
                const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                                      {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
                for (auto &q : nb) {
                    if (q[0] < 0 || q[0] > N - 1 || q[1] < 0 || q[1] > N - 1 || q[2] < 0 || q[2] > N - 1) continue;
                    size_t s = ((size_t)q[0] * N + q[1]) * N + q[2];
                    if (!wall[s]) continue;
                    sum++;
                    for (int a = 0; a < 3; ++a) nr[a] += wn[4 * s + a];
                }
                if (sum > 0) {
                    float d = (float)hypot(hypot((double)nr[0], (double)nr[1]), (double)nr[2]);
                    for (int a = 0; a < 3; ++a) out[4 * c + a] = nr[a] / d;
                    out[4 * c + 3] = 1.0f;
                }
            }
    wn.swap(out);
}

// render/save/objects.dat sits beside the <step>_scene.dat files
std::string objects_path(const char *path)
{
    std::string p(path);
    size_t slash = p.find_last_of('/');
    return (slash == std::string::npos ? std::string() : p.substr(0, slash + 1)) + "objects.dat";
}

// saveState's object list (flip3D.cpp:498-508), in its text format
bool write_objects(const std::string &path, const std::vector<FlipCtx::SceneObject> &obj)
{
    File f;
    f.f = fopen(path.c_str(), "w");
    if (!f.f) return false;
    bool ok = fprintf(f.f, "%d\n", (int)obj.size()) > 0;
    for (const auto &o : obj) {
        ok = ok && fprintf(f.f, "%d %d %d %d\n", o.type, o.shape, o.material, o.visible) > 0;
        ok = ok && fprintf(f.f, "%f %f %f %f %f %f %f %f %f %f\n", o.f[0], o.f[1], o.f[2], o.f[3], o.f[4], o.f[5],
                           o.f[6], o.f[7], o.f[8], o.f[9]) > 0;
    }
    return ok;
}

// loadState's read of it (flip3D.cpp:561-574); false when the file is absent
// or malformed
bool read_objects(const std::string &path, std::vector<FlipCtx::SceneObject> &obj)
{
    File f;
    f.f = fopen(path.c_str(), "r");
    if (!f.f) return false;
    int count = 0;
    if (fscanf(f.f, "%d\n", &count) != 1 || count < 0 || count > 1000000) return false;
    std::vector<FlipCtx::SceneObject> out(count);
    for (auto &o : out) {
        if (fscanf(f.f, "%d %d %d %d\n", &o.type, &o.shape, &o.material, &o.visible) != 4) return false;
        if (fscanf(f.f, "%f %f %f %f %f %f %f %f %f %f\n", &o.f[0], &o.f[1], &o.f[2], &o.f[3], &o.f[4], &o.f[5],
                   &o.f[6], &o.f[7], &o.f[8], &o.f[9]) != 10)
            return false;
    }
    obj.swap(out);
    return true;
}

}  // namespace

extern "C" int flip_gpu_save_state(flip_ctx *ctx, const char *path)
{
    if (!ctx || !path) return flip_internal_fail(FLIP_E_ARG, "null argument");
    FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
    try {
        if (!c->uploaded) return flip_internal_fail(FLIP_E_STATE, "no particles uploaded");
        const int64_t np = c->np;
        std::vector<float4> p4(np), v4(np), n4(np);
        std::vector<int32_t> mt(np);
        std::vector<float> press(c->C);
        FLIP_CHECK(hipStreamSynchronize(c->stream));
        if (np > 0) {
            FLIP_CHECK(hipMemcpy(p4.data(), c->P4(), np * sizeof(float4), hipMemcpyDeviceToHost));
            FLIP_CHECK(hipMemcpy(v4.data(), c->V4(), np * sizeof(float4), hipMemcpyDeviceToHost));
            FLIP_CHECK(hipMemcpy(mt.data(), c->M(), np * sizeof(int32_t), hipMemcpyDeviceToHost));
            FLIP_CHECK(hipMemcpy(n4.data(), c->nrm_id, np * sizeof(float4), hipMemcpyDeviceToHost));
        }
        FLIP_CHECK(hipMemcpy(press.data(), c->press, c->C * sizeof(float), hipMemcpyDeviceToHost));
        std::vector<RefParticle> rec(np);
        for (int64_t s = 0; s < np; ++s) {
            int64_t id = (uint32_t)mt[s] >> 2;
            if (id >= np) return flip_internal_fail(FLIP_E_STATE, "corrupt particle index");
            RefParticle &r = rec[id];
            memset(&r, 0, sizeof(r));
            r.p[0] = p4[s].x;
            r.p[1] = p4[s].y;
            r.p[2] = p4[s].z;
            r.u[0] = v4[s].x;
            r.u[1] = v4[s].y;
            r.u[2] = v4[s].z;
            r.type = (char)(mt[s] & 3);
            r.visible = r.type == FLIP_FLUID;  // fluid objects are visible, walls not (flip3D.cpp:282-327)
            r.m = fabsf(p4[s].w);
            r.dens = v4[s].w;
        }
        for (int64_t i = 0; i < np; ++i) {
            rec[i].n[0] = n4[i].x;
            rec[i].n[1] = n4[i].y;
            rec[i].n[2] = n4[i].z;
        }
        std::vector<float> wn;
        wall_normal_grid(c->n, rec, wn);
        File f;
        f.f = fopen(path, "wb");
        if (!f.f) return flip_internal_fail(FLIP_E_ARG, "cannot open checkpoint for writing");
        int count = (int)np, pour_time = c->pour_time;
        float md = c->max_dens, pour_pos[2] = {c->pour_pos[0], c->pour_pos[1]}, pour_rad = c->pour_rad;
        bool ok = fwrite(&count, 4, 1, f.f) == 1 && fwrite(&md, 4, 1, f.f) == 1 && fwrite(&pour_time, 4, 1, f.f) == 1 &&
                  fwrite(pour_pos, 8, 1, f.f) == 1 && fwrite(&pour_rad, 4, 1, f.f) == 1;
        ok = ok && (np == 0 || fwrite(rec.data(), sizeof(RefParticle), np, f.f) == (size_t)np);
        ok = ok && fwrite(press.data(), 4, c->C, f.f) == (size_t)c->C;
        ok = ok && fwrite(wn.data(), 4, wn.size(), f.f) == wn.size();
        if (!ok) return flip_internal_fail(FLIP_E_ARG, "short write to checkpoint");
        if (!c->objects_saved) {  // the first save of the context, as saveState's firstTime
            if (!write_objects(objects_path(path), c->objects))
                return flip_internal_fail(FLIP_E_ARG, "cannot write objects.dat beside the checkpoint");
            c->objects_saved = true;
        }
        return FLIP_OK;
    } catch (const flip_hip_error &e) {
        return flip_internal_fail(FLIP_E_HIP, e.what.c_str());
    } catch (const flip_arg_error &e) {
        return flip_internal_fail(e.code, e.what.c_str());
    }
}

extern "C" int flip_gpu_load_state(flip_ctx *ctx, const char *path)
{
    if (!ctx || !path) return flip_internal_fail(FLIP_E_ARG, "null argument");
    FlipCtx *c = reinterpret_cast<FlipCtx *>(ctx);
    try {
        File f;
        f.f = fopen(path, "rb");
        if (!f.f) return flip_internal_fail(FLIP_E_ARG, "cannot open checkpoint");
        fseek(f.f, 0, SEEK_END);
        long size = ftell(f.f);
        fseek(f.f, 0, SEEK_SET);
        int count = 0, pour_time = 0;
        float md = 0.0f, pour_pos[2], pour_rad;
        if (fread(&count, 4, 1, f.f) != 1 || fread(&md, 4, 1, f.f) != 1 || fread(&pour_time, 4, 1, f.f) != 1 ||
            fread(pour_pos, 8, 1, f.f) != 1 || fread(&pour_rad, 4, 1, f.f) != 1)
            return flip_internal_fail(FLIP_E_ARG, "truncated checkpoint header");
        if (count < 0) return flip_internal_fail(FLIP_E_ARG, "negative particle count in checkpoint");
        long expect = HEADER + 72L * count + 4L * c->C + 16L * c->C;
        if (size != expect) return flip_internal_fail(FLIP_E_ARG, "checkpoint size does not match the grid size n");
        std::vector<RefParticle> rec(count);
        if (count > 0 && fread(rec.data(), sizeof(RefParticle), count, f.f) != (size_t)count)
            return flip_internal_fail(FLIP_E_ARG, "truncated checkpoint particles");
        std::vector<float> press(c->C);
        if (fread(press.data(), 4, c->C, f.f) != (size_t)c->C)
            return flip_internal_fail(FLIP_E_ARG, "truncated checkpoint pressure");
        std::vector<float> pos(3 * (size_t)count), vel(3 * (size_t)count), nrm(3 * (size_t)count), m(count),
            dens(count);
        std::vector<int8_t> type(count);
        for (int i = 0; i < count; ++i) {
            for (int a = 0; a < 3; ++a) {
                pos[3 * i + a] = rec[i].p[a];
                vel[3 * i + a] = rec[i].u[a];
                nrm[3 * i + a] = rec[i].n[a];
            }
            type[i] = (int8_t)rec[i].type;
            m[i] = rec[i].m;
            dens[i] = rec[i].dens;
        }
        int rc = flip_gpu_upload_particles(ctx, count, pos.data(), vel.data(), nrm.data(), type.data(), m.data(),
                                           dens.data());
        if (rc) return rc;
        c->max_dens = md;
        FLIP_CHECK(hipMemcpy(c->press, press.data(), c->C * sizeof(float), hipMemcpyHostToDevice));
        // the emitter travels with the state (flip3D.cpp:536-538); the caller
        // sets the step number, which the reference takes from the file name
        set_emitter(c, pour_pos[0], pour_pos[1], pour_rad, pour_time);
        // the object list on the context's first load, as loadState's firstTime
        // (the reference reads it unchecked; a missing or malformed file keeps
        // the context's list)
        if (!c->objects_loaded) {
            std::vector<FlipCtx::SceneObject> obj;
            if (read_objects(objects_path(path), obj)) c->objects = obj;
            c->objects_loaded = true;
        }
        return FLIP_OK;
    } catch (const flip_hip_error &e) {
        return flip_internal_fail(FLIP_E_HIP, e.what.c_str());
    } catch (const flip_arg_error &e) {
        return flip_internal_fail(e.code, e.what.c_str());
    }
}
