// export.hip - the reference's per-frame output files from the context's
// state: the point-splat preview (flip3D.cpp:1105-1192 write_frame, rendered
// on the GPU) written as a BMP (write_bmp.cpp:52-118), and the last surface
// mesh as OBJ (exporter.cpp:127-164) and Sunflow (exporter.cpp:72-125) files.
// The files are byte-identical to the reference's for the same state.
#include "flip_device.hpp"
#include "flip_internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

constexpr int PREVIEW_W = 1024, PREVIEW_H = 700;  // write_frame's image

__global__ void k_preview_zero(int32_t *a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

struct PreviewView {
    float R[3][3];
    float winrate, eye, offset;
};

// write_frame's projection of one FLUID particle (FLOAT arithmetic with the
// reference's double literals) and its 2x2 splat, counted per pixel: every
// splat blends the same colour at the same alpha, so a pixel's value depends
// only on how many splats reach it.
__global__ void k_preview_count(int64_t np, const float4 *__restrict__ pos, const int32_t *__restrict__ meta,
                                PreviewView view, int32_t *count)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= np || (meta[n] & 3) != FLIP_FLUID) return;
    const float4 q = pos[n];
    const float p0[3] = {(float)((double)q.x - 0.5), (float)((double)q.y - 0.5), (float)((double)q.z - 0.5)};
    float p[3];
    for (int i = 0; i < 3; ++i) {
        float acc = 0.0f;
        for (int k = 0; k < 3; ++k) acc += view.R[i][k] * p0[k];
        p[i] = acc;
    }
    const float z = (float)(((double)view.offset + 0.5) + (double)p[2]);
    const float fac = view.eye / (view.eye + z);
    const float x = (float)((double)fac * ((double)p[0] + 0.05));
    const float y = (float)((double)fac * ((double)p[1] + 0.25));
    const double wr = (double)view.winrate;
    if (!((double)x > -0.5 && (double)x < 0.5 && (double)y > -wr / 2.0 && (double)y < wr / 2.0)) return;
    const int i = (int)(PREVIEW_W * ((double)x + 0.5)), j = (int)(PREVIEW_W * ((double)y + wr / 2.0));
    for (int ni = i - 1; ni < i + 1; ++ni)
        for (int nj = j - 1; nj < j + 1; ++nj)
            if (ni > 0 && ni < PREVIEW_W && nj > 0 && nj < PREVIEW_H) atomicAdd(&count[ni + nj * PREVIEW_W], 1);
}

// the blend image = a c + (1 - a) image applied count times from black, then
// 255 min(1, image) as bytes; alpha stays 1.  Once a step leaves the value
// unchanged every later one does too.
__global__ void k_preview_pixels(const int32_t *__restrict__ count, float ac0, float ac1, float ac2, double keep,
                                 uint8_t *rgba)
{
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= PREVIEW_W * PREVIEW_H) return;
    const int m = count[px];
    const float ac[3] = {ac0, ac1, ac2};
    for (int c = 0; c < 3; ++c) {
        float v = 0.0f;
        for (int k = 0; k < m; ++k) {
            const float nv = (float)((double)ac[c] + keep * (double)v);
            if (nv == v) break;
            v = nv;
        }
        rgba[4 * px + c] = (uint8_t)(255.0 * fmin(1.0, (double)v));
    }
    rgba[4 * px + 3] = 255;
}

// buildRotationMatrix(-0.2, 0.2) (flip3D.cpp:1105-1121) in FLOAT with the
// float overloads of cos and sin
void rotation(float R[3][3])
{
    const float xr = -0.2f, yr = 0.2f;
    const float Rx[3][3] = {{1, 0, 0}, {0, cosf(xr), -sinf(xr)}, {0, sinf(xr), cosf(xr)}};
    const float Ry[3][3] = {{cosf(yr), 0, sinf(yr)}, {0, 1, 0}, {-sinf(yr), 0, cosf(yr)}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float acc = 0.0f;
            for (int k = 0; k < 3; ++k) acc += Rx[i][k] * Ry[k][j];
            R[i][j] = acc;
        }
}

struct OutFile {
    FILE *f = nullptr;
    ~OutFile()
    {
        if (f) fclose(f);
    }
};

// The Sunflow template writeSunflow writes beside every scene file (the
// scene files include it); its text is the reference's renderer setup.
const char SUNFLOW_TEMPLATE[] = R"(image {
    resolution 900 600
    aa 0 2
    %  filter gaussian
}
trace-depths {
    diff 4
    refl 3
    refr 3
}
gi {
   type igi
   samples 64         % number of virtual photons per set
   sets 1             % number of sets (increase this to translate shadow boundaries into noise)
   b 0.00005          % bias - decrease this values until bright spots dissapear
   bias-samples 0     % set this >0 to make the algorithm unbiased
}
camera {
    type pinhole
    eye    50 -130 100
    target 50 55 20
    up     0 0 1
    fov    45
    aspect 1.5
}
shader {
   name Liquid
   type shiny
   diff { "sRGB nonlinear" 0.8 0.9 1.0 }
   refl 0.5
}
object {
    shader none
    type cornellbox
    corner0 0 0 0
    corner1  100  100 100
    left    0.80 0.25 0.25
    right   0.25 0.25 0.80
    top     0.70 0.70 0.70
    bottom  0.70 0.70 0.70
    back    0.70 0.70 0.70
    emit    15 15 15
    samples 32
}
light {
    type point
    color { "sRGB nonlinear" 1.000 1.000 1.000 }
    power 90000
    p 50 -200 100
}
)";

}  // namespace

void preview_render(FlipCtx *c, uint8_t *rgba_host)
{
    PreviewView view;
    rotation(view.R);
    view.winrate = (float)PREVIEW_H / (float)PREVIEW_W;
    view.eye = 2.0f;
    view.offset = 0.3f;
    const float alpha = (float)std::exp(-c->prm.density * c->n / 23.0);
    const float col[3] = {0.3f, 0.5f, 0.8f};
    int32_t *count = nullptr;
    uint8_t *img = nullptr;
    hipStream_t s = c->stream;
    FLIP_CHECK(hipMalloc((void **)&count, sizeof(int32_t) * PREVIEW_W * PREVIEW_H));
    FLIP_CHECK(hipMalloc((void **)&img, 4 * PREVIEW_W * PREVIEW_H));
    const int npx = PREVIEW_W * PREVIEW_H;
    hipLaunchKernelGGL(k_preview_zero, (npx + 255) / 256, 256, 0, s, count, npx);
    if (c->np > 0)
        hipLaunchKernelGGL(k_preview_count, (unsigned)((c->np + 255) / 256), 256, 0, s, c->np, c->P4(), c->M(), view,
                           count);
    hipLaunchKernelGGL(k_preview_pixels, (npx + 255) / 256, 256, 0, s, count, alpha * col[0], alpha * col[1],
                       alpha * col[2], 1.0 - (double)alpha, img);
    FLIP_CHECK(hipMemcpyAsync(rgba_host, img, 4 * npx, hipMemcpyDeviceToHost, s));
    FLIP_CHECK(hipStreamSynchronize(s));
    (void)hipFree(count);
    (void)hipFree(img);
}

int preview_width() { return PREVIEW_W; }
int preview_height() { return PREVIEW_H; }

// write_bmp(name, buffer, width, height, true): 54-byte little-endian header,
// then rows of BGR bytes padded to (3 (width + 1) / 4) * 4 bytes
bool write_bmp_file(const char *path, const uint8_t *rgba, int width, int height)
{
    const int bpl = (3 * (width + 1) / 4) * 4;
    OutFile f;
    f.f = fopen(path, "wb");
    if (!f.f) return false;
    auto put32 = [&](int32_t v) { return fwrite(&v, 4, 1, f.f) == 1; };
    auto put16 = [&](int16_t v) { return fwrite(&v, 2, 1, f.f) == 1; };
    bool ok = fwrite("BM", 2, 1, f.f) == 1 && put32(54 + bpl * height) && put32(0) && put32(54) && put32(40) &&
              put32(width) && put32(height) && put16(1) && put16(24) && put32(0) && put32(bpl * height) && put32(0) &&
              put32(0) && put32(0) && put32(0);
    std::vector<uint8_t> lines((size_t)bpl * height, 0);
    for (int i = 0; i < height; ++i)
        for (int j = 0; j < width; ++j) {
            const uint8_t *px = rgba + 4 * ((size_t)width * i + j);
            uint8_t *o = &lines[(size_t)i * bpl + 3 * j];
            o[0] = px[2];
            o[1] = px[1];
            o[2] = px[0];
        }
    return ok && fwrite(lines.data(), lines.size(), 1, f.f) == 1;
}

// write_obj: vertices shifted by wall_thick and scaled by 1/(1 - 2
// wall_thick), y and z swapped; normals swapped; 1-based faces with normal
// indices
bool write_obj_file(const char *path, const std::vector<double> &v, const std::vector<double> &n,
                    const std::vector<int32_t> &f, float wall_thick)
{
    OutFile o;
    o.f = fopen(path, "w");
    if (!o.f) return false;
    const double s = 1.0 / (1.0 - 2.0 * wall_thick);
    bool ok = true;
    for (size_t i = 0; i < v.size(); i += 3)
        ok = ok && fprintf(o.f, "v %lf %lf %lf\n", s * (v[i] - wall_thick), s * (v[i + 2] - wall_thick),
                           s * (v[i + 1] - wall_thick)) > 0;
    for (size_t i = 0; i < n.size(); i += 3) ok = ok && fprintf(o.f, "vn %lf %lf %lf\n", n[i], n[i + 2], n[i + 1]) > 0;
    for (size_t i = 0; i < f.size(); i += 3)
        ok = ok && fprintf(o.f, "f %d//%d %d//%d %d//%d\n", f[i] + 1, f[i] + 1, f[i + 1] + 1, f[i + 1] + 1, f[i + 2] + 1,
                           f[i + 2] + 1) > 0;
    return ok;
}

// writeSunflow's scene file: the 100-unit box scaling, positions clamped
// 0.001 inside it, y and z swapped
bool write_sunflow_file(const char *path, const std::vector<double> &v, const std::vector<double> &n,
                        const std::vector<int32_t> &f, float wall_thick)
{
    OutFile o;
    o.f = fopen(path, "w");
    if (!o.f) return false;
    const double s = 100.0 * 1.0 / (1.0 - 2.0 * wall_thick);
    const float h = 0.001f;
    bool ok = fprintf(o.f, "include template.sc\n\n") > 0;
    ok = ok && fprintf(o.f, "object { \n shader Liquid \n type mesh \n name polySurfac \n %d %d \n", (int)(v.size() / 3),
                       (int)(f.size() / 3)) > 0;
    for (size_t i = 0; i < v.size(); i += 3)
        ok = ok && fprintf(o.f, "v %lf %lf %lf %lf %lf %lf 0 0\n", s * fmin(1.0 - h, fmax(h, (v[i] - wall_thick))),
                           s * fmin(1.0 - h, fmax(h, (v[i + 2] - wall_thick))),
                           s * fmin(1.0 - h, fmax(h, (v[i + 1] - wall_thick))), n[i], n[i + 2], n[i + 1]) > 0;
    for (size_t i = 0; i < f.size(); i += 3) ok = ok && fprintf(o.f, "t %d %d %d\n", f[i], f[i + 1], f[i + 2]) > 0;
    return ok && fprintf(o.f, "}\n") > 0;
}

bool write_sunflow_template(const char *path)
{
    OutFile o;
    o.f = fopen(path, "w");
    if (!o.f) return false;
    return fwrite(SUNFLOW_TEMPLATE, 1, strlen(SUNFLOW_TEMPLATE), o.f) == strlen(SUNFLOW_TEMPLATE);
}
