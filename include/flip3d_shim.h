/*
 * flip3d_shim.h - the reference's C++ entry points, served by the GPU library.
 *
 * Declares the namespace functions of mmmovania/flip3d that the per-step path
 * is reached through, with the reference's own names, argument meanings and
 * types, so a build of the reference can link libflip3d_shim.so in place of
 * src/flip3D.cpp's step, src/mapper.cpp and src/solver.cpp:
 *   flip3D::init / simulateStep        src/flip3D.h:9-12
 *   mapper::mapP2G / mapG2P / fetchVelocity   src/mapper.h:11-15
 *   solver::setSubcell / solve         src/solver.h:9-12
 *   corrector::correct                 src/corrector.h:10-13
 * The reference reports no errors (its functions return void); the shim
 * throws std::runtime_error carrying flip_gpu_last_error() instead of
 * continuing on a failed GPU call.
 *
 * Types: the particle record and FLOAT match src/common.h:10,34-45 byte for
 * byte (72-byte record, offsets checked below), so vectors of the reference's
 * own particles pass straight through.  Grids are the reference's jagged
 * alloc3D<T>(w,h,d) arrays ([i][j][k], k contiguous, src/utility.h:40-51).
 * Define FLIP3D_SHIM_REFERENCE_TYPES and include the reference's common.h
 * first to use its definitions instead of these.
 */
#ifndef FLIP3D_SHIM_H
#define FLIP3D_SHIM_H

#include <cstddef>
#include <vector>

#include "flip3d_gpu.h"

#ifndef FLIP3D_SHIM_REFERENCE_TYPES
typedef float FLOAT;

typedef struct _particle {
    FLOAT p[3];
    FLOAT u[3];
    FLOAT n[3];
    char type;
    char visible;
    char remove;
    char thinparticle;
    FLOAT tmp[2][3];
    FLOAT m;
    FLOAT dens;
} particle;
#endif

static_assert(sizeof(particle) == 72, "particle record must match the reference's 72-byte layout");
static_assert(offsetof(particle, type) == 36 && offsetof(particle, tmp) == 40 && offsetof(particle, m) == 64 &&
                  offsetof(particle, dens) == 68,
              "particle field offsets must match the reference");

class sorter;  // the reference's bucket sorter; the GPU path bins particles itself

namespace flip3D {
// Builds the reference dam-break scene at the shim's grid size (or loads
// render/save/<load_step>_scene.dat when load_step > 0) on the GPU.
void init(int load_step = 0);
// One simulateStep on the GPU-resident state.
void simulateStep();
}  // namespace flip3D

namespace mapper {
void mapP2G(sorter *sort, std::vector<particle *> &particles, FLOAT ****grid, int gn);
void mapG2P(std::vector<particle *> &particles, FLOAT ****grid, int gn);
void fetchVelocity(FLOAT p[3], FLOAT u[3], FLOAT ****grid, int gn);
}  // namespace mapper

namespace solver {
void setSubcell(char value);
void solve(char ***A, FLOAT ***L, FLOAT ***x, FLOAT ***b, int n);
}  // namespace solver

namespace corrector {
void correct(sorter *sort, std::vector<particle *> &particle, FLOAT dt, FLOAT re);
}  // namespace corrector

// Shim-only helpers (no reference counterpart).
namespace flip3d_shim {
// Grid size used by flip3D::init (default 32, the reference's #define N).
void set_grid_size(int n);
// Scene flip3D::init(0) builds: FLIP_SCENE_REFERENCE (damBreakTest, the
// default, as placeObjects has it), FLIP_SCENE_SPHERE_POUR, _CLIFF_POUR or
// _WATER_DROP (the scene functions placeObjects leaves commented out).
void set_scene(int scene);
// Per-step output files as simulateStep writes them (off by default): the
// preview render/preview/frame_<step>.bmp (write_frame, flip3D.cpp:1133) and
// the mesh files render/obj/<step>_scene.obj, render/sunflow/<step>_scene.sc
// and template.sc (write_mesh, flip3D.cpp:999, WRITE_FILE).
void set_outputs(bool preview, bool mesh);
// The context flip3D::init created, for callers that render or export.
flip_ctx *context();
// Copies the GPU-resident particles out in reference record layout.
void download(std::vector<particle> &out);
// Releases every context the shim holds.
void shutdown();
}  // namespace flip3d_shim

#endif
