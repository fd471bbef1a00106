"""flip3d_amd - MI355X (gfx950) FLIP step library for the flip3d hot path.

Python binding over the C ABI in include/flip3d_gpu.h (libflip3d_amd.so, built
in-tree by flip3d_amd/build.py).  The binding does no computation of its own:
every operation runs in the HIP library, and importing fails loudly when the
library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflip3d_amd.so")

AIR, FLUID, WALL = 0, 1, 2
GRID_U0, GRID_U1, GRID_U2, GRID_US0, GRID_US1, GRID_US2, GRID_A, GRID_L, GRID_PRESS, GRID_DIV = range(10)
PRECOND_REFERENCE, PRECOND_IC0, PRECOND_LEARNED = 0, 1, 2
(STAGE_DENSITY, STAGE_EXT_FORCE, STAGE_P2G, STAGE_PROJECT, STAGE_EXTRAP, STAGE_G2P, STAGE_ADVECT,
 STAGE_CORRECT) = range(8)
SCENE_REFERENCE, SCENE_BENCHMARK, SCENE_RANDOM, SCENE_RANDOM_BOXES = 0, 1, 2, 3
SCENE_SPHERE_POUR, SCENE_CLIFF_POUR, SCENE_WATER_DROP = 4, 5, 6
BENCH_SEED = 20261203

EXPORTS = [
    "flip_gpu_last_error", "flip_gpu_default_params", "flip_gpu_create", "flip_gpu_destroy",
    "flip_gpu_synchronize", "flip_gpu_upload_particles", "flip_gpu_download_particles",
    "flip_gpu_num_particles", "flip_gpu_set_max_dens", "flip_gpu_get_max_dens", "flip_gpu_srand",
    "flip_gpu_set_dt", "flip_gpu_get_grid", "flip_gpu_set_grid", "flip_gpu_step", "flip_gpu_stage",
    "flip_gpu_map_p2g", "flip_gpu_map_g2p", "flip_gpu_solve", "flip_gpu_sorted_cells",
    "flip_gpu_sort_cells_noalloc",
    "flip_gpu_kernel_times", "flip_gpu_set_profiling", "flip_gpu_build_scene", "flip_gpu_selftest_math",
    "flip_gpu_save_state", "flip_gpu_load_state", "flip_gpu_step_many", "flip_gpu_step_batch",
    "flip_gpu_set_emitter", "flip_gpu_set_step", "flip_gpu_get_step", "flip_gpu_mesh", "flip_gpu_mesh_get",
    "flip_gpu_render_preview", "flip_gpu_write_preview", "flip_gpu_write_obj", "flip_gpu_write_sunflow",
    "flip_gpu_step_begin",
    "flip_gpu_set_learned_weights", "flip_gpu_set_launch", "flip_gpu_set_step_graphs", "flip_unet_create", "flip_unet_destroy", "flip_unet_num_params", "flip_unet_get_params",
    "flip_unet_set_params", "flip_unet_forward_features", "flip_unet_backward", "flip_unet_train_step",
    "flip_unet_predict", "flip_unet_predict_fused", "flip_unet_solve_system", "flip_unet_save", "flip_unet_load",
]


class FlipError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("dt", C.c_double), ("dt_policy", C.c_int), ("cfl", C.c_double),
        ("alpha", C.c_double), ("gravity", C.c_double), ("density", C.c_double),
        ("subcell", C.c_int), ("tol_policy", C.c_int), ("tol", C.c_double), ("precond", C.c_int),
        ("max_iters", C.c_int), ("correction", C.c_int), ("reposition", C.c_int),
        ("rand_seed", C.c_uint), ("refine", C.c_int), ("refine_tol", C.c_double),
        ("solve_fp64", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("error2", C.c_double), ("bnorm2", C.c_double),
        ("error2_0", C.c_double), ("dt", C.c_double), ("n_fluid_cells", C.c_int),
        ("n_reposition", C.c_int), ("n_jitter", C.c_int), ("n_precond_unit", C.c_int),
        ("refine_sweeps", C.c_int), ("ms_step", C.c_double), ("ms_solve", C.c_double),
        ("learned_iterations", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FlipError(f"{LIB_PATH} is missing: run flip3d_amd/build.py (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.flip_gpu_last_error.restype = C.c_char_p
        L.flip_gpu_default_params.argtypes = [C.POINTER(Params), C.c_int]
        L.flip_gpu_create.argtypes = [C.POINTER(Params), C.POINTER(P)]
        L.flip_gpu_destroy.argtypes = [P]
        L.flip_gpu_synchronize.argtypes = [P]
        L.flip_gpu_upload_particles.argtypes = [P, C.c_int64, P, P, P, P, P, P]
        L.flip_gpu_download_particles.argtypes = [P, P, P, P, P, P]
        L.flip_gpu_num_particles.argtypes = [P, C.POINTER(C.c_int64)]
        L.flip_gpu_set_emitter.argtypes = [P, C.c_float, C.c_float, C.c_float, C.c_int]
        L.flip_gpu_set_step.argtypes = [P, C.c_int]
        L.flip_gpu_mesh.argtypes = [P, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.flip_gpu_mesh_get.argtypes = [P, P, P, P]
        L.flip_gpu_render_preview.argtypes = [P, P]
        L.flip_gpu_write_preview.argtypes = [P, C.c_char_p]
        L.flip_gpu_write_obj.argtypes = [P, C.c_char_p]
        L.flip_gpu_write_sunflow.argtypes = [P, C.c_char_p, C.c_char_p]
        L.flip_gpu_step_begin.argtypes = [P]
        L.flip_gpu_get_step.argtypes = [P, C.POINTER(C.c_int)]
        L.flip_gpu_set_max_dens.argtypes = [P, C.c_float]
        L.flip_gpu_get_max_dens.argtypes = [P, C.POINTER(C.c_float)]
        L.flip_gpu_srand.argtypes = [P, C.c_uint]
        L.flip_gpu_set_dt.argtypes = [P, C.c_double]
        L.flip_gpu_get_grid.argtypes = [P, C.c_int, P]
        L.flip_gpu_set_grid.argtypes = [P, C.c_int, P]
        L.flip_gpu_step.argtypes = [P, C.POINTER(Stats)]
        L.flip_gpu_stage.argtypes = [P, C.c_int, C.POINTER(Stats)]
        L.flip_gpu_step_many.argtypes = [C.POINTER(P), C.c_int, C.c_int, C.POINTER(Stats), C.POINTER(C.c_double)]
        L.flip_gpu_step_batch.argtypes = [C.POINTER(P), C.c_int, C.c_int, C.POINTER(Stats), C.POINTER(C.c_double)]
        L.flip_gpu_map_p2g.argtypes = [P, P, P, P]
        L.flip_gpu_map_g2p.argtypes = [P, P, P, P, C.c_int64, P, P]
        L.flip_gpu_solve.argtypes = [P, P, P, P, P, C.c_int, C.POINTER(Stats)]
        L.flip_gpu_sorted_cells.argtypes = [P, P, P]
        L.flip_gpu_sort_cells_noalloc.argtypes = [P, C.c_int64, C.c_int, P, P, P]
        L.flip_gpu_kernel_times.argtypes = [P, C.c_char_p, C.c_int, P, C.c_int]
        L.flip_gpu_set_profiling.argtypes = [P, C.c_int]
        L.flip_gpu_build_scene.argtypes = [P, C.c_int, C.c_uint64, C.POINTER(C.c_int64)]
        L.flip_gpu_selftest_math.argtypes = [C.c_int64, P, P, P]
        L.flip_gpu_save_state.argtypes = [P, C.c_char_p]
        L.flip_gpu_load_state.argtypes = [P, C.c_char_p]
        L.flip_gpu_set_learned_weights.argtypes = [P, C.c_char_p]
        L.flip_gpu_set_launch.argtypes = [P, C.c_int, C.c_int]
        L.flip_gpu_set_step_graphs.argtypes = [P, C.c_int]
        L.flip_unet_create.argtypes = [C.c_int, C.c_int, C.c_uint64, C.POINTER(P)]
        L.flip_unet_destroy.argtypes = [P]
        L.flip_unet_num_params.argtypes = [P]
        L.flip_unet_num_params.restype = C.c_int64
        L.flip_unet_get_params.argtypes = [P, P]
        L.flip_unet_set_params.argtypes = [P, P]
        L.flip_unet_forward_features.argtypes = [P, C.c_int, P, P, P]
        L.flip_unet_backward.argtypes = [P, C.c_int, P, P]
        L.flip_unet_train_step.argtypes = [P, C.c_int, P, P, P, P, C.c_float, C.POINTER(C.c_double)]
        L.flip_unet_predict.argtypes = [P, C.c_int, P, P, P, P, P]
        L.flip_unet_predict_fused.argtypes = [P, P, P, P, P]
        L.flip_unet_solve_system.argtypes = [P, C.c_int, P, P, P, P]
        L.flip_unet_save.argtypes = [P, C.c_char_p]
        L.flip_unet_load.argtypes = [P, C.c_char_p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc < 0:
        raise FlipError(f"flip_gpu error {rc}: {lib().flip_gpu_last_error().decode()}")
    return rc


def default_params(n, **kw):
    p = Params()
    _check(lib().flip_gpu_default_params(C.byref(p), n))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


def selftest_math(x, y):
    """Device rounding self-check: (x/y, sqrtf(x), the glibc hypotf form,
    (float)(x/y in double)) per pair, and the device's sharp kernel of
    r2 = x with h*h = y (for x < y)."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    out = np.empty(5 * len(x), np.float32)
    _check(lib().flip_gpu_selftest_math(len(x), _ptr(x), _ptr(y), _ptr(out)))
    return out[:4 * len(x)].reshape(len(x), 4), out[4 * len(x):]


class FlipContext:
    """One device-resident simulation (a flip_ctx)."""

    def __init__(self, n, **kw):
        self.n = n
        self.params = default_params(n, **kw)
        h = C.c_void_p()
        _check(lib().flip_gpu_create(C.byref(self.params), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().flip_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- particles --
    def upload(self, p, u, nrm, type_, dens=None, mass=None):
        p = np.ascontiguousarray(p, np.float32).reshape(-1)
        u = None if u is None else np.ascontiguousarray(u, np.float32).reshape(-1)
        nrm = None if nrm is None else np.ascontiguousarray(nrm, np.float32).reshape(-1)
        t = np.ascontiguousarray(type_, np.int8)
        d = None if dens is None else np.ascontiguousarray(dens, np.float32)
        m = None if mass is None else np.ascontiguousarray(mass, np.float32)
        _check(lib().flip_gpu_upload_particles(self.h, len(t), _ptr(p), _ptr(u), _ptr(nrm), _ptr(t), _ptr(m),
                                               _ptr(d)))

    def num_particles(self):
        v = C.c_int64()
        _check(lib().flip_gpu_num_particles(self.h, C.byref(v)))
        return v.value

    def download(self):
        n = self.num_particles()
        p = np.empty((n, 3), np.float32)
        u = np.empty((n, 3), np.float32)
        nr = np.empty((n, 3), np.float32)
        t = np.empty(n, np.int8)
        d = np.empty(n, np.float32)
        _check(lib().flip_gpu_download_particles(self.h, _ptr(p), _ptr(u), _ptr(nr), _ptr(t), _ptr(d)))
        return {"p": p, "u": u, "n": nr, "type": t, "dens": d}

    @property
    def max_dens(self):
        v = C.c_float()
        _check(lib().flip_gpu_get_max_dens(self.h, C.byref(v)))
        return v.value

    @max_dens.setter
    def max_dens(self, v):
        _check(lib().flip_gpu_set_max_dens(self.h, v))

    def srand(self, seed):
        _check(lib().flip_gpu_srand(self.h, seed))

    def set_emitter(self, x, z, r, pour_time):
        """pourWater disc (flip3D.cpp:375-402); pour_time < 0 switches it off."""
        _check(lib().flip_gpu_set_emitter(self.h, x, z, r, pour_time))

    def mesh(self, mg=0):
        """mesher::generateMesh through the last sort's buckets (mg <= 0:
        N + 1): vertices and normals float64 (nv, 3), faces int32 (nf, 3)."""
        nv, nf = C.c_int64(), C.c_int64()
        _check(lib().flip_gpu_mesh(self.h, mg, C.byref(nv), C.byref(nf)))
        v = np.empty((nv.value, 3), np.float64)
        nr = np.empty((nv.value, 3), np.float64)
        f = np.empty((nf.value, 3), np.int32)
        _check(lib().flip_gpu_mesh_get(self.h, _ptr(v), _ptr(nr), _ptr(f)))
        return v, nr, f

    def render_preview(self):
        """write_frame's preview image: uint8 (700, 1024, 4) RGBA."""
        img = np.empty((700, 1024, 4), np.uint8)
        _check(lib().flip_gpu_render_preview(self.h, _ptr(img)))
        return img

    def write_preview(self, path):
        _check(lib().flip_gpu_write_preview(self.h, os.fsencode(path)))

    def write_obj(self, path):
        _check(lib().flip_gpu_write_obj(self.h, os.fsencode(path)))

    def write_sunflow(self, path, template_path=None):
        _check(lib().flip_gpu_write_sunflow(self.h, os.fsencode(path),
                                            None if template_path is None else os.fsencode(template_path)))

    def step_begin(self):
        _check(lib().flip_gpu_step_begin(self.h))

    @property
    def step_count(self):
        v = C.c_int()
        _check(lib().flip_gpu_get_step(self.h, C.byref(v)))
        return v.value

    @step_count.setter
    def step_count(self, v):
        _check(lib().flip_gpu_set_step(self.h, v))

    def set_dt(self, dt):
        _check(lib().flip_gpu_set_dt(self.h, dt))

    def build_scene(self, scene=SCENE_BENCHMARK, seed=BENCH_SEED):
        v = C.c_int64()
        _check(lib().flip_gpu_build_scene(self.h, scene, seed, C.byref(v)))
        return v.value

    # -- grids --
    def grid(self, which):
        n = self.n
        if which in (GRID_U0, GRID_US0):
            shape, dt = (n + 1, n, n), np.float32
        elif which in (GRID_U1, GRID_US1):
            shape, dt = (n, n + 1, n), np.float32
        elif which in (GRID_U2, GRID_US2):
            shape, dt = (n, n, n + 1), np.float32
        else:
            shape, dt = (n, n, n), (np.int8 if which == GRID_A else np.float32)
        out = np.empty(shape, dt)
        _check(lib().flip_gpu_get_grid(self.h, which, _ptr(out)))
        return out

    def set_grid(self, which, arr):
        dt = np.int8 if which == GRID_A else np.float32
        a = np.ascontiguousarray(arr, dt)
        _check(lib().flip_gpu_set_grid(self.h, which, _ptr(a)))

    # -- work --
    def step(self):
        st = Stats()
        _check(lib().flip_gpu_step(self.h, C.byref(st)))
        return st.as_dict()

    def stage(self, which):
        st = Stats()
        _check(lib().flip_gpu_stage(self.h, which, C.byref(st)))
        return st.as_dict()

    def map_p2g(self):
        n = self.n
        g = [np.empty((n + 1) * n * n, np.float32) for _ in range(3)]
        _check(lib().flip_gpu_map_p2g(self.h, _ptr(g[0]), _ptr(g[1]), _ptr(g[2])))
        return g[0].reshape(n + 1, n, n), g[1].reshape(n, n + 1, n), g[2].reshape(n, n, n + 1)

    def map_g2p(self, g0, g1, g2, pts):
        g = [np.ascontiguousarray(a, np.float32).reshape(-1) for a in (g0, g1, g2)]
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        out = np.empty_like(pts)
        _check(lib().flip_gpu_map_g2p(self.h, _ptr(g[0]), _ptr(g[1]), _ptr(g[2]), len(pts), _ptr(pts), _ptr(out)))
        return out

    def solve(self, A, L, x, b, subcell=1):
        A = np.ascontiguousarray(A, np.int8)
        L = np.ascontiguousarray(L, np.float32)
        x = np.array(x, np.float32)
        b = np.array(b, np.float32)
        st = Stats()
        _check(lib().flip_gpu_solve(self.h, _ptr(A), _ptr(L), _ptr(x), _ptr(b), subcell, C.byref(st)))
        return x, b, st.as_dict()

    def sorted_cells(self):
        counts = np.empty(self.n ** 3, np.int32)
        ids = np.empty(self.num_particles(), np.int32)
        _check(lib().flip_gpu_sorted_cells(self.h, _ptr(counts), _ptr(ids)))
        return counts, ids

    def set_profiling(self, on=True):
        _check(lib().flip_gpu_set_profiling(self.h, 1 if on else 0))

    def kernel_times(self):
        cap = 1 << 16
        buf = C.create_string_buffer(32 * cap)
        ms = np.zeros(cap, np.float64)
        cnt = _check(lib().flip_gpu_kernel_times(self.h, buf, 32 * cap, _ptr(ms), cap))
        names = buf.value.decode().split(";") if cnt else []
        return list(zip(names, ms[:min(cnt, cap)].tolist()))

    def synchronize(self):
        _check(lib().flip_gpu_synchronize(self.h))

    def set_learned_weights(self, path):
        """Weights of the coarse-level network used by PRECOND_LEARNED."""
        _check(lib().flip_gpu_set_learned_weights(self.h, os.fsencode(path)))

    def set_launch(self, pcg_blocks=0, pcg_batch=0):
        """Launch shape of the solver's FLUID-list kernels (0 = built in):
        their block count and the PCG iterations per captured batch."""
        _check(lib().flip_gpu_set_launch(self.h, int(pcg_blocks), int(pcg_batch)))

    def set_step_graphs(self, on=True):
        """Replay the step's kernels around the pressure solve from captured
        HIP graphs (the default) or launch them one by one."""
        _check(lib().flip_gpu_set_step_graphs(self.h, 1 if on else 0))

    def save_state(self, path):
        """Checkpoint in the reference's render/save/<step>_scene.dat layout."""
        _check(lib().flip_gpu_save_state(self.h, os.fsencode(path)))

    def load_state(self, path):
        _check(lib().flip_gpu_load_state(self.h, os.fsencode(path)))


def step_many(contexts, threads=8):
    """One step of every FlipContext in `contexts`, driven concurrently by
    `threads` host threads (flip_gpu_step_many); returns (stats dicts,
    per-context step wall-clock in ms)."""
    n = len(contexts)
    arr = (C.c_void_p * n)(*[c.h for c in contexts])
    st = (Stats * n)()
    ms = (C.c_double * n)()
    _check(lib().flip_gpu_step_many(arr, n, threads, st, ms))
    return [st[i].as_dict() for i in range(n)], [ms[i] for i in range(n)]


def step_batch(contexts, threads=8):
    """One step of every FlipContext in `contexts` with the kernels around
    the pressure solves of all of them in one graph launch each
    (flip_gpu_step_batch); returns (stats dicts, per-context ms = the
    batch's wall clock)."""
    n = len(contexts)
    arr = (C.c_void_p * n)(*[c.h for c in contexts])
    st = (Stats * n)()
    ms = (C.c_double * n)()
    _check(lib().flip_gpu_step_batch(arr, n, threads, st, ms))
    return [st[i].as_dict() for i in range(n)], [ms[i] for i in range(n)]


class UNet:
    """The learned coarse-level solver (BF16 MFMA U-Net) for R^3 systems."""

    def __init__(self, R, capacity=1, seed=1):
        h = C.c_void_p()
        _check(lib().flip_unet_create(R, capacity, seed, C.byref(h)))
        self.h, self.R, self.capacity = h, R, capacity

    def close(self):
        if self.h:
            lib().flip_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def num_params(self):
        return int(lib().flip_unet_num_params(self.h))

    def params(self):
        w = np.empty(self.num_params, np.float32)
        _check(lib().flip_unet_get_params(self.h, _ptr(w)))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.num_params
        _check(lib().flip_unet_set_params(self.h, _ptr(w)))

    def forward_features(self, F, mask):
        F = np.ascontiguousarray(F, np.float32)
        mask = np.ascontiguousarray(mask, np.float32)
        S = mask.size // self.R ** 3
        out = np.empty(mask.shape, np.float32)
        _check(lib().flip_unet_forward_features(self.h, S, _ptr(F), _ptr(mask), _ptr(out)))
        return out

    def backward(self, gout):
        gout = np.ascontiguousarray(gout, np.float32)
        S = gout.size // self.R ** 3
        g = np.empty(self.num_params, np.float32)
        _check(lib().flip_unet_backward(self.h, S, _ptr(gout), _ptr(g)))
        return g

    @staticmethod
    def _systems(A, L, b, x=None):
        A = np.ascontiguousarray(A, np.int8)
        L = np.ascontiguousarray(L, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        x = None if x is None else np.ascontiguousarray(x, np.float32)
        return A, L, b, x

    def train_step(self, A, L, b, x, lr):
        A, L, b, x = self._systems(A, L, b, x)
        S = A.size // self.R ** 3
        loss = C.c_double()
        _check(lib().flip_unet_train_step(self.h, S, _ptr(A), _ptr(L), _ptr(b), _ptr(x), lr, C.byref(loss)))
        return loss.value

    def predict(self, A, L, b):
        A, L, b, _ = self._systems(A, L, b)
        S = A.size // self.R ** 3
        x = np.empty(b.shape, np.float32)
        _check(lib().flip_unet_predict(self.h, S, _ptr(A), _ptr(L), _ptr(b), _ptr(x), None))
        return x

    def predict_fused(self, A, L, b):
        """predict() of one system through the solver's fused inference
        kernels (bit-identical to predict())."""
        A, L, b, _ = self._systems(A, L, b)
        if A.size != self.R ** 3:
            raise ValueError("predict_fused takes one system")
        x = np.empty(b.shape, np.float32)
        _check(lib().flip_unet_predict_fused(self.h, _ptr(A), _ptr(L), _ptr(b), _ptr(x)))
        return x

    def solve_system(self, A, L, b, path=1):
        """The network on one system as the learned solver runs it (sigma
        from per-256-cell partials of b): path 0 through separate kernels,
        path 1 through the seven fused layer kernels."""
        A, L, b, _ = self._systems(A, L, b)
        if A.size != self.R ** 3:
            raise ValueError("solve_system takes one system")
        x = np.empty(b.shape, np.float32)
        _check(lib().flip_unet_solve_system(self.h, int(path), _ptr(A), _ptr(L), _ptr(b), _ptr(x)))
        return x

    def save(self, path):
        _check(lib().flip_unet_save(self.h, os.fsencode(path)))

    def load(self, path):
        _check(lib().flip_unet_load(self.h, os.fsencode(path)))
