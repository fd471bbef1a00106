"""ctypes bindings for the oracle (liboracle_f32/f64.so) and the reference harness.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg.  The product package never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

AIR, FLUID, WALL = 0, 1, 2
G_U0, G_U1, G_U2, G_US0, G_US1, G_US2, G_A, G_L, G_PRESS, G_DIV = range(10)


class Params(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("dt", C.c_double), ("dt_policy", C.c_int), ("cfl", C.c_double),
        ("alpha", C.c_double), ("gravity", C.c_double), ("density", C.c_double),
        ("subcell", C.c_int), ("tol_policy", C.c_int), ("tol", C.c_double),
        ("precond", C.c_int), ("max_iters", C.c_int), ("correction", C.c_int),
        ("reposition", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("error2", C.c_double), ("bnorm2", C.c_double),
        ("error2_0", C.c_double), ("dt", C.c_double), ("n_fluid_cells", C.c_int),
        ("n_reposition", C.c_int), ("n_jitter", C.c_int), ("n_precond_unit", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build(quiet=True):
    """Compile the oracle libraries (and the reference harness when the
    reference sources are present)."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if os.path.isdir("/root/reference/src"):
        out = subprocess.run(["make", "-C", HERE, "-j8", "ref"], capture_output=True, text=True)
        if out.returncode != 0:
            raise RuntimeError("reference harness build failed:\n" + out.stdout + out.stderr)


def _lib(f64):
    name = "liboracle_f64.so" if f64 else "liboracle_f32.so"
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    P = C.c_void_p
    lib.orc_create.restype = P
    lib.orc_create.argtypes = [C.POINTER(Params)]
    lib.orc_destroy.argtypes = [P]
    lib.orc_default_params.argtypes = [C.POINTER(Params), C.c_int]
    lib.orc_set_particles.argtypes = [P, C.c_int, P, P, P, P, P, P]
    lib.orc_num_particles.argtypes = [P]
    lib.orc_get_particles.argtypes = [P, P, P, P, P, P]
    rt = C.c_double if f64 else C.c_float
    lib.orc_set_max_dens.argtypes = [P, rt]
    lib.orc_get_max_dens.argtypes = [P]
    lib.orc_get_max_dens.restype = rt
    lib.orc_grid_size.argtypes = [P, C.c_int]
    lib.orc_grid_size.restype = C.c_long
    lib.orc_get_grid.argtypes = [P, C.c_int, P]
    lib.orc_set_grid.argtypes = [P, C.c_int, P]
    lib.orc_get_sorted.argtypes = [P, P, P]
    lib.orc_srand.argtypes = [P, C.c_uint]
    lib.orc_rand.argtypes = [P]
    for fn in ("orc_sort", "orc_compute_density", "orc_add_ext_force", "orc_map_p2g",
               "orc_mark_water", "orc_save_grid", "orc_enforce_boundary",
               "orc_extrapolate_velocity", "orc_subtract_grid", "orc_flip_g2p",
               "orc_compute_wall_normals"):
        getattr(lib, fn).argtypes = [P]
    for fn in ("orc_project", "orc_solve_picflip", "orc_advect", "orc_correct", "orc_step"):
        getattr(lib, fn).argtypes = [P, C.POINTER(Stats)]
    lib.orc_current_dt.argtypes = [P]
    lib.orc_current_dt.restype = C.c_double
    lib.orc_set_dt.argtypes = [P, C.c_double]
    lib.orc_solve.argtypes = [C.POINTER(Params), P, P, P, P, C.POINTER(Stats)]
    lib.orc_smooth_kernel.argtypes = [rt, rt]
    lib.orc_smooth_kernel.restype = rt
    lib.orc_sharp_kernel.argtypes = [rt, rt]
    lib.orc_sharp_kernel.restype = rt
    lib.orc_length.argtypes = [P, P]
    lib.orc_length.restype = rt
    lib.orc_length2.argtypes = [P, P]
    lib.orc_length2.restype = rt
    lib.orc_fetch_velocity.argtypes = [P, C.c_int, P, P]
    lib.orc_compute_ax.argtypes = [C.POINTER(Params), P, P, P, P]
    lib.orc_build_precond.argtypes = [C.POINTER(Params), P, P, P]
    lib.orc_apply_precond.argtypes = [C.POINTER(Params), P, P, P, P]
    lib.orc_scene_capacity.argtypes = [C.c_int]
    lib.orc_scene_capacity.restype = C.c_long
    lib.orc_calibrate_max_dens.argtypes = [C.c_int, C.c_double]
    lib.orc_calibrate_max_dens.restype = rt
    lib.orc_scene_reference.argtypes = [P]
    lib.orc_scene_benchmark.argtypes = [P, C.c_uint64]
    lib.orc_scene_test.argtypes = [P, C.c_int]
    lib.orc_set_emitter.argtypes = [P, rt, rt, rt, C.c_int]
    lib.orc_set_step.argtypes = [P, C.c_int]
    lib.orc_get_step.argtypes = [P]
    lib.orc_pour.argtypes = [P]
    lib.orc_mesh.argtypes = [P, C.c_int]
    lib.orc_mesh_get.argtypes = [P, P, P]
    return lib


_LIBS = {}


def lib(f64=False):
    if f64 not in _LIBS:
        _LIBS[f64] = _lib(f64)
    return _LIBS[f64]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_params(n, **kw):
    p = Params()
    lib().orc_default_params(C.byref(p), n)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Sim:
    """One oracle simulation (creation-order particles, flat [i][j][k] grids)."""

    def __init__(self, n, f64=False, **kw):
        self.f64 = f64
        self.L = lib(f64)
        self.dtype = np.float64 if f64 else np.float32
        self.params = default_params(n, **kw)
        self.n = n
        self.h = self.L.orc_create(C.byref(self.params))

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # particles
    def set_particles(self, p, u, nrm, type_, dens=None, m=None):
        np_ = len(type_)
        p = np.ascontiguousarray(p, self.dtype).reshape(-1)
        u = np.ascontiguousarray(u, self.dtype).reshape(-1)
        nrm = np.ascontiguousarray(nrm, self.dtype).reshape(-1)
        t = np.ascontiguousarray(type_, np.int8)
        d = None if dens is None else np.ascontiguousarray(dens, self.dtype)
        mm = None if m is None else np.ascontiguousarray(m, self.dtype)
        self.L.orc_set_particles(self.h, np_, _ptr(p), _ptr(u), _ptr(nrm), _ptr(t), _ptr(mm), _ptr(d))

    def num_particles(self):
        return self.L.orc_num_particles(self.h)

    def particles(self):
        n = self.num_particles()
        p = np.empty((n, 3), self.dtype)
        u = np.empty((n, 3), self.dtype)
        nr = np.empty((n, 3), self.dtype)
        t = np.empty(n, np.int8)
        d = np.empty(n, self.dtype)
        self.L.orc_get_particles(self.h, _ptr(p), _ptr(u), _ptr(nr), _ptr(t), _ptr(d))
        return {"p": p, "u": u, "n": nr, "type": t, "dens": d}

    @property
    def max_dens(self):
        return self.L.orc_get_max_dens(self.h)

    @max_dens.setter
    def max_dens(self, v):
        self.L.orc_set_max_dens(self.h, v)

    def grid(self, which):
        size = self.L.orc_grid_size(self.h, which)
        dt = np.int8 if which == G_A else self.dtype
        out = np.empty(size, dt)
        self.L.orc_get_grid(self.h, which, _ptr(out))
        n = self.n
        if which in (G_U0, G_US0):
            return out.reshape(n + 1, n, n)
        if which in (G_U1, G_US1):
            return out.reshape(n, n + 1, n)
        if which in (G_U2, G_US2):
            return out.reshape(n, n, n + 1)
        return out.reshape(n, n, n)

    def set_grid(self, which, arr):
        dt = np.int8 if which == G_A else self.dtype
        a = np.ascontiguousarray(arr, dt).reshape(-1)
        assert a.size == self.L.orc_grid_size(self.h, which)
        self.L.orc_set_grid(self.h, which, _ptr(a))

    def sorted_cells(self):
        C3 = self.n ** 3
        cs = np.empty(C3 + 1, np.int32)
        ids = np.empty(self.num_particles(), np.int32)
        self.L.orc_get_sorted(self.h, _ptr(cs), _ptr(ids))
        return cs, ids

    def srand(self, seed):
        self.L.orc_srand(self.h, seed)

    def rand(self):
        return self.L.orc_rand(self.h)

    def call(self, name):
        getattr(self.L, "orc_" + name)(self.h)

    def call_stats(self, name):
        st = Stats()
        getattr(self.L, "orc_" + name)(self.h, C.byref(st))
        return st.as_dict()

    def step(self):
        return self.call_stats("step")

    @property
    def dt(self):
        return self.L.orc_current_dt(self.h)

    @dt.setter
    def dt(self, v):
        self.L.orc_set_dt(self.h, v)

    def scene_reference(self):
        return self.L.orc_scene_reference(self.h)

    def scene_benchmark(self, seed=20261203):
        return self.L.orc_scene_benchmark(self.h, seed)

    def scene_test(self, which):
        """The reference's test scenes: 0 damBreak, 1 spherePour, 2 cliffPour,
        3 waterDrop (flip3D.cpp:203-297); pour scenes set the emitter."""
        return self.L.orc_scene_test(self.h, which)

    def set_emitter(self, x, z, r, pour_time):
        self.L.orc_set_emitter(self.h, x, z, r, pour_time)

    def mesh(self, mg=None):
        """mesher::generateMesh through the current buckets (mg = N + 1 as
        write_mesh uses): (vertices, normals) as float64 (nv, 3), faces as
        int32 (nf, 3)."""
        mg = self.n + 1 if mg is None else mg
        nv = self.L.orc_mesh(self.h, mg)
        nf = self.L.orc_mesh_faces()
        v = np.empty((nv, 3), np.float64)
        nr = np.empty((nv, 3), np.float64)
        f = np.empty((nf, 3), np.int32)
        self.L.orc_mesh_get(_ptr(v), _ptr(nr), _ptr(f))
        return v, nr, f

    @property
    def step_count(self):
        return self.L.orc_get_step(self.h)

    @step_count.setter
    def step_count(self, v):
        self.L.orc_set_step(self.h, v)

    def fetch_velocity(self, grid_set, pts):
        pts = np.ascontiguousarray(pts, self.dtype).reshape(-1, 3)
        out = np.empty_like(pts)
        for i in range(len(pts)):
            self.L.orc_fetch_velocity(self.h, grid_set, _ptr(pts[i]), _ptr(out[i]))
        return out


def solve(n, A, L, x, b, f64=False, **kw):
    """Stand-alone solver::solve restatement; returns (x, b, stats)."""
    dt = np.float64 if f64 else np.float32
    prm = default_params(n, **kw)
    A = np.ascontiguousarray(A, np.int8).reshape(-1)
    L = np.ascontiguousarray(L, dt).reshape(-1)
    x = np.array(x, dt).reshape(-1)
    b = np.array(b, dt).reshape(-1)
    st = Stats()
    lib(f64).orc_solve(C.byref(prm), _ptr(A), _ptr(L), _ptr(x), _ptr(b), C.byref(st))
    return x.reshape(n, n, n), b.reshape(n, n, n), st.as_dict()


def compute_ax(n, A, L, x, f64=False, **kw):
    dt = np.float64 if f64 else np.float32
    prm = default_params(n, **kw)
    A = np.ascontiguousarray(A, np.int8).reshape(-1)
    L = np.ascontiguousarray(L, dt).reshape(-1)
    x = np.ascontiguousarray(x, dt).reshape(-1)
    out = np.empty_like(x)
    lib(f64).orc_compute_ax(C.byref(prm), _ptr(A), _ptr(L), _ptr(x), _ptr(out))
    return out.reshape(n, n, n)


def build_precond(n, A, L, f64=False, **kw):
    dt = np.float64 if f64 else np.float32
    prm = default_params(n, **kw)
    A = np.ascontiguousarray(A, np.int8).reshape(-1)
    L = np.ascontiguousarray(L, dt).reshape(-1)
    P = np.empty(n ** 3, np.float64)
    lib(f64).orc_build_precond(C.byref(prm), _ptr(A), _ptr(L), _ptr(P))
    return P.reshape(n, n, n)


def apply_precond(n, A, P, r, f64=False, **kw):
    dt = np.float64 if f64 else np.float32
    prm = default_params(n, **kw)
    A = np.ascontiguousarray(A, np.int8).reshape(-1)
    P = np.ascontiguousarray(P, np.float64).reshape(-1)
    r = np.ascontiguousarray(r, dt).reshape(-1)
    z = np.zeros_like(r)
    lib(f64).orc_apply_precond(C.byref(prm), _ptr(A), _ptr(P), _ptr(r), _ptr(z))
    return z.reshape(n, n, n)


def calibrate_max_dens(n, density=0.5):
    return lib().orc_calibrate_max_dens(n, density)


def kernels(r2, h):
    L = lib()
    r2 = np.asarray(r2, np.float32)
    h = np.asarray(h, np.float32)
    sm = np.array([L.orc_smooth_kernel(float(a), float(b)) for a, b in zip(r2, h)], np.float32)
    sh = np.array([L.orc_sharp_kernel(float(a), float(b)) for a, b in zip(r2, h)], np.float32)
    return sm, sh


def lengths(a, b):
    L = lib()
    a = np.ascontiguousarray(a, np.float32).reshape(-1, 3)
    b = np.ascontiguousarray(b, np.float32).reshape(-1, 3)
    ln = np.array([L.orc_length(_ptr(a[i]), _ptr(b[i])) for i in range(len(a))], np.float32)
    l2 = np.array([L.orc_length2(_ptr(a[i]), _ptr(b[i])) for i in range(len(a))], np.float32)
    return ln, l2


# ---------------------------------------------------------------- reference
REF_PATH = os.path.join(HERE, "_ref", "libflipref.so")


def ref_available():
    return os.path.exists(REF_PATH)


def ref_lib():
    lib = C.CDLL(REF_PATH)
    P = C.c_void_p
    lib.ref_get_max_dens.restype = C.c_float
    lib.ref_set_max_dens.argtypes = [C.c_float]
    lib.ref_srand.argtypes = [C.c_uint]
    lib.ref_get_particles.argtypes = [P, P, P, P, P]
    lib.ref_set_particles.argtypes = [C.c_int, P, P, P, P, P]
    lib.ref_get_grid.argtypes = [C.c_int, P]
    lib.ref_set_press.argtypes = [P]
    lib.ref_kat_sort.argtypes = [C.c_int, C.c_int, P, P, P, P]
    lib.ref_kat_p2g.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P]
    lib.ref_kat_g2p.argtypes = [C.c_int, P, P, P, C.c_int, P, P]
    lib.ref_kat_correct.argtypes = [C.c_int, C.c_int, P, P, P, P, C.c_float, C.c_float]
    lib.ref_kat_solve.argtypes = [C.c_int, P, P, P, P, C.c_int]
    lib.ref_kat_kernels.argtypes = [C.c_int, P, P, P, P]
    lib.ref_kat_length.argtypes = [C.c_int, P, P, P, P]
    lib.ref_save_state.argtypes = [C.c_int]
    lib.ref_init_test.argtypes = [C.c_int, C.c_uint]
    lib.ref_mesh.argtypes = [C.c_int]
    lib.ref_write_outputs.argtypes = [C.c_int]
    lib.ref_mesh_get.argtypes = [P, P, P]
    lib.ref_set_step.argtypes = [C.c_int]
    return lib
