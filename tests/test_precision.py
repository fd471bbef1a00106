"""FP64 reference solve and FP64-residual iterative refinement of the FP32
PCG: argument checks on the CPU; on the GPU, the FP64 solve against the
oracle's FP64 build and the refined solve against the FP64 solve.

Tolerances: the FP64 solves agree to 1e-9 relative (both stop at 1e-12 rel.
residual on systems with condition number < 1e3); the refined solve's FP64
residual meets refine_tol, and its pressure (rounded to FP32 on output)
agrees with the FP64 solve to 1e-6 relative.
"""
import ctypes
import os

import numpy as np
import pytest

import flip3d_amd as fa
from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("kw, msg", [
    (dict(refine=-1), b"refine must be >= 0"),
    (dict(refine=3, tol_policy=0), b"iterative refinement needs tol_policy 1"),
    (dict(refine=3, tol_policy=1, refine_tol=0.0), b"iterative refinement needs tol_policy 1"),
    (dict(solve_fp64=2), b"solve_fp64 must be 0 or 1"),
    (dict(solve_fp64=1, tol_policy=0), b"the FP64 solve needs tol_policy 1"),
])
def test_precision_arguments_rejected(kw, msg):
    lib = fa.lib()
    p = fa.default_params(16, **kw)
    h = ctypes.c_void_p()
    assert lib.flip_gpu_create(ctypes.byref(p), ctypes.byref(h)) == -1  # FLIP_E_ARG, before any device call
    assert msg in lib.flip_gpu_last_error()


def _system():
    d = np.load(os.path.join(GOLD, "solve.npz"), allow_pickle=False)
    return int(d["n"]), d["A"], d["L"], d["x"], d["b"]


@pytest.mark.gpu
def test_fp64_solve_matches_fp64_oracle():
    n, A, L, x0, b0 = _system()
    xo, bo, so = po.solve(n, A, L, x0.copy(), b0.copy(), f64=True, tol_policy=1, tol=1e-12)
    with fa.FlipContext(n, tol_policy=1, tol=1e-12, solve_fp64=1) as g:
        x, b, st = g.solve(A, L, x0, b0)
    assert st["error2"] <= 1e-24 * st["bnorm2"]
    assert np.array_equal(b.reshape(-1), (-b0).reshape(-1))  # b negated in place
    fl = A == po.FLUID
    assert rel_err(x[fl], xo[fl].astype(np.float32)) <= 1e-6  # FP32 output of the FP64 solution
    assert np.all(x[~fl] == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("precond", [fa.PRECOND_REFERENCE, fa.PRECOND_IC0])
def test_refinement_reaches_fp64_accuracy(precond):
    n, A, L, x0, b0 = _system()
    with fa.FlipContext(n, tol_policy=1, tol=1e-12, solve_fp64=1) as g:
        x64, _, _ = g.solve(A, L, x0, b0)
    with fa.FlipContext(n, tol_policy=1, tol=1e-3, refine=20, refine_tol=1e-9, precond=precond) as g:
        x, b, st = g.solve(A, L, x0, b0)
    assert st["refine_sweeps"] >= 2  # 1e-3 per sweep cannot reach 1e-9 in one
    assert st["error2"] <= 1e-18 * st["bnorm2"], st
    assert np.array_equal(b.reshape(-1), (-b0).reshape(-1))
    fl = A == po.FLUID
    assert rel_err(x[fl], x64[fl]) <= 1e-6


@pytest.mark.gpu
def test_refined_step_tracks_fp64_step():
    """One benchmark-scene step at N=32: pressure and face velocities of the
    refined FP32 solve against the FP64 solve, and the plain FP32 solve's
    error, which refinement must not exceed."""
    res = {}
    for name, kw in (("fp64", dict(solve_fp64=1, tol=1e-12)),
                     ("refine", dict(refine=10, refine_tol=1e-9, tol=1e-5)),
                     ("fp32", dict(tol=1e-5))):
        with fa.FlipContext(32, tol_policy=1, dt_policy=1, **kw) as g:
            g.build_scene(fa.SCENE_BENCHMARK)
            g.srand(1)
            st = g.step()
            res[name] = (g.grid(fa.GRID_PRESS), [g.grid(w) for w in (fa.GRID_U0, fa.GRID_U1, fa.GRID_U2)], st)
    p64, u64, _ = res["fp64"]
    pr, ur, sr = res["refine"]
    p32, u32, _ = res["fp32"]
    assert sr["error2"] <= 1e-18 * sr["bnorm2"]
    e_ref = rel_err(pr, p64)
    e_32 = rel_err(p32, p64)
    assert e_ref <= 1e-6 and e_ref <= e_32, (e_ref, e_32)
    for a, b in zip(ur, u64):
        assert rel_err(a, b) <= 1e-5
