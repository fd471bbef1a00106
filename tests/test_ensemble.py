"""flip_gpu_step_many and flip_gpu_step_batch: argument checks on the CPU;
on the GPU, an ensemble stepped concurrently (host threads, or one graph
launch per part of the step) equals the same scenes stepped one by one, bit
for bit (each context is independent and deterministic)."""
import ctypes
import os

import numpy as np
import pytest

import flip3d_amd as fa


def test_step_many_rejects_bad_arguments():
    lib = fa.lib()
    assert lib.flip_gpu_step_many(None, 0, 4, None, None) == -1
    assert b"no contexts" in lib.flip_gpu_last_error()
    arr = (ctypes.c_void_p * 2)(None, None)
    assert lib.flip_gpu_step_many(arr, 2, 4, None, None) == -1
    assert b"null context" in lib.flip_gpu_last_error()


def test_step_batch_rejects_bad_arguments():
    lib = fa.lib()
    assert lib.flip_gpu_step_batch(None, 0, 4, None, None) == -1
    assert b"no contexts" in lib.flip_gpu_last_error()
    arr = (ctypes.c_void_p * 2)(None, None)
    assert lib.flip_gpu_step_batch(arr, 2, 4, None, None) == -1
    assert b"null context" in lib.flip_gpu_last_error()
    assert lib.flip_gpu_step_batch(arr, 2, 0, None, None) == -1
    assert b"threads" in lib.flip_gpu_last_error()


def _scenes(sizes, precond=fa.PRECOND_REFERENCE):
    out = []
    for i, n in enumerate(sizes):
        c = fa.FlipContext(n, tol_policy=1, tol=1e-5, dt_policy=1, precond=precond)
        if precond == fa.PRECOND_LEARNED:
            c.set_learned_weights(os.path.join(os.path.dirname(fa.__file__), "weights", "unet_coarse.bin"))
        c.build_scene(fa.SCENE_RANDOM_BOXES, i)
        c.srand(1)
        out.append(c)
    return out


@pytest.mark.gpu
def test_step_many_equals_single_steps():
    sizes = [24, 31, 40, 48, 57, 64]
    a, b = _scenes(sizes), _scenes(sizes)
    try:
        for _ in range(3):
            st, ms = fa.step_many(a, threads=4)
            assert len(st) == len(a) and all(m > 0 for m in ms)
            for c in b:
                c.step()
        for ca, cb in zip(a, b):
            pa, pb = ca.download(), cb.download()
            for k in ("p", "u", "dens", "type"):
                assert np.array_equal(pa[k], pb[k]), k
            assert np.array_equal(ca.grid(fa.GRID_PRESS), cb.grid(fa.GRID_PRESS))
    finally:
        for c in a + b:
            c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precond", [fa.PRECOND_REFERENCE, fa.PRECOND_LEARNED])
def test_step_batch_equals_single_steps(precond):
    """Four batch steps (the batch's graphs built on the first, replayed
    after) against one-by-one steps with step graphs off: particles and
    pressure bit for bit, the same iteration and reposition counts."""
    sizes = [24, 31, 40, 48, 57, 64]
    a, b = _scenes(sizes, precond), _scenes(sizes, precond)
    for c in b:
        c.set_step_graphs(False)
    try:
        for k in range(4):
            st, ms = fa.step_batch(a, threads=4)
            assert len(st) == len(a) and all(m > 0 for m in ms)
            for i, c in enumerate(b):
                sb = c.step()
                assert st[i]["iterations"] == sb["iterations"], (k, i)
                assert st[i]["n_reposition"] == sb["n_reposition"], (k, i)
        for ca, cb in zip(a, b):
            pa, pb = ca.download(), cb.download()
            for k in ("p", "u", "dens", "type"):
                assert np.array_equal(pa[k], pb[k]), k
            assert np.array_equal(ca.grid(fa.GRID_PRESS), cb.grid(fa.GRID_PRESS))
    finally:
        for c in a + b:
            c.close()


@pytest.mark.gpu
def test_random_boxes_scene_has_liquid():
    for seed in range(4):
        with fa.FlipContext(32) as c:
            assert c.build_scene(fa.SCENE_RANDOM_BOXES, seed) > 0
