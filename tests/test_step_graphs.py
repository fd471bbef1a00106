"""Step graphs: flip_gpu_step replaying its kernels around the pressure solve
from captured HIP graphs gives the same steps as launching them one by one,
and captures again when the particle storage changes."""
import os

import numpy as np
import pytest

import flip3d_amd as fa

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(fa.__file__), "weights", "unet_coarse.bin")


def _ctx(n, precond, graphs):
    g = fa.FlipContext(n, tol_policy=1, tol=1e-5, dt_policy=1, precond=precond)
    if precond == fa.PRECOND_LEARNED:
        g.set_learned_weights(WEIGHTS)
    g.set_step_graphs(graphs)
    g.build_scene(fa.SCENE_BENCHMARK)
    g.srand(1)
    return g


def _same(a, b):
    for k in ("p", "u", "dens", "type"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("precond", [fa.PRECOND_REFERENCE, fa.PRECOND_LEARNED])
def test_step_graphs_equal_direct_launches(precond):
    """Twelve steps at N=32 with and without step graphs: particles, pressure
    and every step statistic bit for bit (replays after the first capture
    included), the reposition and jitter counts among them."""
    with _ctx(32, precond, True) as g, _ctx(32, precond, False) as d:
        moved = 0
        for k in range(12):
            sg, sd = g.step(), d.step()
            for key in ("dt", "iterations", "learned_iterations", "n_reposition", "n_jitter", "n_fluid_cells",
                        "error2"):
                assert sg[key] == sd[key], (k, key, sg[key], sd[key])
            moved += sg["n_reposition"]
            _same(g.download(), d.download())
            assert np.array_equal(g.grid(fa.GRID_PRESS), d.grid(fa.GRID_PRESS))
        print(f"repositioned over 12 steps: {moved}")


def test_step_graphs_recapture_on_new_particles():
    """A context whose particles are replaced (fewer of them, new storage)
    between steps steps exactly like a fresh context given those particles."""
    with _ctx(32, fa.PRECOND_REFERENCE, True) as g:
        for _ in range(3):
            g.step()
        q = g.download()
        keep = np.arange(len(q["type"])) % 3 != 0
        sub = {k: v[keep] for k, v in q.items()}
        g.upload(sub["p"], sub["u"], sub["n"], sub["type"], dens=sub["dens"])
        with fa.FlipContext(32, tol_policy=1, tol=1e-5, dt_policy=1) as f:
            f.build_scene(fa.SCENE_BENCHMARK)
            f.upload(sub["p"], sub["u"], sub["n"], sub["type"], dens=sub["dens"])
            f.max_dens = g.max_dens
            f.set_grid(fa.GRID_PRESS, g.grid(fa.GRID_PRESS))  # the solve's warm start
            f.set_step_graphs(False)
            g.srand(7)
            f.srand(7)
            for k in range(3):
                sg, sf = g.step(), f.step()
                assert sg["iterations"] == sf["iterations"] and sg["n_reposition"] == sf["n_reposition"], k
                _same(g.download(), f.download())
