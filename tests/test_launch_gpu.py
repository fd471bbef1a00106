"""flip_gpu_set_launch: the solver kernels' launch shape (block count, PCG
iterations per captured batch) on the GPU."""
import numpy as np
import pytest

import flip3d_amd as fa

pytestmark = pytest.mark.gpu

CONFIG = dict(dt_policy=1, tol_policy=1, tol=1e-5)


def _steps(n, launch, steps=3, precond=fa.PRECOND_REFERENCE):
    with fa.FlipContext(n, precond=precond, **CONFIG) as g:
        g.set_launch(*launch)
        g.build_scene(fa.SCENE_BENCHMARK)
        g.srand(1)
        sts = [g.step() for _ in range(steps)]
        return sts, g.download(), g.grid(fa.GRID_PRESS)


def test_batch_size_changes_nothing_but_the_launch_pattern():
    """Iterations per captured batch: 4 and 16 give the built-in 8's
    particles and pressure bit for bit (the iterations are the same kernels
    with the same arguments, only grouped into other graph replays)."""
    ref, pr, gr = _steps(32, (0, 0))
    for batch in (4, 16):
        sts, p, g = _steps(32, (0, batch))
        assert [s["iterations"] for s in sts] == [s["iterations"] for s in ref]
        for k in ("p", "u", "dens", "type"):
            assert np.array_equal(p[k], pr[k]), (batch, k)
        assert np.array_equal(g, gr)


@pytest.mark.parametrize("precond", [fa.PRECOND_REFERENCE, fa.PRECOND_LEARNED])
def test_block_cap_meets_the_tolerance(precond):
    """A capped block count (16 of the built-in 108 at N=48) regroups the dot
    products' partials: every solve still meets the tolerance, and step 1's
    flags and iteration count are those of the built-in shape."""
    if precond == fa.PRECOND_LEARNED:
        import os
        w = os.path.join(os.path.dirname(fa.__file__), "weights", "unet_coarse.bin")

        def run(launch):
            with fa.FlipContext(48, precond=precond, **CONFIG) as g:
                g.set_learned_weights(w)
                g.set_launch(*launch)
                g.build_scene(fa.SCENE_BENCHMARK)
                g.srand(1)
                sts = [g.step() for _ in range(3)]
                return sts, g.download(), g.grid(fa.GRID_PRESS)
    else:
        def run(launch):
            return _steps(48, launch, precond=precond)
    ref, pr, _ = run((0, 0))
    sts, p, _ = run((16, 8))
    for s in sts:
        assert s["error2"] <= CONFIG["tol"] ** 2 * s["bnorm2"] * 1.0001, s
    assert sts[0]["n_fluid_cells"] == ref[0]["n_fluid_cells"]
    assert abs(sts[0]["iterations"] - ref[0]["iterations"]) <= 2


def test_launch_shape_rejects_out_of_range_values():
    with fa.FlipContext(16) as g:
        for bad in ((257, 0), (-1, 0), (0, 65), (0, -2)):
            with pytest.raises(fa.FlipError):
                g.set_launch(*bad)
        g.set_launch(256, 64)
        g.set_launch(0, 0)
