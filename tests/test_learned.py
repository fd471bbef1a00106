"""The learned preconditioner (flexible CG with a multigrid V-cycle whose
coarsest solve is the U-Net) on the GPU, with the committed weights.

What it must show:
  * the learned solve meets the same relative tolerance as plain PCG (the
    reference PCG re-checks the true residual at the end);
  * its pressure agrees with the plain solve's on the same system;
  * the trained network is what makes it fast: the same V-cycle with a
    zero-output network (the method without the network) needs clearly more
    flexible-CG iterations.
"""
import os

import numpy as np
import pytest

import flip3d_amd as fa

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "flip3d_amd", "weights",
                       "unet_coarse.bin")
TOL = 1e-5

pytestmark = pytest.mark.gpu


def _steps(n, precond, weights, steps):
    out = []
    with fa.FlipContext(n, tol_policy=1, tol=TOL, dt_policy=1, precond=precond) as g:
        if weights:
            g.set_learned_weights(weights)
        g.build_scene(fa.SCENE_BENCHMARK)
        g.srand(1)
        for _ in range(steps):
            st = g.step()
            out.append((st, g.grid(fa.GRID_PRESS)))
    return out


def _zero_weights(path):
    with fa.UNet(32, capacity=1, seed=1) as net:
        net.set_params(np.zeros(net.num_params, np.float32))
        net.save(path)


def test_learned_meets_tolerance_and_matches_plain():
    plain = _steps(64, fa.PRECOND_REFERENCE, None, 1)
    learned = _steps(64, fa.PRECOND_LEARNED, WEIGHTS, 1)
    (sp, pp), (sl, pl) = plain[0], learned[0]
    assert sl["learned_iterations"] > 0
    assert sl["error2"] <= TOL * TOL * sl["bnorm2"] * 1.0001, sl
    fl = pp != 0
    diff = np.abs(pl[fl] - pp[fl]).max() / np.abs(pp[fl]).max()
    assert diff <= 1e-3, diff  # both within 1e-5 relative residual of the same system


def test_trained_network_beats_zero_network(tmp_path):
    zero = str(tmp_path / "zero.bin")
    _zero_weights(zero)
    trained = _steps(64, fa.PRECOND_LEARNED, WEIGHTS, 3)
    ablated = _steps(64, fa.PRECOND_LEARNED, zero, 3)
    it_t = sum(s["iterations"] for s, _ in trained)  # flexible-CG plus any plain-PCG iterations
    it_z = sum(s["iterations"] for s, _ in ablated)
    for s, _ in trained + ablated:
        assert s["error2"] <= TOL * TOL * s["bnorm2"] * 1.0001, s
    assert it_t <= 0.7 * it_z, (it_t, it_z)


def test_learned_at_reference_stop_rule():
    """With the reference's absolute stop test the learned path still ends
    on the reference criterion (sum r^2 <= 1e-2 n^3)."""
    with fa.FlipContext(64, precond=fa.PRECOND_LEARNED) as g:
        g.set_learned_weights(WEIGHTS)
        g.build_scene(fa.SCENE_BENCHMARK)
        g.srand(1)
        for _ in range(2):
            st = g.step()
            assert np.float32(st["error2"]) <= 1e-2 * 64 ** 3, st
