"""The reference's other test scenes and the pourWater emitter on the GPU:
placement against the oracle bit for bit (wall obstacles, interior-wall
normals, calibrated max_dens), and free-running pour trajectories against
the oracle (which tests/test_oracle_golden.py pins to the reference's own
digests): particle counts and flags exact, fields within 1e-5 relative."""
import numpy as np
import pytest

import flip3d_amd as fa
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5
SCENES = [(fa.SCENE_SPHERE_POUR, 1), (fa.SCENE_CLIFF_POUR, 2), (fa.SCENE_WATER_DROP, 3)]


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)) if a.size else 0.0


def _pair(scene, which, n=32):
    g = fa.FlipContext(n)
    s = po.Sim(n)
    g.build_scene(scene)
    s.scene_test(which)
    g.srand(1)
    s.srand(1)
    return g, s


@pytest.mark.parametrize("scene,which", SCENES)
def test_scene_placement(scene, which):
    g, s = _pair(scene, which)
    try:
        a, b = g.download(), s.particles()
        assert np.array_equal(a["type"], b["type"])
        assert np.array_equal(a["p"], b["p"])
        assert np.array_equal(a["n"], b["n"])
        assert np.float32(g.max_dens) == np.float32(s.max_dens)
        assert g.step_count == 0
    finally:
        g.close()
        s.close()


@pytest.mark.parametrize("scene,which", SCENES[:2])
def test_pour_trajectory(scene, which):
    steps = 12
    g, s = _pair(scene, which)
    try:
        for k in range(1, steps + 1):
            gs = g.step()
            ss = s.step()
            assert g.step_count == s.step_count == k
            a, b = g.download(), s.particles()
            assert len(a["type"]) == len(b["type"]), f"particle count differs at step {k}"
            assert np.array_equal(a["type"], b["type"])
            fl = b["type"] == po.FLUID
            assert np.array_equal(g.grid(fa.GRID_A), s.grid(po.G_A)), f"flags differ at step {k}"
            ep, eu = rel_err(a["p"], b["p"]), rel_err(a["u"][fl], b["u"][fl])
            epr = rel_err(g.grid(fa.GRID_PRESS), s.grid(po.G_PRESS))
            assert max(ep, eu, epr) <= REL_TOL, (k, ep, eu, epr, gs, ss)
            assert gs["n_reposition"] == ss["n_reposition"] and gs["n_jitter"] == ss["n_jitter"]
    finally:
        g.close()
        s.close()


def test_emitter_grows_storage_and_stops(tmp_path):
    """A disc larger than the initial storage: the particle buffers grow on
    the first pour, pouring stops after pour_time steps, and the emitter
    travels with a checkpoint (the step number is the caller's, as the
    reference takes it from the file name)."""
    path = str(tmp_path / "3_scene.dat")
    with fa.FlipContext(32) as g:
        g.build_scene(fa.SCENE_WATER_DROP)
        n0 = g.num_particles()
        g.set_emitter(0.5, 0.5, 0.3, 4)
        counts = []
        for _ in range(3):
            g.step()
            counts.append(g.num_particles())
        per = counts[0] - n0
        assert per > 1000 and counts[1] - counts[0] == per and counts[2] - counts[1] == per
        q = g.download()
        assert np.all(q["type"][n0:] == fa.FLUID)
        g.save_state(path)
        with fa.FlipContext(32) as h:
            h.load_state(path)
            h.step_count = 3
            h.step()  # step 4 <= pour_time: pours
            assert h.num_particles() == counts[2] + per
            h.step()  # step 5: done
            assert h.num_particles() == counts[2] + per
