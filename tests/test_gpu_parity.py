"""HIP path against the oracle (and the reference's golden vectors) on the GPU.

Bar: cell keys, bucket order and flags bit-exact; fields bit-exact where the
arithmetic order is reproduced, and within 1e-5 relative where PCG dot
products are summed in a different (tree) order.
"""
import os

import numpy as np
import pytest

import flip3d_amd as fa
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REL_TOL = 1e-5  # field tolerance at the benchmark's solver stop (relative, max-norm scaled)


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return float(np.abs(a - b).max() / scale) if a.size else 0.0


def test_device_rounding_matches_ieee():
    """Division, sqrtf, glibc hypotf and double division round as on the host."""
    rs = np.random.RandomState(0)
    m = 200000
    x = (rs.uniform(1e-6, 4.0, m) * 10.0 ** rs.randint(-6, 6, m)).astype(np.float32)
    y = (rs.uniform(1e-6, 4.0, m) * 10.0 ** rs.randint(-6, 6, m)).astype(np.float32)
    out, _ = fa.selftest_math(x, y)
    assert np.array_equal(out[:, 0], (x / y).astype(np.float32))
    assert np.array_equal(out[:, 1], np.sqrt(x))
    hyp = np.sqrt(x.astype(np.float64) ** 2 + y.astype(np.float64) ** 2).astype(np.float32)
    assert np.array_equal(out[:, 2], hyp)
    assert np.array_equal(out[:, 3], (x.astype(np.float64) / y.astype(np.float64)).astype(np.float32))


@pytest.mark.parametrize("scene", ["reference", "benchmark"])
def test_scene_matches_oracle(scene):
    with fa.FlipContext(32) as g:
        s = po.Sim(32)
        if scene == "reference":
            g.build_scene(fa.SCENE_REFERENCE)
            s.scene_reference()
        else:
            g.build_scene(fa.SCENE_BENCHMARK)
            s.scene_benchmark()
        a, b = g.download(), s.particles()
        assert np.array_equal(a["p"], b["p"]) and np.array_equal(a["n"], b["n"])
        assert np.array_equal(a["type"], b["type"])
        assert np.float32(g.max_dens) == np.float32(s.max_dens)


@pytest.mark.parametrize("hh", [np.float32(1.4) * np.float32(1.4), np.float32(0.5 / 128) ** 2,
                                np.float32(0.5 / 32) ** 2, np.float32(1.0)])
def test_device_sharp_kernel_matches_reference(hh):
    """kernel.cpp:8-10 sharp_kernel as the device forms it (the FP64 division
    as reciprocal, two Newton steps and one remainder correction, without the
    scaling steps) equals the reference's double formula bit for bit: 4M
    random r2 in [1e-5, hh), every float in the 20000 ulp below hh and above
    1e-5, among them r2 whose value lies within 1e-3 ulp of a float rounding
    midpoint."""
    hh = np.float32(hh)
    rs = np.random.RandomState(int(hh * 1e6) % 1000)
    lo = np.float32(1e-5)
    r2 = [rs.uniform(float(lo), float(hh), 4000000).astype(np.float32)]
    top = hh.view(np.uint32)
    r2.append((top - np.arange(1, 20001, dtype=np.uint32)).view(np.float32))
    r2.append((lo.view(np.uint32) + np.arange(0, 20000, dtype=np.uint32)).view(np.float32))
    r2 = np.concatenate(r2)
    r2 = r2[(r2 >= lo) & (r2 < hh)]
    ref = np.maximum(np.float64(hh) / np.maximum(r2.astype(np.float64), 1e-5) - 1.0, 0.0).astype(np.float32)
    # operands whose double value sits within 1e-3 ulp of a float midpoint
    mid = np.abs((np.maximum(np.float64(hh) / r2.astype(np.float64) - 1.0, 0.0) - ref.astype(np.float64))
                 / np.spacing(ref).astype(np.float64))
    assert (np.abs(mid - 0.5) < 1e-3).sum() > 0
    _, got = fa.selftest_math(r2, np.full_like(r2, hh))
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def _ctx_with(gn, p, t, u=None, nrm=None, dens=None):
    g = fa.FlipContext(gn)
    n = len(t)
    g.upload(p, np.zeros((n, 3)) if u is None else u, np.zeros((n, 3)) if nrm is None else nrm, t,
             dens=np.ones(n, np.float32) if dens is None else dens)
    return g


@pytest.mark.parametrize("gn", [8, 13])
def test_sort_golden(gn):
    d = gold("sort")
    with _ctx_with(gn, d[f"p{gn}"], d[f"t{gn}"]) as g:
        counts, ids = g.sorted_cells()
        assert np.array_equal(counts, d[f"counts{gn}"])
        assert np.array_equal(ids, d[f"ids{gn}"])


@pytest.mark.parametrize("gn", [8, 12])
def test_p2g_golden(gn):
    d = gold("p2g")
    with _ctx_with(gn, d[f"p{gn}"], d[f"t{gn}"], u=d[f"u{gn}"]) as g:
        grids = g.map_p2g()
        for c in range(3):
            assert np.array_equal(grids[c].reshape(-1).view(np.uint32), d[f"g{c}_{gn}"].view(np.uint32)), c


def test_g2p_golden():
    d = gold("g2p")
    gn = int(d["gn"])
    with fa.FlipContext(gn) as g:
        out = g.map_g2p(d["g0"], d["g1"], d["g2"], d["pts"])
        assert np.array_equal(out.view(np.uint32), d["u"].view(np.uint32))


def test_solve_golden():
    d = gold("solve")
    n = int(d["n"])
    with fa.FlipContext(n) as g:
        x, b, st = g.solve(d["A"], d["L"], d["x"], d["b"])
        assert st["iterations"] > 1
        assert np.array_equal(b.reshape(-1), d["b_out"].reshape(-1))
        assert rel_err(x, d["x_out"]) <= REL_TOL


def _pair(n=32, scene="benchmark", **kw):
    g = fa.FlipContext(n, **kw)
    oracle_keys = {name for name, _ in po.Params._fields_}
    s = po.Sim(n, **{k: v for k, v in kw.items() if k in oracle_keys})
    if scene == "benchmark":
        g.build_scene(fa.SCENE_BENCHMARK)
        s.scene_benchmark()
    else:
        g.build_scene(fa.SCENE_REFERENCE)
        s.scene_reference()
    g.srand(1)
    s.srand(1)
    return g, s


@pytest.mark.parametrize("n", [32, 24])
def test_stage_parity_first_step(n):
    """Each stage of step 1 from identical inputs (teacher forced).  N=24
    takes the paths for a kernel support h*h that is not a power of two."""
    g, s = _pair(n=n)
    try:
        g.stage(fa.STAGE_DENSITY)
        s.call("sort")
        s.call("compute_density")
        a, b = g.download(), s.particles()
        assert np.array_equal(a["dens"], b["dens"])

        g.stage(fa.STAGE_EXT_FORCE)
        s.call("add_ext_force")
        assert np.array_equal(g.download()["u"], s.particles()["u"])

        g.stage(fa.STAGE_P2G)
        for nm in ("sort", "map_p2g", "mark_water", "save_grid", "enforce_boundary"):
            s.call(nm)
        assert np.array_equal(g.grid(fa.GRID_A), s.grid(po.G_A))
        for w in (fa.GRID_U0, fa.GRID_U1, fa.GRID_U2, fa.GRID_US0, fa.GRID_US1, fa.GRID_US2):
            assert np.array_equal(g.grid(w), s.grid(w)), w

        gst = g.stage(fa.STAGE_PROJECT)
        sst = s.call_stats("project")
        s.call("enforce_boundary")
        assert np.array_equal(g.grid(fa.GRID_L), s.grid(po.G_L))
        assert gst["iterations"] == sst["iterations"]
        assert rel_err(g.grid(fa.GRID_PRESS), s.grid(po.G_PRESS)) <= REL_TOL
        for w in (fa.GRID_U0, fa.GRID_U1, fa.GRID_U2):
            assert rel_err(g.grid(w), s.grid(w)) <= REL_TOL, w
        # continue from the oracle's projected state so later stages compare exactly
        for w in (fa.GRID_U0, fa.GRID_U1, fa.GRID_U2, fa.GRID_PRESS):
            g.set_grid(w, s.grid(w))

        g.stage(fa.STAGE_EXTRAP)
        s.call("extrapolate_velocity")
        s.call("subtract_grid")
        for w in (fa.GRID_U0, fa.GRID_U1, fa.GRID_U2, fa.GRID_US0, fa.GRID_US1, fa.GRID_US2):
            assert np.array_equal(g.grid(w), s.grid(w)), w

        g.stage(fa.STAGE_G2P)
        s.call("flip_g2p")
        a, b = g.download(), s.particles()
        fl = b["type"] == po.FLUID
        assert np.array_equal(a["u"][fl], b["u"][fl])

        gst = g.stage(fa.STAGE_ADVECT)
        sst = s.call_stats("advect")
        a, b = g.download(), s.particles()
        assert gst["n_reposition"] == sst["n_reposition"]
        assert np.array_equal(a["p"], b["p"])
        assert np.array_equal(a["u"][fl], b["u"][fl])

        gst = g.stage(fa.STAGE_CORRECT)
        sst = s.call_stats("correct")
        a, b = g.download(), s.particles()
        assert gst["n_jitter"] == sst["n_jitter"]
        assert np.array_equal(a["p"], b["p"])
        assert np.array_equal(a["u"][fl], b["u"][fl])
    finally:
        g.close()
        s.close()


@pytest.mark.parametrize("n,deep", [(32, "scene"), (64, "scene"), (32, "one cell")])
def test_reposition_many_stuck(n, deep):
    """flip3D.cpp:932-968 and 405-451 with thousands of stuck particles: the
    fluid cells of the layer above the floor are flagged WALL, so every
    particle in them is stuck (about 3,300 at N=32 and 12,000 at N=64, past
    the plan kernel's LDS shuffle).  With "one cell" the flags leave a
    single deep-water cell: the deep list repeats, every stuck particle lands
    in that cell, and the resample levels form one chain as long as the
    list.  The device plan (creation order, shuffle, positions, levels) must
    give the oracle's positions and velocities bit for bit."""
    g, s = _pair(n=n)
    try:
        g.stage(fa.STAGE_DENSITY)
        s.call("sort")
        s.call("compute_density")
        g.stage(fa.STAGE_EXT_FORCE)
        s.call("add_ext_force")
        g.stage(fa.STAGE_P2G)
        for nm in ("sort", "map_p2g", "mark_water", "save_grid", "enforce_boundary"):
            s.call(nm)
        A = s.grid(po.G_A).copy()
        layer = A[:, 1, :]
        layer[layer == po.FLUID] = po.WALL
        if deep == "one cell":
            A[A == po.FLUID] = po.AIR
            c = n // 2
            A[c - 1:c + 2, c - 1:c + 2, c - 1:c + 2] = po.FLUID
        s.set_grid(po.G_A, A)
        g.set_grid(fa.GRID_A, A)
        gst = g.stage(fa.STAGE_ADVECT)
        sst = s.call_stats("advect")
        assert sst["n_reposition"] > (8192 if n == 64 else 1000), sst["n_reposition"]
        assert gst["n_reposition"] == sst["n_reposition"]
        a, b = g.download(), s.particles()
        assert np.array_equal(a["p"], b["p"])
        fl = b["type"] == po.FLUID
        assert np.array_equal(a["u"][fl], b["u"][fl])
    finally:
        g.close()
        s.close()


# step policies: the reference's (fixed dt 6e-3, absolute stop sum r^2 <= 1e-2 n^3)
# and the benchmark configuration's (CFL dt, relative residual 1e-5)
POLICIES = {"reference": {}, "config": dict(dt_policy=1, tol_policy=1, tol=1e-5)}


@pytest.mark.parametrize("scene,n,policy", [("benchmark", 32, "config"), ("benchmark", 32, "reference"),
                                            ("reference", 32, "reference"), ("reference", 32, "config"),
                                            ("benchmark", 24, "reference")])
def test_trajectory_parity(scene, n, policy):
    """Free-running steps (100 by default): flags bit-exact, the same time
    step, PCG iteration count and rand() events every step, fields within
    1e-5 relative.  Prints the error growth and the bit-identical steps."""
    steps = int(os.environ.get("FLIP_PARITY_STEPS", "100"))
    g, s = _pair(n=n, scene=scene, **POLICIES[policy])
    exact, worst, growth = 0, 0.0, {}
    try:
        for k in range(steps):
            gs = g.step()
            ss = s.step()
            a, b = g.download(), s.particles()
            fl = b["type"] == po.FLUID
            assert np.array_equal(g.grid(fa.GRID_A), s.grid(po.G_A)), f"flags differ at step {k + 1}"
            ep = rel_err(a["p"], b["p"])
            eu = rel_err(a["u"][fl], b["u"][fl])
            epr = rel_err(g.grid(fa.GRID_PRESS), s.grid(po.G_PRESS))
            assert max(ep, eu, epr) <= REL_TOL, (k + 1, ep, eu, epr, gs, ss)
            assert gs["dt"] == ss["dt"], (k + 1, gs["dt"], ss["dt"])
            assert gs["iterations"] == ss["iterations"], (k + 1, gs["iterations"], ss["iterations"])
            exact += int(np.array_equal(a["p"], b["p"]) and np.array_equal(a["u"][fl], b["u"][fl]))
            assert gs["n_reposition"] == ss["n_reposition"] and gs["n_jitter"] == ss["n_jitter"]
            worst = max(worst, ep, eu, epr)
            if k + 1 in (1, 10, 25, 50, 100, steps):
                growth[k + 1] = f"{max(ep, eu, epr):.1e}"
        print(f"{scene} N={n} {policy} policy: {exact}/{steps} steps bit-identical, "
              f"max rel error by step {growth}, worst {worst:.2e}")
    finally:
        g.close()
        s.close()


@pytest.mark.parametrize("speed,wall_speed", [(0.5, 0.0), (20.0, 0.0), (55.5, 0.0), (0.5, 100.0)])
def test_cfl_time_step_on_device(speed, wall_speed):
    """The device's CFL rule (max-speed reduction over FLUID particles):
    dt = min(6e-3, 0.6 / (N max|u|)), equal to the closed form and the
    oracle's."""
    n = 8
    rs = np.random.RandomState(int(speed * 10 + wall_speed))
    m = 5000
    p = rs.uniform(0.3, 0.7, size=(m, 3)).astype(np.float32)
    dirs = rs.normal(size=(m, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    scale = rs.uniform(0.0, 1.0, size=(m, 1))
    scale[4321] = 1.0
    u = (dirs * scale * speed).astype(np.float32)
    t = np.full(m, po.FLUID, np.int8)
    t[5] = po.WALL
    u[5] = np.float32(wall_speed) * dirs[5].astype(np.float32)
    fl = t == po.FLUID
    umax = np.sqrt((u[fl].astype(np.float64) ** 2).sum(1).max())
    want = min(6e-3, 0.6 / (n * umax))
    with fa.FlipContext(n, dt_policy=1) as g:
        g.upload(p, u, np.zeros_like(p), t, dens=np.ones(m, np.float32))
        gs = g.step()
    s = po.Sim(n, dt_policy=1)
    try:
        s.set_particles(p, u, np.zeros_like(p), t, dens=np.ones(m, np.float32))
        ss = s.step()
    finally:
        s.close()
    assert gs["dt"] == want == ss["dt"]


@pytest.mark.parametrize("tol_policy,tol", [(0, 1e-5), (1, 1e-6)])
def test_unit_factor_solve_matches_reference(tol_policy, tol):
    """A system with P == 1.0 cells (a FLUID cell with five WALL neighbours):
    the reference-form value test makes the factor and sweep couplings live,
    so the device runs the hyperplane-wavefront sweeps.  Under the
    reference's stop rule the result equals the reference's own solve
    (golden solve_unit.npz); under a relative tolerance it equals the
    oracle's, iteration for iteration."""
    d = gold("solve_unit")
    n = int(d["n"])
    kw = dict(tol_policy=tol_policy, tol=tol)
    with fa.FlipContext(n, **kw) as g:
        x, b, st = g.solve(d["A"], d["L"], d["x"], d["b"])
    assert st["n_precond_unit"] == 10, st
    xo, bo, so = po.solve(n, d["A"], d["L"], d["x"].copy(), d["b"].copy(), **kw)
    assert so["n_precond_unit"] == 10
    assert st["iterations"] == so["iterations"] > 1
    assert np.array_equal(b.reshape(-1), bo.reshape(-1))
    assert rel_err(x, xo) <= REL_TOL
    if tol_policy == 0:
        assert np.array_equal(b.reshape(-1), d["b_out"].reshape(-1))
        assert rel_err(x, d["x_out"]) <= REL_TOL


def test_pcg_graph_replays_match_eager_launches():
    """The captured PCG batch (replayed many times per step, with eager work
    between steps) against kernel-by-kernel launches: bit-identical fields."""
    ga = fa.FlipContext(48, tol_policy=1, tol=1e-5, dt_policy=1)
    gb = fa.FlipContext(48, tol_policy=1, tol=1e-5, dt_policy=1)
    try:
        ga.build_scene(fa.SCENE_BENCHMARK)
        gb.build_scene(fa.SCENE_BENCHMARK)
        gb.set_profiling(True)  # profiling launches every kernel directly
        for k in range(6):
            sa, sb = ga.step(), gb.step()
            assert sa["iterations"] == sb["iterations"] and sa["iterations"] > 8
            assert np.array_equal(ga.grid(fa.GRID_PRESS), gb.grid(fa.GRID_PRESS)), k
            a, b = ga.download(), gb.download()
            assert np.array_equal(a["p"], b["p"]) and np.array_equal(a["u"], b["u"]), k
    finally:
        ga.close()
        gb.close()


@pytest.mark.parametrize("precond", [fa.PRECOND_REFERENCE, fa.PRECOND_IC0])
def test_solver_modes_match_oracle(precond):
    """Stand-alone solve on a real N=16 system; IC(0) runs the hyperplane
    wavefront sweeps (the reference-form path with a coupling-free factor
    runs per cell)."""
    d = gold("solve")
    n = int(d["n"])
    x0, b0 = d["x"].copy(), d["b"].copy()
    xo, bo, so = po.solve(n, d["A"], d["L"], x0, b0, precond=precond, tol_policy=1, tol=1e-6)
    with fa.FlipContext(n, precond=precond, tol_policy=1, tol=1e-6) as g:
        x, b, st = g.solve(d["A"], d["L"], d["x"], d["b"])
    assert st["iterations"] == so["iterations"] > 1
    assert np.array_equal(b.reshape(-1), bo.reshape(-1))
    assert rel_err(x, xo) <= REL_TOL


def test_corrector_fine_grid():
    """At N=160 the corrector's resample radius satisfies re^2 <= 1e-5, so
    every resample weight is 0 and velocities are kept (the kernel is
    skipped): positions and velocities after the corrector stage equal the
    oracle's, for a dense cluster where springs and resample neighbours
    are plentiful."""
    gn = 160
    rs = np.random.RandomState(5)
    m = 6000
    p = (0.4 + 0.03 * rs.uniform(size=(m, 3))).astype(np.float32)
    u = rs.normal(0, 0.1, size=(m, 3)).astype(np.float32)
    t = np.full(m, po.FLUID, np.int8)
    s = po.Sim(gn)
    s.set_particles(p, u, np.zeros_like(p), t, dens=np.ones(m, np.float32))
    s.srand(1)
    with fa.FlipContext(gn) as g:
        g.upload(p, u, np.zeros_like(p), t, dens=np.ones(m, np.float32))
        g.srand(1)
        gst = g.stage(fa.STAGE_CORRECT)
        a = g.download()
    sst = s.call_stats("correct")
    b = s.particles()
    assert gst["n_jitter"] == sst["n_jitter"]
    assert np.array_equal(a["p"], b["p"])
    assert np.array_equal(a["u"], b["u"])
    assert not np.array_equal(b["p"], p)  # the springs moved particles


def test_density_general_mass():
    """Non-unit masses take the density kernel that reads whole particle
    records (m * kernel, walls skipped): densities equal the oracle's bit
    for bit, for kernel supports h*h that are (gn 8) and are not (gn 12)
    powers of two."""
    d = gold("p2g")
    for gn in (8, 12):
        p, u, t = d[f"p{gn}"], d[f"u{gn}"], d[f"t{gn}"]
        m = (0.5 + np.random.RandomState(gn).uniform(size=len(t))).astype(np.float32)
        s = po.Sim(gn)
        s.set_particles(p, u, np.zeros_like(p), t, dens=np.ones(len(t), np.float32), m=m)
        s.max_dens = 2.0
        s.call("sort")
        s.call("compute_density")
        with fa.FlipContext(gn) as g:
            g.upload(p, u, np.zeros_like(p), t, dens=np.ones(len(t), np.float32), mass=m)
            g.max_dens = 2.0
            g.stage(fa.STAGE_DENSITY)
            got, want = g.download()["dens"], s.particles()["dens"]
        assert np.array_equal(got, want), gn
        assert (want[t == 1] > 0).any()


def test_p2g_general_mass():
    """Non-unit masses take the general P2G kernel (m * sharp weights)."""
    d = gold("p2g")
    gn = 12
    p, u, t = d[f"p{gn}"], d[f"u{gn}"], d[f"t{gn}"]
    m = (0.5 + np.random.RandomState(9).uniform(size=len(t))).astype(np.float32)
    s = po.Sim(gn)
    s.set_particles(p, u, np.zeros_like(p), t, dens=np.ones(len(t), np.float32), m=m)
    s.call("sort")
    s.call("map_p2g")
    with fa.FlipContext(gn) as g:
        g.upload(p, u, np.zeros_like(p), t, dens=np.ones(len(t), np.float32), mass=m)
        grids = g.map_p2g()
    for c, w in enumerate((po.G_U0, po.G_U1, po.G_U2)):
        assert np.array_equal(grids[c], s.grid(w)), c


@pytest.mark.parametrize("where", ["velocity", "position"])
def test_nonfinite_particle_is_reported(where):
    """A NaN velocity or position in the particle state (forced on the host)
    ends the step with FLIP_E_NUMERIC instead of running on: the advection
    checks the values it moves, the corrector the values it writes."""
    with fa.FlipContext(16, dt_policy=0) as g:
        g.build_scene(fa.SCENE_BENCHMARK)
        q = g.download()
        fl = np.flatnonzero(q["type"] == po.FLUID)
        p, u = q["p"].copy(), q["u"].copy()
        if where == "velocity":
            u[fl[7]] = [np.nan, 0.0, np.inf]
        else:
            p[fl[7], 1] = np.nan
        g.upload(p, u, q["n"], q["type"], dens=q["dens"])
        with pytest.raises(fa.FlipError) as e:
            g.step()
        assert "non-finite" in str(e.value) or "NaN" in str(e.value)
