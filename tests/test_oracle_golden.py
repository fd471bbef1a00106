"""The oracle against the reference's own outputs (golden vectors generated by
tests/golden/gen_golden.py from mmmovania/flip3d compiled in place).

Every comparison is bit-exact: the oracle restates the reference's float/double
arithmetic in the reference's order.
"""
import ctypes
import hashlib
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_kernels_and_lengths_bit_exact():
    g = gold("kernels")
    sm, sh = po.kernels(g["r2"], g["h"])
    assert np.array_equal(sm.view(np.uint32), g["smooth"].view(np.uint32))
    assert np.array_equal(sh.view(np.uint32), g["sharp"].view(np.uint32))
    ln, l2 = po.lengths(g["a"], g["b"])
    assert np.array_equal(ln.view(np.uint32), g["length"].view(np.uint32))
    assert np.array_equal(l2.view(np.uint32), g["length2"].view(np.uint32))


def test_rand_matches_glibc():
    libc = ctypes.CDLL("libc.so.6")
    s = po.Sim(4)
    for seed in (1, 2, 12345, 2 ** 31 + 7):
        libc.srand(ctypes.c_uint(seed))
        s.srand(seed)
        mine = [s.rand() for _ in range(2000)]
        ref = [libc.rand() for _ in range(2000)]
        assert mine == ref, seed


def _sim_with(gn, p, t, u=None, nrm=None):
    s = po.Sim(gn)
    n = len(t)
    s.set_particles(p, np.zeros((n, 3)) if u is None else u, np.zeros((n, 3)) if nrm is None else nrm, t,
                    dens=np.ones(n, np.float32))
    return s


@pytest.mark.parametrize("gn", [8, 13])
def test_sort_buckets(gn):
    g = gold("sort")
    s = _sim_with(gn, g[f"p{gn}"], g[f"t{gn}"])
    s.call("sort")
    cs, ids = s.sorted_cells()
    assert np.array_equal(np.diff(cs), g[f"counts{gn}"])
    assert np.array_equal(ids, g[f"ids{gn}"])


@pytest.mark.parametrize("gn", [8, 12])
def test_p2g(gn):
    g = gold("p2g")
    s = _sim_with(gn, g[f"p{gn}"], g[f"t{gn}"], u=g[f"u{gn}"])
    s.call("sort")
    s.call("map_p2g")
    for c, which in enumerate((po.G_U0, po.G_U1, po.G_U2)):
        assert np.array_equal(s.grid(which).reshape(-1).view(np.uint32), g[f"g{c}_{gn}"].view(np.uint32)), c


def test_g2p():
    g = gold("g2p")
    gn = int(g["gn"])
    s = po.Sim(gn)
    s.set_grid(po.G_U0, g["g0"])
    s.set_grid(po.G_U1, g["g1"])
    s.set_grid(po.G_U2, g["g2"])
    out = s.fetch_velocity(0, g["pts"])
    assert np.array_equal(out.view(np.uint32), g["u"].view(np.uint32))


def test_corrector_with_rand_stream():
    g = gold("correct")
    gn = int(g["gn"])
    s = _sim_with(gn, g["p"], g["t"], u=g["u"], nrm=g["nrm"])
    s.srand(1)
    s.dt = float(g["dt"])
    st = s.call_stats("correct")
    assert st["n_jitter"] > 0  # the fixture exercises the rand() branch
    q = s.particles()
    assert np.array_equal(q["p"].view(np.uint32), g["p_out"].view(np.uint32))
    assert np.array_equal(q["u"].view(np.uint32), g["u_out"].view(np.uint32))


def test_solver_known_answer():
    g = gold("solve")
    n = int(g["n"])
    x, b, st = po.solve(n, g["A"], g["L"], g["x"], g["b"])
    assert st["iterations"] > 1
    assert np.array_equal(x.reshape(-1).view(np.uint32), g["x_out"].reshape(-1).view(np.uint32))
    assert np.array_equal(b.reshape(-1).view(np.uint32), g["b_out"].reshape(-1).view(np.uint32))


def test_solver_unit_factor_known_answer():
    """A system with P == 1.0 cells (tests/pocket_system.py): the value-test
    couplings of the factor (solver.cpp:127-135) and of both sweeps
    (:151-156, :167-172) are live.  The oracle's solve equals the
    reference's bit for bit, its factor is exactly 1.0 at the ten tunnel
    cells (1/sqrt(2) at the upper ones if the coupling were dropped), and
    the result differs from the same solve with couplings keyed off."""
    from tests.pocket_system import pocket_system, unit_cells
    g = gold("solve_unit")
    n = int(g["n"])
    A, L, x0, b0 = pocket_system(n)
    assert np.array_equal(A, g["A"]) and np.array_equal(b0, g["b"])
    x, b, st = po.solve(n, g["A"], g["L"], g["x"], g["b"])
    assert st["n_precond_unit"] == 10
    assert np.array_equal(x.reshape(-1).view(np.uint32), g["x_out"].reshape(-1).view(np.uint32))
    assert np.array_equal(b.reshape(-1).view(np.uint32), g["b_out"].reshape(-1).view(np.uint32))
    P = po.build_precond(n, g["A"], g["L"])
    assert sorted(zip(*np.nonzero(P == 1.0))) == sorted(unit_cells(n))
    # the forward sweep's value test on q is live too: with r = 1.0 at a
    # tunnel bottom, the cell above gets q = (r + 1) * 1
    r = np.zeros((n, n, n), np.float32)
    for (i, j, k) in unit_cells(n):
        if k == 3:
            r[i, j, k] = 1.0
    z = po.apply_precond(n, g["A"], P, r)
    i, j = unit_cells(n)[0][:2]
    assert z[i, j, 4] != 0.0  # zero under a per-cell (coupling-free) preconditioner


def test_reference_scene():
    g = gold("scene_ref")
    s = po.Sim(32)
    s.scene_reference()
    q = s.particles()
    assert np.array_equal(q["p"], g["p"]) and np.array_equal(q["n"], g["n"]) and np.array_equal(q["type"], g["t"])
    assert np.float32(s.max_dens) == g["max_dens"]


def _check_traj(s, g, steps):
    mask = None
    for k in range(1, steps + 1):
        st = s.step()
        q = s.particles()
        if mask is None:
            mask = q["type"] == po.FLUID
        dg = g[f"dig{k}"]
        mine = [digest(q["p"]), digest(q["u"][mask]), digest(q["dens"]), digest(s.grid(po.G_PRESS)),
                digest(s.grid(po.G_A)), digest(s.grid(po.G_L))]
        names = ["p", "u", "dens", "press", "A", "L"]
        bad = [nm for nm, a, b in zip(names, mine, dg) if a != b]
        assert not bad, f"step {k}: mismatch in {bad} ({st})"


def test_reference_trajectory_20_steps():
    g = gold("traj_ref")
    s = po.Sim(32)
    s.scene_reference()
    s.srand(1)  # the harness restarts the rand() stream after init
    _check_traj(s, g, int(g["steps"]))


def _bench_sim():
    g0 = gold("traj_bench.init")
    s = po.Sim(32)
    s.scene_benchmark(20261203)
    q = s.particles()
    assert np.array_equal(q["p"], g0["p"]) and np.array_equal(q["n"], g0["n"]) and np.array_equal(q["type"], g0["t"])
    assert np.float32(s.max_dens) == g0["max_dens"]
    s.srand(1)
    return s


def test_benchmark_scene_stage_probes():
    """Step 1 of the benchmark scene, stage by stage (flip3D.cpp:1216-1247)."""
    g = gold("traj_bench")
    s = _bench_sim()
    s.call("sort")
    s.call("compute_density")
    assert np.array_equal(s.particles()["dens"], g["s1_dens"])
    s.call("add_ext_force")
    s.call("sort")
    s.call("map_p2g")
    s.call("mark_water")
    for c, w in enumerate((po.G_U0, po.G_U1, po.G_U2)):
        assert np.array_equal(s.grid(w).reshape(-1), g[f"s1_p2g_u{c}"]), c
    assert np.array_equal(s.grid(po.G_A).reshape(-1), g["s1_A"])
    s.call("save_grid")
    s.call("enforce_boundary")
    s.call_stats("project")
    s.call("enforce_boundary")
    s.call("extrapolate_velocity")
    s.call("subtract_grid")
    s.call("flip_g2p")
    assert np.array_equal(s.grid(po.G_L).reshape(-1), g["s1_L"])
    assert np.array_equal(s.grid(po.G_PRESS).reshape(-1), g["s1_press"])
    for c, (w, ws) in enumerate(((po.G_U0, po.G_US0), (po.G_U1, po.G_US1), (po.G_U2, po.G_US2))):
        assert np.array_equal(s.grid(w).reshape(-1), g[f"s1_u{c}"]), c
        assert np.array_equal(s.grid(ws).reshape(-1), g[f"s1_us{c}"]), c
    q = s.particles()
    mask = q["type"] == po.FLUID
    assert np.array_equal(q["u"][mask], g["s1_flip_u"])
    st = s.call_stats("advect")
    q = s.particles()
    assert np.array_equal(q["p"], g["s1_adv_p"])
    assert np.array_equal(q["u"][mask], g["s1_adv_u"])
    assert st["n_reposition"] == int(g["s1_stuck"])


def test_benchmark_trajectory():
    """Free-running trajectory digests; FLIP_ORACLE_FULL=1 runs all 100 steps."""
    g = gold("traj_bench")
    steps = int(g["steps"]) if os.environ.get("FLIP_ORACLE_FULL") else 12
    s = _bench_sim()
    _check_traj(s, g, steps)
    if steps >= 10:
        q = s.particles() if steps == 10 else None
    full = _bench_sim()
    for _ in range(10):
        full.step()
    q = full.particles()
    assert np.array_equal(q["p"], g["full10_p"])
    assert np.array_equal(q["dens"], g["full10_dens"])
    assert np.array_equal(full.grid(po.G_PRESS).reshape(-1), g["full10_press"])


TEST_SCENES = [(1, "sphere_pour"), (2, "cliff_pour"), (3, "water_drop")]


@pytest.mark.parametrize("which,name", TEST_SCENES)
def test_reference_test_scenes(which, name):
    """spherePourTest / cliffPourTest / waterDropTest placement: obstacles
    (a wall sphere, a wall box), the interior-wall normals of
    compute_wall_normal (flip3D.cpp:107-119) and the calibrated max_dens."""
    g = gold("scenes_test")
    s = po.Sim(32)
    assert s.scene_test(which) == len(g[f"{name}_init_t"])
    q = s.particles()
    assert np.array_equal(q["type"], g[f"{name}_init_t"])
    assert np.array_equal(q["p"], g[f"{name}_init_p"])
    assert np.array_equal(q["n"], g[f"{name}_init_n"])
    assert np.float32(s.max_dens) == g[f"{name}_max_dens"]


@pytest.mark.parametrize("which,name", TEST_SCENES[:2])
def test_pour_trajectories(which, name):
    """Steps of the pour scenes in simulateStep order: step counter, pourWater
    (flip3D.cpp:375-402) appending a disc of particles, then the step; every
    step bit-exact against the reference's digests."""
    g = gold("scenes_test")
    s = po.Sim(32)
    s.scene_test(which)
    s.srand(1)
    steps = int(g[f"{name}_steps"])
    for k in range(1, steps + 1):
        n0 = s.num_particles()
        st = s.step()
        q = s.particles()
        assert s.num_particles() - n0 == int(g[f"{name}_added{k}"])
        fl = q["type"] == po.FLUID
        mine = [digest(q["p"]), digest(q["u"][fl]), digest(q["dens"]), digest(s.grid(po.G_PRESS)),
                digest(s.grid(po.G_A)), digest(s.grid(po.G_L))]
        bad = [nm for nm, a, b in zip(["p", "u", "dens", "press", "A", "L"], mine, g[f"{name}_dig{k}"]) if a != b]
        assert not bad, f"{name} step {k}: mismatch in {bad} ({st})"


def test_mesh_after_five_steps():
    """mesher::generateMesh (implicit field, marching cubes, edge filter, two
    smoothing passes) on the dam-break scene after 5 steps, through the
    buckets of the corrector's sort: vertices, normals and faces bit-exact."""
    g = gold("mesh")
    s = po.Sim(32)
    s.scene_reference()
    s.srand(1)
    for _ in range(5):
        s.step()
    v, n, f = s.mesh()
    assert np.array_equal(f, g["f"])
    assert np.array_equal(v.view(np.uint64), g["v"].view(np.uint64))
    assert np.array_equal(n.view(np.uint64), g["n"].view(np.uint64))


def test_mc_table_edges_follow_the_corners():
    """Every edge a case's triangles use is cut (its corners lie on opposite
    sides of the iso level), and every cut edge is used."""
    import re
    src = open(os.path.join(os.path.dirname(GOLD), "..", "flip3d_amd", "csrc", "mc_table.h")).read()
    cases = re.findall(r'"([0-9a-b]*)"', src[src.index("MC_CASE_LIST"):src.index("MC_EDGE_CORNERS")])
    assert len(cases) == 256
    corners = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    for c, tri in enumerate(cases):
        assert len(tri) % 3 == 0
        used = {int(ch, 16) for ch in tri}
        cut = {e for e, (a, b) in enumerate(corners) if ((c >> a) ^ (c >> b)) & 1}
        assert used == cut, c


def test_output_files_after_five_steps():
    """write_obj, writeSunflow and write_frame + write_bmp restated
    (oracle/exporters.py): the files of frame 6 of the dam-break scene match
    the reference's byte for byte."""
    from oracle import exporters as ex
    g = gold("exports")
    s = po.Sim(32)
    s.scene_reference()
    s.srand(1)
    for _ in range(5):
        s.step()
    v, n, f = s.mesh()
    q = s.particles()
    files = {"obj": ex.obj_bytes(v, n, f, 32), "sc": ex.sunflow_bytes(v, n, f, 32),
             "bmp": ex.bmp_bytes(ex.render_preview(q["p"], q["type"], 32))}
    for key, raw in files.items():
        assert len(raw) == int(g[f"{key}_size"]), key
        assert hashlib.sha256(raw).hexdigest() == str(g[f"{key}_sha"]), key
    assert files["bmp"][:54] == g["bmp_header"].tobytes()
