"""Generate golden vectors from the reference flip3d (mmmovania/flip3d).

Runs ONLY in a container where /root/reference exists: it builds
oracle/_ref/libflipref.so from the reference sources in place (oracle/Makefile)
and drives it through ctypes.  Every fixture is data: inputs and the
reference's outputs (or SHA-256 digests of them for long trajectories).

    python tests/golden/gen_golden.py            # all sections
    python tests/golden/gen_golden.py kernels    # one section

Each section runs in its own process: solver.cpp keeps static scratch sized at
its first call, and flip3D.cpp keeps its scene in file-static state.
The reference runs single-threaded (OMP_NUM_THREADS=1) so its rand() draws
happen in particle order.
"""
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SECTIONS = ["kernels", "sort", "p2g", "g2p", "correct", "solve", "scene_ref", "traj_ref", "traj_bench",
            "checkpoint", "scenes_test", "mesh", "exports", "scene_bench_ref",
            "solve_unit", "objects0", "objects1", "objects2", "objects3", "objects4"]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


class Quiet:
    """Silence the reference's stdout logging (utility.h:64-76)."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)

    def __exit__(self, *a):
        os.dup2(self.saved, 1)
        os.close(self.null)
        os.close(self.saved)


def P(a):
    from oracle import pyoracle as po
    return po._ptr(a)


def ref():
    from oracle import pyoracle as po
    return po.ref_lib()


def rng_particles(rs, gn, np_, spread=1.0, frac_wall=0.1):
    p = rs.uniform(-0.02, 1.02, size=(np_, 3)).astype(np.float32) * np.float32(spread)
    t = np.where(rs.uniform(size=np_) < frac_wall, 2, 1).astype(np.int8)
    u = rs.normal(size=(np_, 3)).astype(np.float32)
    return p, u, t


def sec_kernels(out):
    R = ref()
    rs = np.random.RandomState(1)
    m = 20000
    h = np.concatenate([np.full(m // 4, 1.4, np.float32), np.full(m // 4, 2.0 / 32, np.float32),
                        np.full(m // 4, 0.5 / 32, np.float32), rs.uniform(0.01, 2, m // 4).astype(np.float32)])
    r2 = (rs.uniform(0, 1.3, m).astype(np.float32) * h * h).astype(np.float32)
    r2[:50] = 0.0
    r2[50:100] = np.float32(1e-5)
    r2[100:150] = (h[100:150] * h[100:150]).astype(np.float32)
    sm = np.empty(m, np.float32)
    sh = np.empty(m, np.float32)
    R.ref_kat_kernels(m, P(r2), P(h), P(sm), P(sh))
    a = rs.uniform(0, 1, (m, 3)).astype(np.float32)
    b = (a + rs.normal(scale=0.03, size=(m, 3))).astype(np.float32)
    b[:20] = a[:20]
    ln = np.empty(m, np.float32)
    l2 = np.empty(m, np.float32)
    R.ref_kat_length(m, P(a), P(b), P(ln), P(l2))
    np.savez_compressed(out, r2=r2, h=h, smooth=sm, sharp=sh, a=a, b=b, length=ln, length2=l2)


def sec_sort(out):
    R = ref()
    rs = np.random.RandomState(2)
    res = {}
    for gn, np_ in ((8, 3000), (13, 6000)):
        p, _, t = rng_particles(rs, gn, np_)
        p[:5] = np.nan
        p[5:10] = [1.0, 0.0, 0.999999]
        counts = np.empty(gn ** 3, np.int32)
        ids = np.empty(np_, np.int32)
        R.ref_kat_sort(gn, np_, P(p), P(t), P(counts), P(ids))
        res.update({f"p{gn}": p, f"t{gn}": t, f"counts{gn}": counts, f"ids{gn}": ids})
    np.savez_compressed(out, **res)


def sec_p2g(out):
    R = ref()
    rs = np.random.RandomState(3)
    res = {}
    for gn, np_ in ((8, 4000), (12, 9000)):
        p, u, t = rng_particles(rs, gn, np_)
        g0 = np.empty((gn + 1) * gn * gn, np.float32)
        g1 = np.empty_like(g0)
        g2 = np.empty_like(g0)
        R.ref_kat_p2g(gn, np_, P(p), P(u), P(t), P(g0), P(g1), P(g2))
        res.update({f"p{gn}": p, f"u{gn}": u, f"t{gn}": t, f"g0_{gn}": g0, f"g1_{gn}": g1, f"g2_{gn}": g2})
    np.savez_compressed(out, **res)


def sec_g2p(out):
    R = ref()
    rs = np.random.RandomState(4)
    gn = 10
    F = (gn + 1) * gn * gn
    g = [rs.normal(size=F).astype(np.float32) for _ in range(3)]
    # points whose z stays below the top half cell (mapper.cpp:102 reads past
    # grid[1]'s rows above it), plus clamped out-of-range points
    pts = rs.uniform(-0.1, 1.0 - 0.5 / gn - 1e-3, (5000, 3)).astype(np.float32)
    uo = np.empty_like(pts)
    R.ref_kat_g2p(gn, P(g[0]), P(g[1]), P(g[2]), len(pts), P(pts), P(uo))
    np.savez_compressed(out, gn=gn, g0=g[0], g1=g[1], g2=g[2], pts=pts, u=uo)


def sec_correct(out):
    R = ref()
    rs = np.random.RandomState(5)
    gn = 8
    np_ = 2500
    p, u, t = rng_particles(rs, gn, np_)
    p = np.clip(p, 0.0, 0.999).astype(np.float32)
    # near-coincident fluid pairs and wall pairs trigger the jitter branches
    p[100:140] = (p[60:100] + 1e-4).astype(np.float32)
    t[100:120] = 1
    t[60:80] = 1
    nrm = rs.normal(size=(np_, 3)).astype(np.float32)
    nrm[t == 1] = 0
    pp, uu = p.copy(), u.copy()
    R.ref_srand(1)
    R.ref_kat_correct(gn, np_, P(pp), P(uu), P(nrm), P(t), np.float32(6e-3), np.float32(0.5 / gn))
    np.savez_compressed(out, gn=gn, p=p, u=u, nrm=nrm, t=t, p_out=pp, u_out=uu,
                        dt=np.float32(6e-3), re=np.float32(0.5 / gn))


def _solve_system(n, seed):
    """A realistic system: the oracle's benchmark scene at n after one step's
    P2G + projection inputs (flags, level set, divergence)."""
    from oracle import pyoracle as po
    S = po.Sim(n)
    S.scene_benchmark(seed)
    S.call("sort")
    S.call("compute_density")
    S.call("add_ext_force")
    S.call("sort")
    S.call("map_p2g")
    S.call("mark_water")
    S.call("save_grid")
    S.call("enforce_boundary")
    A = S.grid(po.G_A)
    # divergence and level set as project() computes them (flip3D.cpp:815-826)
    S.call_stats("project")
    L = S.grid(po.G_L)
    rs = np.random.RandomState(seed)
    b = rs.normal(size=(n, n, n)).astype(np.float32)
    x = (rs.normal(size=(n, n, n)) * 1e-2).astype(np.float32)
    return A, L, x, b


def sec_solve(out):
    R = ref()
    n = 16
    A, L, x, b = _solve_system(n, 7)
    xo, bo = x.copy(), b.copy()
    with Quiet():
        R.ref_kat_solve(n, P(A), P(L), P(xo), P(bo), 1)
    np.savez_compressed(out, n=n, A=A, L=L, x=x, b=b, x_out=xo, b_out=bo)


def _ref_particles(R):
    np_ = R.ref_num_particles()
    p = np.empty((np_, 3), np.float32)
    u = np.empty((np_, 3), np.float32)
    nr = np.empty((np_, 3), np.float32)
    t = np.empty(np_, np.int8)
    d = np.empty(np_, np.float32)
    R.ref_get_particles(P(p), P(u), P(nr), P(t), P(d))
    return p, u, nr, t, d


def _ref_grids(R):
    n = R.ref_grid_n()
    F = (n + 1) * n * n
    g = {}
    for i, nm in enumerate(["u0", "u1", "u2", "us0", "us1", "us2"]):
        a = np.empty(F, np.float32)
        R.ref_get_grid(i, P(a))
        g[nm] = a
    A = np.empty(n ** 3, np.int8)
    R.ref_get_grid(6, P(A))
    L = np.empty(n ** 3, np.float32)
    R.ref_get_grid(7, P(L))
    pr = np.empty(n ** 3, np.float32)
    R.ref_get_grid(8, P(pr))
    g.update(A=A, L=L, press=pr)
    return g


def sec_scene_ref(out):
    R = ref()
    with Quiet():
        R.ref_init()
    p, u, nr, t, d = _ref_particles(R)
    np.savez_compressed(out, p=p, n=nr, t=t, max_dens=np.float32(R.ref_get_max_dens()))


def _trajectory(R, steps, full_at, out):
    """Run reference steps; record per-step digests of particle state and
    grids, full arrays at `full_at`, and stage probes of step 1."""
    R.ref_srand(1)
    rec = {}
    mask = None
    for s in range(1, steps + 1):
        if s == 1:
            with Quiet():
                R.ref_stage_density()
            p, u, nr, t, d = _ref_particles(R)
            rec["s1_dens"] = d
            with Quiet():
                R.ref_stage_ext_force()
                R.ref_probe_p2g()
            g = _ref_grids(R)
            rec.update(s1_p2g_u0=g["u0"], s1_p2g_u1=g["u1"], s1_p2g_u2=g["u2"], s1_A=g["A"])
            with Quiet():
                R.ref_stage_solve_picflip()
            g = _ref_grids(R)
            p, u, nr, t, d = _ref_particles(R)
            mask = t == 1
            rec.update(s1_L=g["L"], s1_press=g["press"], s1_u0=g["u0"], s1_u1=g["u1"], s1_u2=g["u2"],
                       s1_us0=g["us0"], s1_us1=g["us1"], s1_us2=g["us2"], s1_flip_u=u[mask])
            with Quiet():
                R.ref_stage_advect()
            p, u, nr, t, d = _ref_particles(R)
            rec.update(s1_adv_p=p, s1_adv_u=u[mask], s1_stuck=np.int32(R.ref_num_stuck()))
            with Quiet():
                R.ref_stage_correct()
        else:
            with Quiet():
                R.ref_step()
        p, u, nr, t, d = _ref_particles(R)
        g = _ref_grids(R)
        rec[f"dig{s}"] = np.array([digest(p), digest(u[mask]), digest(d), digest(g["press"]), digest(g["A"]),
                                   digest(g["L"])])
        if s in full_at:
            rec.update({f"full{s}_p": p, f"full{s}_u": u, f"full{s}_dens": d, f"full{s}_press": g["press"],
                        f"full{s}_A": g["A"], f"full{s}_L": g["L"]})
    rec["steps"] = np.int32(steps)
    np.savez_compressed(out, **rec)


def sec_traj_ref(out):
    R = ref()
    with Quiet():
        R.ref_init()
    p, u, nr, t, d = _ref_particles(R)
    _trajectory(R, 20, (1,), out)


def sec_traj_bench(out):
    from oracle import pyoracle as po
    R = ref()
    with Quiet():
        R.ref_init()
    S = po.Sim(32)
    m = S.scene_benchmark(20261203)
    q = S.particles()
    R.ref_set_particles(m, P(q["p"]), None, None, P(q["type"]), None)
    with Quiet():
        R.ref_compute_wall_normal()
    p, u, nr, t, d = _ref_particles(R)
    np.savez_compressed(out + ".init.npz", p=p, n=nr, t=t, max_dens=np.float32(R.ref_get_max_dens()))
    _trajectory(R, 100, (1, 10), out)


def sec_objects(which):
    """render/save/objects.dat as saveState writes it on its first save
    (flip3D.cpp:498-508), for the scene `which` of ref_init_test (0 damBreak,
    1 spherePour, 2 cliffPour, 3 waterDrop, 4 the benchmark objects): the
    reference's own text, kept as tests/golden/objects_<which>.txt."""
    R = ref()
    with Quiet():
        R.ref_init_test(which, 1)
        R.ref_save_state(1)
    text = open(os.path.join("render", "save", "objects.dat")).read()
    with open(os.path.join(HERE, f"objects_{which}.txt"), "w") as fh:
        fh.write(text)


def sec_solve_unit(out):
    """solver::solve on a system with P == 1.0 cells (tests/pocket_system.py),
    where the reference's value-test couplings are live: its own result
    under its stop rule, and its factor P."""
    from tests.pocket_system import pocket_system
    R = ref()
    n = 12
    A, L, x, b = pocket_system(n)
    xo, bo = x.copy(), b.copy()
    with Quiet():
        R.ref_kat_solve(n, P(A), P(L), P(xo), P(bo), 1)
    np.savez_compressed(out, n=n, A=A, L=L, x=x, b=b, x_out=xo, b_out=bo)


def sec_scene_bench_ref(out):
    """The benchmark scene's liquid objects placed by the reference's own
    placement loop and object test (flip3D.cpp:580-633, 711-761) at N=32:
    the sub-lattice index of every FLUID particle in creation order (the
    reference's +-0.002/N rand() jitter is far below the half spacing, so
    rounding recovers the index) and the WALL particle positions (no
    jitter).  The oracle's benchmark placement is checked against these."""
    R = ref()
    with Quiet():
        R.ref_init_test(4, 1)
    p, u, nr, t, d = _ref_particles(R)
    n = R.ref_grid_n()
    w = 0.5 / n
    fl = t == 1
    idx = np.rint(p[fl].astype(np.float64) / w - 0.5).astype(np.int16)
    assert np.abs(p[fl] - (idx + 0.5) * w).max() < 0.01 / n
    np.savez_compressed(out, n=np.int32(n), fluid_lattice=idx, wall_p=p[~fl], wall_n=nr[~fl],
                        n_fluid=np.int32(fl.sum()), n_wall=np.int32((~fl).sum()))


def sec_scenes_test(out):
    """The reference's other test scenes (spherePourTest, cliffPourTest,
    waterDropTest: obstacles, interior wall normals and the pourWater
    emitter) at N=32: the initial particles, then 12 steps of simulateStep
    order (step counter, pourWater, the five stages) with per-step digests
    and full arrays at steps 1 and 12."""
    R = ref()
    rec = {}
    for which, name, steps in ((1, "sphere_pour", 12), (2, "cliff_pour", 12), (3, "water_drop", 0)):
        with Quiet():
            R.ref_init_test(which, 1)
        p, u, nr, t, d = _ref_particles(R)
        rec.update({f"{name}_init_p": p, f"{name}_init_n": nr, f"{name}_init_t": t,
                    f"{name}_max_dens": np.float32(R.ref_get_max_dens())})
        R.ref_srand(1)
        for s in range(1, steps + 1):
            R.ref_set_step(s)
            added = R.ref_pour()
            with Quiet():
                R.ref_step()
            p, u, nr, t, d = _ref_particles(R)
            g = _ref_grids(R)
            fl = t == 1
            rec[f"{name}_dig{s}"] = np.array([digest(p), digest(u[fl]), digest(d), digest(g["press"]),
                                              digest(g["A"]), digest(g["L"])])
            rec[f"{name}_added{s}"] = np.int32(added)
            if s in (1, steps):
                rec.update({f"{name}_full{s}_p": p, f"{name}_full{s}_u": u, f"{name}_full{s}_t": t,
                            f"{name}_full{s}_press": g["press"], f"{name}_full{s}_A": g["A"]})
        rec[f"{name}_steps"] = np.int32(steps)
    np.savez_compressed(out, **rec)


def sec_mesh(out):
    """mesher::generateMesh (mg = N + 1, as write_mesh calls it) after 5 steps
    of the reference dam-break scene: smoothed vertices and normals, faces."""
    R = ref()
    with Quiet():
        R.ref_init()
    R.ref_srand(1)
    with Quiet():
        for _ in range(5):
            R.ref_step()
        nv = R.ref_mesh(R.ref_grid_n() + 1)
    nf = R.ref_mesh_faces()
    v = np.empty((nv, 3), np.float64)
    nr = np.empty((nv, 3), np.float64)
    f = np.empty((nf, 3), np.int32)
    R.ref_mesh_get(P(v), P(nr), P(f))
    np.savez_compressed(out, v=v, n=nr, f=f)


def sec_exports(out):
    """The files simulateStep writes for frame 6 of the dam-break scene after
    5 steps (preview BMP, OBJ, Sunflow scene and template): sizes and
    SHA-256 digests, plus the BMP header and the first OBJ lines as text."""
    R = ref()
    with Quiet():
        R.ref_init()
    R.ref_srand(1)
    with Quiet():
        for _ in range(5):
            R.ref_step()
    os.makedirs(os.path.join("render", "preview"), exist_ok=True)
    os.makedirs(os.path.join("render", "obj"), exist_ok=True)
    os.makedirs(os.path.join("render", "sunflow"), exist_ok=True)
    with Quiet():
        R.ref_write_outputs(6)
    rec = {}
    for key, path in (("bmp", "render/preview/frame_6.bmp"), ("obj", "render/obj/6_scene.obj"),
                      ("sc", "render/sunflow/6_scene.sc"), ("template", "render/sunflow/template.sc")):
        raw = open(path, "rb").read()
        rec[f"{key}_size"] = np.int64(len(raw))
        rec[f"{key}_sha"] = np.array(hashlib.sha256(raw).hexdigest())
    rec["bmp_header"] = np.frombuffer(open("render/preview/frame_6.bmp", "rb").read()[:54], np.uint8)
    rec["obj_head"] = np.array(open("render/obj/6_scene.obj").read()[:400])
    np.savez_compressed(out, **rec)


def sec_checkpoint(out):
    """The reference dam-break scene after 5 steps, saved by saveState."""
    R = ref()
    with Quiet():
        R.ref_init()
        R.ref_srand(1)
        for _ in range(5):
            R.ref_step()
        R.ref_save_state(5)
    raw = open(os.path.join("render", "save", "5_scene.dat"), "rb").read()
    from tests.checkpoint_format import checkpoint_fields, field_digests
    f = checkpoint_fields(raw, R.ref_grid_n())
    rec = {"n": np.int32(R.ref_grid_n()), "size": np.int64(f["size"]), "count": np.int32(f["count"]),
           "max_dens": f["max_dens"], "pour_time": f["pour_time"], "pour": f["pour"]}
    for k, v in field_digests(f).items():
        rec["sha_" + k] = np.array(v)
    np.savez_compressed(out, **rec)


def run_section(name):
    out = os.path.join(HERE, f"{name}.npz")
    cwd = tempfile.mkdtemp(prefix="flipref_")
    os.chdir(cwd)  # the reference writes render/ into its working directory
    os.makedirs(os.path.join(cwd, "render", "log"), exist_ok=True)  # solver.cpp:209 logs there
    if name.startswith("objects"):
        sec_objects(int(name[len("objects"):]))
    else:
        globals()["sec_" + name](out if not name.startswith("traj_bench") else os.path.join(HERE, name))
    print("wrote", name)


def main():
    if not os.path.isdir("/root/reference/src"):
        raise SystemExit("reference sources not present; fixtures are generated only where they are")
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8", "all", "ref"], check=True,
                   capture_output=True)
    names = sys.argv[1:] or SECTIONS
    env = dict(os.environ, OMP_NUM_THREADS="1")
    for nm in names:
        subprocess.run([sys.executable, os.path.abspath(__file__), "--section", nm], check=True, env=env)


if __name__ == "__main__":
    if len(sys.argv) > 2 and sys.argv[1] == "--section":
        run_section(sys.argv[2])
    else:
        main()
