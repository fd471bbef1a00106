"""Pins for oracle parts beyond the golden vectors: closed forms, invariants,
brute force on tiny inputs, and the rounding arguments the HIP kernels rely on.
Each check is chosen so a plausible slip (wrong operand order, dropped term,
swapped index, wrong clamp) fails it."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po


# ---------------------------------------------------------------- rounding
def test_smooth_kernel_single_float_subtraction_is_exact():
    """kernel.cpp:4-6 computes (float)fmax(1.0 - (double)q, 0.0); the device
    uses max(1.0f - q, 0): equal for every float q >= 0 (dense samples of
    every binade from 2^-40 to 2^2 plus the endpoints)."""
    rs = np.random.RandomState(0)
    qs = [np.float32(0.0), np.float32(1.0), np.nextafter(np.float32(1.0), np.float32(0))]
    for e in range(-40, 3):
        m = rs.uniform(1.0, 2.0, 20000)
        qs.extend((m * 2.0 ** e).astype(np.float32))
    q = np.array(qs, np.float32)
    ref = np.maximum(1.0 - q.astype(np.float64), 0.0).astype(np.float32)
    dev = np.maximum(np.float32(1.0) - q, np.float32(0.0))
    assert np.array_equal(ref.view(np.uint32), dev.view(np.uint32))


def test_sharp_kernel_zero_beyond_support():
    """sharp_kernel is exactly 0 for r2 >= h*h (the kernels skip those pairs)."""
    rs = np.random.RandomState(1)
    h = np.float32(1.4)
    hh = h * h
    r2 = (hh * rs.uniform(1.0, 3.0, 5000)).astype(np.float32)
    r2[:10] = hh
    _, sh = po.kernels(r2, np.full_like(r2, h))
    assert np.all(sh == 0.0)


def test_glibc_hypotf_is_double_sqrt():
    """utility.cpp:16 hypotf == (float)sqrt((double)x*x + (double)y*y), the
    form the device evaluates."""
    libm = ctypes.CDLL("libm.so.6")
    libm.hypotf.restype = ctypes.c_float
    libm.hypotf.argtypes = [ctypes.c_float, ctypes.c_float]
    rs = np.random.RandomState(2)
    x = rs.normal(size=20000).astype(np.float32) * np.float32(10.0) ** rs.randint(-3, 3, 20000).astype(np.float32)
    y = rs.normal(size=20000).astype(np.float32)
    ref = np.array([libm.hypotf(float(a), float(b)) for a, b in zip(x, y)], np.float32)
    dev = np.sqrt(x.astype(np.float64) ** 2 + y.astype(np.float64) ** 2).astype(np.float32)
    assert np.array_equal(ref, dev)


# ---------------------------------------------------------------- sorter
def test_cell_keys_closed_form_and_stable_order():
    """sorter.cpp:35-37 keys = trunc(clamp(float(gn)*p, 0, gn-1)), NaN -> gn-1;
    buckets list particles in ascending index."""
    gn = 7
    rs = np.random.RandomState(3)
    p = rs.uniform(-0.2, 1.2, (500, 3)).astype(np.float32)
    p[:3] = np.nan
    s = po.Sim(gn)
    s.set_particles(p, np.zeros_like(p), np.zeros_like(p), np.full(500, 1, np.int8))
    s.call("sort")
    cs, ids = s.sorted_cells()
    v = np.float32(gn) * p
    k = np.where(np.isnan(v), gn - 1, np.clip(v, 0, gn - 1)).astype(np.int64)
    key = (k[:, 0] * gn + k[:, 1]) * gn + k[:, 2]
    expect = np.argsort(key, kind="stable")
    assert np.array_equal(ids, expect)
    assert np.array_equal(np.diff(cs), np.bincount(key, minlength=gn ** 3))


# ------------------------------------------------------------------- P2G
def test_p2g_constant_velocity_is_reproduced():
    """A power-of-two constant velocity scales every product exactly, so every
    face with weight gets exactly that value and faces without get 0."""
    gn = 6
    rs = np.random.RandomState(4)
    p = rs.uniform(0.05, 0.95, (800, 3)).astype(np.float32)
    u = np.full((800, 3), [0.5, -2.0, 0.25], np.float32)
    s = po.Sim(gn)
    s.set_particles(p, u, np.zeros_like(p), np.full(800, 1, np.int8))
    s.call("sort")
    s.call("map_p2g")
    for c, w in enumerate((po.G_U0, po.G_U1, po.G_U2)):
        g = s.grid(w)
        vals = set(np.unique(g).tolist())
        assert vals <= {0.0, float(u[0, c])}, (c, vals)
        assert (g == u[0, c]).sum() > g.size // 2


def test_p2g_brute_force_tiny():
    """Face value = weighted mean over the reference window, recomputed by
    brute force (float32 in window order) for a handful of particles."""
    gn = 4
    rs = np.random.RandomState(5)
    p = rs.uniform(0, 1, (40, 3)).astype(np.float32)
    u = rs.normal(size=(40, 3)).astype(np.float32)
    t = np.where(rs.uniform(size=40) < 0.2, 2, 1).astype(np.int8)
    s = po.Sim(gn)
    s.set_particles(p, u, np.zeros_like(p), t)
    s.call("sort")
    s.call("map_p2g")
    g0 = s.grid(po.G_U0)
    cell = np.clip(np.float32(gn) * p, 0, gn - 1).astype(int)
    f32 = np.float32
    nonzero = 0
    for i in range(gn + 1):
        for j in range(gn):
            for k in range(gn):
                sw = f32(0)
                sx = f32(0)
                for si in (i - 1, i):
                    for sj in range(j - 2, j + 2):
                        for sk in range(k - 2, k + 2):
                            for q in range(40):
                                if tuple(cell[q]) != (si, sj, sk) or t[q] != 1:
                                    continue
                                pos = np.clip(f32(gn) * p[q], f32(0), f32(gn)).astype(f32)
                                d = pos - np.array([i, j + 0.5, k + 0.5], f32)
                                r2 = f32(f32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
                                hh = f32(f32(1.4) * f32(1.4))
                                w = f32(max(float(hh) / max(float(r2), 1e-5) - 1.0, 0.0))
                                sx = f32(sx + f32(w * u[q, 0]))
                                sw = f32(sw + w)
                expect = f32(sx / sw) if sw != 0 else f32(0)
                nonzero += expect != 0
                assert g0[i, j, k] == expect, (i, j, k)
    assert nonzero > g0.size // 2


# ------------------------------------------------------------------- G2P
def test_g2p_reproduces_linear_field():
    """Trilinear interpolation of a linear field is the field (to float
    rounding) in the interior."""
    gn = 8
    s = po.Sim(gn, f64=True)
    i, j, k = np.meshgrid(np.arange(gn + 1), np.arange(gn), np.arange(gn), indexing="ij")
    s.set_grid(po.G_U0, 0.3 * i - 0.2 * (j + 0.5) + 0.1 * (k + 0.5))
    pts = np.random.RandomState(6).uniform(0.15, 0.8, (300, 3))
    out = s.fetch_velocity(0, pts)
    x, y, z = gn * pts[:, 0], gn * pts[:, 1], gn * pts[:, 2]
    assert np.allclose(out[:, 0], 0.3 * x - 0.2 * y + 0.1 * z, atol=1e-12)


# ---------------------------------------------------------------- solver
def _box_system(n, air=True):
    A = np.full((n, n, n), po.WALL, np.int8)
    A[1:-1, 1:-1, 1:-1] = po.FLUID
    L = np.full((n, n, n), -0.5, np.float32)
    L[A == po.WALL] = 1.0
    if air:
        A[1:-1, n - 2:, 1:-1] = po.AIR
        L[A == po.AIR] = 0.7
    return A, L


def test_laplacian_annihilates_constants_without_air():
    """solver.cpp:29-41: walls and the domain edge are Neumann, so A*1 = 0
    on every FLUID cell of a closed box."""
    n = 8
    A, L = _box_system(n, air=False)
    ax = po.compute_ax(n, A, L, np.ones((n, n, n)), f64=True)
    assert np.all(ax[A == po.FLUID] == 0.0)


def test_laplacian_stencil_closed_form():
    """One FLUID cell next to AIR: (6x - sum x_ref)/h^2 with the ghost value
    L_air / fmin(1e-6, L_self) * x_self (solver.cpp:19-26)."""
    n = 8
    A, L = _box_system(n, air=True)
    x = np.random.RandomState(7).normal(size=(n, n, n))
    ax = po.compute_ax(n, A, L, x, f64=True)
    c = (3, n - 3, 4)  # its +j neighbour is AIR
    i, j, k = c
    ghost = float(L[i, j + 1, k]) / min(1e-6, float(L[c])) * x[c]
    nb = [x[i + 1, j, k], x[i - 1, j, k], ghost, x[i, j - 1, k], x[i, j, k + 1], x[i, j, k - 1]]
    expect = (6 * x[c] - sum(nb)) * n * n
    assert abs(ax[c] - expect) <= 1e-12 * abs(expect)


def test_preconditioner_closed_form():
    """With no P == 1 the reference preconditioner is z = (r P) P with
    P = 1/sqrtf(diag) (solver.cpp:125-180)."""
    n = 8
    A, L = _box_system(n)
    P = po.build_precond(n, A, L)
    r = np.random.RandomState(8).normal(size=(n, n, n)).astype(np.float32)
    z = po.apply_precond(n, A, P, r)
    f = A == po.FLUID
    assert not np.any(P[f] == 1.0)
    expect = ((r.astype(np.float64) * P) * P).astype(np.float32)
    assert np.array_equal(z[f], expect[f])
    # diag: 6 minus walls (and ghost terms); corner cell of the box has 3 walls
    assert P[1, 1, 1] == 1.0 / np.float64(np.sqrt(np.float32(3.0)))


@pytest.mark.parametrize("precond", [0, 1])
def test_pcg_matches_direct_solve(precond):
    """PCG converges to the solution of the same SPD system; the operator is
    assembled column by column from compute_Ax and solved densely."""
    n = 6
    A, L = _box_system(n)
    f = np.flatnonzero(A.reshape(-1) == po.FLUID)
    M = np.zeros((len(f), len(f)))
    for col, c in enumerate(f):
        e = np.zeros(n ** 3)
        e[c] = 1.0
        M[:, col] = po.compute_ax(n, A, L, e.reshape(n, n, n), f64=True).reshape(-1)[f]
    assert np.allclose(M, M.T)
    rhs = np.random.RandomState(9).normal(size=len(f))
    direct = np.linalg.solve(M, rhs)
    b = np.zeros(n ** 3)
    b[f] = -rhs  # solve() negates b (solver.cpp:231)
    x, _, st = po.solve(n, A, L, np.zeros(n ** 3), b, f64=True, precond=precond, tol_policy=1, tol=1e-12)
    assert st["iterations"] < len(f) + 5
    assert np.allclose(x.reshape(-1)[f], direct, rtol=1e-9, atol=1e-12)


def test_ic0_needs_fewer_iterations_than_reference_form():
    """The flag-keyed IC(0) factor couples cells; the reference's value-test
    form reduces to a diagonal scaling, which needs more CG iterations."""
    n = 16
    A, L = _box_system(n)
    b = np.zeros((n, n, n))
    b[A == po.FLUID] = np.random.RandomState(10).normal(size=(A == po.FLUID).sum())
    _, _, s0 = po.solve(n, A, L, np.zeros_like(b), b.copy(), f64=True, precond=0, tol_policy=1, tol=1e-8)
    _, _, s1 = po.solve(n, A, L, np.zeros_like(b), b.copy(), f64=True, precond=1, tol_policy=1, tol=1e-8)
    assert s1["iterations"] < s0["iterations"]


# ------------------------------------------------------------- reposition
def test_random_shuffle_matches_libstdcxx():
    """flip3D.cpp:433 std::random_shuffle with rand(): the oracle's and the
    library's restatement (swap a[i], a[rand() % (i+1)]) against libstdc++."""
    src = r'''
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>
int main() { srand(1); for (int n = 1; n <= 40; n += 13) { std::vector<int> v(n);
  for (int i = 0; i < n; ++i) v[i] = i; std::random_shuffle(v.begin(), v.end());
  for (int x : v) printf("%d ", x); printf("\n"); } return 0; }'''
    d = tempfile.mkdtemp()
    with open(os.path.join(d, "s.cpp"), "w") as fh:
        fh.write(src)
    subprocess.run(["g++", "-O1", "-std=c++11", "-o", os.path.join(d, "s"), os.path.join(d, "s.cpp")], check=True,
                   capture_output=True)
    out = subprocess.run([os.path.join(d, "s")], capture_output=True, text=True, check=True).stdout.split("\n")
    s = po.Sim(4)
    s.srand(1)
    for line, n in zip(out, range(1, 41, 13)):
        a = list(range(n))
        for i in range(1, n):
            j = s.rand() % (i + 1)
            a[i], a[j] = a[j], a[i]
        assert a == [int(x) for x in line.split()]


# ----------------------------------------------------------------- flags
def test_flags_closed_form():
    """sorter.cpp:70-93: WALL if any wall particle; otherwise FLUID iff the
    float sum of dens in bucket order exceeds 1.6 (L = 1.6 - sum < 0)."""
    gn = 3
    p = np.array([[0.5, 0.5, 0.5]] * 3 + [[0.1, 0.1, 0.1]] * 2 + [[0.9, 0.9, 0.9]], np.float32)
    t = np.array([1, 1, 1, 1, 2, 1], np.int8)
    dens = np.array([0.6, 0.6, 0.4000001, 5.0, 1.0, 1.5], np.float32)
    s = po.Sim(gn)
    s.set_particles(p, np.zeros_like(p), np.zeros_like(p), t, dens=dens)
    s.call("sort")
    s.call("mark_water")
    A = s.grid(po.G_A)
    assert A[1, 1, 1] == po.FLUID  # 0.6+0.6+0.4000001 > 1.6
    assert A[0, 0, 0] == po.WALL
    assert A[2, 2, 2] == po.AIR  # 1.5 < 1.6
    assert A[0, 1, 2] == po.AIR  # empty


def test_f64_oracle_tracks_f32_oracle():
    """The FP64 build runs the same step; over a few steps the two stay close
    (the f64 run is the error reference for the float path)."""
    a, b = po.Sim(16), po.Sim(16, f64=True)
    a.scene_benchmark()
    b.scene_benchmark()
    for _ in range(3):
        a.step()
        b.step()
    pa, pb = a.particles()["p"], b.particles()["p"]
    assert np.abs(pa - pb).max() < 1e-3


# ------------------------------------------------------- benchmark scene
_M64 = (1 << 64) - 1


def splitmix64_stream(seed, count):
    """SplitMix64 (Steele, Lea and Flood, OOPSLA 2014; Vigna's splitmix64.c):
    state += 0x9E3779B97F4A7C15, then the two xor-shift-multiply rounds."""
    out, x = [], seed
    for _ in range(count):
        x = (x + 0x9E3779B97F4A7C15) & _M64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        out.append(z ^ (z >> 31))
    return out


def splitmix64_at(seed, k):
    """Value k (0-based) of the stream seeded with `seed`, vectorised over k."""
    k = np.asarray(k, np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (k + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def test_splitmix64_published_values():
    """The generator's published outputs: seed 1234567 starts 6457827717110365317,
    3203168211198807973, ... (splitmix64.c test vector) and seed 0 starts
    0xE220A8397B1DCDAF; the vectorised form equals the stream."""
    assert splitmix64_stream(1234567, 5) == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                             4593380528125082431, 16408922859458223821]
    assert splitmix64_stream(0, 1) == [0xE220A8397B1DCDAF]
    assert [int(v) for v in splitmix64_at(20261203, np.arange(64))] == splitmix64_stream(20261203, 64)


def _hypotf(a, b):
    # glibc hypotf == (float)sqrt((double)a^2 + (double)b^2) (pinned above)
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.sqrt(a * a + b * b).astype(np.float32)


def benchmark_placement(n, seed=20261203, density=0.5):
    """The benchmark scene written independently of the oracle: the
    reference's sub-lattice loop (flip3D.cpp:711-727) with its object test
    (pushParticle, :580-633) for the liquid box {x<0.4, y<0.6} (FLOAT
    corners compared in double) and sphere (float hypotf distance to the
    float centre, < 0.12f), jitter U(-0.25, 0.25)*dx per axis from SplitMix64
    value 3*fluid_index + axis (53-bit uniform), then one WALL particle at
    the centre of every boundary cell (:730-740)."""
    w = density * (1.0 / n)
    lim = int(n / density)
    c = np.arange(lim) * w + w / 2.0
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    X, Y, Z = X.reshape(-1), Y.reshape(-1), Z.reshape(-1)
    wt = 1.0 / n
    inside = (X > wt) & (X < 1 - wt) & (Y > wt) & (Y < 1 - wt) & (Z > wt) & (Z < 1 - wt)
    box = (X > 0) & (X < np.float64(np.float32(0.4))) & (Y > 0) & (Y < np.float64(np.float32(0.6))) & (Z > 0) & (Z < 1)
    f32 = np.float32
    d0 = X.astype(f32) - f32(0.6)
    d1 = Y.astype(f32) - f32(0.7)
    d2 = Z.astype(f32) - f32(0.5)
    sphere = _hypotf(_hypotf(d0, d1), d2) < f32(0.12)
    keep = inside & (box | sphere)
    lat = np.stack([X[keep], Y[keep], Z[keep]], 1)
    m = len(lat)
    v = splitmix64_at(seed, np.arange(3 * m, dtype=np.uint64)).reshape(m, 3)
    uu = (v >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    jit = (uu * 0.5 - 0.25) * (1.0 / n)
    fluid = (lat + jit).astype(np.float32)
    cc = np.arange(n) * wt + wt / 2.0
    A, B, Cc = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    edge = (np.minimum(np.minimum(A, B), Cc) == 0) | (np.maximum(np.maximum(A, B), Cc) == n - 1)
    wall = np.stack([cc[A[edge]], cc[B[edge]], cc[Cc[edge]]], 1).astype(np.float32)
    return fluid, lat, wall


def test_benchmark_placement_matches_reference_lattice():
    """At N=32 the reference's own placement of the benchmark objects
    (tests/golden/scene_bench_ref.npz, from the compiled reference) keeps
    exactly the sub-lattice points and wall cells of the independent
    placement above, in the same order: 53,730 liquid + 5,768 wall."""
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "scene_bench_ref.npz"), allow_pickle=False)
    fluid, lat, wall = benchmark_placement(32)
    assert (len(fluid), len(wall)) == (53730, 5768) == (int(d["n_fluid"]), int(d["n_wall"]))
    w = 0.5 / 32
    assert np.array_equal(np.rint(lat / w - 0.5).astype(np.int16), d["fluid_lattice"])
    assert np.array_equal(wall, d["wall_p"])


@pytest.mark.parametrize("n,counts", [(32, (53730, 5768)), (128, (3951862, 96776))])
def test_oracle_benchmark_scene_is_pinned(n, counts):
    """The oracle's benchmark scene equals the independent placement bit for
    bit (liquid then walls, creation order), with the particle counts of the
    reference's placement (53,730 + 5,768 at N=32; 3,951,862 + 96,776 at
    N=128); every jitter stays within +-0.25 dx of its sub-point.  In exact
    arithmetic that keeps each particle in its sub-point's grid cell; the
    float rounding of position + jitter can land a particle exactly on the
    cell face above (23 particles at N=128, none at N=32), and only there."""
    fluid, lat, wall = benchmark_placement(n)
    assert (len(fluid), len(wall)) == counts
    s = po.Sim(n)
    try:
        m = s.scene_benchmark(20261203)
        q = s.particles()
    finally:
        s.close()
    assert m == sum(counts)
    t = q["type"]
    assert np.all(t[:counts[0]] == po.FLUID) and np.all(t[counts[0]:] == po.WALL)
    assert np.array_equal(q["p"][:counts[0]], fluid)
    assert np.array_equal(q["p"][counts[0]:], wall)
    off = np.abs(fluid.astype(np.float64) - lat) * n
    assert off.max() <= 0.25 + 1e-6 and off.max() > 0.24
    cell = lambda a: np.floor(np.asarray(a, np.float64) * n)  # noqa: E731
    moved = cell(fluid) != cell(lat)
    on_face = fluid.astype(np.float64) * n == cell(fluid)
    assert np.all(on_face[moved])
    assert moved.sum() == {32: 0, 128: 23}[n]


# ------------------------------------------------------------- CFL dt
@pytest.mark.parametrize("speed,wall_speed", [(0.0, 0.0), (0.5, 0.0), (20.0, 0.0), (55.5, 0.0), (0.5, 100.0)])
def test_cfl_time_step_closed_form(speed, wall_speed):
    """dt_policy 1: dt = min(dt_cap, cfl / (N max|u|)) over FLUID particles
    only (walls do not move), with |u| the double norm of the float
    components; dt_cap = 6e-3 (flip3D.cpp:31) and cfl = 0.6."""
    n = 8
    rs = np.random.RandomState(int(speed * 10 + wall_speed))
    m = 200
    p = rs.uniform(0.3, 0.7, size=(m, 3)).astype(np.float32)
    dirs = rs.normal(size=(m, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    scale = rs.uniform(0.0, 1.0, size=(m, 1))
    scale[17] = 1.0
    u = (dirs * scale * speed).astype(np.float32)
    t = np.full(m, po.FLUID, np.int8)
    t[5] = po.WALL
    u[5] = np.float32(wall_speed) * dirs[5].astype(np.float32)
    fl = t == po.FLUID
    umax = np.sqrt((u[fl].astype(np.float64) ** 2).sum(1).max())
    want = 6e-3 if umax == 0 else min(6e-3, 0.6 / (n * umax))
    s = po.Sim(n, dt_policy=1)
    try:
        s.set_particles(p, u, np.zeros_like(p), t, dens=np.ones(m, np.float32))
        st = s.step()
    finally:
        s.close()
    assert st["dt"] == want
    if speed >= 20.0:
        assert want < 6e-3  # the CFL branch is the one taken (0.6 / (8 * 12.5) = 6e-3)
