"""Train the coarsest-level network of FLIP_PRECOND_LEARNED on pressure
systems that the plain GPU solver produces.

    python -m flip3d_amd.train_unet --out weights.bin [--scenes 2000] [--steps 20000]

The network solves the coarsest level of the learned preconditioner's
multigrid hierarchy: a 32^3 system (16^3 for n = 32) rediscretised from the
pressure system by the hierarchy's rule (a coarse cell is FLUID if any child
is, WALL if all are, AIR otherwise; its level set is the mean over the FLUID
children, or over the non-WALL children of an AIR cell; its right-hand side is
the sum over the FLUID children).

Data (every target x solved by the library's reference-preconditioned PCG to
relative residual 1e-6 at the coarse size):
  * `--scenes` N=64 scenes of 1-4 random liquid boxes and spheres
    (SCENE_RANDOM, seed = scene index; every 50th the benchmark scene),
    stepped a random 2-24 steps; each sampled step's pressure system
    (A, L, b = -div) is coarsened once to 32^3;
  * `--coarse-scenes` N=128 scenes, coarsened twice to 32^3;
  * for every coarse system also a second right-hand side: random noise on
    the FLUID cells smoothed by masked neighbour averaging;
  * the right-hand sides the network actually receives: for every
    `--vcycle-every`-th N=64 scene and `--vcycle-coarse-every`-th N=128
    scene system, the first
    `--vcycle-iters` flexible-CG iterations of the preconditioner are
    replayed in numpy (the same hierarchy, damped-Jacobi smoothing and
    transfers as the GPU V-cycle) with the GPU solver as the coarsest
    solve, and each coarsest-level right-hand side becomes a system.
Training: Adam with the gradient clipped to norm 1, cosine learning rate,
random axis and sign flips, loss = relative A-norm error
(flip_unet_train_step).
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import flip3d_amd as fa  # noqa: E402

R = 32
TOL = 1e-6


def coarsen(A, L, b):
    """One level of the preconditioner's hierarchy: (n/2)^3 flags, level
    set and summed right-hand side."""
    n = A.shape[0]
    h = n // 2

    def blocks(a):
        return a.reshape(h, 2, h, 2, h, 2).transpose(0, 2, 4, 1, 3, 5).reshape(h, h, h, 8)

    Ar, Lr, br = blocks(A), blocks(L), blocks(b)
    fluid = Ar == fa.FLUID
    Ac = np.where(fluid.any(-1), fa.FLUID, np.where((Ar == fa.WALL).all(-1), fa.WALL, fa.AIR)).astype(np.int8)
    nonwall = Ar != fa.WALL
    cnt = nonwall.sum(-1)
    nfl = fluid.sum(-1)
    l_air = np.where(cnt > 0, (Lr * nonwall).sum(-1, dtype=np.float32) / np.maximum(cnt, 1), 1.0)
    l_fluid = (Lr * fluid).sum(-1, dtype=np.float32) / np.maximum(nfl, 1)
    Lc = np.where(nfl > 0, l_fluid, l_air).astype(np.float32)
    bc = (br * fluid).sum(-1, dtype=np.float32)
    return Ac, Lc, bc


def coarsen_to(A, L, b, side):
    while A.shape[0] > side:
        A, L, b = coarsen(A, L, b)
    return A, L, b


def smooth_noise(A, rs, passes=4):
    """Random right-hand side on the FLUID cells, smoothed by masked
    neighbour averaging."""
    fl = (A == fa.FLUID).astype(np.float32)
    v = rs.normal(size=A.shape).astype(np.float32) * fl
    for _ in range(passes):
        s = v.copy()
        c = np.ones_like(v)
        for ax in range(3):
            for d in (1, -1):
                s += np.roll(v, d, axis=ax)
                c += np.roll(fl, d, axis=ax)
        v = s / c * fl
    return v


# ---------------------------------------------------------------- V-cycle replay
MG_W = 0.9   # damped Jacobi weight of the GPU V-cycle (the shipped weights were trained with 0.8)
MG_S = 10.0  # its coarse-correction scale


def level_operator(A, L):
    """Sparse ghost-fluid operator of a level (spacing 1/n) on its FLUID
    cells, with the FLUID index map."""
    import scipy.sparse as sp

    n = A.shape[0]
    M = unit_operator(A, L, sp) * float(n * n)
    idx = -np.ones(A.shape, np.int64)
    idx[A == fa.FLUID] = np.arange(int((A == fa.FLUID).sum()))
    return M.tocsr(), idx


class Hierarchy:
    """The GPU V-cycle's levels for one system: fine level, then halvings
    down to side R (the network's level)."""

    def __init__(self, A, L):
        import scipy.sparse as sp

        self.levels = []
        while True:
            M, idx = level_operator(A, L)
            self.levels.append({"A": A, "L": L, "M": M, "D": M.diagonal(), "idx": idx})
            if A.shape[0] <= R:
                break
            A, L, _ = coarsen(A, L, np.zeros(A.shape, np.float32))
        for f, c in zip(self.levels[:-1], self.levels[1:]):
            fl = np.argwhere(f["A"] == fa.FLUID)
            par = fl // 2
            ci = c["idx"][par[:, 0], par[:, 1], par[:, 2]]
            f["R"] = sp.csr_matrix((np.ones(len(fl)), (ci, np.arange(len(fl)))), shape=(c["M"].shape[0], len(fl)))

    def vcycle(self, r, coarse):
        def rec(l, b):
            lv = self.levels[l]
            if l == len(self.levels) - 1:
                return coarse(lv, b)
            x = np.zeros_like(b)
            for _ in range(2):
                x = x + MG_W * (b - lv["M"] @ x) / lv["D"]
            xc = rec(l + 1, lv["R"] @ (b - lv["M"] @ x))
            x = x + (lv["R"].T @ xc) / MG_S
            for _ in range(2):
                x = x + MG_W * (b - lv["M"] @ x) / lv["D"]
            return x

        return rec(0, r)


def vcycle_systems(A, L, b, solver, iters):
    """Coarsest-level systems met by the first `iters` flexible-CG
    iterations on (A, L, b), the coarsest solves done by the GPU solver."""
    h = Hierarchy(A, L)
    out = []
    K = h.levels[-1]

    def coarse(lv, bk):
        g = np.zeros(lv["A"].shape, np.float32)
        g[lv["A"] == fa.FLUID] = bk
        x, _, _ = solver.solve(lv["A"], lv["L"], np.zeros_like(g), -g)
        out.append((lv["A"], lv["L"], g, x.astype(np.float32)))
        return x[lv["A"] == fa.FLUID].astype(np.float64)

    M = h.levels[0]["M"]
    r = b[A == fa.FLUID].astype(np.float64)
    x = np.zeros_like(r)
    z = h.vcycle(r, coarse)
    s, zr = z.copy(), z @ r
    for _ in range(iters - 1):
        q = M @ s
        a = zr / (s @ q)
        x += a * s
        r -= a * q
        zn = h.vcycle(r, coarse)
        beta = -a * (zn @ q) / zr
        zr = zn @ r
        s = zn + beta * s
    assert K["A"].shape[0] == R
    return out


def scene_systems(n, scene, seed, steps_at):
    out = []
    with fa.FlipContext(n, tol_policy=1, tol=TOL, dt_policy=1) as ctx:
        ctx.build_scene(scene, seed)
        ctx.srand(1)
        for s in range(1, max(steps_at) + 1):
            st = ctx.step()
            if s in steps_at and st["n_fluid_cells"] > 0:
                out.append((ctx.grid(fa.GRID_A), ctx.grid(fa.GRID_L), ctx.grid(fa.GRID_DIV)))
    return out


def generate(n_scenes, n_coarse_scenes, vc_every, vc_coarse_every, vc_iters, log):
    """(A, L, b, x) systems at side R: b as restricted (or as the V-cycle
    restricts it), x = A^-1 b."""
    rs = np.random.RandomState(12345)
    data = []
    t0 = time.time()
    with fa.FlipContext(R, tol_policy=1, tol=TOL) as solver:
        def add(A, L, div):
            # the step leaves div negated in place: the system's b is div
            Ac, Lc, bc = coarsen_to(A, L, div, R)
            if not (Ac == fa.FLUID).any():
                return
            for rhs in (bc, smooth_noise(Ac, rs)):
                if not np.any(rhs):
                    continue
                x, _, _ = solver.solve(Ac, Lc, np.zeros_like(rhs), -rhs)
                data.append((Ac, Lc, rhs.astype(np.float32), x.astype(np.float32)))

        nvc = 0
        for sc in range(n_scenes):
            steps = sorted(set(rs.randint(2, 25, size=2).tolist()))
            scene = fa.SCENE_BENCHMARK if sc % 50 == 0 else fa.SCENE_RANDOM
            for A, L, div in scene_systems(2 * R, scene, sc, steps):
                add(A, L, div)
                if vc_every > 0 and sc % vc_every == 0:
                    got = vcycle_systems(A, L, div, solver, vc_iters)
                    data += got
                    nvc += len(got)
            if sc % 100 == 0:
                log(f"N=64 scenes {sc}/{n_scenes}: {len(data)} systems ({nvc} V-cycle), {time.time() - t0:.0f} s")
        n64 = len(data)
        for sc in range(n_coarse_scenes):
            scene = fa.SCENE_BENCHMARK if sc % 4 == 0 else fa.SCENE_RANDOM
            steps = sorted(set(rs.randint(2, 30, size=3).tolist()))
            for A, L, div in scene_systems(4 * R, scene, 100000 + sc, steps):
                add(A, L, div)
                if vc_coarse_every > 0 and sc % vc_coarse_every == 0:
                    got = vcycle_systems(A, L, div, solver, vc_iters)
                    data += got
                    nvc += len(got)
            if sc % 20 == 0:
                log(f"N=128 scenes {sc}/{n_coarse_scenes}: {len(data) - n64} systems, {time.time() - t0:.0f} s")
    log(f"{n64} systems from N=64 scenes, {len(data) - n64} from N=128 scenes, {nvc} of them V-cycle right-hand sides")
    return data


def flip_axes(a, axes):
    for ax in axes:
        a = np.flip(a, axis=ax)
    return np.ascontiguousarray(a)


def train(data, steps, batch, lr0, seed, log, out):
    rs = np.random.RandomState(seed)
    net = fa.UNet(R, capacity=batch, seed=seed)
    A_all = np.stack([d[0] for d in data])
    L_all = np.stack([d[1] for d in data])
    b_all = np.stack([d[2] for d in data])
    x_all = np.stack([d[3] for d in data])
    t0 = time.time()
    ema = None
    for it in range(steps):
        idx = rs.randint(0, len(data), batch)
        axes = [ax for ax in (1, 2, 3) if rs.uniform() < 0.5]
        sgn = np.float32(1.0 if rs.uniform() < 0.5 else -1.0)
        A = flip_axes(A_all[idx], axes)
        L = flip_axes(L_all[idx], axes)
        b = flip_axes(b_all[idx], axes) * sgn
        x = flip_axes(x_all[idx], axes) * sgn
        lr = lr0 * 0.5 * (1.0 + np.cos(np.pi * it / steps)) + 1e-6
        loss = net.train_step(A, L, b, x, lr)
        ema = loss if ema is None else 0.98 * ema + 0.02 * loss
        if it % 500 == 0 or it == steps - 1:
            log(f"step {it}: loss {loss:.4f} (ema {ema:.4f}) lr {lr:.2e} {time.time() - t0:.0f} s")
        if it % 5000 == 4999:
            net.save(out)
    return net


def evaluate(net, data, log, what):
    """Relative A-norm error of the prediction on held-out systems."""
    import scipy.sparse as sp

    errs = []
    for A, L, b, x in data:
        pred = net.predict(A[None], L[None], b[None])[0]
        e = (pred - x)[A == fa.FLUID].astype(np.float64)
        t = x[A == fa.FLUID].astype(np.float64)
        M = unit_operator(A, L, sp)
        errs.append(float(np.sqrt((e @ (M @ e)) / max(t @ (M @ t), 1e-300))))
    log(f"held out ({what}): relative A-norm error mean {np.mean(errs):.3f}, max {np.max(errs):.3f} "
        f"over {len(errs)} systems")
    return errs


def unit_operator(A, L, sp):
    """The unit-spacing ghost-fluid operator on the FLUID cells."""
    n = A.shape[0]
    fl = np.argwhere(A == fa.FLUID)
    idx = -np.ones(A.shape, np.int64)
    idx[A == fa.FLUID] = np.arange(len(fl))
    rows, cols, vals = [], [], []
    diag = np.full(len(fl), 6.0)
    Ls = L[A == fa.FLUID].astype(np.float64)
    for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        q = fl + np.array(d)
        out = ((q < 0) | (q >= n)).any(1)
        qc = np.clip(q, 0, n - 1)
        Aq = A[qc[:, 0], qc[:, 1], qc[:, 2]]
        nbr = ~out & (Aq == fa.FLUID)
        rows.append(np.flatnonzero(nbr))
        cols.append(idx[qc[nbr, 0], qc[nbr, 1], qc[nbr, 2]])
        vals.append(-np.ones(nbr.sum()))
        diag[out | (Aq == fa.WALL)] -= 1.0
        air = ~out & (Aq == fa.AIR)
        diag[air] -= L[qc[air, 0], qc[air, 1], qc[air, 2]].astype(np.float64) / np.minimum(1e-6, Ls[air])
    rows.append(np.arange(len(fl)))
    cols.append(np.arange(len(fl)))
    vals.append(diag)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(len(fl),) * 2)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--coarse-scenes", type=int, default=60)
    ap.add_argument("--vcycle-every", type=int, default=8)
    ap.add_argument("--vcycle-coarse-every", type=int, default=3)
    ap.add_argument("--vcycle-iters", type=int, default=6)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args()

    def log(msg):
        print(msg, flush=True)

    data = generate(args.scenes, args.coarse_scenes, args.vcycle_every, args.vcycle_coarse_every, args.vcycle_iters,
                    log)
    rs = np.random.RandomState(7)
    data = [data[i] for i in rs.permutation(len(data))]
    nv = max(8, len(data) // 50)
    valid, data = data[:nv], data[nv:]
    log(f"{len(data)} training systems, {nv} held out")
    net = train(data, args.steps, args.batch, args.lr, args.seed, log, args.out)
    evaluate(net, valid, log, "coarsened scene and smoothed-noise systems")
    net.save(args.out)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
