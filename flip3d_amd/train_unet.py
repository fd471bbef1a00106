"""Train the coarse-level network of FLIP_PRECOND_LEARNED on pressure systems
that the plain GPU solver produces.

    python -m flip3d_amd.train_unet --out weights.bin [--scenes 2000] [--steps 4000]

Data (all solved by the library's reference-preconditioned PCG to rel 1e-6):
  * N=64 scenes of 1-4 random liquid boxes and spheres (SCENE_RANDOM, seed =
    scene index) stepped a random 2-24 steps; each snapshot is the step's
    pressure system (flags A, level set L, right-hand side b = -div, x);
  * coarse systems of N=128 scenes (benchmark and random, random steps): the (n/2)^3 rediscretisation the
    preconditioner builds (FLUID if any child is FLUID, WALL if all are; level
    set = mean over FLUID children, or over non-WALL children of an AIR cell)
    with right-hand side the sum of the children's b, solved at n=64.
Training: Adam on batches of 8 systems (half from each set), random axis
flips and sign flips,
loss = relative A-norm error (flip_unet_train_step), cosine learning rate.
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

R = 64
TOL = 1e-6


def coarsen(A, L, b):
    """(n/2)^3 system of the learned preconditioner's coarse level."""
    n = A.shape[0]
    h = n // 2
    Ar = A.reshape(h, 2, h, 2, h, 2).transpose(0, 2, 4, 1, 3, 5).reshape(h, h, h, 8)
    Lr = L.reshape(h, 2, h, 2, h, 2).transpose(0, 2, 4, 1, 3, 5).reshape(h, h, h, 8)
    br = b.reshape(h, 2, h, 2, h, 2).transpose(0, 2, 4, 1, 3, 5).reshape(h, h, h, 8)
    fluid = (Ar == fa.FLUID)
    Ac = np.where(fluid.any(-1), fa.FLUID, np.where((Ar == fa.WALL).all(-1), fa.WALL, fa.AIR)).astype(np.int8)
    nonwall = Ar != fa.WALL
    cnt = nonwall.sum(-1)
    nfl = fluid.sum(-1)
    l_air = np.where(cnt > 0, (Lr * nonwall).sum(-1, dtype=np.float32) / np.maximum(cnt, 1), 1.0)
    l_fluid = (Lr * fluid).sum(-1, dtype=np.float32) / np.maximum(nfl, 1)
    Lc = np.where(nfl > 0, l_fluid, l_air).astype(np.float32)
    rc = (br * fluid).sum(-1, dtype=np.float32)
    return Ac, Lc, rc


def scene_systems(n, scene, seed, steps_at, log):
    out = []
    with fa.FlipContext(n, tol_policy=1, tol=TOL, dt_policy=1) as ctx:
        ctx.build_scene(scene, seed)
        ctx.srand(1)
        last = max(steps_at)
        for s in range(1, last + 1):
            st = ctx.step()
            if s in steps_at and st["n_fluid_cells"] > 0:
                A = ctx.grid(fa.GRID_A)
                L = ctx.grid(fa.GRID_L)
                b = ctx.grid(fa.GRID_DIV)
                x = ctx.grid(fa.GRID_PRESS)
                out.append((A, L, b, x))
    return out


def generate(n_scenes, n_coarse_scenes, log):
    rs = np.random.RandomState(12345)
    data = []
    t0 = time.time()
    for sc in range(n_scenes):
        steps = sorted(set(rs.randint(2, 25, size=2).tolist()))
        scene = fa.SCENE_BENCHMARK if sc % 50 == 0 else fa.SCENE_RANDOM
        data += scene_systems(R, scene, sc, steps, log)
        if sc % 100 == 0:
            log(f"scenes {sc}/{n_scenes}: {len(data)} systems, {time.time() - t0:.0f} s")
    # coarse systems of N=128 scenes
    coarse = []
    with fa.FlipContext(R, tol_policy=1, tol=TOL) as solver:
        for sc in range(n_coarse_scenes):
            scene = fa.SCENE_BENCHMARK if sc % 4 == 0 else fa.SCENE_RANDOM
            steps = sorted(set(rs.randint(2, 30, size=3).tolist()))
            for A, L, b, _ in scene_systems(2 * R, scene, 100000 + sc, steps, log):
                Ac, Lc, rc = coarsen(A, L, b)
                if not (Ac == fa.FLUID).any():
                    continue
                x, _, st = solver.solve(Ac, Lc, np.zeros_like(rc), -rc)
                coarse.append((Ac, Lc, rc, x.astype(np.float32)))
            if sc % 20 == 0:
                log(f"coarse scenes {sc}/{n_coarse_scenes}: {len(coarse)} systems, {time.time() - t0:.0f} s")
        log(f"coarse systems from N={2 * R}: {len(coarse)}")
    return data, coarse


def flip_axes(a, axes):
    for ax in axes:
        a = np.flip(a, axis=ax)
    return np.ascontiguousarray(a)


def train(fine, coarse, steps, batch, lr0, seed, log, out):
    """Batches draw half their systems from each set."""
    rs = np.random.RandomState(seed)
    net = fa.UNet(R, capacity=batch, seed=seed)
    data = fine + coarse
    A_all = np.stack([d[0] for d in data])
    L_all = np.stack([d[1] for d in data])
    b_all = np.stack([d[2] for d in data])
    x_all = np.stack([d[3] for d in data])
    nf = len(fine)
    t0 = time.time()
    ema = None
    for it in range(steps):
        half = batch // 2 if coarse else batch
        idx = np.concatenate([rs.randint(0, nf, half), nf + rs.randint(0, max(len(coarse), 1), batch - half)]) \
            if coarse else rs.randint(0, nf, batch)
        axes = [ax for ax in (1, 2, 3) if rs.uniform() < 0.5]
        sgn = np.float32(1.0 if rs.uniform() < 0.5 else -1.0)
        A = flip_axes(A_all[idx], axes)
        L = flip_axes(L_all[idx], axes)
        b = flip_axes(b_all[idx], axes) * sgn
        x = flip_axes(x_all[idx], axes) * sgn
        lr = lr0 * 0.5 * (1.0 + np.cos(np.pi * it / steps)) + 1e-5
        loss = net.train_step(A, L, b, x, lr)
        ema = loss if ema is None else 0.98 * ema + 0.02 * loss
        if it % 500 == 0 or it == steps - 1:
            log(f"step {it}: loss {loss:.4f} (ema {ema:.4f}) lr {lr:.2e} {time.time() - t0:.0f} s")
        if it % 5000 == 4999:
            net.save(out)
    return net


def evaluate(net, data, log):
    """Relative A-norm error of the prediction on held-out systems, with
    the unit operator applied through the training loss of a zero-lr step
    is avoided: computed here from the prediction directly."""
    errs = []
    for A, L, b, x in data:
        pred = net.predict(A[None], L[None], b[None])[0]
        e = (pred - x)[A == fa.FLUID].astype(np.float64)
        t = x[A == fa.FLUID].astype(np.float64)
        errs.append(float(np.linalg.norm(e) / max(np.linalg.norm(t), 1e-30)))
    log(f"held-out relative L2 error: mean {np.mean(errs):.3f}, max {np.max(errs):.3f} over {len(errs)} systems")
    return errs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--coarse-scenes", type=int, default=60)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args()

    def log(msg):
        print(msg, flush=True)

    fine, coarse = generate(args.scenes, args.coarse_scenes, log)
    rs = np.random.RandomState(7)
    fine = [fine[i] for i in rs.permutation(len(fine))]
    coarse = [coarse[i] for i in rs.permutation(len(coarse))]
    nvf, nvc = max(4, len(fine) // 50), max(2, len(coarse) // 20)
    valid_f, fine = fine[:nvf], fine[nvf:]
    valid_c, coarse = coarse[:nvc], coarse[nvc:]
    log(f"{len(fine)} + {len(coarse)} training systems, {nvf} + {nvc} held out")
    net = train(fine, coarse, args.steps, args.batch, args.lr, args.seed, log, args.out)
    log("held out, N=64 scene systems:")
    evaluate(net, valid_f, log)
    log("held out, coarse systems of N=128 scenes:")
    evaluate(net, valid_c, log)
    net.save(args.out)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
