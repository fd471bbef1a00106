"""Batched ensemble of independent scenes on one GPU.

    python -m flip3d_amd.ensemble_bench [--scenes 64] [--steps 20] [--threads 8]

64 scenes with N drawn uniformly from {24..96} (seeded), each holding 1-3
random liquid boxes (FLIP_SCENE_RANDOM_BOXES, seed = scene index) at 8
particles per cell, CFL dt, pressure solve to relative residual 1e-5.
Every scene uses the learned preconditioner (--no-learned: the
reference-preconditioned PCG), with the solver launch shape the autotuner
chose for its size (flip3d_amd/tuning.json; --no-tuning: the built-in one).
Every step advances all scenes at once.  --mode threads (the default):
flip_gpu_step_many, host threads stepping the contexts concurrently, each on
its own HIP stream, every step replaying its own captured parts around the
solve.  --mode batch: flip_gpu_step_batch, the kernels before the pressure
solves of all scenes in one graph launch (each scene's captured part a
child graph), the solves concurrently on the scenes' streams, then the
kernels after the solves in one more graph launch (slower here: every
scene waits at each phase for the largest).

Prints one JSON line: scene-steps/s over the timed steps, and the p50/p99 of
the per-scene step latency (wall clock from the start of a scene's step to
its end, inside the concurrent batch).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import flip3d_amd as fa  # noqa: E402
from flip3d_amd import autotune  # noqa: E402

WEIGHTS = os.path.join(ROOT, "flip3d_amd", "weights", "unet_coarse.bin")
SEED = 20261203


def ensemble_sizes(scenes, seed=SEED):
    """The ensemble's grid sizes: N uniform in {24..96}, seeded."""
    return [int(n) for n in np.random.RandomState(seed).randint(24, 97, size=scenes)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--seed", type=int, default=SEED)
    ap.add_argument("--mode", choices=["threads", "batch"], default="threads")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-tuning", action="store_true",
                    help="built-in solver launch shapes instead of flip3d_amd/tuning.json's")
    args = ap.parse_args()
    table = None if args.no_tuning else autotune.load()

    sizes = np.array(ensemble_sizes(args.scenes, args.seed))
    ctxs, particles, learned = [], 0, 0
    for i, n in enumerate(sizes):
        use_learned = not args.no_learned
        pc = fa.PRECOND_LEARNED if use_learned else fa.PRECOND_REFERENCE
        launch = autotune.launch_for(table, "learned" if use_learned else "reference", n)
        c = fa.FlipContext(int(n), tol_policy=1, tol=1e-5, dt_policy=1, precond=pc)
        c.set_launch(*launch)
        if use_learned:
            c.set_learned_weights(WEIGHTS)
            learned += 1
        particles += c.build_scene(fa.SCENE_RANDOM_BOXES, i)
        c.srand(1)
        ctxs.append(c)
    step = fa.step_batch if args.mode == "batch" else fa.step_many
    for _ in range(args.warmup):
        step(ctxs, args.threads)
    lat = []
    iters = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st, ms = step(ctxs, args.threads)
        lat += ms
        iters += [s["iterations"] for s in st]
    elapsed = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    out = {
        "metric": "ensemble scene-steps/s", "value": round(args.scenes * args.steps / elapsed, 2),
        "unit": "scene-steps/s", "scenes": args.scenes, "steps": args.steps, "threads": args.threads,
        "mode": args.mode,
        "ms_per_batch_step": round(1e3 * elapsed / args.steps, 3),
        "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
        "particles": int(particles), "n_min": int(sizes.min()), "n_max": int(sizes.max()),
        "learned_scenes": learned, "pcg_iterations_mean": round(float(np.mean(iters)), 1),
        "tuned": table is not None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
