"""Launch-shape autotuner for the pressure solver's FLUID-list kernels.

    python -m flip3d_amd.autotune [--sizes 24,32,...] [--precond learned,reference]
                                  [--steps 4] [--out flip3d_amd/tuning.json]

For every grid size N and solver, a scene is stepped from the same initial
state once per candidate launch shape: the benchmark dam break at N = 128
and above (the metric's workload), elsewhere seeded random liquid boxes
(the ensemble's kind of scene).  The shapes are the kernels' block count
(the built-in min(ceil(N^3 / 1024), 256) and smaller caps) times the PCG
iterations per captured batch.  Each candidate runs 2 warm-up steps and `steps` timed ones
(wall clock around synchronized steps); the shape with the lowest mean step
time is written to the table with the default's time beside it.  The table
(flip3d_amd/tuning.json) is what bench.py and the ensemble bench apply
through FlipContext.set_launch; launch_for() looks a size up.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "flip3d_amd", "tuning.json")
WEIGHTS = os.path.join(ROOT, "flip3d_amd", "weights", "unet_coarse.bin")
BLOCK_CAPS = (0, 128, 64, 32)   # 0: the built-in count
BATCHES = (8, 4, 16)            # 8: the built-in batch
SIZES = (24, 32, 40, 48, 56, 64, 72, 80, 88, 96, 128)


def default_blocks(n):
    return min((n ** 3 + 1023) // 1024, 256)


def candidates(n):
    """(pcg_blocks, pcg_batch) pairs for size n: caps below the built-in
    block count only (0 stands for the built-in count)."""
    d = default_blocks(n)
    return [(b, t) for b in BLOCK_CAPS if b == 0 or b < d for t in BATCHES]


def load(path=TABLE):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def launch_for(table, precond, n):
    """(pcg_blocks, pcg_batch) for a context of size n from the table: the
    entry of the nearest tuned size for that solver; (0, 0) without one."""
    if not table or precond not in table.get("shapes", {}):
        return 0, 0
    entries = table["shapes"][precond]
    if not entries:
        return 0, 0
    best = min(entries, key=lambda k: (abs(int(k) - n), int(k)))
    e = entries[best]
    return int(e["pcg_blocks"]), int(e["pcg_batch"])


def scene_for(n):
    return "benchmark" if n >= 128 else "random_boxes"


def time_shape(fa, n, precond, shape, steps, seed):
    code = fa.PRECOND_LEARNED if precond == "learned" else fa.PRECOND_REFERENCE
    with fa.FlipContext(n, precond=code, dt_policy=1, tol_policy=1, tol=1e-5) as g:
        if precond == "learned":
            g.set_learned_weights(WEIGHTS)
        g.set_launch(*shape)
        if scene_for(n) == "benchmark":
            g.build_scene(fa.SCENE_BENCHMARK)
        else:
            g.build_scene(fa.SCENE_RANDOM_BOXES, seed)
        g.srand(1)
        for _ in range(2):
            g.step()
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.step()
        g.synchronize()
        return 1e3 * (time.perf_counter() - t0) / steps


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--sizes", default=",".join(str(s) for s in SIZES))
    ap.add_argument("--precond", default="learned,reference")
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--merge", action="store_true", help="update the entries of an existing --out table")
    a = ap.parse_args(argv)
    sys.path.insert(0, ROOT)
    import flip3d_amd as fa
    sizes = [int(s) for s in a.sizes.split(",") if s]
    cmd = "python -m flip3d_amd.autotune " + " ".join(argv if argv is not None else sys.argv[1:])
    table = load(a.out) if a.merge else None
    if table:
        table["command"] = table.get("command", "") + " ; " + cmd
    else:
        table = {"command": cmd, "steps": a.steps, "seed": a.seed, "shapes": {}}
    for precond in a.precond.split(","):
        out = table["shapes"].setdefault(precond, {})
        for n in sizes:
            times = {}
            for shape in candidates(n):
                times[shape] = time_shape(fa, n, precond, shape, a.steps, a.seed)
            best = min(times, key=times.get)
            blocks = best[0] if best[0] else default_blocks(n)
            out[str(n)] = {"scene": scene_for(n), "pcg_blocks": blocks, "pcg_batch": best[1],
                           "ms": round(times[best], 4),
                           "default_ms": round(times[(0, 8)], 4),
                           "swept": {f"{b or default_blocks(n)}x{t}": round(v, 4) for (b, t), v in times.items()}}
            print(f"[autotune] {precond} N={n}: blocks {blocks} batch {best[1]} {times[best]:.3f} ms "
                  f"(built-in {times[(0, 8)]:.3f} ms)", file=sys.stderr, flush=True)
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps({k: v for k, v in table.items() if k != "shapes"}))


if __name__ == "__main__":
    main()
