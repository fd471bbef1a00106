"""FLIP steps/s at 128^3 (~4M particles) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]

Workload: the benchmark dam-break scene (box x<0.4, y<0.6 plus sphere r=0.12 at
(0.6,0.7,0.5); 8 particles per liquid cell, SplitMix64 jitter seed 20261203),
CFL-capped dt, pressure solve to relative residual 1e-5, fp32 fields with fp64
PCG reductions.  A step is one full simulateStep (sort, density, P2G, flags,
projection, extrapolation, FLIP G2P, advection with wall push-out and
repositioning, corrector), with all state resident in HBM.

Prints one JSON line with the throughput, the roofline of the dominant kernel
(live HIP-event durations of its launches) and the CPU oracle timed on this
host on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TOL = 1e-5
METRIC = "FLIP steps/s at 128^3 (~4M particles), learned vs plain PCG, rel resid 1e-5"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pcg_a_n128_traffic.txt")
DEFAULT_WEIGHTS = os.path.join(ROOT, "flip3d_amd", "weights", "unet_coarse.bin")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the same counter on a copy of
    known size); None when no summary for it exists."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return int(float(parts[2]))
    return None


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB through the HIP runtime's copy
    kernel: the known-size reference a counter run scales FETCH_SIZE and
    WRITE_SIZE by."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    size = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    for rc in (hip.hipMalloc(C.byref(a), C.c_size_t(size)), hip.hipMalloc(C.byref(b), C.c_size_t(size)),
               hip.hipMemset(a, 1, C.c_size_t(size))):
        if rc != 0:
            raise RuntimeError(f"copy probe: HIP error {rc}")
    for _ in range(3):
        if hip.hipMemcpy(b, a, C.c_size_t(size), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("copy probe: hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def kernel_bytes(name, n_particles, n_cells, n_faces, n_fluid):
    """Algorithmic bytes per launch (unique reads + writes, neighbour re-reads
    cached) of the kernels bench reports on; per-unit figures in DESIGN.md."""
    per = {
        # PCG iteration kernels, per FLUID cell
        "pcg_a": 25 * n_fluid,    # fl 4, A 1, L 4, z 4, s_old 4 -> s_new 4, q 4
        "pcg_b": 41 * n_fluid,    # fl 4, A 1, x 4, r 4, s 4, q 4, P 8 -> x 4, r 4, z 4
        # particle gathers, per particle (float4 pos/vel = 16 B)
        "density": 16 * n_particles + 4 * n_particles + 4 * n_cells,
        "p2g": 32 * n_particles + 4 * n_cells + 2 * 4 * n_faces + n_cells,
        "spring": 16 * n_particles + 4 * n_particles + 16 * n_particles + 4 * n_cells,
        "resample": 16 * n_particles * 3 + 16 * n_particles + 4 * n_cells,
        "g2p_flip": 32 * n_particles + 16 * n_particles + 24 * n_faces,
        "advect": 32 * n_particles + 32 * n_particles + 12 * n_faces,
    }
    return per.get(name)


def cpu_baseline(n, seconds_budget):
    """The oracle (C restatement of the reference path, OpenMP where the
    reference uses it) on this host: the first step of the same scene."""
    from oracle import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    s = po.Sim(n, tol_policy=1, tol=TOL, dt_policy=1)
    s.scene_benchmark()
    s.srand(1)
    t0 = time.perf_counter()
    st = s.step()
    dt = time.perf_counter() - t0
    s.close()
    return {
        "value": 1.0 / dt, "unit": "steps/s", "cores": threads, "kind": "port",
        "sample": f"oracle step 1 of the N={n} benchmark scene from its initial state "
                  f"({st['iterations']} PCG iterations, {dt:.1f} s)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--precond", default="reference", choices=["reference", "ic0", "learned"])
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS,
                    help="coarse-level network weights for --precond learned")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the slab decomposition is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import flip3d_amd as fa

    precond = {"reference": fa.PRECOND_REFERENCE, "ic0": fa.PRECOND_IC0, "learned": fa.PRECOND_LEARNED}[args.precond]
    ctx = fa.FlipContext(args.n, tol_policy=1, tol=TOL, dt_policy=1, precond=precond)
    if args.precond == "learned":
        ctx.set_learned_weights(args.weights)
    n_particles = ctx.build_scene(fa.SCENE_BENCHMARK)
    ctx.srand(1)
    for _ in range(args.warmup):
        ctx.step()
    ctx.synchronize()
    iters, solve_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = ctx.step()
        iters.append(st["iterations"])
        solve_ms.append(st["ms_solve"])
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    n_fluid = st["n_fluid_cells"]

    # one more (untimed) step with every launch bracketed by HIP events
    ctx.set_profiling(True)
    pst = ctx.step()
    ktimes = ctx.kernel_times()
    ctx.set_profiling(False)
    ctx.close()
    agg = {}
    for name, ms in ktimes:
        tot, cnt = agg.get(name, (0.0, 0))
        agg[name] = (tot + ms, cnt + 1)
    n = args.n
    n_cells, n_faces = n ** 3, (n + 1) * n * n
    ranked = sorted(((tot, name, cnt) for name, (tot, cnt) in agg.items()), reverse=True)
    dom = next((r for r in ranked if kernel_bytes(r[1], n_particles, n_cells, n_faces, n_fluid)), None)
    roofline = None
    if dom:
        tot, name, cnt = dom
        avg_ms = tot / cnt
        b = kernel_bytes(name, n_particles, n_cells, n_faces, n_fluid)
        achieved = b / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(name),
                    "bytes_per_launch": b, "avg_launch_ms": round(avg_ms, 5), "launches_per_step": cnt}

    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(n, 30)

    value = args.steps / elapsed
    out = {
        "metric": METRIC,
        "value": round(value, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"benchmark dam-break scene N={n}, {n_particles} particles "
                               f"({n_fluid} fluid cells), CFL dt, PCG to rel 1e-5, {args.precond} preconditioner",
                   "n": n, "particles": n_particles, "precond": args.precond, "tol": TOL},
        "particles_per_s": round(value * n_particles, 1),
        "pcg_iterations_mean": round(sum(iters) / len(iters), 1),
        "solve_ms_mean": round(sum(solve_ms) / len(solve_ms), 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "profile_step_ms": {k: round(v[0], 4) for k, v in sorted(agg.items(), key=lambda kv: -kv[1][0])},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
