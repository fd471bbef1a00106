"""FLIP steps/s at 128^3 (~4M particles) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]

Workload: the benchmark dam-break scene (box x<0.4, y<0.6 plus sphere r=0.12 at
(0.6,0.7,0.5); 8 particles per liquid cell, SplitMix64 jitter seed 20261203),
CFL-capped dt, pressure solve to relative residual 1e-5, fp32 fields with fp64
PCG reductions.  A step is one full simulateStep (sort, density, P2G, flags,
projection, extrapolation, FLIP G2P, advection with wall push-out and
repositioning, corrector), with all state resident in HBM.

The reported value is the learned solver's (FLIP_PRECOND_LEARNED, flexible CG
with the multigrid/U-Net preconditioner; --precond selects another), with
the plain reference-preconditioned PCG timed beside it from the same initial
state.  Prints one JSON line with both throughputs, the roofline of the
dominant kernel of the reported run (live HIP-event durations of its
launches) and the CPU oracle timed on this host on a bounded sample of the
same workload.
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
PROFILES = os.path.join(ROOT, "profiles")
DEFAULT_WEIGHTS = os.path.join(ROOT, "flip3d_amd", "weights", "unet_coarse.bin")


def measured_traffic(kernel, n=128):
    """HBM bytes per launch of `kernel` from the committed counter summary
    profiles/<kernel>_n<n>_traffic.txt (FETCH_SIZE + WRITE_SIZE, each scaled
    by the same counter on a copy of known size); None when there is none."""
    path = os.path.join(PROFILES, f"{kernel}_n{n}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
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
        # flexible-CG kernels of the learned solver, per FLUID cell
        "fcg_a": 28 * n_fluid,    # fl 4, meta 4, L 4 (AIR-adjacent), (z, s) 8 -> (z, s) 8, q 4
        "fcg_b": 28 * n_fluid,    # fl 4, x 4, s 4, r 4, q 4 -> x 4, r 4
        # particle gathers, per particle (float4 pos/vel = 16 B); density reads
        # its windows from the sort's 12-byte (x, y, z) records
        "density": 16 * n_particles + 12 * n_particles + 4 * n_particles + 4 * n_cells,
        # density with add_ExtForce fused in: the velocity is read and written whole
        "density_force": 16 * n_particles + 12 * n_particles + 32 * n_particles + 4 * n_cells,
        "p2g": 32 * n_particles + 4 * n_cells + 2 * 4 * n_faces + n_cells,
        "spring": 16 * n_particles + 4 * n_particles + 16 * n_particles + 4 * n_cells,
        # own pos, pseudo-position, window records (12 B), window velocities -> tmp1
        "resample": 16 * n_particles * 2 + 12 * n_particles + 16 * n_particles + 16 * n_particles + 4 * n_cells,
        "g2p_flip": 32 * n_particles + 16 * n_particles + 24 * n_faces,
        "advect": 32 * n_particles + 32 * n_particles + 12 * n_faces,
    }
    return per.get(name)


def mesh_leg(fa, n, weights, cpu=True):
    """The surface mesher and the preview on the state after 3 steps of the
    same scene (outside every timed step): device mesh time (median of 5
    builds, counts read back), its kernels from one profiled build, the
    preview render, and the oracle's mesher on the same particle state
    (bucketed by the current positions) on this host's cores."""
    ctx = fa.FlipContext(n, tol_policy=1, tol=TOL, dt_policy=1, precond=fa.PRECOND_LEARNED)
    ctx.set_learned_weights(weights)
    ctx.build_scene(fa.SCENE_BENCHMARK)
    ctx.srand(1)
    for _ in range(3):
        ctx.step()
    ctx.synchronize()
    ms = []
    for _ in range(5):
        t0 = time.perf_counter()
        v, nr, f = ctx.mesh()
        ms.append(1e3 * (time.perf_counter() - t0))
    ctx.set_profiling(True)
    ctx.mesh()
    kern = {}
    for name, t in ctx.kernel_times():
        kern[name] = round(kern.get(name, 0.0) + t, 4)
    ctx.set_profiling(False)
    pv = []
    for _ in range(5):
        t0 = time.perf_counter()
        ctx.render_preview()
        pv.append(1e3 * (time.perf_counter() - t0))
    q = ctx.download()
    ctx.close()
    out = {"n": n, "lattice": n + 1, "vertices": len(v), "faces": len(f),
           "gpu_ms_median": round(sorted(ms)[2], 3), "kernels_ms": kern,
           "preview_ms_median": round(sorted(pv)[2], 3)}
    if cpu:
        from oracle import pyoracle as po
        s = po.Sim(n)
        s.set_particles(q["p"], q["u"], q["n"], q["type"], dens=q["dens"])
        s.call("sort")
        t0 = time.perf_counter()
        vo, _, fo = s.mesh()
        out["cpu_oracle_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
        out["cpu_threads"] = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        out["cpu_sample"] = f"oracle mesher on the same particles, {len(vo)} vertices, {len(fo)} faces"
        s.close()
    return out


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


def run_mode(fa, n, precond, weights, steps, warmup, profile_step=True):
    """Timed steps of one solver mode from the initial scene, then (unless
    profile_step is off) one untimed step with every kernel launch timed."""
    code = {"reference": fa.PRECOND_REFERENCE, "ic0": fa.PRECOND_IC0, "learned": fa.PRECOND_LEARNED}[precond]
    ctx = fa.FlipContext(n, tol_policy=1, tol=TOL, dt_policy=1, precond=code)
    if precond == "learned":
        ctx.set_learned_weights(weights)
    n_particles = ctx.build_scene(fa.SCENE_BENCHMARK)
    ctx.srand(1)
    for _ in range(warmup):
        ctx.step()
    ctx.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(ctx.step())
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ktimes = []
    if profile_step:
        ctx.set_profiling(True)
        ctx.step()
        ktimes = ctx.kernel_times()
        ctx.set_profiling(False)
    ctx.close()
    agg = {}
    for name, ms in ktimes:
        tot, cnt = agg.get(name, (0.0, 0))
        agg[name] = (tot + ms, cnt + 1)
    mean = lambda k: sum(st[k] for st in stats) / len(stats)  # noqa: E731
    return {
        "precond": precond, "steps_per_s": steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
        "particles": n_particles, "n_fluid": stats[-1]["n_fluid_cells"],
        "iterations_mean": mean("iterations"), "learned_iterations_mean": mean("learned_iterations"),
        "solve_ms_mean": mean("ms_solve"), "repositioned_mean": mean("n_reposition"),
        "rel_residual_max": max((st["error2"] / max(st["bnorm2"], 1e-300)) ** 0.5 for st in stats),
        "agg": agg,
    }


def roofline_of(agg, n, n_particles, n_fluid):
    """HBM roofline of the kernel with the largest share of the profiled
    step among those with an algorithmic-bytes model."""
    n_cells, n_faces = n ** 3, (n + 1) * n * n
    ranked = sorted(((tot, name, cnt) for name, (tot, cnt) in agg.items()), reverse=True)
    dom = next((r for r in ranked if kernel_bytes(r[1], n_particles, n_cells, n_faces, n_fluid)), None)
    if not dom:
        return None
    tot, name, cnt = dom
    avg_ms = tot / cnt
    b = kernel_bytes(name, n_particles, n_cells, n_faces, n_fluid)
    achieved = b / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(name, n),
            "bytes_per_launch": b, "avg_launch_ms": round(avg_ms, 5), "launches_per_step": cnt}


def summary(r):
    return {k: (float(f"{v:.4g}") if isinstance(v, float) else v) for k, v in r.items() if k != "agg"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--precond", default="learned", choices=["reference", "ic0", "learned"],
                    help="solver of the reported value; the plain (reference) PCG is always timed beside it")
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS,
                    help="coarse-level network weights for --precond learned")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile-step", action="store_true",
                    help="skip the per-launch timed step (for runs under an external profiler)")
    ap.add_argument("--no-mesh", action="store_true", help="skip the mesher / preview leg")
    ap.add_argument("--direct-launch", action="store_true",
                    help="launch solver iterations kernel by kernel instead of replaying HIP graphs "
                         "(for kernel tracers that cannot follow graph launches; slower)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the slab decomposition is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.direct_launch:
        os.environ["FLIP_DIRECT_LAUNCH"] = "1"
    import flip3d_amd as fa

    n = args.n
    prof = not args.no_profile_step
    plain = run_mode(fa, n, "reference", None, args.steps, args.warmup, prof)
    main_run = plain if args.precond == "reference" else run_mode(fa, n, args.precond, args.weights, args.steps,
                                                                 args.warmup, prof)
    n_particles = main_run["particles"]
    roofline = roofline_of(main_run["agg"], n, n_particles, main_run["n_fluid"])

    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(n, 30)
    mesh = None if args.no_mesh else mesh_leg(fa, n, args.weights, cpu=not args.no_cpu_baseline)

    value = main_run["steps_per_s"]
    out = {
        "metric": METRIC,
        "value": round(value, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"benchmark dam-break scene N={n}, {n_particles} particles "
                               f"({main_run['n_fluid']} fluid cells), CFL dt, pressure solve to rel 1e-5, "
                               f"{args.precond} preconditioner; plain reference-preconditioned PCG timed beside it",
                   "n": n, "particles": n_particles, "precond": args.precond, "tol": TOL,
                   "solver_launch": "direct" if args.direct_launch else "graph"},
        "particles_per_s": round(value * n_particles, 1),
        "speedup_vs_plain_step": round(value / plain["steps_per_s"], 3),
        "speedup_vs_plain_solve": round(plain["solve_ms_mean"] / main_run["solve_ms_mean"], 3),
        "solver": summary(main_run),
        "plain": summary(plain),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mesh": mesh,
        "profile_step_ms": {k: round(v[0], 4) for k, v in sorted(main_run["agg"].items(), key=lambda kv: -kv[1][0])},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
