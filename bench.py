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

from flip3d_amd import passes as P  # noqa: E402  (pure Python: loads no library)

HBM_PEAK_GBS = 8000.0
TOL = 1e-5
METRIC = "FLIP steps/s at 128^3 (~4M particles), learned vs plain PCG, rel resid 1e-5"
PROFILES = os.path.join(ROOT, "profiles")
DEFAULT_WEIGHTS = os.path.join(ROOT, "flip3d_amd", "weights", "unet_coarse.bin")


def log(msg):
    """progress on stderr (the JSON line alone goes to stdout)"""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(pass_name, n=128):
    """HBM bytes per invocation of a pass from the committed counter summary
    profiles/traffic_n<n>.txt (FETCH_SIZE + WRITE_SIZE of its kernels, each
    scaled by the same counter on a copy of known size); None when absent."""
    path = os.path.join(PROFILES, f"traffic_n{n}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_per_invocation" and parts[1] == pass_name:
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


def cpu_model():
    """The host CPU's model name (lscpu's "Model name", from /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def oracle_steps(n, k):
    """Seconds and PCG iterations of the oracle's first k steps of the
    benchmark scene (the reference's timed region: density through the
    corrector, no I/O), from the same initial state as the GPU runs."""
    from oracle import pyoracle as po
    s = po.Sim(n, tol_policy=1, tol=TOL, dt_policy=1)
    s.scene_benchmark()
    s.srand(1)
    secs, its = [], []
    for _ in range(k):
        t0 = time.perf_counter()
        st = s.step()
        secs.append(time.perf_counter() - t0)
        its.append(int(st["iterations"]))
        log(f"oracle N={n} step {len(secs)}/{k}: {secs[-1]:.2f} s, {its[-1]} PCG iterations")
    s.close()
    return secs, its


def cpu_baseline(n, steps, one_thread_steps, one_thread_timeout):
    """The oracle (C restatement of the reference path, OpenMP where the
    reference uses it) on this host's cores: the first `steps` steps on all
    the threads OpenMP is given (mean and spread per step), then the first
    `one_thread_steps` on one thread in a child process.  Runs before the GPU
    is touched, so the child starts from a process without a HIP runtime."""
    import statistics
    import subprocess
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    secs, its = oracle_steps(n, steps)
    mean = statistics.fmean(secs)
    out = {
        "value": round(1.0 / mean, 5), "unit": "steps/s", "cores": threads, "kind": "port",
        "sample": f"oracle, first {steps} steps of the N={n} benchmark scene from the GPU runs' initial "
                  f"state, {threads} OpenMP threads; mean and spread per step",
        "steps": steps, "s_per_step_mean": round(mean, 3),
        "s_per_step_min": round(min(secs), 3), "s_per_step_max": round(max(secs), 3),
        "s_per_step_stdev": round(statistics.stdev(secs), 3) if len(secs) > 1 else 0.0,
        "s_per_step": [round(x, 3) for x in secs], "pcg_iterations": its,
        "nproc": os.cpu_count(), "cpu_model": cpu_model(),
    }
    if one_thread_steps > 0:
        env = dict(os.environ, OMP_NUM_THREADS="1")
        cmd = [sys.executable, os.path.abspath(__file__), "--oracle-worker", str(n), str(one_thread_steps)]
        try:
            r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True, timeout=one_thread_timeout)
            w = json.loads(r.stdout.strip().splitlines()[-1]) if r.returncode == 0 else None
        except (subprocess.TimeoutExpired, ValueError, IndexError):
            w = None
        if w:
            m1 = sum(w["secs"]) / len(w["secs"])
            out["one_thread"] = {"steps": one_thread_steps, "s_per_step_mean": round(m1, 3),
                                 "s_per_step": [round(x, 3) for x in w["secs"]], "pcg_iterations": w["its"],
                                 "steps_per_s": round(1.0 / m1, 5)}
        else:
            out["one_thread"] = {"steps": one_thread_steps, "error": f"no result within {one_thread_timeout} s"}
    return out


def run_mode(fa, n, precond, weights, steps, warmup, profile_step=True, launch=(0, 0)):
    """Timed steps of one solver mode from the initial scene, then (unless
    profile_step is off) one untimed step with every kernel launch timed.
    launch: the solver kernels' (block count, PCG batch), 0 = built in."""
    code = {"reference": fa.PRECOND_REFERENCE, "ic0": fa.PRECOND_IC0, "learned": fa.PRECOND_LEARNED}[precond]
    ctx = fa.FlipContext(n, tol_policy=1, tol=TOL, dt_policy=1, precond=code)
    if precond == "learned":
        ctx.set_learned_weights(weights)
    ctx.set_launch(*launch)
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
    pst = None
    if profile_step:
        ctx.set_profiling(True)
        pst = ctx.step()
        ktimes = ctx.kernel_times()
        ctx.set_profiling(False)
    ctx.close()
    agg = {}
    for name, ms in ktimes:
        tot, cnt = agg.get(name, (0.0, 0))
        agg[name] = (tot + ms, cnt + 1)
    mean = lambda k: sum(st[k] for st in stats) / len(stats)  # noqa: E731
    log(f"{precond}: {steps / elapsed:.2f} steps/s over {steps} steps")
    return {
        "precond": precond, "steps_per_s": steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
        "particles": n_particles, "n_fluid": stats[-1]["n_fluid_cells"],
        "iterations_mean": mean("iterations"), "learned_iterations_mean": mean("learned_iterations"),
        "solve_ms_mean": mean("ms_solve"), "repositioned_mean": mean("n_reposition"),
        "rel_residual_max": max((st["error2"] / max(st["bnorm2"], 1e-300)) ** 0.5 for st in stats),
        "agg": agg, "profiled_fluid": pst["n_fluid_cells"] if pst else None,
        "launch": {"pcg_blocks": launch[0], "pcg_batch": launch[1]},
    }


def roofline_of(passes, n):
    """HBM roofline of the modelled pass with the largest share of the
    profiled step: algorithmic bytes per invocation over the pass's average
    invocation time (HIP events around its launches on the library's
    stream)."""
    name, p = P.dominant(passes)
    if name is None:
        return None
    achieved = p["bytes_per_invocation"] / (p["avg_ms"] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(name, n),
            "pass": name, "kernels": P.by_name(name)["regions"], "bytes_per_invocation": p["bytes_per_invocation"],
            "avg_invocation_ms": p["avg_ms"], "invocations_per_step": p["invocations"]}


def step_roofline(passes):
    """All modelled passes of the step together: their algorithmic bytes
    over their summed time."""
    b = sum(p["bytes_per_invocation"] * p["invocations"] for p in passes.values())
    ms = sum(p["ms"] for p in passes.values())
    if ms <= 0:
        return None
    gbs = b / (ms * 1e-3) / 1e9
    return {"bytes": b, "ms": round(ms, 4), "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}


def summary(r):
    return {k: (float(f"{v:.4g}") if isinstance(v, float) else v) for k, v in r.items()
            if k not in ("agg", "profiled_fluid")}


def zero_network(fa, path):
    """Weights of a network whose output is zero: the learned solver's
    V-cycle with a zero coarsest-level correction."""
    import numpy as np
    with fa.UNet(32, capacity=1, seed=1) as net:
        net.set_params(np.zeros(net.num_params, np.float32))
        net.save(path)


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
    ap.add_argument("--skip-plain", action="store_true",
                    help="do not time the plain PCG beside the reported solver (counter runs)")
    ap.add_argument("--leg-steps", type=int, default=10,
                    help="timed steps of the IC(0) and zero-network legs (0: skip them)")
    ap.add_argument("--no-tuning", action="store_true",
                    help="built-in solver launch shapes instead of flip3d_amd/tuning.json's")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=5, help="oracle steps timed on all threads")
    ap.add_argument("--cpu-one-thread-steps", type=int, default=1, help="oracle steps timed on one thread")
    ap.add_argument("--cpu-one-thread-timeout", type=float, default=170.0)
    ap.add_argument("--no-profile-step", action="store_true",
                    help="skip the per-launch timed step (for runs under an external profiler)")
    ap.add_argument("--no-mesh", action="store_true", help="skip the mesher / preview leg")
    ap.add_argument("--direct-launch", action="store_true",
                    help="launch solver iterations kernel by kernel instead of replaying HIP graphs "
                         "(for kernel tracers that cannot follow graph launches; slower)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device (counter calibration)")
    ap.add_argument("--oracle-worker", nargs=2, type=int, metavar=("N", "STEPS"), help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.oracle_worker:
        secs, its = oracle_steps(*args.oracle_worker)
        print(json.dumps({"secs": secs, "its": its}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the slab decomposition is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    n = args.n
    # the CPU baseline first: its one-thread leg starts a child process,
    # which must not be forked from a process that holds the GPU
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(n, args.cpu_steps, args.cpu_one_thread_steps, args.cpu_one_thread_timeout)

    if args.direct_launch:
        os.environ["FLIP_DIRECT_LAUNCH"] = "1"
    import flip3d_amd as fa

    prof = not args.no_profile_step
    from flip3d_amd import autotune
    table = None if args.no_tuning else autotune.load()
    plain = None
    if not args.skip_plain or args.precond == "reference":
        plain = run_mode(fa, n, "reference", None, args.steps, args.warmup, prof,
                         autotune.launch_for(table, "reference", n))
    main_run = plain if args.precond == "reference" else run_mode(fa, n, args.precond, args.weights, args.steps,
                                                                 args.warmup, prof,
                                                                 autotune.launch_for(table, args.precond, n))
    n_particles = main_run["particles"]
    passes = P.step_passes(main_run["agg"], P.units(n, n_particles, main_run["profiled_fluid"] or main_run["n_fluid"]))
    roofline = roofline_of(passes, n)

    legs = {}
    if args.leg_steps > 0:
        if args.precond != "ic0":
            legs["ic0"] = summary(run_mode(fa, n, "ic0", None, args.leg_steps, 2, False))
        if args.precond == "learned":
            import tempfile
            with tempfile.TemporaryDirectory() as td:
                zpath = os.path.join(td, "zero.bin")
                zero_network(fa, zpath)
                legs["learned_zero_network"] = summary(run_mode(fa, n, "learned", zpath, args.leg_steps, 2, False))

    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)

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
        "speedup_vs_plain_step": round(value / plain["steps_per_s"], 3) if plain else None,
        "speedup_vs_plain_solve": round(plain["solve_ms_mean"] / main_run["solve_ms_mean"], 3) if plain else None,
        "solver": summary(main_run),
        "plain": summary(plain) if plain else None,
        "legs": legs,
        "roofline": roofline,
        "step_roofline": step_roofline(passes),
        "passes": passes,
        "cpu_baseline": cpu,
        "mesh": mesh,
        "profile_step_ms": {k: round(v[0], 4) for k, v in sorted(main_run["agg"].items(), key=lambda kv: -kv[1][0])},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
