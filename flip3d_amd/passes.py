"""Algorithmic bytes of the FLIP step, pass by pass, and the HBM traffic the
hardware counters measure for the same passes.

One model serves the bench line, the counter summaries and the kernel
table: every pass of the step is credited the unique input bytes it reads
plus the output bytes it writes, per unit of work (a particle, a cell, a face
of the staggered grid), with neighbour re-reads assumed cached.  Fields are
fp32, flags int8, the solver's P fp64.  A kernel that fuses passes is
credited the bytes of the passes it replaces.

Units: Np particles the pass visits (fluid and wall), C = N^3 cells,
F = 3 N^2 (N+1) faces (all three components), Cf fluid cells (the solver's
kernels visit the compacted fluid list, so the solver passes are counted
over Cf, not over all N^3 cells: crediting N^3 would count bytes no kernel
moves).  With Cf = C the totals are 328 Np + 150 C + 72 F per step plus
117 C per solver iteration.

Each pass names the timed regions of the library's per-launch profile
(FlipContext.kernel_times) that implement it and the kernel symbols they
launch (for matching rocprofv3 output).
"""
import csv
import os
import re
from collections import defaultdict

def _p(name, credited, fn, regions, kernels, count_region, count_kernel):
    return {"name": name, "credited": credited, "bytes": fn, "regions": regions, "kernels": kernels,
            "count_region": count_region, "count_kernel": count_kernel}


# bytes per invocation as a function of the units; the timed regions that
# implement the pass; the kernel-symbol pattern of its launches; the region
# and the kernel whose launches count its invocations (None: one a step)
PASSES = [
    _p("sort", "sort (twice a step: after the corrector and after advect)",
       lambda u: 68 * u["Np"] + 8 * u["C"],
       ("sort_count", "sort_scan", "sort_scatter", "sort_gather"),
       r"^k_fill_i32$|^k_sort_", "sort_count", r"^k_sort_gather$"),
    _p("density", "density",
       lambda u: 16 * u["Np"],
       ("density_force", "density", "ext_force"), r"^k_density|^k_ext_force$", None, r"^k_density_lds<true>$|^k_density<true, true>$"),
    _p("flags", "flags + L",
       lambda u: 4 * u["Np"] + 10 * u["C"],
       ("flags",), r"^k_flags$", "flags", r"^k_flags$"),
    _p("p2g", "P2G + save_grid + first enforce_boundary",
       lambda u: 24 * u["Np"] + 4 * u["C"] + 4 * u["F"] + 8 * u["F"] + (8 * u["F"] + u["C"]),
       ("p2g_prep", "p2g"), r"^k_p2g", "p2g", r"^k_p2g_lds3$|^k_p2g_lds<0>"),
    _p("divergence", "div",
       lambda u: 4 * u["F"] + 5 * u["C"],
       ("divergence",), r"^k_divergence$", "divergence", r"^k_divergence$"),
    _p("solver_setup", "IC build + PCG setup (over the fluid list)",
       lambda u: 13 * u["Cf"] + 90 * u["Cf"],
       ("pcg_setup", "learned_setup"), None, None, None),
    _p("solver_iteration", "PCG iteration (over the fluid list; flexible CG for the learned solver)",
       lambda u: 117 * u["Cf"],
       ("pcg_a", "pcg_b", "fcg_a", "fcg_b"), r"^k_pcg_a$|^k_pcg_b(<|$)|^k_fcg_a$|^k_fcg_b_pre0$", "pcg_a|fcg_a",
       r"^k_pcg_a$|^k_fcg_a$"),
    _p("gradient", "grad-sub + second enforce_boundary",
       lambda u: 8 * u["C"] + 8 * u["F"] + (8 * u["F"] + u["C"]),
       ("gradient",), r"^k_gradient", "gradient", r"^k_gradient3$|^k_gradient<0>"),
    _p("extrapolate", "extrap + subtract_grid",
       lambda u: u["C"] + 8 * u["F"] + 12 * u["F"],
       ("extrapolate",), r"^k_extrapolate", "extrapolate", r"^k_extrapolate3$|^k_extrapolate<0>"),
    _p("g2p_flip", "G2P + FLIP",
       lambda u: 36 * u["Np"] + 8 * u["F"],
       ("g2p_flip",), r"^k_g2p_flip", "g2p_flip", r"^k_g2p_flip"),
    _p("advect", "advect + stuck flags",
       lambda u: 24 * u["Np"] + 4 * u["F"] + (16 * u["Np"] + u["C"]),
       ("advect", "stuck"), r"^k_advect$|^k_stuck$", "advect", r"^k_advect$"),
    _p("corrector", "corrector (springs, jitter, resample, update)",
       lambda u: 72 * u["Np"],
       ("spring", "jitter_order", "jitter_draw", "jitter_fix", "resample", "correct_update"),
       r"^k_spring|^k_jitter_|^k_resample$|^k_correct_update$", "spring", r"^k_spring$"),
]

# per-unit coefficients of each pass (the documentation table): Np, C, F, Cf
COEFFS = {
    "sort": (68, 8, 0, 0), "density": (16, 0, 0, 0), "flags": (4, 10, 0, 0), "p2g": (24, 5, 20, 0),
    "divergence": (0, 5, 4, 0), "solver_setup": (0, 0, 0, 103), "solver_iteration": (0, 0, 0, 117),
    "gradient": (0, 9, 16, 0), "extrapolate": (0, 1, 20, 0), "g2p_flip": (36, 0, 8, 0),
    "advect": (40, 1, 4, 0), "corrector": (72, 0, 0, 0),
}


def units(n, n_particles, n_fluid):
    return {"Np": n_particles, "C": n ** 3, "F": 3 * n * n * (n + 1), "Cf": n_fluid}


def by_name(name):
    for p in PASSES:
        if p["name"] == name:
            return p
    raise KeyError(name)


def pass_bytes(name, u):
    return by_name(name)["bytes"](u)


def kernel_base(symbol):
    """A kernel's name with its template arguments, from rocprofv3's
    demangled form ("void (anonymous namespace)::k_x<1>(int, ...)" -> "k_x<1>")."""
    s = symbol.strip()
    if s.startswith("void "):
        s = s[5:]
    s = s.replace("(anonymous namespace)::", "")
    s = s.split("(")[0].strip()
    return re.sub(r"^(?:[A-Za-z_][A-Za-z_0-9]*::)+", "", s)


def kernel_pass(symbol):
    """The pass a kernel symbol (bare, or rocprofv3's demangled name)
    belongs to, or None."""
    base = kernel_base(symbol)
    for p in PASSES:
        if p["kernels"] and re.search(p["kernels"], base):
            return p["name"]
    return None


def step_passes(agg, u):
    """Per-pass figures of one profiled step.  agg maps a timed region to
    (total ms, launches) over the step; a pass's invocations are the
    launches of its counting region (one a step when it has none)."""
    out = {}
    for p in PASSES:
        ms = sum(agg[r][0] for r in p["regions"] if r in agg)
        if ms <= 0.0:
            continue
        if p["count_region"] is None:
            inv = 1
        else:
            inv = sum(agg.get(r, (0.0, 0))[1] for r in p["count_region"].split("|"))
        if inv <= 0:
            continue
        b = p["bytes"](u)
        out[p["name"]] = {"bytes_per_invocation": b, "invocations": inv, "ms": round(ms, 4),
                          "avg_ms": round(ms / inv, 5), "achieved_gbs": round(b * inv / (ms * 1e-3) / 1e9, 1)}
    return out


def dominant(passes):
    """The pass with the largest share of the step among those modelled."""
    return max(passes.items(), key=lambda kv: kv[1]["ms"]) if passes else (None, None)


# ---------------------------------------------------------------- counters

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def read_counter_csv(path):
    """rows of a rocprofv3 counter_collection.csv: (dispatch id, kernel, counter, value)"""
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], r["Counter_Name"], float(r["Counter_Value"])))
    return rows


def find_csv(root, suffix="counter_collection.csv"):
    for d, _, files in os.walk(root):
        for f in sorted(files):
            if f.endswith(suffix):
                return os.path.join(d, f)
    raise FileNotFoundError(f"no *{suffix} under {root}")


def per_dispatch(rows, counter):
    """dispatch id -> (kernel, value) of one counter, the value summed over
    the counter's instances"""
    val = defaultdict(float)
    name = {}
    for d, k, c, v in rows:
        if c == counter:
            val[d] += v
            name[d] = k
    return {d: (name[d], val[d]) for d in val}


def traffic_summary(fetch_csv, write_csv, copy_kib, u):
    """Scaled HBM traffic per kernel and per pass from a FETCH_SIZE run and a
    WRITE_SIZE run of the same command.  Each counter is scaled by copy_kib
    over what it reads for the three largest dispatches of the runtime's copy
    kernel (copies of copy_kib KiB read and written).  A pass's traffic per
    invocation is the traffic of all its dispatches over the dispatches of
    its counting kernel."""
    out = {"kernels": {}, "passes": {}}
    dk = {}
    for key, path, counter in (("fetch", fetch_csv, "FETCH_SIZE"), ("write", write_csv, "WRITE_SIZE")):
        disp = per_dispatch(read_counter_csv(path), counter)
        copies = sorted((v for k, v in disp.values() if k.startswith(COPY_KERNEL)), reverse=True)[:3]
        ref = sum(copies) / len(copies) if copies else None
        scale = copy_kib / ref if ref else None
        out["copy_" + key + "_kib"] = ref
        out["scale_" + key] = scale
        for k, v in disp.values():
            if k.startswith(COPY_KERNEL):
                continue
            base = kernel_base(k)
            e = dk.setdefault(base, {"fetch": [0, 0.0], "write": [0, 0.0]})
            e[key][0] += 1
            e[key][1] += v
    for base, e in sorted(dk.items()):
        nf, sf = e["fetch"]
        nw, sw = e["write"]
        mf = sf / nf if nf else 0.0
        mw = sw / nw if nw else 0.0
        b = None
        if out["scale_fetch"] and out["scale_write"]:
            b = 1024.0 * (mf * out["scale_fetch"] + mw * out["scale_write"])
        out["kernels"][base] = {"dispatches": max(nf, nw), "fetch_kib": mf, "write_kib": mw, "bytes": b,
                                "pass": kernel_pass(base)}
    for p in PASSES:
        ks = [k for k, v in out["kernels"].items() if v["pass"] == p["name"]]
        if not ks or any(out["kernels"][k]["bytes"] is None for k in ks) or not p["count_kernel"]:
            continue
        inv = sum(out["kernels"][k]["dispatches"] for k in ks if re.search(p["count_kernel"], k))
        if inv == 0:
            continue
        total = sum(out["kernels"][k]["bytes"] * out["kernels"][k]["dispatches"] for k in ks)
        out["passes"][p["name"]] = {"kernels": ks, "invocations": inv, "traffic_per_invocation": total / inv,
                                    "algorithmic_per_invocation": p["bytes"](u)}
    return out


def read_stats_csv(path):
    """kernel -> (calls, total ns) from a rocprofv3 --stats kernel_stats.csv"""
    out = {}
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            base = kernel_base(r["Name"])
            calls, tot = out.get(base, (0, 0.0))
            out[base] = (calls + int(r["Calls"]), tot + float(r["TotalDurationNs"]))
    return out


def bench_line(path):
    """The JSON line a bench run printed (the last line starting with '{')."""
    import json
    line = [x for x in open(path) if x.startswith("{")][-1]
    return json.loads(line)


def _fmt_mb(b):
    return f"{b / 1e6:10.2f}"


def traffic_text(summary, steps_note):
    """The committed text form of traffic_summary's result."""
    lines = [f"calibration: copy FETCH_SIZE {summary['copy_fetch_kib']:.0f} KiB -> scale {summary['scale_fetch']:.5f}; "
             f"copy WRITE_SIZE {summary['copy_write_kib']:.0f} KiB -> scale {summary['scale_write']:.5f}", "",
             "per kernel, mean per dispatch:", "",
             f"  {'kernel':<34} {'disp':>6} {'FETCH KiB':>12} {'WRITE KiB':>12} {'scaled MB':>10}  pass"]
    for k, v in sorted(summary["kernels"].items(), key=lambda kv: -(kv[1]["bytes"] or 0) * kv[1]["dispatches"]):
        lines.append(f"  {k:<34} {v['dispatches']:>6} {v['fetch_kib']:>12.1f} {v['write_kib']:>12.1f} "
                     f"{_fmt_mb(v['bytes'] or 0)}  {v['pass'] or '-'}")
    lines += ["", f"per pass, per invocation ({steps_note}):", "",
              f"  {'pass':<18} {'inv':>5} {'traffic MB':>11} {'algorithmic MB':>15} {'ratio':>7}"]
    for name, v in summary["passes"].items():
        lines.append(f"  {name:<18} {v['invocations']:>5} {_fmt_mb(v['traffic_per_invocation'])} "
                     f"{v['algorithmic_per_invocation'] / 1e6:15.2f} "
                     f"{v['traffic_per_invocation'] / v['algorithmic_per_invocation']:7.2f}")
    lines.append("")
    for name, v in summary["passes"].items():
        lines.append(f"traffic_per_invocation {name} {v['traffic_per_invocation']:.0f}")
    return "\n".join(lines)


def stats_text(stats, steps, passes_ms=None):
    """Per-kernel and per-pass durations from a kernel_stats.csv over
    `steps` steps of one solver run; passes_ms: the bench's own per-pass
    figures (ms per step) to set beside them."""
    lines = [f"  {'kernel':<40} {'calls':>6} {'total ms':>9} {'avg us':>8}  pass"]
    total = sum(t for _, t in stats.values())
    for k, (c, t) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        if t < 0.001 * total:
            continue
        lines.append(f"  {k:<40} {c:>6} {t / 1e6:9.3f} {t / c / 1e3:8.2f}  {kernel_pass(k) or '-'}")
    lines += ["", f"  {'pass':<18} {'ms/step (trace)':>16} {'ms/step (bench events)':>23}"]
    per = defaultdict(float)
    for k, (c, t) in stats.items():
        p = kernel_pass(k)
        if p:
            per[p] += t / 1e6 / steps
    for p in PASSES:
        if p["name"] in per:
            b = passes_ms.get(p["name"]) if passes_ms else None
            lines.append(f"  {p['name']:<18} {per[p['name']]:16.4f} {b if b is not None else '-':>23}")
    return "\n".join(lines)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="summaries of rocprofv3 runs of bench.py, pass by pass")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("traffic", help="FETCH_SIZE / WRITE_SIZE counter runs -> traffic per pass")
    t.add_argument("--fetch", required=True, help="output directory of the FETCH_SIZE run")
    t.add_argument("--write", required=True, help="output directory of the WRITE_SIZE run")
    t.add_argument("--copy-mib", type=int, required=True, help="the run's --copy-probe-mib")
    t.add_argument("--bench-log", required=True, help="stdout of one of the runs (its JSON line)")
    s = sub.add_parser("stats", help="kernel-trace --stats run -> durations per kernel and per pass")
    s.add_argument("--dir", required=True)
    s.add_argument("--steps", type=int, required=True, help="steps of the solver run the trace covers")
    s.add_argument("--bench-log", help="a bench run's stdout, for its per-pass figures")
    a = ap.parse_args(argv)
    if a.cmd == "traffic":
        d = bench_line(a.bench_log)
        u = units(d["config"]["n"], d["config"]["particles"], d["solver"]["n_fluid"])
        summ = traffic_summary(find_csv(a.fetch), find_csv(a.write), a.copy_mib * 1024, u)
        print(traffic_text(summ, f"N={d['config']['n']}, {u['Np']} particles, {u['Cf']} fluid cells, "
                                 f"{d['config']['precond']} solver"))
    else:
        pm = None
        if a.bench_log:
            pm = {k: v["ms"] for k, v in bench_line(a.bench_log)["passes"].items()}
        print(stats_text(read_stats_csv(find_csv(a.dir, "kernel_stats.csv")), a.steps, pm))


if __name__ == "__main__":
    main()
