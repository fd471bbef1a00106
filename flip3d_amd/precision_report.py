"""Per-step error of the FP32 pressure solve (plain PCG, or PCG inside
FP64-residual iterative refinement) against a full-FP64 GPU solve.

    python -m flip3d_amd.precision_report --n 128 --steps 200
    python -m flip3d_amd.precision_report --n 256 --steps 50 --refine 8 --refine-tol 1e-7

The trajectory is the FP32 run's own.  Before every step its state
(particles, max_dens, the pressure warm start, the rand() stream) is copied
into a second context whose solver is the FP64 Jacobi-preconditioned CG at
relative residual 1e-12; both take the step, and the report compares:
  * pressure: max |p32 - p64| / max |p64| over the grid;
  * face velocities after projection: max over the three components of
    max |u32 - u64| / max |u64|;
  * particle velocities of FLUID particles after the step, the same norm;
and gives the FP32 solve's iterations, sweeps and final FP64 relative
residual.  One line per step, then a summary.
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


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--refine", type=int, default=0)
    ap.add_argument("--refine-tol", type=float, default=1e-7)
    ap.add_argument("--json", default=None, help="also write the per-step records here")
    args = ap.parse_args()

    kw32 = dict(tol_policy=1, tol=args.tol, dt_policy=1)
    if args.refine > 0:
        kw32.update(refine=args.refine, refine_tol=args.refine_tol)
    a = fa.FlipContext(args.n, **kw32)
    b = fa.FlipContext(args.n, tol_policy=1, tol=1e-12, dt_policy=1, solve_fp64=1)
    np_ = a.build_scene(fa.SCENE_BENCHMARK)
    a.srand(1)
    print(f"N={args.n}: {np_} particles, FP32 solve tol {args.tol}"
          + (f", refinement to {args.refine_tol} (<= {args.refine} sweeps)" if args.refine else ""), flush=True)
    recs = []
    t0 = time.time()
    for k in range(args.steps):
        st = a.download()
        press = a.grid(fa.GRID_PRESS)
        b.upload(st["p"], st["u"], st["n"], st["type"], dens=st["dens"])
        b.max_dens = a.max_dens
        b.set_grid(fa.GRID_PRESS, press)
        a.srand(1000 + k)
        b.srand(1000 + k)
        sa = a.step()
        sb = b.step()
        ua = [a.grid(w) for w in (fa.GRID_U0, fa.GRID_U1, fa.GRID_U2)]
        ub = [b.grid(w) for w in (fa.GRID_U0, fa.GRID_U1, fa.GRID_U2)]
        pa, pb = a.download(), b.download()
        fl = pa["type"] == fa.FLUID
        rec = {
            "step": k + 1,
            "press_err": rel(a.grid(fa.GRID_PRESS), b.grid(fa.GRID_PRESS)),
            "face_u_err": max(rel(x, y) for x, y in zip(ua, ub)),
            "particle_u_err": rel(pa["u"][fl], pb["u"][fl]),
            "iterations": sa["iterations"], "refine_sweeps": sa["refine_sweeps"],
            "rel_residual": float(np.sqrt(sa["error2"] / max(sa["bnorm2"], 1e-300))),
            "fp64_iterations": sb["iterations"], "ms_step": round(sa["ms_step"], 3), "ms_solve": round(sa["ms_solve"], 3),
        }
        recs.append(rec)
        print(json.dumps(rec), flush=True)
    keys = ("press_err", "face_u_err", "particle_u_err", "rel_residual", "ms_step", "ms_solve", "iterations")
    summ = {k: {"max": max(r[k] for r in recs), "median": float(np.median([r[k] for r in recs]))} for k in keys}
    summ["steps"] = len(recs)
    summ["seconds"] = round(time.time() - t0, 1)
    print("summary " + json.dumps(summ), flush=True)
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"records": recs, "summary": summ}, f)
    a.close()
    b.close()


if __name__ == "__main__":
    main()
