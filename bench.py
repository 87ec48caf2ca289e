#!/usr/bin/env python3
"""Benchmark of the MOLOCH dynamical-core step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form fp64]
                    [--workload cfg2] [--no-cpu]

A step is one moloch() call (Main/mod_moloch.F90:312-446) with physics off:
mo_nadv x (mo_nsound acoustic sub-steps + WAF advection of every prognostic
field), the dycore tail, diagnostics and status_update, on a synthetic
domain whose state is resident in HBM before the timed region starts.
The metric is grid-point steps per second with grid points
(jx-1)(iy-1)kz.  The default workload is 256x256x41, dx 5 km, dt 15 s,
100 steps (8 random hills, jet winds, 5 water species).

Prints one JSON line.  `roofline` reports the dominant kernel (largest
share of step time): its algorithmic bytes per launch (byte model below)
over its mean launch duration from HIP events recorded on the library's
stream during the timed region.  `cpu_baseline` times the CPU oracle on
this host's cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "grid-pt·steps/s, plain FP64 vs mixed/batched at stated L2 error; % HBM peak"


def kernel_bytes_per_launch(name: str, g, nadvf: int) -> float | None:
    """Algorithmic HBM bytes of one launch: every 3-D array the kernel's
    reference loop nests read or write, counted once over N cross points
    (2-D and 1-D metrics ignored).  F = number of advected fields."""
    n = g.npoints
    word = 8.0
    F = nadvf
    arrays = {
        # zdiv2 from u, v, rfmzu, rfmzv, fmz; s from u, v     -> zdiv2, s
        "sound_div": 5 + 2,
        # zdiv2, s, tetav, tetavf, w, fmzf, pai, fmz, bdywtw -> tetavf, w, pai
        "sound_column": 9 + 3,
        # u, v, zdiv2, tetav, pai, bdywtu, bdywtv          -> u, v
        "sound_momentum": 7 + 2,
        # u, v, w, s, fmzf                                  -> ux, vx, wx, s
        "adv_stagger_in": 5 + 4,
        # F fields, s, fmz, fmzf                            -> F half-step fields
        "waf_vertical": 2 * F + 3,
        # F half-step fields, F fields, u, v, rfmzu, fmz    -> F fields
        "waf_horizontal": 3 * F + 4,
        "adv_curvature": 2 + 2,
        "adv_stagger_out": 3 + 3,
        "uv_stagger": 2 + 2,
    }
    if name not in arrays:
        return None
    return arrays[name] * n * word


def step_bytes_floor(g, nadvf: int) -> float:
    """Fused floor per step (B_min): every acoustic sub-step reads state and
    metrics once (13 arrays) and writes 5, advection reads and writes each
    field once plus 10 shared arrays, the tail ~10 + nqx."""
    F = nadvf
    per_pt = g.nadv * (18 * g.nsound + 1) + g.nadv * (2 * F + 10) + 10 + g.nqx
    return per_pt * 8.0 * g.npoints


def nadv_fields(g) -> int:
    return 6 + (g.nqx - 1 if g.ipptls > 0 else 0)


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), if any."""
    path = ROOT / "profiles" / f"pmc_traffic_{workload}_fp64.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        m = re.match(r"\|\s*" + re.escape(kernel) + r"\s*\|.*\|\s*([0-9.eE+]+)\s*\|\s*$", line)
        if m:
            return float(m.group(1))
    return None


def cpu_baseline(workload: str, budget_s: float = 12.0) -> dict:
    """The CPU oracle (a C restatement of the reference step, OpenMP over
    points) timed on this host, on a bounded number of steps."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_api as O   # checker / baseline only
    from regcm_amd import synth

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    g, f, _ = synth.config(workload)
    t0 = time.perf_counter()
    O.moloch_steps(g, f, 1, nthreads=cores)
    t1 = time.perf_counter() - t0
    nsteps = int(max(2, min(50, round(budget_s / max(t1, 1e-3)))))
    t0 = time.perf_counter()
    O.moloch_steps(g, f, nsteps, nthreads=cores)
    dt = time.perf_counter() - t0
    return {"value": g.npoints * nsteps / dt, "unit": "grid-pt·steps/s",
            "cores": cores, "kind": "port",
            "sample": f"{workload}: {nsteps} steps of the C oracle (gcc -O2 "
                      f"-ffp-contract=off, OpenMP {cores} threads) after 1 warm-up "
                      f"step, {dt:.1f} s"}


WORKLOADS = {
    "cfg2": "plain FP64 MOLOCH step, 256x256x41, dx 5 km, dt 15 s, nadv 2 x nsound 5, "
            "8 random hills (seed 0), jet + N(0,1) winds, 5 water species (ipptls 2)",
    "cfg1": "plain FP64 MOLOCH step, 64x64x23, dx 12 km, dt 30 s, one 1 km hill, "
            "u = 10 + N(0, 0.25) m/s, 2 water species (ipptls 0)",
    "cfg3": "plain FP64 MOLOCH step, 512x512x41, dx 3 km, dt 10 s, cfg2 recipe + 3 K bubble",
}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cfg2", choices=sorted(WORKLOADS))
    ap.add_argument("--form", default="fp64", choices=["fp64"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the single-domain MOLOCH "
              "step has no multi-GPU path); run with --gpus 1", file=sys.stderr)
        return 2

    from regcm_amd import synth
    from regcm_amd.moloch import Moloch, MLC_FP64

    g, f, _ = synth.config(args.workload)
    F = nadv_fields(g)
    with Moloch(g, MLC_FP64) as m:
        m.load(f)
        m.init_moloch()
        del f
        m.moloch(args.warmup)
        m.synchronize()
        m.profile(True)
        t0 = time.perf_counter()
        m.moloch(args.steps)
        m.synchronize()
        dt = time.perf_counter() - t0
        stats = m.kernel_stats()
        m.profile(False)
        psmax, psmin = m.check_finite()
        npts = m.npoints

    value = npts * args.steps / dt
    ms_step = dt / args.steps * 1e3
    # dominant kernel = largest total time over the timed region
    dom = max(stats.items(), key=lambda kv: kv[1][1])
    dname, (dlaunch, dms) = dom
    avg_s = dms / dlaunch * 1e-3
    bpl = kernel_bytes_per_launch(dname, g, F)
    achieved = (bpl / avg_s / 1e9) if bpl else None
    traffic = read_traffic(dname, args.workload)
    kernels = {k: {"launches": n, "avg_us": round(ms / n * 1e3, 2),
                   "share": round(ms / sum(v[1] for v in stats.values()), 4),
                   **({"GBps": round(kernel_bytes_per_launch(k, g, F) / (ms / n * 1e-3) / 1e9, 1)}
                      if kernel_bytes_per_launch(k, g, F) else {})}
               for k, (n, ms) in sorted(stats.items(), key=lambda kv: -kv[1][1])}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-pt·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": WORKLOADS[args.workload],
                   "form": args.form, "jx": g.jx, "iy": g.iy, "kz": g.kz,
                   "grid_points": npts, "advected_fields": F, "dtsec": g.dtsec,
                   "dx": g.dx, "nadv": g.nadv, "nsound": g.nsound},
        "roofline": {"bound": "hbm", "kernel": dname,
                     "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic,
                     "algorithmic_bytes_per_launch": bpl,
                     "avg_launch_us": round(avg_s * 1e6, 2)},
        "step_roofline": {"floor_bytes_per_step": step_bytes_floor(g, F),
                          "frac": round(step_bytes_floor(g, F) / (dt / args.steps) / 1e9
                                        / HBM_PEAK_GBS, 4)},
        "kernels": kernels,
        "ps_range_pa": [psmin, psmax],
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.workload)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
