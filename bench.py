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
over its mean launch duration from the HIP events its launches record
(hipExtLaunchKernel start / stop) on the library's stream during the timed
region, in one timed step of every EVENT_EVERY; `frac_of_attainable` divides the same rate
by what a plain streaming kernel (torch add, two reads and one write over
1 GiB vectors, `attainable`) reaches on this GPU.  `cpu_baseline` times the
CPU oracle on this host's cores on a bounded sample of the same workload.
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
EVENT_EVERY = 5   # timed steps per step whose dominant-kernel launches record events
METRIC = "grid-pt·steps/s, plain FP64 vs mixed/batched at stated L2 error; % HBM peak"


# Arrays each kernel slot reads or writes per call, as (bulk, double): bulk
# arrays are double in the FP64 form and float in the mixed form; the
# Exner function, theta_v, w, tetavf and the solve's metrics are double in
# both.  F advected fields (2 of them, pai and theta_v, double), nqx species.
# `terrain` is the MLC_VAR_TERRAIN mask: the stages it names form fmz, fmzf,
# rfmzu, rfmzv from the 2-D orography and read none of them.
def kernel_arrays(name: str, F: int, nqx: int, terrain: int = 0):
    t = lambda bit, formed, streamed: formed if terrain & bit else streamed
    return {
        # u, v, rfmzu, rfmzv, fmz                           -> zdiv2, s
        "sound_div": (t(1, 2, 5) + 2, 0),
        # zdiv2, s | tetav, tetavf, w, fmzf, pai, fmz         -> tetavf, w, pai
        # (the boundary weights come from a 2-D index into a column table)
        "sound_column": (2, t(2, 4, 6) + 3),
        # u, v, zdiv2 | tetav, pai                          -> u, v
        "sound_momentum": (3 + 2, 2),
        # u, v, s, fmzf | w                                 -> ux, vx, wx, s
        "adv_stagger_in": (t(4, 3, 4) + 4, 1),
        # F fields, s, zrfmu, zrfmd                         -> F half-step fields
        "waf_vertical": (2 * (F - 2) + 3, 2 * 2),
        # F half-step fields, F fields, u, v, rfmzu, fmz    -> F fields
        "waf_horizontal": (3 * (F - 2) + t(8, 2, 4), 3 * 2),
        "adv_curvature": (2 + 2, 0),
        # ux, vx, wx                                        -> u, v | w
        "adv_stagger_out": (3 + 2, 1),
        "uv_stagger": (2 + 2, 0),
        # nqx species | tetav, pai -> t, tvirt, p, rho, qsat | tetav (one pass)
        "tail": (nqx + 5, 3),
    }.get(name)


def kernel_bytes_per_launch(name: str, g, nadvf: int, form: str = "fp64",
                            terrain: int = 0) -> float | None:
    """Algorithmic HBM bytes of one call of a kernel slot: every 3-D array
    it reads or writes, counted once over N cross points at the element
    size the form stores it in (2-D and 1-D metrics ignored)."""
    arr = kernel_arrays(name, nadvf, g.nqx, terrain)
    if arr is None:
        return None
    bulk = 8.0 if form == "fp64" else 4.0
    return (arr[0] * bulk + arr[1] * 8.0) * g.npoints


def calls_per_step(name: str, g) -> int:
    if name.startswith("sound_"):
        return g.nadv * g.nsound
    if name.startswith("adv_") or name.startswith("waf_"):
        return g.nadv
    return 1


def step_bytes_floor(g, nadvf: int) -> float:
    """Fused floor per step (B_min): every acoustic sub-step reads state and
    metrics once (13 arrays) and writes 5, advection reads and writes each
    field once plus 10 shared arrays, the tail ~10 + nqx."""
    F = nadvf
    per_pt = g.nadv * (18 * g.nsound + 1) + g.nadv * (2 * F + 10) + 10 + g.nqx
    return per_pt * 8.0 * g.npoints


def nadv_fields(g) -> int:
    return 6 + (g.nqx - 1 if g.ipptls > 0 else 0)


def kernel_struct(slot: str, variants: dict, form: str, kz: int = 41,
                  crm: int = 0) -> str | None:
    """The argument-struct name rocprofv3 shows for a kernel slot's launch
    in the given launch shapes (as regcm_amd.profsum.short_name prints it);
    crm = 1 on a doubly periodic domain (the horizontal WAF's own
    instantiation)."""
    r = "double" if form == "fp64" else "float"
    per = 1 if crm else 0
    if slot == "sound_column":
        if variants["column"] == 2:
            cj = variants.get("column_j", 32)
            ck = 256 // cj
            cc = 6 if cj == 32 else 3          # levels per thread when kz <= 48
            return f"SoundColumnChunk<{r},{cj},{ck},{cc if kz <= ck * cc else 8}>"
        base = {0: "SoundColumn", 1: "SoundColumnStream"}[variants["column"]]
        return f"{base}<{r}>"
    if slot == "waf_horizontal":
        if variants["waf_h"] == 0:
            return f"WafRows<{r},{r},{variants['waf_group']}>"
        if (form == "mixed" and variants.get("mixed_merge", 0) & 4
                and variants.get("waf_hgroup", 0) == 0):
            return f"WafHorizontalFused<{r},{variants.get('waf_tile', 4)},{per}>"
        if form == "mixed" and variants.get("mixed_merge", 0) & 2:
            return f"WafHorizontalMixed<{r},{variants.get('waf_tile', 4)},{per}>"
        return f"WafHorizontal<{r},{r},{variants.get('waf_tile', 4)},{per}>"
    if slot == "waf_vertical":
        if form == "mixed" and variants.get("mixed_merge", 0) & 1:
            return f"WafVerticalMixed<{r},{variants.get('wafv_group', 4)},2>"
        return f"WafVertical<{r},{r},{variants.get('wafv_group', 4)}>"
    names = {"sound_div": "SoundDiv", "sound_momentum": "SoundMomentum",
             "adv_stagger_in": "StaggerIn", "adv_curvature": "Curvature",
             "adv_stagger_out": "StaggerOut", "uv_stagger": "StaggerOut", "tail": "Tail"}
    return f"{names[slot]}<{r}>" if slot in names else None


def read_traffic(slot: str, workload: str, variants: dict, form: str, kz: int = 41,
                 crm: int = 0):
    """HBM bytes per launch of the kernel of `slot` from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, copy-kernel
    calibrated) for this workload and form, if there is one."""
    path = ROOT / "profiles" / f"pmc_traffic_{workload}_{form}.md"
    name = kernel_struct(slot, variants, form, kz, crm)
    if not path.exists() or name is None:
        return None
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == name:
            try:
                return float(cells[4])
            except ValueError:
                return None
    return None


def cpu_model() -> str:
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload: str, run_s: float = 3.0, runs: int = 3) -> dict:
    """The CPU oracle (a C restatement of the reference step, gcc -O3
    -ffp-contract=off, OpenMP over points) timed on this host: on every core
    of this job's CPU share and on 1 core, each the median of `runs` timed
    runs of about `run_s` seconds of steps after one warm-up step."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_api as O   # the oracle, timed here as the CPU baseline
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    # a GPU box gives one GPU's job 16 CPUs (OMP_NUM_THREADS is set to that)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or affinity
    cores = max(1, min(affinity, share))
    g, f0 = workload_config(workload)
    res = {}
    for threads in (cores, 1):
        f = O.copy_fields(f0)
        t0 = time.perf_counter()
        O.moloch_steps(g, f, 1, nthreads=threads)
        t1 = time.perf_counter() - t0
        nsteps = int(max(1, min(50, round(run_s / max(t1, 1e-3)))))
        rates = []
        for _ in range(runs):
            t0 = time.perf_counter()
            O.moloch_steps(g, f, nsteps, nthreads=threads)
            rates.append(g.npoints * nsteps / (time.perf_counter() - t0))
        res[threads] = (sorted(rates)[len(rates) // 2], nsteps, rates)
        progress(f"cpu baseline {threads} thread(s): {res[threads][0]:.3e} pt-steps/s")
    v, n, rates = res[cores]
    v1, n1, rates1 = res[1]
    return {"value": v, "unit": "grid-pt·steps/s", "cores": cores, "kind": "port",
            "sample": f"{workload}: median of {runs} runs of {n} steps of the C oracle "
                      f"(gcc -O3 -ffp-contract=off, OpenMP {cores} threads) after 1 warm-up "
                      f"step; 1 core: median of {runs} runs of {n1} step(s)",
            "cpu_model": cpu_model(), "affinity_cpus": affinity,
            "runs": [round(x) for x in rates],
            "one_core": {"value": v1, "cores": 1, "runs": [round(x) for x in rates1]}}


WORKLOADS = {
    "cfg2": "MOLOCH step, 256x256x41, dx 5 km, dt 15 s, nadv 2 x nsound 5, "
            "8 random hills (seed 0), jet + N(0,1) winds, 5 water species (ipptls 2)",
    "cfg1": "MOLOCH step, 64x64x23, dx 12 km, dt 30 s, one 1 km hill, "
            "u = 10 + N(0, 0.25) m/s, 2 water species (ipptls 0)",
    "cfg3": "MOLOCH step, 512x512x41, dx 3 km, dt 10 s, 256x256x41 recipe + 3 K "
            "warm bubble (R 10 km, 1.5 km altitude)",
    "isc24": "MOLOCH step on the doubly periodic domain of Testing/isc24.in, 500x100x60, "
             "dx 2 km, dt 30 s, two hills and a 3 K warm bubble across the seams",
}
FORMS = {"fp64": 0, "mixed": 1}
ERROR_FIELDS = ("u", "v", "w", "pai", "tetav", "t", "qx", "ps")


def _owned(name, a, g):
    """The extents the reference step defines (cross fields 1..jx-1, 1..iy-1;
    u on dot j, v on dot i; every point on a doubly periodic domain)."""
    if getattr(g, "crm", 0):
        return a
    if name == "u":
        return a[..., : g.iy - 1, :]
    if name == "v":
        return a[..., :, : g.jx - 1]
    return a[..., : g.iy - 1, : g.jx - 1]


_CONFIGS: dict = {}
_T0 = time.perf_counter()
_TUNE = True


def progress(msg: str) -> None:
    """One line per finished phase on stderr (the JSON line goes to stdout)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def workload_config(workload: str):
    """(grid, initial fields) of a workload, generated once per process."""
    from regcm_amd import synth

    if workload not in _CONFIGS:
        g, f, _ = synth.config(workload)
        _CONFIGS[workload] = (g, f)
        progress(f"generated {workload}")
    return _CONFIGS[workload]


def time_form(workload: str, precision: int, steps: int, warmup: int) -> dict:
    """Timed run of one form: upload, init, autotune the launch shapes
    (state restored), W warm-up steps with every kernel timed by a pair of
    HIP events its launch records (per-kernel table; the dominant kernel is
    the one with the largest total), then K timed steps with events on the
    dominant kernel's launches in one step of every EVENT_EVERY: a pair of
    events around every launch of it would cost the step about 2%."""
    from regcm_amd.moloch import Moloch

    g, f = workload_config(workload)
    with Moloch(g, precision) as m:
        m.load(f)
        m.init_moloch()
        tuned_ms = m.autotune(8) if _TUNE else None
        variants = m.variants()
        m.profile(True)
        m.moloch(warmup)
        m.synchronize()
        warm = m.kernel_stats()
        dom = max(warm.items(), key=lambda kv: kv[1][1])[0]
        n_ev, ms_ev, sampled = 0, 0.0, 0
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            m.profile(True, slots=[dom])
            m.moloch(1)
            n, ms = m.kernel_stats()[dom]     # waits for the sampled step
            n_ev, ms_ev, sampled = n_ev + n, ms_ev + ms, sampled + 1
            m.profile(False)
            rest = min(EVENT_EVERY, steps - done) - 1
            if rest > 0:
                m.moloch(rest)
            done += rest + 1
        m.synchronize()
        dt = time.perf_counter() - t0
        psmax, psmin = m.check_finite()
        npts = m.npoints
    progress(f"timed {workload} precision={precision}: {dt / steps * 1e3:.3f} ms/step "
             f"(tuned {variants})")
    return {"g": g, "value": npts * steps / dt, "ms_per_step": dt / steps * 1e3,
            "stats": warm, "dominant": (dom, (n_ev, ms_ev)), "dominant_steps": sampled,
            "variants": variants,
            "autotune_ms_per_step": tuned_ms,
            "ps_range_pa": [psmin, psmax], "npoints": npts}


def final_state(workload: str, precision: int, nsteps: int, round_inputs=False) -> dict:
    """State after `nsteps` steps from the workload's initial state."""
    import numpy as np
    from regcm_amd.moloch import Moloch

    g, f = workload_config(workload)
    if round_inputs:
        f = {k: (a.astype(np.float32).astype(np.float64)
                 if isinstance(a, np.ndarray) and k not in ("pai", "w") else a)
             for k, a in f.items()}
    with Moloch(g, precision) as m:
        m.load(f)
        m.init_moloch()
        m.moloch(nsteps)
        m.synchronize()
        out = {n: _owned(n, m.get(n), g) for n in ERROR_FIELDS}
    progress(f"state of {workload} precision={precision} after {nsteps} steps")
    return out


def field_errors(ref: dict, got: dict) -> dict:
    """Relative L2 and max abs error per prognostic field (each species of qx)."""
    import numpy as np
    from regcm_amd.moloch import relative_l2

    out = {}
    for n in ERROR_FIELDS:
        pairs = ([(f"qx{q + 1}", ref[n][q], got[n][q]) for q in range(ref[n].shape[0])]
                 if n == "qx" else [(n, ref[n], got[n])])
        for name, a, b in pairs:
            if not np.any(a) and not np.any(b):
                continue
            out[name] = {"rel_l2": float(f"{relative_l2(a, b):.3e}"),
                         "max_abs": float(f"{float(np.max(np.abs(b - a))):.3e}")}
    return out


def kernel_table(stats: dict, g, F: int, form: str, steps: int, terrain: int = 0) -> dict:
    """Per-kernel-slot launches, mean launch time, share of the timed kernel
    time and achieved GB/s: the slot's algorithmic bytes per step (every call
    of the step, whatever number of launches a call takes) times `steps`,
    over the slot's total time."""
    total = sum(v[1] for v in stats.values())
    out = {}
    for k, (n, ms) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        row = {"launches": n, "avg_us": round(ms / n * 1e3, 2), "share": round(ms / total, 4)}
        b = kernel_bytes_per_launch(k, g, F, form, terrain)
        if b and ms > 0:
            row["GBps"] = round(b * calls_per_step(k, g) * steps / (ms * 1e-3) / 1e9, 1)
        out[k] = row
    return out


def mixed_report(workload: str, steps: int, warmup: int, fp64_value: float | None,
                 err_steps: int) -> dict:
    """The mixed form on `workload`: throughput, speed-up over plain FP64 on
    the same workload, and per-field error after `err_steps` steps against
    the FP64 form (which is bit-identical to the FP64 oracle).  The
    `input_rounding_floor` is the error of an FP64 run whose initial state
    was rounded to float (all but pai and w, which the mixed form keeps in
    double): how far rounding alone moves this flow."""
    r = time_form(workload, FORMS["mixed"], steps, warmup)
    if fp64_value is None:
        fp64_value = time_form(workload, FORMS["fp64"], steps, warmup)["value"]
    ref = final_state(workload, FORMS["fp64"], err_steps)
    got = final_state(workload, FORMS["mixed"], err_steps)
    rounded = final_state(workload, FORMS["fp64"], err_steps, round_inputs=True)
    F = nadv_fields(r["g"])
    return {"value": r["value"], "ms_per_step": r["ms_per_step"],
            "speedup_vs_fp64": round(r["value"] / fp64_value, 3),
            "fp64_value": fp64_value,
            "error_after_steps": err_steps,
            "error_vs_fp64": field_errors(ref, got),
            "input_rounding_floor": field_errors(ref, rounded),
            "kernels": kernel_table(r["stats"], r["g"], F, "mixed", warmup,
                                    r["variants"].get("terrain", 0))}


BATCHES = {
    "cfg4": (32, "ensemble of 32 members of 128x128x30, dx 12 km, dt 30 s, shared "
                 "terrain (8 hills, seed 0), T += N(0, 0.1^2), u, v += N(0, 0.2^2) per member"),
    "cfg5": (24, "24 uneven domains, jx, iy in {96..384}, kz in {23, 30, 41}, "
                 "1500 km wide, dt = 2.5 dx[km] s"),
}


def _member_fields(workload: str, m: int):
    from regcm_amd import synth
    g, f, _ = synth.config(workload, member=m)
    return g, f


def batch_run(workload: str, precision: int, steps: int, warmup: int,
              final_steps: int | None = None) -> dict:
    """The batched form on every member of `workload`: autotune (launch
    shapes per member, then the number of member groups), W warm-up steps,
    then K timed steps (one graph launch per group and step).  With
    final_steps, instead returns the members' states after that many steps
    (one group, no tuning: the grouping does not change results)."""
    from regcm_amd.moloch import MolochBatch

    n = BATCHES[workload][0]
    if workload not in _CONFIGS:
        _CONFIGS[workload] = [_member_fields(workload, m) for m in range(n)]
        progress(f"generated {n} members of {workload}")
    members = _CONFIGS[workload]
    with MolochBatch([g for g, _ in members], precision) as b:
        for i, (g, f) in enumerate(members):
            b.member(i).load(f)
            b.member(i).init_moloch()
        progress(f"batch {workload} precision={precision} loaded")
        if final_steps is not None:
            b.step(final_steps)
            b.synchronize()
            out = [{k: _owned(k, b.member(i).get(k), g) for k in ERROR_FIELDS}
                   for i, (g, _) in enumerate(members)]
            progress(f"batch {workload} precision={precision} state after {final_steps} steps")
            return out
        if _TUNE:
            b.autotune(8)
            progress(f"batch {workload} precision={precision} tuned: {b.streams} group(s)")
        b.step(warmup)
        b.synchronize()
        t0 = time.perf_counter()
        b.step(steps)
        b.synchronize()
        dt = time.perf_counter() - t0
        progress(f"batch {workload} precision={precision} groups={b.streams}: "
                 f"{dt / steps * 1e3:.3f} ms/step")
        # FP64 fused-floor bytes of every member over the step time
        floor = sum(step_bytes_floor(g, nadv_fields(g)) for g, _ in members)
        return {"value": b.npoints * steps / dt, "ms_per_step": dt / steps * 1e3,
                "npoints": b.npoints, "members": n, "groups": b.streams,
                "step_roofline_frac_fp64_floor": round(floor / (dt / steps) / 1e9
                                                      / HBM_PEAK_GBS, 4)}


def singles_run(workload: str, precision: int, steps: int, warmup: int) -> dict:
    """The same members as independent domains (one handle and stream each),
    stepped one after the other every step: the plain single-domain path."""
    from regcm_amd.moloch import Moloch

    members = _CONFIGS[workload]
    hs = []
    try:
        for g, f in members:
            m = Moloch(g, precision)
            hs.append(m)
            m.load(f)
            m.init_moloch()
            if _TUNE:
                m.autotune(8)
        for _ in range(warmup):
            for m in hs:
                m.moloch(1)
        for m in hs:
            m.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            for m in hs:
                m.moloch(1)
        for m in hs:
            m.synchronize()
        dt = time.perf_counter() - t0
        npts = sum(m.npoints for m in hs)
        progress(f"single domains {workload}: {dt / steps * 1e3:.3f} ms/step")
    finally:
        for m in hs:
            m.close()
    return {"value": npts * steps / dt, "ms_per_step": dt / steps * 1e3}


def batch_report(workload: str, steps: int, warmup: int, err_steps: int) -> dict:
    """Batched FP64 and mixed forms on a batch workload (member grouping
    chosen by the batch autotuner), the single-domain FP64 path on the same
    members, and the mixed form's per-field error (all members pooled) after
    err_steps steps against the FP64 form."""
    import numpy as np

    fp64 = batch_run(workload, FORMS["fp64"], steps, warmup)
    mixed = batch_run(workload, FORMS["mixed"], steps, warmup)
    single = singles_run(workload, FORMS["fp64"], steps, warmup)
    ref = batch_run(workload, FORMS["fp64"], 0, 0, final_steps=err_steps)
    got = batch_run(workload, FORMS["mixed"], 0, 0, final_steps=err_steps)
    pooled_ref = {k: np.concatenate([r[k].ravel() if k != "qx" else r[k].reshape(r[k].shape[0], -1)
                                     for r in ref], axis=-1) for k in ERROR_FIELDS}
    pooled_got = {k: np.concatenate([r[k].ravel() if k != "qx" else r[k].reshape(r[k].shape[0], -1)
                                     for r in got], axis=-1) for k in ERROR_FIELDS}
    return {"description": BATCHES[workload][1], "members": fp64["members"],
            "grid_points": fp64["npoints"],
            "single_domain_fp64": single,
            "batched_fp64": {**fp64, "speedup_vs_single_fp64":
                             round(fp64["value"] / single["value"], 3)},
            "batched_mixed": {**mixed, "speedup_vs_single_fp64":
                              round(mixed["value"] / single["value"], 3),
                              "speedup_vs_batched_fp64": round(mixed["value"] / fp64["value"], 3),
                              "error_after_steps": err_steps,
                              "error_vs_fp64": field_errors(pooled_ref, pooled_got)}}


def attainable_bandwidth(gib: float = 1.0, reps: int = 10) -> dict:
    """What streaming kernels reach on this GPU: torch's elementwise add
    (two reads, one write) and fill (one write) over `gib` GiB float64
    vectors, HIP events around `reps` launches.  A reference point for the
    kernels' fractions of the 8 TB/s peak, not a bound on them.

    Measured in a child process: torch carries its own HIP runtime under the
    same soname as the system one, and once torch has loaded it, the
    library's kernels in this process run on it too (on MI355X about 3.5%
    slower per step than on the system runtime the library is built with)."""
    import subprocess
    code = ("import json, bench; "
            f"print(json.dumps(bench._attainable_here({gib!r}, {reps!r})))")
    r = subprocess.run([sys.executable, "-c", code], cwd=str(ROOT), capture_output=True,
                       text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"streaming-bandwidth probe failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def _attainable_here(gib: float, reps: int) -> dict:
    import torch

    n = int(gib * 2 ** 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    c = torch.empty(n, dtype=torch.float64, device="cuda")
    out = {}
    for name, fn, nbytes in (("add_2r1w", lambda: torch.add(a, b, out=c), 24 * n),
                             ("fill_1w", lambda: c.fill_(2.0), 8 * n)):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name + "_gbs"] = round(nbytes / ms / 1e6, 1)
    del a, b, c
    torch.cuda.empty_cache()
    return out


def operators_report(workload: str = "cfg2", reps: int = 50) -> dict:
    """The operators beside the step, at the workload's shape.

    mkslice (Main/mod_slice.F90:112-382) on the step's state: mean kernel
    time from HIP events on the library stream over `reps` calls.
    Algorithmic bytes per call: reads pai, w, zetaf, t, p, rho, qsat and the
    nqx species, writes pf, th3d, rhb3d, wpx3d, i.e. (11 + nqx) x 8 B per
    grid point.

    MM-core horizontal diffusion (Main/mod_diffusion.F90), idiffu = 1,
    non-hydrostatic coefficients, nqx species: mean wall time per call over
    `reps` back-to-back calls (stream synchronized at both ends).  Bytes per
    grid point: calc_coeff 6 arrays (ud, vd, wx in; xkc, xkd, xkcf out),
    diffu_d 7 (u, v, xkd, uten, vten in; uten, vten out), diffu_x3d 4 (f,
    ften, xkc in; ften out), diffu_x4d over n species 3n + 1, diffu_x3df 4
    (over kz+1 levels)."""
    from regcm_amd import synth
    from regcm_amd.moloch import (MLC_DIFFU_X3D, MLC_DIFFU_X3DF, MLC_DIFFU_X4D, Diffusion,
                                  Moloch)

    g, f = workload_config(workload)
    npts = (g.jx - 1) * (g.iy - 1) * g.kz
    out = {"workload": workload, "grid_points": npts}
    si = synth.slice_inputs(g, f)
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        m.load_slice(si)
        m.mkslice(si)
        m.synchronize()
        m.profile(True, slots=["mkslice"])
        for _ in range(reps):
            m.mkslice(si)
        m.synchronize()
        n, ms = m.kernel_stats()["mkslice"]
        m.profile(False)
    us = ms / n * 1e3
    b = (11 + g.nqx) * 8.0 * npts
    out["mkslice"] = {"us_per_call": round(us, 2), "bytes_per_call": b,
                      "achieved_gbs": round(b / us / 1e3, 1),
                      "frac": round(b / us / 1e3 / HBM_PEAK_GBS, 4)}
    progress(f"mkslice {workload}: {us:.1f} us/call")
    dg = synth.DiffGrid(g.jx, g.iy, g.kz, dx=g.dx, dtsec=g.dtsec, idiffu=1, idynamic=2,
                        nspec=g.nqx)
    inp = synth.diffusion_inputs(dg)
    ns = g.nqx
    npf = (g.jx - 1) * (g.iy - 1) * (g.kz + 1)
    ops = [("calc_coeff", lambda d: d.calc_coeff(), 6 * npts),
           ("diffu_d", lambda d: d.diffu_d(), 7 * npts),
           ("diffu_x3d", lambda d: d.diffu_x(MLC_DIFFU_X3D, 1, 1), 4 * npts),
           (f"diffu_x4d_{ns}species", lambda d: d.diffu_x(MLC_DIFFU_X4D, 1, ns, 1.0),
            (3 * ns + 1) * npts),
           ("diffu_x3df", lambda d: d.diffu_x(MLC_DIFFU_X3DF, 1, 1, 1.0), 4 * npf)]
    diff = {"idiffu": 1, "idynamic": 2, "timing": "wall time over back-to-back calls"}
    with Diffusion(dg) as d:
        d.load(inp)
        d.initialize()
        d.calc_coeff()
        d.synchronize()
        for name, fn, words in ops:
            fn(d)
            d.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn(d)
            d.synchronize()
            us = (time.perf_counter() - t0) / reps * 1e6
            b = words * 8.0
            diff[name] = {"us_per_call": round(us, 2), "bytes_per_call": b,
                          "achieved_gbs": round(b / us / 1e3, 1),
                          "frac": round(b / us / 1e3 / HBM_PEAK_GBS, 4)}
    out["diffusion"] = diff
    progress(f"diffusion operators {workload} timed")
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cfg2", choices=sorted(WORKLOADS))
    ap.add_argument("--form", default="fp64", choices=sorted(FORMS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the report of the other forms")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the library's default launch shapes (profiling runs)")
    ap.add_argument("--operators-only", action="store_true",
                    help="only time mkslice and the diffusion operators (one JSON line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also launch a copy-like kernel (dst = -src, 2 GiB of float64) "
                         "that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only (the single-domain MOLOCH "
              "step has no multi-GPU path); run with --gpus 1", file=sys.stderr)
        return 2

    if args.operators_only:
        print(json.dumps(operators_report(args.workload)), flush=True)
        return 0
    if args.copy_calibration:
        # copy-like kernel of known traffic: dst = -src over 2 GiB of float64,
        # with the source pushed out of the 256 MB last-level cache first by
        # writing a 1 GiB buffer
        import torch
        n = 2 * 1024 * 1024 * 1024 // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        evict = torch.zeros(n // 2, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        del evict
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
        progress("copy calibration kernel launched")
    global _TUNE
    _TUNE = not args.no_tune
    attainable = attainable_bandwidth()
    progress(f"attainable streaming bandwidth {attainable}")
    r = time_form(args.workload, FORMS[args.form], args.steps, args.warmup)
    g, stats, npts = r["g"], r["stats"], r["npoints"]
    F = nadv_fields(g)
    dt_step = r["ms_per_step"] * 1e-3
    # dominant kernel (largest total in the warm-up profile), timed with
    # events on its own launches over the timed region
    dname, (dlaunch, dms) = r["dominant"]
    avg_s = dms / dlaunch * 1e-3
    # bytes per call; the FP64 dominant kernels are one launch per call
    bpl = kernel_bytes_per_launch(dname, g, F, args.form, r["variants"].get("terrain", 0))
    per_call = dlaunch / (calls_per_step(dname, g) * r["dominant_steps"])
    if bpl:
        bpl = bpl / per_call
    achieved = (bpl / avg_s / 1e9) if bpl else None
    traffic = read_traffic(dname, args.workload, r["variants"], args.form, g.kz,
                           getattr(g, "crm", 0))
    out = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "grid-pt·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": WORKLOADS[args.workload],
                   "form": args.form, "jx": g.jx, "iy": g.iy, "kz": g.kz,
                   "grid_points": npts, "advected_fields": F, "dtsec": g.dtsec,
                   "dx": g.dx, "nadv": g.nadv, "nsound": g.nsound},
        "roofline": {"bound": "hbm", "kernel": dname,
                     "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "frac_of_attainable": round(achieved / attainable["add_2r1w_gbs"], 4)
                     if achieved else None,
                     "traffic": traffic,
                     "algorithmic_bytes_per_launch": bpl,
                     "avg_launch_us": round(avg_s * 1e6, 2),
                     "launches_timed": dlaunch, "timed_steps_with_events": r["dominant_steps"]},
        "attainable": attainable,
        "step_roofline": {"floor_bytes_per_step": step_bytes_floor(g, F),
                          "frac": round(step_bytes_floor(g, F) / dt_step / 1e9
                                        / HBM_PEAK_GBS, 4)},
        "kernels": kernel_table(stats, g, F, args.form, args.warmup,
                                r["variants"].get("terrain", 0)),
        "kernels_note": "per-kernel means over the warm-up steps, every launch timed by the events it records",
        "launch_shapes": r["variants"],
        "ps_range_pa": r["ps_range_pa"],
    }
    if not args.no_forms and args.form == "fp64":
        forms = {}
        forms["mixed_" + args.workload] = mixed_report(
            args.workload, args.steps, args.warmup, r["value"], 100)
        if args.workload == "cfg2":
            forms["mixed_cfg3"] = mixed_report("cfg3", 50, 3, None, 100)
            forms["periodic_isc24"] = mixed_report("isc24", 30, 3, None, 100)
            forms["batched_cfg4"] = batch_report("cfg4", 50, 3, 200)
            forms["batched_cfg5"] = batch_report("cfg5", 20, 2, 100)
        out["forms"] = forms
        out["operators"] = operators_report(args.workload)
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(args.workload)
    progress("cpu baseline done")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
