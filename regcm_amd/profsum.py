"""Summaries of rocprofv3 output for the profiles/ tables.

    python -m regcm_amd.profsum stats <kernel_stats.csv> <out.md> <title> <command>
    python -m regcm_amd.profsum pmc <fetch counter_collection.csv> <write counter_collection.csv>
                                    <out.md> <title> <command>

`stats` turns the `--kernel-trace --stats --output-format csv` kernel
statistics into a table (kernel, calls, total, mean, share, min, max).
`pmc` averages FETCH_SIZE and WRITE_SIZE (KiB per dispatch) per kernel from
two `--pmc` runs and scales both by the factor that makes the reference
copy-like kernel (dst = -src over CALIBRATION_BYTES bytes of float64, the
dispatch bench.py --copy-calibration launches, its source first evicted
from the last-level cache) read and write exactly its known bytes; the
calibrated HBM bytes per dispatch are in the last column.
"""
from __future__ import annotations

import csv
import re
import sys
from collections import defaultdict

CALIBRATION_BYTES = 2 * 1024 * 1024 * 1024   # bytes read (and written) by the copy-like kernel
CALIBRATION_PATTERN = re.compile(r"neg_kernel")
TORCH_PATTERN = re.compile(r"at::native::")      # torch's own allocation fills


def short_name(name: str) -> str:
    """mlc::k_one<mlc::SoundColumnChunk<double> >(...) -> SoundColumnChunk<double>"""
    m = re.search(r"k_(?:one_xcd|one|many)<mlc::([A-Za-z]+)<([^>]*)>", name)
    if m:
        return f"{m.group(1)}<{m.group(2).replace(' ', '')}>" + \
            (" (batched)" if "k_many" in name else "")
    m = re.search(r"mlc::(?:diff::)?(k_\w+)", name)
    if m:
        return m.group(1)
    return re.sub(r"\(.*", "", name)[:60]


def stats(src: str, dst: str, title: str, cmd: str) -> None:
    rows = []
    with open(src, newline="") as fh:
        for r in csv.DictReader(fh):
            if TORCH_PATTERN.search(r["Name"]):
                continue
            rows.append((short_name(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]),
                         float(r["AverageNs"]), 0.0, float(r["MinNs"]), float(r["MaxNs"])))
    total = sum(x[2] for x in rows) or 1.0
    rows = [(n, c, t, a, 100.0 * t / total, mn, mx) for n, c, t, a, _, mn, mx in rows]
    rows.sort(key=lambda x: -x[2])
    with open(dst, "w") as fh:
        fh.write(f"# {title}\n\nCommand (repository root, one MI355X):\n\n```\n{cmd}\n```\n\n")
        fh.write("torch's own kernels (input fills, the streaming-bandwidth reference) are left "
                 "out; shares are of the remaining kernel time.\n\n")
        fh.write("| kernel | calls | total (us) | mean (us) | share (%) | min (us) | max (us) |\n"
                 "|---|---:|---:|---:|---:|---:|---:|\n")
        for n, c, t, a, p, mn, mx in rows:
            fh.write(f"| {n} | {c} | {t / 1e3:.1f} | {a / 1e3:.2f} | {p:.2f} | "
                     f"{mn / 1e3:.2f} | {mx / 1e3:.2f} |\n")


def _counter_means(src: str, counter: str) -> dict:
    acc = defaultdict(list)
    with open(src, newline="") as fh:
        for r in csv.DictReader(fh):
            if r.get("Counter_Name", counter) != counter:
                continue
            acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def pmc(fetch_csv: str, write_csv: str, dst: str, title: str, cmd: str) -> None:
    fetch = _counter_means(fetch_csv, "FETCH_SIZE")
    write = _counter_means(write_csv, "WRITE_SIZE")
    cal = [k for k in fetch if CALIBRATION_PATTERN.search(k)]
    if not cal:
        raise SystemExit("no calibration kernel in the FETCH_SIZE run")
    known_kib = CALIBRATION_BYTES / 1024.0
    f_fac = known_kib / fetch[cal[0]][0]
    w_fac = known_kib / write[cal[0]][0] if cal[0] in write else 1.0
    rows = []
    for k, (fk, n) in fetch.items():
        if k in cal or TORCH_PATTERN.search(k):
            continue
        wk = write.get(k, (0.0, 0))[0]
        rows.append((short_name(k), n, fk, wk, (fk * f_fac + wk * w_fac) * 1024.0))
    rows.sort(key=lambda x: -x[4])
    with open(dst, "w") as fh:
        fh.write(f"# {title}\n\nCommands (repository root, one MI355X):\n\n```\n{cmd}\n```\n\n")
        fh.write(f"Calibration: the copy-like kernel (torch neg over float64) moves "
                 f"{CALIBRATION_BYTES} B each way; FETCH_SIZE read "
                 f"{fetch[cal[0]][0]:.1f} KiB (factor {f_fac:.4f}), WRITE_SIZE read "
                 f"{write.get(cal[0], (float('nan'), 0))[0]:.1f} KiB (factor {w_fac:.4f}).\n\n")
        fh.write("| kernel | dispatches | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | "
                 "HBM bytes per dispatch (calibrated) |\n|---|---:|---:|---:|---:|\n")
        for n, c, fk, wk, b in rows:
            fh.write(f"| {n} | {c} | {fk:.1f} | {wk:.1f} | {b:.4e} |\n")


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "stats" and len(sys.argv) == 6:
        stats(*sys.argv[2:6])
    elif len(sys.argv) >= 2 and sys.argv[1] == "pmc" and len(sys.argv) == 7:
        pmc(*sys.argv[2:7])
    else:
        raise SystemExit(__doc__)
