"""Writes tests/golden/ref_leaf.json.gz from the reference's own leaf modules.

Builds oracle/ref (Share/mod_constants.F90, Share/mod_zita.F90 and
Share/pfwsat.inc compiled verbatim by amdflang, plus a probe driver) into
oracle/_ref/, runs the probe and stores its bit patterns as a fixture:

    {"const": {name: hex}, "zita": {"tag kz m [ig] k": hex},
     "pfwsat": {"it ip": hex}, "temps": [...], "press": [...],
     "ztops": [...], "zhs": [...], "a0s": [...], "geos": [...]}

Run in a container that has the reference checkout:
    python tests/golden/gen_ref_leaf.py [/root/reference]
"""
import gzip
import json
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parents[2]


def main():
    refdir = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle" / "ref"), f"REFDIR={refdir}"],
                   check=True)
    text = (ROOT / "oracle" / "_ref" / "ref_probe.txt").read_text()
    out = {"const": {}, "zita": {}, "pfwsat": {},
           # sample points, as set in oracle/ref/ref_probe.F90
           "ztops": [30000.0, 28000.0, 25000.0], "zhs": [8000.0, 7500.0, 6000.0],
           "a0s": [0.0, 0.0, 0.5], "geos": [0.0, 4903.325, 19613.3],
           "temps": [180.0, 198.15, 210.5, 230.0, 250.25, 263.0, 273.15, 273.16,
                     285.0, 300.0, 320.0, 380.0],
           "press": [500.0, 2000.0, 20000.0, 50000.0, 85000.0, 101325.0]}
    for line in text.splitlines():
        parts = line.split()
        tag, hexv = parts[0], parts[-1].lower()
        idx = parts[1:-1]
        if not idx:
            out["const"][tag] = hexv
        elif tag == "pfwsat":
            out["pfwsat"][f"{idx[0]} {idx[1]}"] = hexv
        else:
            out["zita"][" ".join([tag] + idx)] = hexv
    dst = ROOT / "tests" / "golden" / "ref_leaf.json.gz"
    with gzip.open(dst, "wt") as fh:
        json.dump(out, fh, sort_keys=True)
    print(f"wrote {dst} ({len(out['const'])} constants, {len(out['zita'])} zita values, "
          f"{len(out['pfwsat'])} pfwsat values)")


if __name__ == "__main__":
    main()
