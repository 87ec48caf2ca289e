"""Pins the oracle's constants, zita levels, gzita and pfwsat bit for bit
against the reference's own leaf modules (Share/mod_constants.F90,
Share/mod_zita.F90, Share/pfwsat.inc compiled verbatim; fixture written by
tests/golden/gen_ref_leaf.py), and the host-side metric builder's closed
forms (md_fmz, md_zeta with the C library's exp) bit for bit: the values the
library's formed terrain metrics must reproduce."""
import gzip
import json
import pathlib
import struct

import numpy as np
import pytest

import oracle_api as O
from regcm_amd import synth

GOLD = pathlib.Path(__file__).parent / "golden" / "ref_leaf.json.gz"


@pytest.fixture(scope="module")
def gold():
    with gzip.open(GOLD, "rt") as fh:
        return json.load(fh)


def h2d(h: str) -> float:
    return struct.unpack(">d", bytes.fromhex(h))[0]


def d2h(x: float) -> str:
    return struct.pack(">d", float(x)).hex()


def test_constants_bit_exact(gold):
    for name, hexv in gold["const"].items():
        assert d2h(O.constant(name)) == hexv, name


def test_python_constants_bit_exact(gold):
    c = gold["const"]
    for name, val in [("egrav", synth.EGRAV), ("rgas", synth.RGAS), ("cpd", synth.CPD),
                      ("cvd", synth.CVD), ("cpovr", synth.CPOVR), ("rovcp", synth.ROVCP),
                      ("govcp", synth.GOVCP), ("ep1", synth.EP1), ("ep2", synth.EP2),
                      ("tzero", synth.TZERO), ("p00", synth.P00), ("lrate", synth.LRATE),
                      ("eomeg2", synth.EOMEG2), ("degrad", synth.DEGRAD),
                      ("regrav", synth.REGRAV), ("mathpi", synth.MATHPI)]:
        assert d2h(val) == c[name], name


@pytest.mark.parametrize("kz", [10, 23, 41])
@pytest.mark.parametrize("m", [1, 2, 3])
def test_zita_levels_and_gzita_bit_exact(gold, kz, m):
    ztop, a0 = gold["ztops"][m - 1], gold["a0s"][m - 1]
    zf, zh, gf, gh = O.zita_levels(kz, ztop, a0)
    z = gold["zita"]
    for k in range(1, kz + 2):
        assert d2h(zf[k - 1]) == z[f"zitaf {kz} {m} {k}"]
        assert d2h(gf[k - 1]) == z[f"gzitaf {kz} {m} {k}"]
    for k in range(1, kz + 1):
        assert d2h(zh[k - 1]) == z[f"zitah {kz} {m} {k}"]
        assert d2h(gh[k - 1]) == z[f"gzitah {kz} {m} {k}"]
    # the host-side builder uses the same recurrences
    assert np.array_equal(synth.model_zitaf(kz, ztop), zf)
    assert np.array_equal(synth.model_zitah(kz, ztop), zh)
    assert np.array_equal(synth.gzita(zf, ztop, a0), gf)


@pytest.mark.parametrize("kz", [10, 23, 41])
def test_metric_closed_forms_bit_exact(gold, kz):
    z = gold["zita"]
    for m in (1, 2, 3):
        ztop, zh_, a0 = gold["ztops"][m - 1], gold["zhs"][m - 1], gold["a0s"][m - 1]
        zf = synth.model_zitaf(kz, ztop)
        zh = synth.model_zitah(kz, ztop)
        for ig, geo in enumerate(gold["geos"], start=1):
            for tag, lev, fn in (("fmzf", zf, synth.md_fmz), ("zetaf", zf, synth.md_zeta),
                                 ("fmzh", zh, synth.md_fmz), ("zetah", zh, synth.md_zeta)):
                got = fn(lev, geo, ztop, zh_, a0)
                ref = np.array([h2d(z[f"{tag} {kz} {m} {ig} {k}"]) for k in range(1, len(lev) + 1)])
                assert np.array_equal(got, ref), (tag, m, ig)


def test_pfwsat_bit_exact(gold):
    for it, t in enumerate(gold["temps"], start=1):
        for ip, p in enumerate(gold["press"], start=1):
            assert d2h(O.pfwsat(t, p)) == gold["pfwsat"][f"{it} {ip}"], (t, p)


def test_pfwsat_known_answer():
    # es(0 C) = a0 (Share/pfwsat.inc:10): ws = ep2*a0/(p-a0)
    p = 100000.0
    a0 = 0.611213476e+03
    assert O.pfwsat(273.15, p) == synth.EP2 * (a0 / (p - a0))
    # clamp es <= 0.15 p at low pressure
    es = 0.15 * 500.0
    assert O.pfwsat(320.0, 500.0) == synth.EP2 * (es / (500.0 - es))
