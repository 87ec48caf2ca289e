"""The terrain metrics formed from the orography (MLC_VAR_TERRAIN).

compute_moloch_static forms fmz, fmzf, rfmzu and rfmzv from the surface
geopotential (Main/mod_params.F90:3356-3385, md_fmz of
Share/mod_zita.F90:135-141, 165-170).  Given mo_h and ht, htu, htv, mlc_init
forms them the same way on the device and compares them with the uploaded
arrays bit for bit; only where all agree may the kernels form them instead
of streaming the arrays (per stage: MLC_VAR_TERRAIN is a mask of the
divergence 1, the column solve 2, the advection staggering 4 and the
horizontal WAF 8).  Bar: the check passes on the synthetic domains (whose
metrics equal the reference build's md_fmz bit for bit,
tests/test_oracle_ref_leaf.py), formed and streamed metrics give identical
bits in both forms, the formed path matches the oracle, and a single
differing metric value turns the formed path off.
"""
import sys

import numpy as np
import pytest

import oracle_api as O
from regcm_amd import synth
from regcm_amd.moloch import (MLC_FP64, MLC_MIXED, MLC_VAR_COLUMN, MLC_VAR_COLUMN_J,
                               MLC_VAR_TERRAIN, MLC_VAR_WAF_TILE, Moloch, MolochError)

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "pai", "tetav", "t", "tvirt", "qx", "ux", "vx", "p", "rho",
          "qsat", "ps")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(g, f, precision, nsteps, terrain=None, variants=()):
    with Moloch(g, precision) as m:
        for k, v in variants:
            m.set_variant(k, v)
        m.load(f)
        m.init_moloch()
        verified = m.get_variant(MLC_VAR_TERRAIN)
        if terrain is not None:
            m.set_variant(MLC_VAR_TERRAIN, terrain)
        m.moloch(nsteps)
        m.synchronize()
        return verified, {n: m.get(n) for n in FIELDS}


CASES = [
    ("small", {}),
    ("small-rotllr", dict(lrotllr=1)),
    ("small-kz70", dict(kz=70)),
    ("crm-small", {}),
]


@pytest.mark.parametrize("precision", [MLC_FP64, MLC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name,over", CASES, ids=[c[0] for c in CASES])
def test_formed_and_streamed_metrics_give_the_same_bits(name, over, precision):
    say(f"[terrain] {name} precision={precision}: formed vs streamed metrics")
    base = "crm-small" if name == "crm-small" else "small"
    g, f, _ = synth.config(base, nadv=2, nsound=3, **over)
    verified, formed = run(g, f, precision, 2, terrain=15)
    assert verified == (7 if precision == MLC_FP64 else 6), \
        "mlc_init did not verify the formed terrain metrics"
    _, streamed = run(g, f, precision, 2, terrain=0)
    for n in FIELDS:
        assert np.array_equal(formed[n], streamed[n]), n


@pytest.mark.parametrize("cj", [16, 32, 64])
def test_formed_metrics_match_the_oracle(cj):
    """FP64 with the formed metrics, column blocks of 16, 32 and 64
    columns and 8-row WAF tiles, bit-identical to the oracle."""
    say(f"[terrain] small fp64 formed metrics vs oracle, column_j {cj}")
    g, f, _ = synth.config("small", nadv=2, nsound=3)
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, 2)
    verified, got = run(g, f, MLC_FP64, 2, terrain=15,
                        variants=((MLC_VAR_COLUMN, 2), (MLC_VAR_COLUMN_J, cj),
                                  (MLC_VAR_WAF_TILE, 8)))
    assert verified == 7
    own = (Ellipsis, slice(0, g.iy - 1), slice(0, g.jx - 1))
    for n in ("pai", "tetav", "w", "t", "qx"):
        assert np.array_equal(ref[n][own], got[n][own]), n
    assert np.array_equal(ref["u"][..., : g.iy - 1, :], got["u"][..., : g.iy - 1, :])
    assert np.array_equal(ref["v"][..., :, : g.jx - 1], got["v"][..., :, : g.jx - 1])


@pytest.mark.parametrize("name", ["fmz", "fmzf", "rfmzu", "rfmzv"])
def test_one_differing_value_keeps_the_streamed_metrics(name):
    """One metric value one ulp off the formula: mlc_init leaves the formed
    path off, refuses to switch it on, and the step streams the uploaded
    (perturbed) arrays, as the oracle does."""
    say(f"[terrain] {name} perturbed by one ulp")
    g, f, _ = synth.config("small", nadv=2, nsound=3)
    f[name][3, 5, 7] = np.nextafter(f[name][3, 5, 7], np.inf)
    with Moloch(g, MLC_FP64) as m:
        m.load(f)
        m.init_moloch()
        assert m.get_variant(MLC_VAR_TERRAIN) == 0
        with pytest.raises(MolochError):
            m.set_variant(MLC_VAR_TERRAIN, 1)
        m.moloch(1)
        m.synchronize()
        got = {n: m.get(n) for n in ("pai", "w", "u", "v")}
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, 1)
    own = (Ellipsis, slice(0, g.iy - 1), slice(0, g.jx - 1))
    assert np.array_equal(ref["pai"][own], got["pai"][own])
    assert np.array_equal(ref["w"][own], got["w"][own])


def test_without_orography_the_metrics_are_streamed():
    say("[terrain] no ht/htu/htv uploaded")
    g, f, _ = synth.config("small", nadv=2, nsound=3)
    for n in ("ht", "htu", "htv"):
        del f[n]
    with Moloch(g, MLC_FP64) as m:
        m.load(f)
        m.init_moloch()
        assert m.get_variant(MLC_VAR_TERRAIN) == 0
