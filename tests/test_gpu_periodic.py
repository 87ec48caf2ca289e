"""Doubly periodic domains (i_crm = 1) on the device against the oracle.

Every extent covers the whole domain, no edge is a lateral boundary, and a
neighbour past an edge is its periodic partner -- what the reference's
exchanges put in the halos on one process that is its own neighbour
(Main/mpplib/mod_mppparam.F90:1290-1309).  Bar: the prognostic fields
bit-identical to the oracle, p, rho, qsat, ps within a few ulp, over the
24 x 20 x 10 stressed-metric case in several configurations and launch
shapes, and the Testing/isc24_small.in domain (100 x 50 x 30, ds 2 km,
dt 30 s) over 10 steps.  Translation invariance of the oracle itself is
pinned by tests/test_oracle_periodic.py.
"""
import math
import sys

import numpy as np
import pytest

import oracle_api as O
from regcm_amd import synth
from regcm_amd.moloch import (MLC_FP64, MLC_MIXED, MLC_VAR_COLUMN, MLC_VAR_COLUMN_J,
                               MLC_VAR_LEVEL_CHUNK, MLC_VAR_WAF_H, MLC_VAR_WAF_HGROUP,
                               MLC_VAR_WAF_TILE, MLC_VAR_WAFV_GROUP, MLC_VAR_XCD_MAP,
                               MLC_VAR_STAGGER_CHUNK, MLC_VAR_MOM_CHUNK,
                               Moloch, MolochBatch, MolochError, relative_l2)

pytestmark = pytest.mark.gpu

PROGNOSTIC = ("u", "v", "w", "pai", "tetav", "t", "tvirt", "qx", "ux", "vx")
TRANSCENDENTAL = ("p", "rho", "qsat", "ps")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compare(ref, got, label, extra=()):
    """Every point is owned on a periodic domain: whole arrays compared."""
    bad = []
    for n in PROGNOSTIC + tuple(extra):
        if not np.array_equal(ref[n], got[n]):
            nd = int(np.count_nonzero(ref[n] != got[n]))
            bad.append(f"{n}: {nd} values differ, max |d| = {np.max(np.abs(ref[n] - got[n])):.3e}")
    for n in TRANSCENDENTAL:
        a, b = ref[n], got[n]
        rel = np.max(np.abs(a - b) / np.maximum(np.abs(a), np.finfo(float).tiny))
        if not rel <= 8e-16:
            bad.append(f"{n}: max relative difference {rel:.3e}")
    assert not bad, f"{label}: " + "; ".join(bad)


def run_gpu(g, f, nsteps, precision=MLC_FP64, variants=None, extra=()):
    with Moloch(g, precision) as m:
        for stage, value in (variants or {}).items():
            m.set_variant(stage, value)
        m.load(f)
        m.init_moloch()
        m.moloch(nsteps)
        m.synchronize()
        return {n: m.get(n) for n in PROGNOSTIC + TRANSCENDENTAL + tuple(extra)}


CASES = [
    ("crm-small", dict(), 3),
    ("crm-small-default-split", dict(nadv=2, nsound=5), 2),
    ("crm-small-ipptls0", dict(ipptls=0, nqx=2), 2),
    ("crm-small-nodamp-nofilter", dict(do_divdamp=0, do_divfilter=0), 2),
    ("crm-small-odd-shape", dict(jx=70, iy=9, kz=7), 2),
    ("crm-small-tke-tracers", dict(ibltyp=2, ntr=3, tkemin=0.01, nadv=2, nsound=3), 2),
    ("crm-small-kz70", dict(kz=70), 2),
]


@pytest.mark.parametrize("name,over,nsteps", CASES, ids=[c[0] for c in CASES])
def test_periodic_small_cases_bit_exact(name, over, nsteps):
    say(f"[periodic] {name} steps={nsteps}")
    g, f, _ = synth.config("crm-small", **over)
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, nsteps)
    extra = (("tke",) if g.ibltyp == 2 else ()) + (("trac",) if g.ntr else ())
    got = run_gpu(g, O.copy_fields(f), nsteps, extra=extra)
    compare(ref, got, name, extra)


@pytest.mark.parametrize("variants", [
    {MLC_VAR_COLUMN: 1},
    {MLC_VAR_COLUMN: 0},
    {MLC_VAR_COLUMN_J: 16},
    {MLC_VAR_COLUMN_J: 64},
    {MLC_VAR_WAF_TILE: 8, MLC_VAR_XCD_MAP: 0},
    {MLC_VAR_WAF_HGROUP: 1, MLC_VAR_LEVEL_CHUNK: 3},
    {MLC_VAR_WAFV_GROUP: 1},
    {MLC_VAR_WAF_H: 0},              # row-streaming form: the tile form runs instead
    {MLC_VAR_STAGGER_CHUNK: 3, MLC_VAR_MOM_CHUNK: 7},
], ids=["column-stream", "column-three-phase", "column16", "column64", "tile8-plain", "hgroup1-chunk3",
        "wafv-g1", "rows-requested", "stagger-chunk3-mom-chunk7"])
@pytest.mark.parametrize("over", [dict(nadv=2, nsound=5), dict(jx=70, iy=9, kz=7)],
                         ids=["default", "odd-shape"])
def test_periodic_launch_shapes_bit_exact(over, variants):
    say(f"[periodic] crm-small {over} variants {variants}")
    g, f, _ = synth.config("crm-small", **over)
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, 2)
    got = run_gpu(g, O.copy_fields(f), 2, variants=variants)
    compare(ref, got, f"crm-small {over} {variants}")


def test_isc24_small_domain_ten_steps_bit_exact():
    say("[periodic] crm 100x50x30 (Testing/isc24_small.in) 10 steps")
    g, f, nsteps = synth.config("crm")
    assert (g.jx, g.iy, g.kz, g.dx, g.dtsec, g.crm) == (100, 50, 30, 2000.0, 30.0, 1)
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, nsteps)
    got = run_gpu(g, O.copy_fields(f), nsteps)
    compare(ref, got, "crm 10 steps")
    # the seam moved: the bubble and the hills straddle it
    assert np.max(np.abs(got["w"][:, :, 0])) > 1e-3


def test_periodic_mixed_form_close_to_fp64():
    say("[periodic] crm 100x50x30 mixed vs FP64, 10 steps")
    g, f, nsteps = synth.config("crm")
    a = run_gpu(g, O.copy_fields(f), nsteps, MLC_FP64)
    b = run_gpu(g, O.copy_fields(f), nsteps, MLC_MIXED)
    errs = {n: relative_l2(a[n], b[n]) for n in ("u", "v", "pai", "tetav", "t", "ps")}
    errs["qv"] = relative_l2(a["qx"][0], b["qx"][0])
    errs["w"] = relative_l2(a["w"], b["w"])
    say("[periodic] mixed relative L2: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    bounds = {n: 1e-4 for n in ("u", "v", "w", "pai", "tetav", "t", "qv", "ps")}   # the target
    for n, bound in bounds.items():
        assert errs[n] <= bound, (n, errs[n], bound)


def test_periodic_massck_has_no_boundary_flux():
    say("[periodic] massck crm-small after 1 step")
    g, f, _ = synth.config("crm-small")
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        m.moloch(1)
        m.synchronize()
        got = m.massck()
        state = {n: m.get(n) for n in ("u", "v", "rho", "qx")}
    ff = dict(f, **state)
    ref = O.massck(g, ff)
    assert got["dryadv"] == 0.0 and got["qadv"] == 0.0
    assert ref["dryadv"] == 0.0 and ref["qadv"] == 0.0
    terms = {"drymass": [(g.dx * g.dx * ff["dz"]) * ff["rho"]],
             "qmass": [((q * (g.dx * g.dx)) * ff["dz"]) * ff["rho"] for q in ff["qx"]]}
    for key, parts in terms.items():
        flat = np.concatenate([p.ravel() for p in parts])
        exact = math.fsum(flat.tolist())
        assert abs(got[key] - exact) <= 4 * np.finfo(float).eps * abs(exact), key
        n = flat.size
        assert abs(ref[key] - exact) <= n * np.finfo(float).eps * float(np.abs(flat).sum()), key


def test_periodic_batch_with_bounded_members():
    say("[periodic] batch: two periodic and one bounded member vs single handles")
    doms = [synth.config("crm-small")[:2], synth.config("crm-small", jx=30, iy=26)[:2],
            synth.config("small")[:2]]
    alone = [run_gpu(g, O.copy_fields(f), 2) for g, f in doms]
    with MolochBatch([g for g, _ in doms], MLC_FP64) as b:
        for m, (g, f) in enumerate(doms):
            b.member(m).load(O.copy_fields(f))
            b.member(m).init_moloch()
        b.step(2)
        b.synchronize()
        for m, (g, _) in enumerate(doms):
            got = {n: b.member(m).get(n) for n in PROGNOSTIC + TRANSCENDENTAL}
            for n in PROGNOSTIC + TRANSCENDENTAL:
                assert np.array_equal(alone[m][n], got[n]), (m, n)


def test_periodic_refuses_lateral_boundaries_and_rotllr():
    g, f, _ = synth.config("crm-small")
    with Moloch(g) as m:
        with pytest.raises(MolochError, match="no lateral boundary"):
            m.lbc_configure()
    g.lrotllr = 1
    with pytest.raises(MolochError, match="ROTLLR"):
        Moloch(g)


def test_periodic_mkslice_matches_oracle():
    """mkslice on a periodic domain: every column is an interior column."""
    from regcm_amd.moloch import SLICE_FIELDS
    say("[periodic] mkslice crm-small")
    g, f, _ = synth.config("crm-small")
    f["w"][1:-1] = synth.normal(91, f["w"][1:-1].shape, 0.05)
    f["qx"][1][2, 3, 0] = 0.5e-16                  # below qxcheckval on the seam column
    si = synth.slice_inputs(g, f)
    ref = O.copy_fields(f)
    want = O.mkslice(g, ref, si)
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        m.load_slice(si)
        m.mkslice(si)
        got = {n: m.get(n) for n in SLICE_FIELDS}
        got["qx"] = m.get("qx")
    for n in ("rhb3d", "wpx3d", "rhox2d", "ktrop", "kmxpbl"):
        assert np.array_equal(want[n], got[n]), n
    for n in ("pf", "th3d", "tp2d", "th700", "ptrop"):
        a, c = want[n], got[n]
        assert np.max(np.abs(a - c) / np.maximum(np.abs(a), 1e-300)) <= 8.9e-16, n
    assert np.array_equal(ref["qx"], got["qx"])
    assert ref["qx"][1][2, 3, 0] == 0.0           # the floor reached the seam column


def test_isc24_domain_first_step_bit_exact():
    """The doubly periodic domain of Testing/isc24.in and ideal.in: 500 x 100
    x 60, ds 2 km, dt 30 s (60 levels: 8-level chunks of the column solve)."""
    say("[periodic] isc24 500x100x60 step 1 vs oracle")
    g, f, _ = synth.config("isc24")
    assert (g.jx, g.iy, g.kz, g.dx, g.dtsec, g.crm) == (500, 100, 60, 2000.0, 30.0, 1)
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, 1)
    got = run_gpu(g, O.copy_fields(f), 1)
    compare(ref, got, "isc24 step 1")
