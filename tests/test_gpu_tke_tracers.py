"""TKE (ibltyp 2) and chemical tracers (ntr > 0) on the device against the
oracle: the step (reset of tketen / chiten, tke staggered to half levels,
advected and staggered back, the tracers advected, status_update with the
tkemin and zero floors) bit-identical in FP64; the TKE boundary rows with
lateral boundaries; more advected fields than one WAF launch holds; the
mixed form close; restart from the SAV variables and the batched form
unchanged."""
import numpy as np
import pytest

import oracle_api as O
from regcm_amd import synth
from regcm_amd.moloch import MLC_MIXED, Moloch, MolochBatch, MolochError

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "pai", "tetav", "t", "tvirt", "qx", "ux", "vx", "tke", "trac")


def say(msg):
    print(msg, flush=True)


def owned(name, a):
    if name == "u":
        return a[..., : a.shape[-2] - 1, :]
    if name == "v":
        return a[..., :, : a.shape[-1] - 1]
    return a[..., : a.shape[-2] - 1, : a.shape[-1] - 1]


def names(g):
    return [n for n in FIELDS if (n != "tke" or g.ibltyp == 2) and (n != "trac" or g.ntr > 0)]


def compare(g, ref, got, label):
    bad = []
    for n in names(g):
        a, b = owned(n, ref[n]), owned(n, got[n])
        if not np.array_equal(a, b):
            bad.append(f"{n}: {int(np.count_nonzero(a != b))} differ")
    assert not bad, f"{label}: " + "; ".join(bad)


CASES = [
    ("tke-tracers", dict(ibltyp=2, ntr=3, tkemin=0.01, nadv=2, nsound=3)),
    ("tke-only-ipptls0", dict(ibltyp=2, tkemin=0.02, ipptls=0, nqx=2)),
    ("tracers-only-rotllr", dict(ntr=2, lrotllr=1)),
    # 5 + 5 species + tkex + 20 tracers = 31 advected fields: two WAF launches
    ("31-fields", dict(ibltyp=2, ntr=20, tkemin=0.01)),
]


@pytest.mark.parametrize("name,over", CASES, ids=[c[0] for c in CASES])
def test_steps_bit_exact(name, over):
    say(f"[tke/tracers] {name} 3 steps")
    g, f, _ = synth.config("small", **over)
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, 3)
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        m.moloch(3)
        got = {n: m.get(n) for n in names(g)}
    compare(g, ref, got, name)


def test_status_update_with_tendencies_bit_exact():
    say("[tke/tracers] status_update with tendencies and floors")
    g, f, _ = synth.config("small", ibltyp=2, ntr=2, tkemin=0.05)
    f["tketen"][:] = synth.normal(31, f["tketen"].shape, 2e-3)
    f["chiten"][:] = synth.normal(32, f["chiten"].shape, 1e-10)
    ref = O.copy_fields(f)
    O.status_update(g, ref, g.dtsec)
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        m.put("tketen", f["tketen"])
        m.put("chiten", f["chiten"])
        m.status_update(g.dtsec)
        got = {n: m.get(n) for n in ("tke", "trac")}
    for n in ("tke", "trac"):
        assert np.array_equal(owned(n, ref[n]), owned(n, got[n])), n
    assert np.any(owned("tke", got["tke"]) == 0.05)          # the floor acted somewhere
    assert np.any(owned("trac", got["trac"]) == 0.0)


def test_tke_with_lateral_boundaries_bit_exact():
    say("[tke/tracers] TKE with lateral boundaries, 3 steps")
    g, f, _ = synth.config("small", ibltyp=2, tkemin=0.01, nadv=2, nsound=3)
    f["w"][1:-1] = synth.normal(91, f["w"][1:-1].shape, 0.05)
    b = synth.lbc_inputs(g, f, nspgx=5, nspgd=6)
    x1 = [0.1, 0.2, 0.3]
    ref = O.copy_fields(f)
    O.moloch_steps_lbc(g, ref, b, x1)
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        m.load_lbc(b)
        m.moloch_lbc(x1)
        got = {n: m.get(n) for n in names(g)}
    compare(g, ref, got, "lbc")


def test_lbc_steps_refuse_tracers():
    say("[tke/tracers] mlc_step_lbc refuses tracers")
    g, f, _ = synth.config("small", ntr=1)
    b = synth.lbc_inputs(g, f)
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        m.load_lbc(b)
        with pytest.raises(MolochError, match="morelax_chiten"):
            m.moloch_lbc([0.5])


def test_mixed_form_close():
    say("[tke/tracers] mixed form vs FP64, 3 steps")
    g, f, _ = synth.config("small", ibltyp=2, ntr=2, tkemin=0.01, nadv=2, nsound=3)
    out = []
    for prec in (0, MLC_MIXED):
        with Moloch(g, prec) as m:
            m.load(f)
            m.init_moloch()
            m.moloch(3)
            out.append({n: owned(n, m.get(n)) for n in ("tke", "trac")})
    for n in ("tke", "trac"):
        a, b = out[0][n], out[1][n]
        rel = np.linalg.norm(a - b) / np.linalg.norm(a)
        assert rel < 1e-5, (n, rel)


def test_restart_with_tke_and_tracers():
    """The SAV variables with tke (Main/mod_savefile.F90:618-628) and the
    chemistry module's tracers, put into a fresh handle, then
    mlc_init_restart: with ipptls 0 the continued run is the uninterrupted
    one bit for bit."""
    say("[tke/tracers] restart from the SAV variables")
    g, f, _ = synth.config("small", ibltyp=2, ntr=2, tkemin=0.01, ipptls=0, nqx=2)
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        m.moloch(2)
        sav = m.save_variables()
        assert "tke" in sav
        sav["trac"] = m.get("trac")
        m.moloch(2)
        ref = {n: m.get(n) for n in names(g)}
    r = O.copy_fields(f)
    for n in ("tetav", "tvirt", "p", "rho", "qsat"):
        owned(n, r[n])[...] = np.nan
    r["ux"][...] = 0.0
    r["vx"][...] = 0.0
    for n, a in sav.items():
        r[n][...] = a
    with Moloch(g) as m:
        m.load(r)
        m.init_moloch()
        m.init_restart()
        m.moloch(2)
        got = {n: m.get(n) for n in names(g)}
    compare(g, ref, got, "restart")


def test_batch_matches_single_domains():
    say("[tke/tracers] batch of 2 vs single domains")
    doms = [synth.config("small", ibltyp=2, ntr=2, tkemin=0.01, **o)[:2]
            for o in (dict(), dict(jx=30, iy=26))]
    alone = []
    for g, f in doms:
        with Moloch(g) as m:
            m.load(f)
            m.init_moloch()
            m.moloch(3)
            alone.append({n: m.get(n) for n in names(g)})
    with MolochBatch([g for g, _ in doms]) as b:
        for i, (g, f) in enumerate(doms):
            b.member(i).load(f)
            b.member(i).init_moloch()
        b.step(3)
        b.synchronize()
        for i, (g, _) in enumerate(doms):
            got = {n: b.member(i).get(n) for n in names(g)}
            compare(g, alone[i], got, f"member {i}")
