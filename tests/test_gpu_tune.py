"""The autotuner (mlc_autotune, mlc_batch_autotune): it only picks among
bit-identical launch shapes and restores the state it measured on, so a
tuned domain must step to exactly the same bits as an untuned one."""
import numpy as np
import pytest

from regcm_amd import synth
from regcm_amd.moloch import MLC_FP64, MLC_MIXED, Moloch, MolochBatch

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "pai", "tetav", "t", "qx", "ps")


def _run(g, f, precision, tune, nsteps=3):
    with Moloch(g, precision) as m:
        m.load(f)
        m.init_moloch()
        if tune:
            ms = m.autotune(2)
            assert ms > 0
            print("[tune] chose", m.variants(), f"{ms:.3f} ms/step", flush=True)
        m.moloch(nsteps)
        m.synchronize()
        return {n: m.get(n) for n in FIELDS}


@pytest.mark.parametrize("precision", [MLC_FP64, MLC_MIXED], ids=["fp64", "mixed"])
def test_autotune_leaves_results_unchanged(precision):
    print(f"[tune] cfg1 precision={precision}", flush=True)
    g, f, _ = synth.config("cfg1")
    a = _run(g, f, precision, tune=False)
    b = _run(g, f, precision, tune=True)
    for n in FIELDS:
        assert np.array_equal(a[n], b[n]), n


def test_batch_with_members_on_different_variants():
    print("[tune] batch of 3 with per-member variants", flush=True)
    doms = [synth.config("small", **o)[:2] for o in (dict(), dict(jx=30, iy=26),
                                                     dict(jx=70, iy=22, kz=12))]
    alone = []
    for g, f in doms:
        with Moloch(g) as m:
            m.load(f)
            m.init_moloch()
            m.moloch(3)
            alone.append({n: m.get(n) for n in FIELDS})
    with MolochBatch([g for g, _ in doms]) as b:
        for i, (g, f) in enumerate(doms):
            b.member(i).load(f)
            b.member(i).init_moloch()
        b.member(1).set_variant(0, 0)       # row-streaming WAF on member 1 only
        b.member(2).set_variant(2, 2)       # two fields per lane on member 2
        b.member(0).set_variant(4, 1)       # streamed column solve on member 0
        b.step(1)
        b.autotune(1)                        # retune all, state unchanged
        assert b.streams in (1, 2)           # the tuned member grouping
        b.step(2)
        b.synchronize()
        for i in range(3):
            for n in FIELDS:
                assert np.array_equal(b.member(i).get(n), alone[i][n]), (i, n)


def test_autotune_with_tke_and_tracers_leaves_results_unchanged():
    """The tuner's snapshot holds tke, its tendency and the tracers' (the
    tracers themselves are advected slots), so tuning leaves them as found."""
    print("[tune] small with TKE and 3 tracers", flush=True)
    g, f, _ = synth.config("small", ibltyp=2, ntr=3, tkemin=0.01, nadv=2, nsound=3)
    names = FIELDS + ("tke", "trac")
    out = []
    for tune in (False, True):
        with Moloch(g) as m:
            m.load(f)
            m.init_moloch()
            if tune:
                assert m.autotune(1) > 0
            m.moloch(3)
            m.synchronize()
            out.append({n: m.get(n) for n in names})
    for n in names:
        assert np.array_equal(out[0][n], out[1][n]), n


def test_mixed_batch_autotune_agrees_on_launch_counts():
    """After mlc_batch_autotune the members of a mixed batch agree on the
    knobs that set how many launches a WAF stage records (MLC_VAR_WAF_H,
    MLC_VAR_WAF_HGROUP, MLC_VAR_MIXED_MERGE), whatever each member's own
    tuning chose, and step to the single-domain bits."""
    from regcm_amd.moloch import MLC_VAR_MIXED_MERGE, MLC_VAR_WAF_H, MLC_VAR_WAF_HGROUP
    print("[tune] mixed batch of 3: autotune, then agreement", flush=True)
    doms = [synth.config("small", **o)[:2] for o in (dict(), dict(jx=30, iy=26),
                                                     dict(jx=70, iy=22, kz=12))]
    alone = []
    for g, f in doms:
        with Moloch(g, MLC_MIXED) as m:
            m.load(f)
            m.init_moloch()
            m.moloch(2)
            alone.append({n: m.get(n) for n in FIELDS})
    with MolochBatch([g for g, _ in doms], MLC_MIXED) as b:
        for i, (g, f) in enumerate(doms):
            b.member(i).load(f)
            b.member(i).init_moloch()
        # members start on different launch counts of the horizontal WAF
        b.member(0).set_variant(MLC_VAR_MIXED_MERGE, 4)
        b.member(1).set_variant(MLC_VAR_WAF_HGROUP, 2)
        b.autotune(1)
        for knob in (MLC_VAR_WAF_H, MLC_VAR_WAF_HGROUP, MLC_VAR_MIXED_MERGE):
            vals = {b.member(i).get_variant(knob) for i in range(3)}
            assert len(vals) == 1, (knob, vals)
        b.step(2)
        b.synchronize()
        for i in range(3):
            for n in FIELDS:
                assert np.array_equal(b.member(i).get(n), alone[i][n]), (i, n)


def test_batch_sees_a_parameter_changed_after_its_first_step():
    """A batch step's recorded launches copy each member's parameters; a
    member's mlc_set_param after the first batch step must reach the next
    steps (here the TKE floor of status_update)."""
    from regcm_amd.moloch import MLC_PARAM_TKEMIN
    print("[tune] batch of 2 with TKE: tkemin changed between steps", flush=True)
    doms = [synth.config("small", ibltyp=2, tkemin=0.01, **o)[:2]
            for o in (dict(), dict(jx=30, iy=26))]
    alone = []
    for g, f in doms:
        with Moloch(g) as m:
            m.load(f)
            m.set_param(MLC_PARAM_TKEMIN, 0.01)
            m.init_moloch()
            m.moloch(1)
            m.set_param(MLC_PARAM_TKEMIN, 0.5)
            m.moloch(1)
            alone.append({n: m.get(n) for n in FIELDS + ("tke",)})
    with MolochBatch([g for g, _ in doms]) as b:
        for i, (g, f) in enumerate(doms):
            b.member(i).load(f)
            b.member(i).set_param(MLC_PARAM_TKEMIN, 0.01)
            b.member(i).init_moloch()
        b.step(1)
        for i in range(2):
            b.member(i).set_param(MLC_PARAM_TKEMIN, 0.5)
        b.step(1)
        b.synchronize()
        for i in range(2):
            assert float(b.member(i).get("tke")[1:, 1:-2, 1:-2].min()) >= 0.5
            for n in FIELDS + ("tke",):
                assert np.array_equal(b.member(i).get(n), alone[i][n]), (i, n)
