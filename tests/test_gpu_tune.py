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
