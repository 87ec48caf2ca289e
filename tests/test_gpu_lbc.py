"""Real-case lateral boundaries on the device (mlc_lbc_configure,
mlc_boundary, mlc_step_lbc) against the oracle's boundary()
(Main/mod_moloch.F90:448-529): bit-identical in the FP64 form."""
import numpy as np
import pytest

import oracle_api as O
from regcm_amd import synth
from regcm_amd.moloch import MLC_MIXED, MLC_VAR_STAGGER_CHUNK, Moloch, relative_l2

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "pai", "tetav", "t", "tvirt", "qx", "ux", "vx", "ps")
STEP_FIELDS = FIELDS[:-1]
# p, rho, qsat and the interior ps go through pow / exp (device vs host libm)
TRANSC = ("p", "rho", "qsat", "ps")


def owned(name, a):
    if name == "u":
        return a[..., : a.shape[-2] - 1, :]
    if name == "v":
        return a[..., :, : a.shape[-1] - 1]
    return a[..., : a.shape[-2] - 1, : a.shape[-1] - 1]


def case(over, **kw):
    g, f, _ = synth.config("small", **over)
    f["w"][1:-1] = synth.normal(91, f["w"][1:-1].shape, 0.05)
    b = synth.lbc_inputs(g, f, nspgx=5, nspgd=6, **kw)
    return g, f, b


def compare(ref, got, label, names=FIELDS):
    bad = [n for n in names if not np.array_equal(owned(n, ref[n]), owned(n, got[n]))]
    assert not bad, f"{label}: {bad} differ"


@pytest.mark.parametrize("over,kw", [
    (dict(), dict()),
    (dict(), dict(top_nudge=True)),
    (dict(), dict(present_qc=False, present_qi=False)),
    (dict(ipptls=1, nqx=2), dict(present_qi=False)),
    (dict(jx=70, iy=22, kz=12), dict(top_nudge=True)),
], ids=["default", "top-nudge", "no-qc-qi-data", "ipptls1", "odd-shape"])
def test_boundary_stage_bit_exact(over, kw):
    print(f"[lbc] boundary() {over} {kw}", flush=True)
    g, f, b = case(over, **kw)
    ref = O.copy_fields(f)
    O.boundary(g, ref, b, 0.3)
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        m.load_lbc(b)
        m.boundary(0.3)
        got = {n: m.get(n) for n in FIELDS}
    compare(ref, got, "boundary()")


@pytest.mark.parametrize("schunk", [1, 4], ids=["stagger-chunk1", "stagger-chunk4"])
@pytest.mark.parametrize("kw", [dict(), dict(top_nudge=True)], ids=["plain", "top-nudge"])
def test_steps_with_lateral_boundaries_bit_exact(kw, schunk):
    print(f"[lbc] 3 steps {kw} stagger chunk {schunk}", flush=True)
    g, f, b = case(dict(nadv=2, nsound=3), **kw)
    x1 = [0.1, 0.2, 0.3]
    ref = O.copy_fields(f)
    O.moloch_steps_lbc(g, ref, b, x1)
    with Moloch(g) as m:
        m.set_variant(MLC_VAR_STAGGER_CHUNK, schunk)
        m.load(f)
        m.init_moloch()
        m.load_lbc(b)
        m.moloch_lbc(x1)
        got = {n: m.get(n) for n in STEP_FIELDS + TRANSC}
    compare(ref, got, "3 lbc steps", STEP_FIELDS)
    for n in TRANSC:
        a, c = owned(n, ref[n]), owned(n, got[n])
        assert np.max(np.abs(a - c) / np.maximum(np.abs(a), 1e-300)) <= 8e-16, n


def test_mixed_form_with_lateral_boundaries_is_close():
    print("[lbc] mixed form 3 steps", flush=True)
    g, f, b = case(dict(nadv=2, nsound=3))
    x1 = [0.1, 0.2, 0.3]
    ref = O.copy_fields(f)
    O.moloch_steps_lbc(g, ref, b, x1)
    with Moloch(g, MLC_MIXED) as m:
        m.load(f)
        m.init_moloch()
        m.load_lbc(b)
        m.moloch_lbc(x1)
        got = {n: m.get(n) for n in ("u", "v", "pai", "t")}
    for n, tol in (("u", 1e-5), ("v", 1e-4), ("pai", 1e-7), ("t", 1e-6)):
        assert relative_l2(owned(n, ref[n]), owned(n, got[n])) < tol, n


@pytest.mark.parametrize("km,lm", [(1, 1), (2, 3)])
def test_spectral_nudging_bit_exact(km, lm):
    """mospectral_nudge inside the boundary stage, applied on every second
    step (dtrad = 2 dtsec), against the oracle."""
    print(f"[lbc] spectral nudging km={km} lm={lm}", flush=True)
    g, f, b = case(dict(nadv=2, nsound=3), top_nudge=True)
    b["spectral"] = synth.spectral_inputs(g, f, dtrad=2 * g.dtsec, km=km, lm=lm)
    x1 = [0.1, 0.2, 0.3, 0.4]
    ref = O.copy_fields(f)
    bo = dict(b, spectral=dict(b["spectral"]))
    O.moloch_steps_lbc(g, ref, bo, x1)
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        m.load_lbc(b)
        m.moloch_lbc(x1)
        got = {n: m.get(n) for n in STEP_FIELDS}
    compare(ref, got, "spectral nudging", STEP_FIELDS)
