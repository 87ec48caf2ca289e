"""Parity of the HIP FP64 path (through the C-ABI) with the CPU oracle.

Bar: bit-identical prognostic fields (u, v, w, pai, tetav, t, tvirt, qx, ux,
vx) after every tested number of steps -- the step uses only IEEE + - * /,
min, max and abs, compiled without FMA contraction on both sides.  p, rho,
qsat and ps go through pow/exp (device libm vs host libm) and are held to
a relative 8e-16 (a few ulp) instead.
"""
import sys

import numpy as np
import pytest

import oracle_api as O
from regcm_amd import synth
from regcm_amd.moloch import (MLC_VAR_COLUMN, MLC_VAR_LEVEL_CHUNK, MLC_VAR_WAF_GROUP,
                               MLC_VAR_WAF_H, MLC_VAR_WAF_ROWS, MLC_VAR_WAF_TILE,
                               MLC_VAR_WAF_HGROUP, MLC_VAR_WAFV_GROUP, MLC_VAR_XCD_MAP,
                               MLC_VAR_COLUMN_J, MLC_VAR_MOM_CHUNK, MLC_VAR_STAGGER_CHUNK,
                               Moloch)

pytestmark = pytest.mark.gpu

PROGNOSTIC = ("u", "v", "w", "pai", "tetav", "t", "tvirt", "qx", "ux", "vx")
TRANSCENDENTAL = ("p", "rho", "qsat", "ps")
ULP_REL = 8e-16


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_gpu(g, f, nsteps, mode="step", variants=None):
    with Moloch(g) as m:
        for stage, value in (variants or {}).items():
            m.set_variant(stage, value)
        m.load(f)
        m.init_moloch()
        if mode == "step":
            m.moloch(nsteps)
        else:   # the split path a caller with host physics would use
            for _ in range(nsteps):
                m.reset_tendencies()
                m.dynamical_core()
                m.diagnostics()
                m.status_update(g.dtsec)
        m.synchronize()
        return {n: m.get(n) for n in PROGNOSTIC + TRANSCENDENTAL}


def owned(name, a):
    """The part of a full-extent array that the reference's array holds
    (Main/mod_atm_interface.F90:579-624 bounds on a single domain)."""
    if name == "u":
        return a[..., : a.shape[-2] - 1, :]          # (jde, ice)
    if name == "v":
        return a[..., :, : a.shape[-1] - 1]          # (jce, ide)
    return a[..., : a.shape[-2] - 1, : a.shape[-1] - 1]   # (jce, ice)


def compare(ref, got, label):
    bad = []
    for n in PROGNOSTIC:
        a, b = owned(n, ref[n]), owned(n, got[n])
        if not np.array_equal(a, b):
            nd = int(np.count_nonzero(a != b))
            bad.append(f"{n}: {nd} values differ, max |d| = {np.max(np.abs(a - b)):.3e}")
    for n in TRANSCENDENTAL:
        a, b = owned(n, ref[n]), owned(n, got[n])
        scale = np.maximum(np.abs(a), np.finfo(float).tiny)
        rel = np.max(np.abs(a - b) / scale)
        if rel > ULP_REL:
            bad.append(f"{n}: max relative difference {rel:.3e}")
    assert not bad, f"{label}: " + "; ".join(bad)


CASES = [
    # name, config overrides, steps
    ("small-nadv1-nsound2", dict(), 3),
    ("small-default-split", dict(nadv=2, nsound=5), 2),
    ("small-ipptls0", dict(ipptls=0, nqx=2), 2),
    ("small-ipptls1", dict(ipptls=1, nqx=2), 2),
    ("small-nodamp-nofilter", dict(do_divdamp=0, do_divfilter=0), 2),
    ("small-rotllr", dict(lrotllr=1), 2),
    ("small-odd-shape", dict(jx=70, iy=9, kz=7), 2),
    # column heights at the edges of the column-solve variants: the fewest
    # levels accepted, past the 48 levels of the 6-level chunks, past the 64
    # levels of the 32-column chunked solve, and past the 128 levels of the
    # 16-column one (streamed solve)
    ("small-kz4", dict(kz=4), 2),
    ("small-kz56", dict(kz=56), 2),         # 8-level chunks of the 32-column solve
    ("small-kz70", dict(kz=70), 2),
    ("small-kz140", dict(kz=140), 2),
]


@pytest.mark.parametrize("name,over,nsteps", CASES, ids=[c[0] for c in CASES])
def test_small_cases_bit_exact(name, over, nsteps):
    say(f"[parity] case={name} steps={nsteps}")
    g, f, _ = synth.config("small", **over)
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, nsteps)
    mode = "split" if "split" in name else "step"
    got = run_gpu(g, O.copy_fields(f), nsteps, mode)
    compare(ref, got, name)


def test_cfg1_ten_steps_bit_exact():
    say("[parity] case=cfg1 64x64x23 steps=10")
    g, f, nsteps = synth.config("cfg1")
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, nsteps)
    got = run_gpu(g, O.copy_fields(f), nsteps)
    compare(ref, got, "cfg1")


def test_cfg2_first_step_bit_exact():
    say("[parity] case=cfg2 256x256x41 steps=1")
    g, f, _ = synth.config("cfg2")
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, 1)
    got = run_gpu(g, O.copy_fields(f), 1)
    compare(ref, got, "cfg2 step 1")


def test_cfg2_hundred_steps_bit_exact():
    """The 100-step workload: the prognostic fields stay bit-identical to
    the oracle (the pow / exp diagnostics do not feed back into them), and
    p, rho, qsat, ps agree to a few ulp."""
    say("[parity] case=cfg2 256x256x41 steps=100")
    g, f, nsteps = synth.config("cfg2")
    assert nsteps == 100
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, nsteps)
    got = run_gpu(g, O.copy_fields(f), nsteps)
    compare(ref, got, "cfg2 step 100")


def _mass_terms(g, f):
    """The massck terms of Main/mod_massck.F90:77-160, one array per sum."""
    jx, iy, kz = g.jx, g.iy, g.kz
    dx, dt = g.dx, g.dtsec
    dxsq = dx * dx
    dz, rho = f["dz"], f["rho"]
    ci = (slice(None), slice(1, iy - 2), slice(1, jx - 2))
    dry = (dxsq * dz[ci]) * rho[ci]
    wat = [((q[ci] * dxsq) * dz[ci]) * rho[ci] for q in f["qx"]]
    u, v = f["u"], f["v"]
    L = (slice(None), slice(0, iy - 1))
    adv = [(((u[L + (0,)] * dt) * dx) * dz[L + (0,)]) * rho[L + (0,)],
           -((((u[L + (jx - 1,)] * dt) * dx) * dz[L + (jx - 2,)]) * rho[L + (jx - 2,)]),
           (((v[:, 0, 1:jx - 2] * dt) * dx) * dz[:, 0, 1:jx - 2]) * rho[:, 0, 1:jx - 2],
           -((((v[:, iy - 1, 1:jx - 2] * dt) * dx) * dz[:, iy - 2, 1:jx - 2]) * rho[:, iy - 2, 1:jx - 2])]
    qadv = []
    for q in f["qx"]:
        qadv += [((((q[L + (0,)] * u[L + (0,)]) * dt) * dx) * dz[L + (0,)]) * rho[L + (0,)],
                 -(((((q[L + (jx - 2,)] * u[L + (jx - 1,)]) * dt) * dx) * dz[L + (jx - 2,)]) * rho[L + (jx - 2,)]),
                 ((((q[:, 0, 1:jx - 2] * v[:, 0, 1:jx - 2]) * dt) * dx) * dz[:, 0, 1:jx - 2]) * rho[:, 0, 1:jx - 2],
                 -(((((q[:, iy - 2, 1:jx - 2] * v[:, iy - 1, 1:jx - 2]) * dt) * dx) * dz[:, iy - 2, 1:jx - 2]) * rho[:, iy - 2, 1:jx - 2])]
    return {"drymass": [dry], "qmass": wat, "dryadv": adv, "qadv": qadv}


def test_massck_compensated_sums_match_exact_sums():
    import math
    say("[parity] case=massck cfg1 after 1 step")
    g, f, _ = synth.config("cfg1")
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        m.moloch(1)
        m.synchronize()
        got = m.massck()
        state = {n: m.get(n) for n in ("u", "v", "rho", "qx")}
    ff = dict(f, **state)
    terms = _mass_terms(g, ff)
    ref = O.massck(g, ff)             # the reference's serial FP64 order
    for key, parts in terms.items():
        flat = np.concatenate([p.ravel() for p in parts])
        exact = math.fsum(flat.tolist())
        absum = float(np.abs(flat).sum())
        # device double-double sums: within a few ulp of the exact sum
        assert abs(got[key] - exact) <= 4 * np.finfo(float).eps * max(abs(exact), absum * 1e-3), key
        # the serial FP64 sum agrees to round-off: |err| <= gamma_N * sum|t|
        n = flat.size
        gamma = n * np.finfo(float).eps / (1 - n * np.finfo(float).eps)
        assert abs(ref[key] - exact) <= gamma * absum, key


def test_check_finite_reports_nan_as_cfl_violation():
    from regcm_amd.moloch import MolochError
    say("[parity] case=check_finite small")
    g, f, _ = synth.config("small")
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        m.moloch(1)
        pmax, pmin = m.check_finite()
        assert 5e4 < pmin <= pmax < 1.2e5
        ps = m.get("ps")
        ps[5, 7] = np.nan
        m.put("ps", ps)
        with pytest.raises(MolochError, match="CFL VIOLATION"):
            m.check_finite()


@pytest.mark.parametrize("name,over", [("small-default", dict(nadv=2, nsound=5)),
                                       ("small-rotllr", dict(lrotllr=1)),
                                       ("small-odd-shape", dict(jx=70, iy=9, kz=7))])
def test_row_streaming_waf_variant_bit_exact(name, over):
    """The row-streaming form of the horizontal WAF pass gives the same bits
    as the default LDS-tile form (and the oracle)."""
    say(f"[parity] {name} with MLC_VAR_WAF_H=0")
    g, f, _ = synth.config("small", **over)
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, 2)
    got = run_gpu(g, f, 2, variants={MLC_VAR_WAF_H: 0})
    compare(ref, got, name + " row-streaming")


@pytest.mark.parametrize("variants", [
    {MLC_VAR_WAF_H: 0, MLC_VAR_WAF_ROWS: 4, MLC_VAR_WAF_GROUP: 2, MLC_VAR_LEVEL_CHUNK: 3},
    {MLC_VAR_WAF_H: 0, MLC_VAR_WAF_ROWS: 5, MLC_VAR_WAF_GROUP: 4, MLC_VAR_LEVEL_CHUNK: 1},
    {MLC_VAR_WAF_H: 0, MLC_VAR_WAF_ROWS: 64, MLC_VAR_LEVEL_CHUNK: 7},
    {MLC_VAR_COLUMN: 1},
    {MLC_VAR_COLUMN: 1, MLC_VAR_WAF_H: 1, MLC_VAR_LEVEL_CHUNK: 2},
    {MLC_VAR_COLUMN: 0},
    {MLC_VAR_WAFV_GROUP: 1},
    {MLC_VAR_WAFV_GROUP: 2},
    {MLC_VAR_WAF_TILE: 8},
    {MLC_VAR_WAF_TILE: 8, MLC_VAR_XCD_MAP: 0, MLC_VAR_COLUMN: 0},
    {MLC_VAR_WAF_HGROUP: 1},
    {MLC_VAR_WAF_HGROUP: 4, MLC_VAR_WAF_TILE: 8},
    {MLC_VAR_WAF_HGROUP: 3, MLC_VAR_XCD_MAP: 0},
    {MLC_VAR_COLUMN_J: 16},
    {MLC_VAR_COLUMN_J: 64},
    {MLC_VAR_MOM_CHUNK: 7},
    {MLC_VAR_LEVEL_CHUNK: 2, MLC_VAR_MOM_CHUNK: 5},
    {MLC_VAR_STAGGER_CHUNK: 3},
    {MLC_VAR_STAGGER_CHUNK: 7, MLC_VAR_WAF_TILE: 8},
], ids=["rows4-g2-chunk3", "rows5-g4-chunk1", "rows64-chunk7", "column-stream",
        "column-stream-tile-chunk2", "column-three-phase", "wafv-g1", "wafv-g2", "tile8",
        "tile8-plain-order", "hgroup1", "hgroup4-tile8", "hgroup3-plain-order", "column16",
        "column64", "mom-chunk7", "chunk2-mom-chunk5", "stagger-chunk3", "stagger-chunk7-tile8"])
@pytest.mark.parametrize("name,over", [("small-default", dict(nadv=2, nsound=5)),
                                       ("small-odd-shape", dict(jx=70, iy=9, kz=7))])
def test_launch_shape_variants_bit_exact(name, over, variants):
    say(f"[parity] {name} variants {variants}")
    g, f, _ = synth.config("small", **over)
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, 2)
    got = run_gpu(g, f, 2, variants=variants)
    compare(ref, got, name + f" {variants}")


@pytest.mark.parametrize("variants", [{MLC_VAR_COLUMN: 0}, {MLC_VAR_COLUMN: 1},
                                      {MLC_VAR_WAF_TILE: 8, MLC_VAR_XCD_MAP: 0},
                                      {MLC_VAR_COLUMN_J: 64}],
                         ids=["three-phase", "streamed", "tile8-plain-order", "column64"])
def test_variants_cfg1_ten_steps_bit_exact(variants):
    say(f"[parity] cfg1 10 steps, variants {variants}")
    g, f, n = synth.config("cfg1")
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, n)
    got = run_gpu(g, f, n, variants=variants)
    compare(ref, got, f"cfg1 {variants}")


@pytest.mark.parametrize("variants", [{MLC_VAR_COLUMN: 0}, {MLC_VAR_COLUMN: 1},
                                      {MLC_VAR_COLUMN: 2, MLC_VAR_COLUMN_J: 32},
                                      {MLC_VAR_COLUMN: 2, MLC_VAR_COLUMN_J: 64}],
                         ids=["three-phase", "streamed", "chunked32", "chunked64"])
def test_column_variants_on_a_tall_column(variants):
    """At kz = 70 the three-phase solve would need 71 KB of LDS and the
    32-column chunked solve holds 64 levels: both requests run the 16-column
    chunked solve instead, with the oracle's bits."""
    say(f"[parity] small kz=70, variants {variants}")
    g, f, _ = synth.config("small", kz=70)
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, 2)
    got = run_gpu(g, f, 2, variants=variants)
    compare(ref, got, f"kz=70 {variants}")


@pytest.mark.parametrize("kz", [41, 42, 43])
def test_column64_at_the_edge_of_its_levels(kz):
    """The 64-column chunked solve in 11-level chunks holds 42 levels in 64 KB
    of LDS; at 43 the request runs the 32-column solve.  Both with the
    oracle's bits."""
    say(f"[parity] small kz={kz} with 64-column blocks")
    g, f, _ = synth.config("small", kz=kz, jx=150, iy=12)
    ref = O.copy_fields(f)
    O.moloch_steps(g, ref, 2)
    got = run_gpu(g, f, 2, variants={MLC_VAR_COLUMN_J: 64})
    compare(ref, got, f"kz={kz} column64")
