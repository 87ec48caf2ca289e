"""Restart through RegCM's own save path on the device.

A save is mlc_get_state of the SAV variables that write_savefile takes for
MOLOCH (Main/mod_savefile.F90:618-628): u, v, w, t, pai, qx, ps.  A restart
puts them into a fresh handle and runs init's restart branch
(Main/mod_init.F90:941-955) with mlc_init_restart, which forms p, qsat, rho,
tvirt = t*(1 + ep1*qv) and tetav = tvirt/pai.  The continued run matches the
oracle continuing from the same SAV variables; where the state holds no
condensate that tvirt counts, the FP64 form's restart is the uninterrupted
run bit for bit."""
import numpy as np
import pytest

import oracle_api as O
from regcm_amd import synth
from regcm_amd.moloch import MLC_FP64, MLC_MIXED, Moloch

pytestmark = pytest.mark.gpu

PROGNOSTIC = ("u", "v", "w", "pai", "tetav", "t", "tvirt", "qx", "ux", "vx")
TRANSCENDENTAL = ("p", "rho", "qsat", "ps")


def say(msg):
    print(msg, flush=True)


def owned(name, a):
    if name == "u":
        return a[..., : a.shape[-2] - 1, :]
    if name == "v":
        return a[..., :, : a.shape[-1] - 1]
    return a[..., : a.shape[-2] - 1, : a.shape[-1] - 1]


def restarted_fields(f, sav):
    """A fresh restart's fields: the statics of f and the SAV variables; the
    derived state NaN on the cross points (init_restart must form it), ux and
    vx zero as getmem leaves them."""
    r = O.copy_fields(f)
    for n in ("tetav", "tvirt", "p", "rho", "qsat"):
        owned(n, r[n])[...] = np.nan
    r["ux"][...] = 0.0
    r["vx"][...] = 0.0
    for n, a in sav.items():
        r[n][...] = a
    return r


def run_from(g, precision, fields, nsteps, restart):
    with Moloch(g, precision) as m:
        m.load(fields)
        m.init_moloch()
        if restart:
            m.init_restart()
        m.moloch(nsteps)
        m.synchronize()
        return {n: m.get(n) for n in PROGNOSTIC + TRANSCENDENTAL}


def saved_after(g, precision, f, nsteps):
    with Moloch(g, precision) as m:
        m.load(f)
        m.init_moloch()
        m.moloch(nsteps)
        m.synchronize()
        return m.save_variables()


def test_restart_matches_oracle_restart_with_condensate():
    """ipptls 2 with cloud water and rain: restart's Tv ignores the
    condensate, so the continued run differs from the uninterrupted one;
    the device and the oracle agree on it."""
    say("[restart] small ipptls=2 with condensate, 2 steps + restart + 3 steps vs oracle")
    g, f, _ = synth.config("small", nadv=2, nsound=3)
    sav = saved_after(g, MLC_FP64, f, 2)
    assert np.any(owned("qx", sav["qx"])[1] > 0)
    r = restarted_fields(f, sav)
    ref = O.copy_fields(r)
    O.init_restart(g, ref)
    O.moloch_steps(g, ref, 3)
    got = run_from(g, MLC_FP64, r, 3, restart=True)
    bad = []
    for n in PROGNOSTIC:
        a, b = owned(n, ref[n]), owned(n, got[n])
        if not np.array_equal(a, b):
            bad.append(f"{n}: {int(np.count_nonzero(a != b))} differ")
    for n in TRANSCENDENTAL:
        a, b = owned(n, ref[n]), owned(n, got[n])
        rel = np.max(np.abs(a - b) / np.maximum(np.abs(a), np.finfo(float).tiny))
        if not rel <= 8e-16:
            bad.append(f"{n}: max relative difference {rel:.3e}")
    assert not bad, "; ".join(bad)


@pytest.mark.parametrize("precision", [MLC_FP64, MLC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("case", ["ipptls0", "ipptls2-dry"])
def test_restart_without_condensate_continues_bit_identically(precision, case):
    """FP64: restart == uninterrupted, bit for bit.  The mixed form keeps
    theta_v in double through its step but saves T in float, so its restart
    starts from theta_v re-derived from the float T: close, not identical."""
    say(f"[restart] {case} precision={precision}: restart == uninterrupted")
    if case == "ipptls0":
        g, f, _ = synth.config("small", nadv=2, nsound=3, ipptls=0, nqx=2)
    else:
        g, f, _ = synth.config("small", nadv=2, nsound=3)
        f["qx"][1:] = 0.0
        f["tvirt"][:] = f["t"] * synth._tv_factor(g, f["qx"])
        f["tetav"][:] = f["tvirt"] / f["pai"]
    with Moloch(g, precision) as m:
        m.load(f)
        m.init_moloch()
        m.moloch(2)
        m.synchronize()
        sav = m.save_variables()
        m.moloch(3)
        m.synchronize()
        ref = {n: m.get(n) for n in PROGNOSTIC + TRANSCENDENTAL}
    got = run_from(g, precision, restarted_fields(f, sav), 3, restart=True)
    if precision == MLC_FP64:
        for n in PROGNOSTIC + TRANSCENDENTAL:
            assert np.array_equal(owned(n, ref[n]), owned(n, got[n])), n
        return
    errs = {n: float(np.linalg.norm(owned(n, ref[n]) - owned(n, got[n])) /
                     np.linalg.norm(owned(n, ref[n])))
            for n in PROGNOSTIC + TRANSCENDENTAL}
    say("[restart] mixed restart vs uninterrupted: " +
        ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    for n, e in errs.items():
        assert e <= (1e-4 if n == "w" else 3e-5), (n, e)
