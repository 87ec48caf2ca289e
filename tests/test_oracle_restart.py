"""Pins of the oracle's restart branch (orc_init_restart, after
Main/mod_init.F90:941-955).

A restart reads the SAV variables u, v, w, t, pai, qx, ps
(Main/mod_savefile.F90:618-628) and forms p, qsat, rho, tvirt and tetav from
them.  Two facts fix it without restating it:
  - with no condensate in tvirt (ipptls 0, or every condensate zero), the
    restart state is the uninterrupted state -- whose tvirt, tetav, p, rho
    and qsat come from status_update, pinned on its own -- bit for bit, and
    so is every later step;
  - with condensate, restart's tvirt leaves it out: the ratio to the
    uninterrupted tvirt is (1 + ep1*qv) / (1 + ep1*qv - qc - qi - qr - qs),
    and the result does not depend on ipptls.
"""
import copy

import numpy as np

import oracle_api as O
from regcm_amd import synth

DERIVED = ("tetav", "tvirt", "p", "rho", "qsat")


def owned(name, a):
    if name == "u":
        return a[..., : a.shape[-2] - 1, :]
    if name == "v":
        return a[..., :, : a.shape[-1] - 1]
    return a[..., : a.shape[-2] - 1, : a.shape[-1] - 1]


def restarted(f):
    r = O.copy_fields(f)
    for n in DERIVED:
        owned(n, r[n])[...] = np.nan
    r["ux"][...] = 0.0
    r["vx"][...] = 0.0
    return r


def test_restart_without_condensate_is_the_uninterrupted_state():
    for over in (dict(ipptls=0, nqx=2), dict()):
        g, f, _ = synth.config("small", **over)
        if g.ipptls > 0:
            f["qx"][1:] = 0.0
            f["tvirt"][:] = f["t"] * synth._tv_factor(g, f["qx"])
            f["tetav"][:] = f["tvirt"] / f["pai"]
        O.moloch_steps(g, f, 2)
        r = restarted(f)
        O.init_restart(g, r)
        for n in DERIVED:
            assert np.array_equal(owned(n, r[n]), owned(n, f[n])), (over, n)
        O.moloch_steps(g, f, 2)
        O.moloch_steps(g, r, 2)
        for n in ("u", "v", "w", "pai", "tetav", "t", "tvirt", "qx", "p", "rho", "qsat", "ps"):
            assert np.array_equal(owned(n, r[n]), owned(n, f[n])), (over, n)


def test_restart_tvirt_leaves_condensate_out():
    g, f, _ = synth.config("small")          # ipptls 2, qc and qr non-zero
    O.moloch_steps(g, f, 2)
    r = restarted(f)
    O.init_restart(g, r)
    qx = owned("qx", f["qx"])
    ep1 = O.constant("ep1")
    want = (1.0 + ep1 * qx[0]) / (1.0 + ep1 * qx[0] - qx[1] - qx[2] - qx[3] - qx[4])
    got = owned("tvirt", r["tvirt"]) / owned("tvirt", f["tvirt"])
    assert np.max(np.abs(got / want - 1.0)) < 4e-16
    assert np.max(np.abs(want - 1.0)) > 1e-5          # the condensate is there
    # tetav follows the restart's tvirt
    assert np.array_equal(owned("tetav", r["tetav"]),
                          owned("tvirt", r["tvirt"]) / owned("pai", f["pai"]))
    # the diagnostics are the uninterrupted ones: they do not read tvirt
    for n in ("p", "rho", "qsat"):
        assert np.array_equal(owned(n, r[n]), owned(n, f[n])), n
    # whatever ipptls says
    g0 = copy.copy(g)
    g0.ipptls = 0
    r0 = restarted(f)
    O.init_restart(g0, r0)
    for n in DERIVED:
        assert np.array_equal(owned(n, r0[n]), owned(n, r[n])), n
