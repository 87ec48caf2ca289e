"""Pins of the oracle's doubly periodic mode (i_crm = 1).

On a periodic domain the reference's exchanges fill every halo with the
periodic partner's current values (Main/mpplib/mod_mppparam.F90:1298-1309),
no edge is a physical boundary, and every extent covers the whole domain.
What fixes the oracle there, without restating it:
  - translation invariance: the step commutes with a cyclic shift of every
    input by any (sj, si), bit for bit -- a wrong wrap, a missed extent or a
    boundary formula left on would break it at the seam;
  - WAF at Courant 1 is an exact cyclic shift in j and in i, both signs,
    across the seam;
  - with uniform wind, flat terrain and unit metrics the WAF is in flux
    form with nothing leaving the torus: the domain sum of the advected
    field is conserved to round-off;
  - a resting, discretely hydrostatic atmosphere over flat terrain stays at
    rest.
"""
import numpy as np
import pytest

import oracle_api as O
from regcm_amd import synth

STEP_FIELDS = ("u", "v", "w", "pai", "tetav", "t", "tvirt", "qx", "ux", "vx",
               "p", "rho", "qsat", "ps")


def rolled(f, g, sj, si):
    out = {}
    for k, v in f.items():
        if isinstance(v, np.ndarray) and v.ndim >= 2 and v.shape[-2:] == (g.iy, g.jx):
            out[k] = np.ascontiguousarray(np.roll(np.roll(v, sj, axis=-1), si, axis=-2))
        elif isinstance(v, np.ndarray):
            out[k] = v.copy()
        else:
            out[k] = v
    return out


@pytest.mark.parametrize("shift", [(1, 0), (0, 1), (7, 5), (23, 19)])
@pytest.mark.parametrize("over", [dict(), dict(ibltyp=2, ntr=2, tkemin=0.01),
                                  dict(ipptls=1, nqx=2, do_divdamp=0)],
                         ids=["plain", "tke-tracers", "ipptls1-nodamp"])
def test_step_commutes_with_cyclic_shifts(shift, over):
    g, f, _ = synth.config("crm-small", **over)
    sj, si = shift
    a = O.copy_fields(f)
    O.moloch_steps(g, a, 3)
    b = rolled(f, g, sj, si)
    O.moloch_steps(g, b, 3)
    names = STEP_FIELDS + (("tke",) if g.ibltyp == 2 else ()) + (("trac",) if g.ntr else ())
    for n in names:
        assert np.array_equal(rolled({n: a[n]}, g, sj, si)[n], b[n]), n
    # and the step did move the state
    assert np.max(np.abs(a["u"] - f["u"])) > 0.1


def periodic_flat(jx=24, iy=20, kz=6):
    g = synth.Grid(jx, iy, kz, 10000.0, 60.0, nqx=2, ipptls=0, nadv=1, nsound=1, crm=1)
    f = synth.empty_fields(g)
    for k in ("mx", "mu", "mv", "fmz", "fmzf", "rfmzu", "rfmzv"):
        f[k][:] = 1.0
    return g, f


@pytest.mark.parametrize("axis,sign", [(2, 1), (2, -1), (1, 1), (1, -1)])
def test_waf_courant_one_is_a_cyclic_shift(axis, sign):
    g, f = periodic_flat()
    dta = g.dtsec / g.nadv
    wind = "u" if axis == 2 else "v"
    f[wind][:] = sign * g.dx / dta
    s = np.zeros((g.kz + 1, g.iy, g.jx))
    pp = np.random.default_rng(11 + axis).normal(size=(g.kz, g.iy, g.jx))
    new = pp.copy()
    O.wafone(g, f, s, new, dta)
    assert np.allclose(new, np.roll(pp, sign, axis=axis), rtol=0, atol=1e-13)


@pytest.mark.parametrize("cu,cv", [(0.37, 0.0), (0.0, -0.61), (0.45, 0.3)])
def test_waf_conserves_the_domain_sum(cu, cv):
    g, f = periodic_flat(jx=30, iy=22, kz=4)
    dta = g.dtsec / g.nadv
    f["u"][:] = cu * g.dx / dta
    f["v"][:] = cv * g.dx / dta
    s = np.zeros((g.kz + 1, g.iy, g.jx))
    pp = np.random.default_rng(12).uniform(1, 2, (g.kz, g.iy, g.jx))
    new = pp.copy()
    O.wafone(g, f, s, new, dta)
    assert np.max(np.abs(new - pp)) > 1e-2
    for k in range(g.kz):
        assert abs(new[k].sum() - pp[k].sum()) <= 1e-12 * pp[k].sum()


def test_rotllr_is_refused():
    g, f, _ = synth.config("crm-small", lrotllr=1)
    with pytest.raises(RuntimeError):
        O.moloch_steps(g, f, 1)


def test_resting_atmosphere_stays_at_rest():
    g = synth.Grid(24, 20, 10, 10000.0, 30.0, nqx=2, ipptls=0, crm=1)
    f = synth.build_domain(g, [])
    synth.initial_state(g, f, resting=True, qv_floor=1e-8)
    a = O.copy_fields(f)
    O.moloch_steps(g, a, 3)
    assert np.max(np.abs(a["w"])) < 1e-12
    assert np.max(np.abs(a["u"])) < 1e-12 and np.max(np.abs(a["v"])) < 1e-12
    assert np.max(np.abs(a["pai"] - f["pai"]) / f["pai"]) < 1e-14


def test_massck_counts_every_point_and_no_boundary_flux():
    g, f = periodic_flat(jx=12, iy=10, kz=4)
    f["u"][:] = 3.0
    f["v"][:] = -2.0
    f["dz"][:] = 250.0
    f["rho"][:] = 1.1
    f["qx"][0] = 0.004
    f["qx"][1] = 0.001
    m = O.massck(g, f)
    n = g.jx * g.iy * g.kz
    assert m["drymass"] == pytest.approx(n * g.dx ** 2 * 250.0 * 1.1, rel=1e-14)
    assert m["qmass"] == pytest.approx(n * g.dx ** 2 * 250.0 * 1.1 * 0.005, rel=1e-13)
    assert m["dryadv"] == 0.0 and m["qadv"] == 0.0
