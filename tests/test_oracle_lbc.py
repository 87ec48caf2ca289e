"""Pins of the oracle's lateral-boundary stage (boundary(),
Main/mod_moloch.F90:448-529; bdyval MOLOCH branch Main/mod_bdycod.F90:
1638-1866; morelax :3962-4031; motopnudge :4049-4081) by what the
algorithm fixes: the time interpolation's end points, the inflow/outflow
switch, the Davies relaxation's convexity and its fixed points."""
import numpy as np
import pytest

import oracle_api as O
from regcm_amd import synth

QXZERO = [1e-8, 0.0, 0.0, 0.0, 0.0]


def setup(**over):
    g, f, _ = synth.config("small", **over)
    f["w"][1:-1] = synth.normal(91, f["w"][1:-1].shape, 0.05)
    b = synth.lbc_inputs(g, f, nspgx=5, nspgd=6)
    return g, f, b


def edges(g):
    """(field, [(k, i, j) index arrays of the bdyval-written points])"""
    jx, iy = g.jx, g.iy
    cross_we = (slice(None), slice(1, iy - 2), [0, jx - 2])      # i in ici, j = jce1/jce2
    cross_sn = (slice(None), [0, iy - 2], slice(0, jx - 1))      # i = ice1/ice2, j in jce
    return cross_we, cross_sn


@pytest.mark.parametrize("x1,which", [(0.0, "0"), (1.0, "1")])
def test_bdyval_end_points_are_the_boundary_states(x1, which):
    g, f, b = setup()
    O.boundary(g, f, b, x1)
    jx, iy = g.jx, g.iy
    we, sn = edges(g)
    for name, key in (("t", "t"), ("pai", "pai")):
        for sl in (we, sn):
            assert np.array_equal(f[name][sl], b[key + which][sl]), (name, sl)
    for sl in (we, sn):
        assert np.array_equal(f["qx"][0][sl], b["qv" + which][sl])
        assert np.array_equal(f["qx"][1][sl], b["qc" + which][sl])
        assert np.array_equal(f["qx"][2][sl], b["qi" + which][sl])
    # u: west/east columns j = 1, jx over ici; south/north rows over all jde
    assert np.array_equal(f["u"][:, 1:iy - 2, [0, jx - 1]], b["u" + which][:, 1:iy - 2, [0, jx - 1]])
    assert np.array_equal(f["u"][:, [0, iy - 2], :], b["u" + which][:, [0, iy - 2], :])
    # v: west/east columns over idi, south/north rows i = 1, iy over jce
    assert np.array_equal(f["v"][:, 1:iy - 1, [0, jx - 2]], b["v" + which][:, 1:iy - 1, [0, jx - 2]])
    assert np.array_equal(f["v"][:, [0, iy - 1], : jx - 1], b["v" + which][:, [0, iy - 1], : jx - 1])
    assert np.array_equal(f["ps"][1:iy - 2, [0, jx - 2]], b["ps" + which][1:iy - 2, [0, jx - 2]])
    assert np.array_equal(f["ps"][[0, iy - 2], : jx - 1], b["ps" + which][[0, iy - 2], : jx - 1])


def test_inflow_outflow_switch_for_species_without_boundary_data_and_w():
    g, f, b = setup()
    qr_in = f["qx"][3].copy()
    qr_in[:] = np.abs(synth.normal(5, qr_in.shape, 1e-4))
    f["qx"][3][:] = qr_in
    w_in = f["w"].copy()
    O.boundary(g, f, b, 0.3)
    jx, iy = g.jx, g.iy
    u, v = f["u"], f["v"]
    for k in range(g.kz):
        for i in range(1, iy - 2):                  # west/east, i in ici
            for jc, jd, jn, sign in ((0, 0, 1, 1.0), (jx - 2, jx - 1, jx - 3, -1.0)):
                inflow = sign * u[k, i, jd] > 0.0
                want_q = QXZERO[3] if inflow else qr_in[k, i, jn]
                want_w = 0.0 if inflow else w_in[k, i, jn]
                assert f["qx"][3][k, i, jc] == want_q
                assert f["w"][k, i, jc] == want_w
        for j in range(1, jx - 2):                  # south/north rows away from corners
            for ic, idd, inn, sign in ((0, 0, 1, 1.0), (iy - 2, iy - 1, iy - 3, -1.0)):
                inflow = sign * v[k, idd, j] > 0.0
                assert f["qx"][3][k, ic, j] == (QXZERO[3] if inflow else f["qx"][3][k, inn, j])


def test_davies_relaxation_is_convex_and_fixes_untouched_points():
    g, f, b = setup()
    t_old = f["t"].copy()
    pai_old = f["pai"].copy()
    w_old = f["w"].copy()
    x1 = 0.25
    O.boundary(g, f, b, x1)
    xf = b["xfc"]
    interior = np.zeros_like(xf, dtype=bool)
    interior[:, 1:g.iy - 2, 1:g.jx - 2] = True
    relaxed = interior & (xf > 0.0)
    fext = (1.0 - x1) * b["t0"] + x1 * b["t1"]
    lo, hi = np.minimum(t_old, fext), np.maximum(t_old, fext)
    tn = f["t"]
    assert np.all(tn[relaxed] >= lo[relaxed] - 1e-12 * np.abs(lo[relaxed]))
    assert np.all(tn[relaxed] <= hi[relaxed] + 1e-12 * np.abs(hi[relaxed]))
    # hefc(nspgx) = 0 rings and points outside every ring keep their values
    fixed = interior & ((xf == 0.0) | (xf < 0.0))
    assert np.array_equal(tn[fixed], t_old[fixed])
    assert np.array_equal(f["pai"][fixed], pai_old[fixed])
    # w relaxes towards zero (morelax_fraction with frac = 0)
    wk = w_old[: g.kz]
    assert np.array_equal(f["w"][: g.kz][relaxed], ((1.0 - xf) * wk)[relaxed])
    # theta_v follows from the relaxed T and pai
    assert np.allclose(f["tetav"] * f["pai"], f["tvirt"], rtol=4e-16, atol=0)


def test_top_nudging_moves_upper_levels_only():
    g, f, b = setup()
    b["top_nudge"] = True
    assert b["nztop"] >= 1
    f2 = O.copy_fields(f)
    b2 = dict(b, top_nudge=False)
    O.boundary(g, f, b, 0.5)
    O.boundary(g, f2, b2, 0.5)
    nz = b["nztop"]
    ii, jj = slice(6, g.iy - 7), slice(6, g.jx - 7)            # outside every ring
    assert not np.array_equal(f["t"][:nz, ii, jj], f2["t"][:nz, ii, jj])
    assert np.array_equal(f["t"][nz:, ii, jj], f2["t"][nz:, ii, jj])
    assert not np.array_equal(f["w"][1, ii, jj], f2["w"][1, ii, jj])   # w(:,:,2)
    assert np.array_equal(f["w"][2:, ii, jj], f2["w"][2:, ii, jj])


def test_lbc_steps_stay_finite_and_differ_from_fixed_boundaries():
    g, f, b = setup()
    f2 = O.copy_fields(f)
    O.moloch_steps_lbc(g, f, b, [0.1, 0.2, 0.3])
    O.moloch_steps(g, f2, 3)
    for n in ("u", "v", "w", "pai", "t"):
        assert np.all(np.isfinite(f[n]))
    assert not np.array_equal(f["t"], f2["t"])
