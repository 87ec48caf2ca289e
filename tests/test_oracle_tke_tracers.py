"""The oracle's TKE (ibltyp 2) and chemical-tracer (ntr > 0) paths of the
step (Main/mod_moloch.F90:781-834, :1066-1075, :1419-1430 and
Main/mod_bdycod.F90:1684-1868), pinned by special cases and by the oracle's
own stages, each of which other tests pin:

- a tracer started equal to qv (kept above qv's 1e-8 floor) ends equal to
  qv bit for bit: both are advected by wafone and neither feeds back; a
  tracer started at twice qv ends at exactly twice it;
- over one advection, tke becomes htozstag(wafone(zstagtoh(tke))) on levels
  2..kz and keeps levels 1 and kz+1, which a tracer started at
  zstagtoh(tke) checks with the oracle's stagger routines;
- status_update's floors, on exactly representable values;
- the TKE boundary rows: tkemin on level 1, level 2 untouched, and levels
  3..kz+1 at tkemin under inflow or copied from the interior neighbour."""
import numpy as np

import oracle_api as O
from regcm_amd import synth


def test_tracer_equal_to_qv_stays_equal():
    g, f, _ = synth.config("small", ntr=2, nadv=2, nsound=3)
    f["qx"][0] = np.maximum(f["qx"][0], 1e-6)
    q0 = f["qx"][0].copy()
    f["trac"][0] = f["qx"][0]
    f["trac"][1] = 2.0 * f["qx"][0]
    O.moloch_steps(g, f, 3)
    assert f["qx"][0].min() > 1e-8            # qv's floor never acted
    assert np.array_equal(f["trac"][0], f["qx"][0])
    assert not np.array_equal(f["trac"][0], q0)
    # wafone is homogeneous of degree one (limiter ratios are scale-free and
    # every flux is linear in the field), and a factor 2 rounds exactly
    assert np.array_equal(f["trac"][1], 2.0 * f["trac"][0])


def test_tke_is_advected_on_half_levels():
    g, f, _ = synth.config("small", ibltyp=2, ntr=1, nadv=1, nsound=2)
    iy, jx = g.iy, g.jx
    tke0 = f["tke"].copy()
    # trac(1) := zstagtoh(tke), through the oracle's own zstagtoh
    tmp = O.copy_fields(f)
    tmp["w"][:] = tke0
    O.stagger(g, tmp, 2)
    f["trac"][0] = tmp["tvirt"]
    O.moloch_steps(g, f, 1)
    assert f["trac"][0][:, : iy - 1, : jx - 1].min() > 0.0      # the zero floor never acted
    # expected tke: htozstag of the advected half-level field over tke0
    exp = O.copy_fields(f)
    exp["tvirt"][:] = f["trac"][0]
    exp["w"][:] = tke0
    O.stagger(g, exp, 3)
    cross = (slice(None), slice(0, iy - 1), slice(0, jx - 1))
    assert np.array_equal(f["tke"][cross], exp["w"][cross])
    assert np.array_equal(f["tke"][0], tke0[0])
    assert np.array_equal(f["tke"][g.kz], tke0[g.kz])
    assert not np.array_equal(f["tke"][cross], tke0[cross])


def test_status_update_floors():
    g, f, _ = synth.config("small", ibltyp=2, ntr=1, tkemin=0.0625)
    kz, iy, jx = g.kz, g.iy, g.jx
    f["tke"][:] = 0.125
    f["tketen"][:] = np.where(np.arange(kz + 1)[:, None, None] % 2 == 0, -0.125, 0.25)
    f["trac"][0][:] = 0.25
    f["chiten"][0][:] = np.where(np.arange(kz)[:, None, None] % 2 == 0, -0.25, 0.125)
    O.status_update(g, f, 2.0)
    inner = (slice(None), slice(1, iy - 2), slice(1, jx - 2))
    tke = f["tke"][inner]
    assert np.all(tke[0::2] == 0.0625)       # 0.125 - 0.25 < tkemin
    assert np.all(tke[1::2] == 0.625)
    tr = f["trac"][0][inner]
    assert np.all(tr[0::2] == 0.0)           # 0.25 - 0.5 < 0
    assert np.all(tr[1::2] == 0.5)
    # jci x ici only
    assert np.all(f["tke"][:, 0, :] == 0.125) and np.all(f["tke"][:, :, jx - 2] == 0.125)
    assert np.all(f["trac"][0][:, iy - 2, :] == 0.25)


def test_tke_boundary_rows():
    g, f, _ = synth.config("small", ibltyp=2, tkemin=0.03125)
    kz, iy, jx = g.kz, g.iy, g.jx
    b = synth.lbc_inputs(g, f)
    # u > 0 on both west and east columns: inflow west, outflow east;
    # v < 0 on both south and north rows: outflow south, inflow north
    for n in ("u0", "u1"):
        b[n][:, :, 0] = 3.0
        b[n][:, :, jx - 1] = 2.0
    for n in ("v0", "v1"):
        b[n][:, 0, :] = -3.0
        b[n][:, iy - 1, :] = -2.0
    tke0 = f["tke"].copy()
    O.boundary(g, f, b, 0.5)
    tk = f["tke"]
    ici = slice(1, iy - 2)
    jin = slice(1, jx - 2)                       # south/north columns off the corners
    assert np.all(tk[0, ici, 0] == g.tkemin) and np.all(tk[0, ici, jx - 2] == g.tkemin)
    assert np.array_equal(tk[1, ici, 0], tke0[1, ici, 0])          # level 2 is not set
    assert np.all(tk[2:, ici, 0] == g.tkemin)                      # west inflow
    assert np.array_equal(tk[2:, ici, jx - 2], tke0[2:, ici, jx - 3])   # east outflow
    assert np.all(tk[0, 0, :jx - 1] == g.tkemin) and np.all(tk[0, iy - 2, :jx - 1] == g.tkemin)
    assert np.array_equal(tk[2:, 0, jin], tke0[2:, 1, jin])        # south outflow
    assert np.all(tk[2:, iy - 2, jin] == g.tkemin)                 # north inflow
    # the interior is untouched
    assert np.array_equal(tk[:, 1:iy - 2, 1:jx - 2], tke0[:, 1:iy - 2, 1:jx - 2])


def test_lbc_steps_refuse_tracers():
    g, f, _ = synth.config("small", ntr=1)
    b = synth.lbc_inputs(g, f)
    try:
        O.moloch_steps_lbc(g, f, b, [0.5])
    except RuntimeError:
        return
    raise AssertionError("the oracle stepped tracers with lateral boundaries")
