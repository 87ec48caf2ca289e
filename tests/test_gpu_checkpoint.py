"""Checkpoint / restart of the device state (mlc_checkpoint, mlc_restart):
the SAV variables of Main/mod_savefile.F90:618-628 come first in the file
over their mo_atm bounds, and a fresh handle restored from the file steps
to exactly the bits of an uninterrupted run."""
import numpy as np
import pytest

from regcm_amd import synth
from regcm_amd.moloch import MLC_FP64, MLC_MIXED, Moloch

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "pai", "tetav", "t", "tvirt", "qx", "ux", "vx", "p", "rho", "qsat", "ps")


@pytest.mark.parametrize("precision", [MLC_FP64, MLC_MIXED], ids=["fp64", "mixed"])
def test_restart_continues_bit_identically(tmp_path, precision):
    print(f"[checkpoint] small precision={precision}", flush=True)
    g, f, _ = synth.config("small", nadv=2, nsound=3)
    path = tmp_path / "moloch.sav"
    with Moloch(g, precision) as m:
        m.load(f)
        m.init_moloch()
        m.moloch(2)
        m.checkpoint(path)
        u_ck = m.get("u")
        m.moloch(3)
        ref = {n: m.get(n) for n in FIELDS}
    # the SAV section: header, then u over (jde x ice x kz)
    raw = path.read_bytes()
    assert raw[:8] == b"MLCSAV2\0"
    hdr = np.frombuffer(raw[8:36], dtype=np.int32)
    assert list(hdr[:4]) == [g.jx, g.iy, g.kz, g.nqx]
    assert list(hdr[5:]) == [0, 0]                 # no TKE, no tracers
    nu = g.jx * (g.iy - 1) * g.kz
    u_file = np.frombuffer(raw[36:36 + 8 * nu], dtype=np.float64).reshape(g.kz, g.iy - 1, g.jx)
    assert np.array_equal(u_file, u_ck[:, : g.iy - 1, :])
    with Moloch(g, precision) as m:
        m.load(f)              # statics (and a state that the restart replaces)
        m.init_moloch()
        m.restart(path)
        m.moloch(3)
        got = {n: m.get(n) for n in FIELDS}
    for n in FIELDS:
        a, b = ref[n], got[n]
        if n == "u":
            a, b = a[..., : g.iy - 1, :], b[..., : g.iy - 1, :]
        elif n == "v":
            a, b = a[..., : g.jx - 1], b[..., : g.jx - 1]
        else:
            a, b = a[..., : g.iy - 1, : g.jx - 1], b[..., : g.iy - 1, : g.jx - 1]
        assert np.array_equal(a, b), n
