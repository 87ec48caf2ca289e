"""mlc_massck_noalloc, the allocation-free budget entry, through the C-ABI:
the same four totals as mlc_massck, bit for bit, on two calls in a row
whose second domain is smaller (the shared static scratch keeps the first
call's partials, which the second must not read); a domain past the scratch
limit is refused before any launch."""
import ctypes

import numpy as np
import pytest

from regcm_amd import synth
from regcm_amd.moloch import MLC_FP64, MLC_MIXED, GridDesc, Moloch, load_library

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def budgets(m):
    a = (ctypes.c_double * 4)()
    b = (ctypes.c_double * 4)()
    assert m.lib.mlc_massck(m.h, a) == 0
    assert m.lib.mlc_massck_noalloc(m.h, b) == 0
    return np.array(a[:]), np.array(b[:])


@pytest.mark.parametrize("precision", [MLC_FP64, MLC_MIXED], ids=["fp64", "mixed"])
def test_noalloc_budget_equals_massck_on_two_calls(precision):
    for name in ("cfg1", "small"):        # 64x64x23, then 24x20x10
        say(f"[no-alloc] {name} precision={precision}")
        g, f, _ = synth.config(name)
        with Moloch(g, precision) as m:
            m.load(f)
            m.init_moloch()
            m.moloch(2)
            m.synchronize()
            ref, got = budgets(m)
            assert np.all(np.isfinite(ref)) and ref[0] > 0
            assert np.array_equal(ref, got), (name, ref, got)
            # a second call on the same state returns the same bits again
            _, again = budgets(m)
            assert np.array_equal(got, again)


def test_noalloc_refuses_a_domain_past_the_scratch():
    say("[no-alloc] 1040x520x64 is refused")
    lib = load_library()
    g = GridDesc(jx=1040, iy=520, kz=64, nqx=1, ipptls=0, nadv=1, nsound=1, do_divdamp=1,
                 do_divfilter=1, lrotllr=0, precision=MLC_MIXED, ibltyp=0, ntr=0,
                 dx=3000.0, dtsec=10.0, mo_ztop=30000.0, mo_a0=0.0)
    h = ctypes.c_void_p()
    assert lib.mlc_create(ctypes.byref(g), ctypes.byref(h)) == 0
    try:
        out = (ctypes.c_double * 4)()
        rc = lib.mlc_massck_noalloc(h, out)
        assert rc != 0
        assert b"too large" in lib.mlc_last_error()
    finally:
        lib.mlc_destroy(h)
