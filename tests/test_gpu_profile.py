"""Per-kernel timing (mlc_profile / mlc_kernel_stats).  A profiled launch
goes through hipExtLaunchKernel with a start and a stop event, so the
profiled step must be the same step bit for bit; the launch counts follow
the step's structure (nadv x nsound acoustic sub-steps per step), every
time is positive, and the kernel times of one stream fit inside the wall
time around them."""
import time

import numpy as np
import pytest

from regcm_amd import synth
from regcm_amd.moloch import Moloch

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "pai", "tetav", "t", "qx")
STEPS = 2


def say(msg):
    print(msg, flush=True)


def run(g, f, mode):
    """mode: None (profiling off), "all", or a list of kernel names."""
    with Moloch(g) as m:
        m.load(f)
        m.init_moloch()
        if mode == "all":
            m.profile(True)
        elif mode is not None:
            m.profile(True, slots=mode)
        m.synchronize()
        t0 = time.perf_counter()
        m.moloch(STEPS)
        m.synchronize()
        wall_ms = (time.perf_counter() - t0) * 1e3
        stats = m.kernel_stats()
        m.profile(False)
        after = m.kernel_stats()
        out = {n: m.get(n) for n in FIELDS}
    return out, stats, after, wall_ms


@pytest.mark.parametrize("over", [dict(), dict(ibltyp=2, ntr=2, tkemin=0.01)],
                         ids=["plain", "tke-tracers"])
def test_profiled_step_is_the_same_step(over):
    say(f"[profile] {over or 'plain'}: profiled vs unprofiled, {STEPS} steps")
    g, f, _ = synth.config("small", **over)
    ref, none_stats, _, _ = run(g, f, None)
    got, stats, after, wall_ms = run(g, f, "all")
    for n in FIELDS:
        assert np.array_equal(ref[n], got[n]), n
    assert none_stats == {}
    assert after == {}                                   # profile(False) resets
    sub = STEPS * g.nadv * g.nsound
    for name in ("sound_div", "sound_column", "sound_momentum"):
        assert stats[name][0] == sub, (name, stats[name])
    assert stats["waf_horizontal"][0] >= STEPS * g.nadv
    assert stats["waf_vertical"][0] >= STEPS * g.nadv
    assert all(n > 0 and ms > 0.0 for n, ms in stats.values()), stats
    total = sum(ms for _, ms in stats.values())
    assert total <= wall_ms * 1.01 + 0.1, (total, wall_ms)


def test_slot_filter_times_only_the_named_kernels():
    say("[profile] slot filter")
    g, f, _ = synth.config("small")
    ref, _, _, _ = run(g, f, None)
    got, stats, _, _ = run(g, f, ["sound_column", "waf_horizontal"])
    for n in FIELDS:
        assert np.array_equal(ref[n], got[n]), n
    assert set(stats) == {"sound_column", "waf_horizontal"}
    assert stats["sound_column"][0] == STEPS * g.nadv * g.nsound
