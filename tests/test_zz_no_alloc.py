"""moist_diagnostics_noalloc (static device scratch, no allocation) against
moist_diagnostics (its partials in memory the library allocates) through the
C ABI: two calls in a row, the second on a smaller handle, so that partials
the first call left in the shared scratch are not read by the second; each
bit for bit equal to the allocating entry point."""
import ctypes

import numpy as np
import pytest

from moist2d_amd import core, ic

pytestmark = pytest.mark.gpu


def _both(eng, member=0):
    L = core.lib()
    a = (ctypes.c_double * 5)()
    b = (ctypes.c_double * 5)()
    assert L.moist_diagnostics_noalloc(eng._h, member, a) == 0
    assert L.moist_diagnostics(eng._h, member, b) == 0
    return list(a), list(b)


def test_noalloc_diagnostics_match_the_allocating_entry_point():
    print("no-alloc diagnostics: 8192x256 mixed, then 64x32 FP64", flush=True)
    p1, s1 = ic.coldpool(np.float64)
    e1 = core.Engine(p1, core.MIXED)
    e1.upload(s1)
    e1.step(2)
    a, b = _both(e1)
    e1.close()
    assert a == b and a[1] > 0 and a[4] == 0.0
    # second call: a smaller handle (fewer tiles than the first left in scratch)
    p2, s2 = ic.bubble(np.float64, nx=64, nz=32, d=312.5)
    e2 = core.Engine(p2, core.FP64)
    e2.upload(s2)
    e2.step(3)
    a2, b2 = _both(e2)
    assert a2 == b2 and a2[1] > 0
    # the first-call-only path of the new entry point: a fresh handle whose
    # first diagnostics call is the no-alloc one
    e2.close()
    e3 = core.Engine(p2, core.FP64)
    e3.upload(s2)
    e3.step(3)
    a3 = e3.diagnostics_noalloc()
    e3.close()
    assert [a3[k] for k in core.DIAG_NAMES] == a2


def test_noalloc_diagnostics_ensemble_members():
    print("no-alloc diagnostics: 3-member ensemble", flush=True)
    members = ic.ensemble(np.float64, members=3, nx_choices=[64, 128, 192], nz_choices=(16, 24), steps=5)
    eng = core.Engine([p for p, _ in members], core.MIXED)
    for m, (p, s) in enumerate(members):
        eng.upload(s, member=m)
    eng.step(2)
    for m in range(3):
        a, b = _both(eng, m)
        assert a == b, (m, a, b)
    eng.close()
