"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

FP64 form: per-field relative L2 error <= 1e-11 after 10 large steps of the
128x64 bubble (snapshot scalar order), the bar of the drop-in.  Mixed form:
error against the FP64 oracle stated per field (<= 1e-4 after 10 steps).
"""
import numpy as np
import pytest

import oracle as O
from conftest import PROG, field_errors, interior, rel_l2
from moist2d_amd import core, ic
from moist2d_amd.state import allocate_state

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-11


def _print(tag, errs):
    print(tag, " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)


def test_fp64_bubble_10_steps():
    p, s = ic.bubble(np.float64)
    ref = s.copy()
    assert O.rk3_step(ref, p, 10) == 0
    eng = core.Engine(p, core.FP64)
    eng.upload(s)
    eng.step(10)
    out = s.copy()
    eng.download(out)
    errs = field_errors(p, out, ref)
    _print("fp64 bubble 10 steps", errs)
    for n, e in errs.items():
        assert e <= TOL_FP64, (n, e)
    np.testing.assert_allclose(out.prec, ref.prec, rtol=1e-11, atol=1e-300)
