"""Device output_history (moist_history) against the oracle's restatement,
and the run_sim driver with its history file and restart."""
import numpy as np
import pytest

import oracle as O
from moist2d_amd import core, ic
from moist2d_amd.history import read_restart, run_sim
from scipy.io import netcdf_file

pytestmark = pytest.mark.gpu


def _rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / n if n > 0 else np.linalg.norm(a - b)


@pytest.mark.parametrize("precision", [core.FP64, core.MIXED])
def test_history_matches_oracle(precision):
    p, s = ic.bubble(np.float64, nx=96, nz=40, d=250.0)
    eng = core.Engine(p, precision)
    eng.upload(s)
    eng.step(40)
    out = s.copy()
    eng.download(out)
    f, st = eng.history()
    eng.close()
    fo, sto = O.output_history(out.copy(), p)
    # the device reads the same state the oracle is given (the mixed form's
    # FP32 deviations plus FP64 base state, downloaded as doubles)
    tol = 1e-12 if precision == core.FP64 else 1e-9
    errs = {n: _rel(f[n], fo[n]) for n in fo}
    print(f"history prec={precision}", " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
    for n, e in errs.items():
        assert e <= tol, (n, e)
    np.testing.assert_allclose(st["Etot"], sto[0], rtol=1e-13)
    for j, k in enumerate(("wmin", "wmax", "qcmax", "qrmax", "RHmin", "RHmax"), start=1):
        assert abs(st[k] - sto[j]) <= tol * max(1.0, abs(sto[j])), (k, st[k], sto[j])


def test_history_of_each_ensemble_member():
    mem = ic.ensemble(np.float64, members=2, nx_choices=[64, 96], nz_choices=(24, 32), steps=5)
    eng = core.Engine([p for p, _ in mem], core.FP64)
    for m, (_, s) in enumerate(mem):
        eng.upload(s, member=m)
    eng.step(3)
    for m, (p, s) in enumerate(mem):
        out = s.copy()
        eng.download(out, member=m)
        f, st = eng.history(member=m)
        fo, sto = O.output_history(out.copy(), p)
        for n in fo:
            assert _rel(f[n], fo[n]) <= 1e-12, (m, n)
        np.testing.assert_allclose(st["Etot"], sto[0], rtol=1e-13)
    eng.close()


def test_run_sim_history_and_restart(tmp_path):
    from moist2d_amd.params import make_params
    p = make_params(np.float64, Nx=64, Nz=32, Lx=64 * 312.5, H=32 * 312.5, dz0=312.5, z_fact=1.0, ns=3,
                    tout_int=2.0, t_end=6.0, do_sf=False, do_rd=False)
    path = str(tmp_path / "history.nc")
    s = run_sim(p, path=path)
    with netcdf_file(path, "r", mmap=False) as nc:
        t = np.array(nc.variables["time"][:])
        assert len(t) == 1 + p.N_steps // p.Nint_steps
        np.testing.assert_allclose(t, np.arange(len(t)) * 2.0)
        th = np.array(nc.variables["theta"][-1])
    h = p.halo
    np.testing.assert_allclose(th, (s.rho_theta / s.rho)[:, h:h + p.Nx], rtol=1e-15)
    r = read_restart(path, p)
    for n in ("rho", "rho_u", "rho_w"):
        np.testing.assert_array_equal(getattr(r, n)[:, h:h + p.Nx], getattr(s, n)[:, h:h + p.Nx])
    # a restarted run continues from the last record
    s2 = run_sim(p.replace(t_end=2.0), path=str(tmp_path / "h2.nc"), restart=path)
    assert np.isfinite(s2.rho).all()
