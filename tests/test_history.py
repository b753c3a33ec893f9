"""History file and restart (moist.jl:1718-1919) on the CPU, with the derived
fields from the oracle's restatement of output_history."""
import numpy as np

import oracle as O
from moist2d_amd import ic
from moist2d_amd.history import VARS_1D, VARS_2D, VARS_PHYS_2D, History, read_restart
from scipy.io import netcdf_file


def test_history_roundtrip_and_restart(tmp_path):
    p, s = ic.bubble(np.float64, nx=32, nz=16, d=312.5)
    O.rk3_step(s, p, 3)
    path = str(tmp_path / "history.nc")
    h = History(path, p)
    f0, st0 = O.output_history(s.copy(), p)
    stats0 = dict(zip(("Etot", "wmin", "wmax", "qcmax", "qrmax", "RHmin", "RHmax"), st0))
    h.write_fields(0.0, f0, stats0, s)
    s1 = s.copy()
    O.rk3_step(s1, p, 2)
    f1, st1 = O.output_history(s1.copy(), p)
    out = h.write_fields(2.0, f1, dict(stats0, Etot=st1[0]), s1)
    assert out["Etot"] == st1[0] - st0[0]
    h.close()
    with netcdf_file(path, "r", mmap=False) as nc:
        assert nc.dimensions["x"] == p.Nx and nc.dimensions["z"] == p.Nz
        assert nc.variables["time"].shape == (2,)
        for name, units in VARS_2D + VARS_PHYS_2D:
            v = nc.variables[name]
            assert v.dimensions == ("time", "z", "x") and v.units.decode() == units
        for name, units in VARS_1D:
            assert nc.variables[name].dimensions == ("time", "x")
        np.testing.assert_array_equal(nc.variables["theta"][1], f1["theta"])
        np.testing.assert_array_equal(nc.variables["pres"][0], f0["pres"])
    r = read_restart(path, p)
    h_ = p.halo
    inner = np.s_[:, h_:h_ + p.Nx]
    for n in ("rho", "rho_u", "rho_w"):
        np.testing.assert_array_equal(getattr(r, n)[inner], getattr(s1, n)[inner])
    for n in ("rho_theta", "rho_qv", "rho_qc", "rho_qr"):
        np.testing.assert_allclose(getattr(r, n)[inner], getattr(s1, n)[inner], rtol=1e-14, atol=1e-300)
    np.testing.assert_array_equal(r.prec[h_:h_ + p.Nx], s1.prec[h_:h_ + p.Nx])
    # halos are periodic copies
    np.testing.assert_array_equal(r.rho[:, :h_], r.rho[:, p.Nx:p.Nx + h_])


def test_output_history_oracle_closed_forms():
    """The oracle's derived fields against their definitions at a rest state:
    u = w = 0, pres = P0 (r rho theta / P0)^(cp/cv) / 100, potential energy
    sum rho g z dx-free, RH = 100 qv / qsat."""
    from moist2d_amd.params import Constants
    c = Constants(np.float64)
    p, s = ic.bubble(np.float64, nx=16, nz=12, d=200.0, amp=0.0, noise=0.0)
    f, st = O.output_history(s.copy(), p)
    assert np.all(f["u"] == 0) and np.all(f["w"] == 0)
    h = p.halo
    inner = np.s_[:, h:h + p.Nx]
    rho = s.rho[inner]
    qv = s.rho_qv[inner] / rho
    r = c.Rd * (1 - qv) + c.Rv * qv
    cp = c.Cpd * (1 - qv) + c.Cpv * qv
    pres = c.P0 * (r * s.rho_theta[inner] / c.P0) ** (cp / (cp - r))
    np.testing.assert_allclose(f["pres"], pres / 100.0, rtol=1e-13)
    temp = s.rho_theta[inner] / rho * (pres / c.P0) ** (r / cp)
    np.testing.assert_allclose(f["temp"], temp, rtol=1e-13)
    Tc = temp - c.T0
    es = c.ES0 * np.exp(c.A_SAT * Tc / (c.B_SAT + Tc))
    qs = c.EPSvap * es / (pres - (1 - c.EPSvap) * es)
    np.testing.assert_allclose(f["RH"], qv / qs * 100.0, rtol=1e-12)
    ep = np.sum(rho * c.GRAV * p.z[:, None])
    ei = np.sum(rho * ((cp - r) * temp + qv * c.Lv))
    np.testing.assert_allclose(st[0], ep + ei, rtol=1e-13)
    assert st[1] == 0 and st[2] == 0 and st[3] == 0 and st[4] == 0
