"""Host side without a GPU: the Params/State mirror, the portable ICs, and
that libmoist2d_hip.so loads and exports every entry point declared in
include/moist2d.h (no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from moist2d_amd import core, ic
from moist2d_amd.params import make_params
from moist2d_amd.state import STATE_ORDER, allocate_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported():
    hdr = open(os.path.join(ROOT, "include", "moist2d.h")).read()
    names = set(re.findall(r"^(?:int|void|const char \*)\s*\*?\s*(moist_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 14
    lib = ctypes.CDLL(core.LIB_PATH)
    for n in sorted(names):
        assert hasattr(lib, n), n
    core.lib()  # the binding's signatures resolve


def test_params_defaults_and_dt():
    p = make_params(np.float32)
    assert (p.Nx, p.Nz, p.halo, p.ia, p.ka) == (640, 250, 2, 644, 250)
    assert p.dx == np.float32(250.0)
    # dt = tout_int / ceil(tout_int / (0.6 dx ns / 340)) (moist.jl:192-194)
    assert p.dt == np.float32(300.0 / np.ceil(300.0 / (0.6 * 250.0 * 3 / 340.0)))
    assert p.zf[-1] >= p.H and p.dz[0] == np.float32(50.0)
    np.testing.assert_allclose(p.z, 0.5 * (np.concatenate([[0.0], p.zf[:-1]]) + p.zf), rtol=1e-6)


def test_params_grid_error():
    with pytest.raises(ValueError, match="z_fact is too small"):
        make_params(np.float64, Nz=10, H=25e3, dz0=50.0, z_fact=1.02)


@pytest.mark.parametrize("maker,dt,ns", [(ic.bubble, 1.0, 6), (ic.coldpool, 0.5, 8), (ic.adjoint_bubble, 0.5, 6)])
def test_workload_params(maker, dt, ns):
    kw = dict(nx=64, nz=32) if maker is not ic.bubble else {}
    p, s = maker(np.float64, **kw)
    assert p.dt == dt and p.ns == ns and p.zf[-1] == p.H
    assert np.all(np.isfinite(s.rho)) and s.rho.flags["F_CONTIGUOUS"]


def test_ic_balanced_and_deterministic():
    p, s = ic.coldpool(np.float64, nx=256, nz=64)
    p2, s2 = ic.coldpool(np.float64, nx=256, nz=64)
    for n in ("rho", "rho_theta", "rho_qv"):
        assert np.array_equal(getattr(s, n), getattr(s2, n))
    # the environment column is in the model's discrete hydrostatic balance
    from moist2d_amd.params import Constants
    C = Constants(np.float64)
    rho, rt, qv = s.rho[:, 2], s.rho_theta[:, 2], s.rho_qv[:, 2] / s.rho[:, 2]
    qd = 1 - qv
    r = C.Rd * qd + C.Rv * qv
    cp = C.Cpd * qd + C.Cpv * qv
    pres = C.P0 * (r * rt / C.P0) ** (cp / (cp - r))
    res = (pres[1:] - pres[:-1]) / np.diff(p.z) + C.GRAV * 0.5 * (rho[1:] + rho[:-1])
    assert np.abs(res[p.z[:-1] > 1000]).max() < 1e-9


def test_rng_is_a_pure_function():
    a = ic.rng_uniform(3, 1, 0, np.arange(5, dtype=np.uint64), 7)
    b = ic.rng_uniform(3, 1, 0, np.arange(5, dtype=np.uint64), 7)
    assert np.array_equal(a, b) and np.all((a >= 0) & (a < 1))
    g = ic.rng_normal(0, 0, 0, np.arange(200000, dtype=np.uint64), 0)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01


def test_ensemble_sizes():
    members = ic.ensemble(np.float64, members=256)
    nxs = {p.Nx for p, _ in members}
    nzs = {p.Nz for p, _ in members}
    assert nxs <= set(range(256, 2049, 64)) and nzs <= {64, 128}
    cells = sum(p.Nx * p.Nz for p, _ in members)
    assert 2.0e7 < cells < 3.6e7


def test_state_layout():
    p = make_params(np.float64, Nx=8, Nz=6, H=600.0, dz0=100.0, z_fact=1.0)
    s = allocate_state(p)
    assert [f for f in STATE_ORDER][:7] == ["rho", "rho_u", "rho_w", "rho_theta", "rho_qv", "rho_qc", "rho_qr"]
    assert s.rho.shape == (p.ka, p.ia) and s.rho.flags["F_CONTIGUOUS"]
    assert s.prec.shape == (p.ia,) and s.theta_ref.shape == (p.ka,)


def test_bf2002_initial_state():
    """init_state_BF2002! (moist.jl:453-551): saturated, qt = 0.02, saturated
    theta_e = 320 K where unperturbed, discrete hydrostatic balance, and the
    model's equation of state returns the column pressure."""
    from moist2d_amd import ic
    from moist2d_amd.params import Constants, make_params
    from moist2d_amd.state import allocate_state
    c = Constants(np.float64)
    p = make_params(np.float64, Nx=64, Nz=50, Lx=20e3, H=10e3, dz0=200.0, z_fact=1.0, ns=6, tout_int=1.0)
    s = ic.init_state_bf2002(allocate_state(p), p)
    h = p.halo
    rho, pc, z = s.rho[:, h], s._bf2002_pressure, p.z
    res = pc[1:] - pc[:-1] + c.GRAV * (z[1:] - z[:-1]) * (rho[1:] + rho[:-1]) / 2
    assert np.abs(res).max() <= 1e-9 * pc.max()
    inner = np.s_[:, h:h + p.Nx]
    qv = s.rho_qv[inner] / s.rho[inner]
    qc = s.rho_qc[inner] / s.rho[inner]
    np.testing.assert_allclose(qv + qc, 0.02, rtol=1e-13)
    assert (qc >= 0).all() and (s.rho_qr == 0).all()
    # model EOS: pres = P0 (r rho theta / P0)^(cp/cv)
    qd = 1.0 - qv - qc
    r = c.Rd * qd + c.Rv * qv
    cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * qc
    pres = c.P0 * (r * s.rho_theta[inner] / c.P0) ** (cp / (cp - r))
    far = np.abs(p.x[h:h + p.Nx] - p.Lx / 2) > 2.5e3  # outside the bubble
    np.testing.assert_allclose(pres[:, far], pc[:, None] * np.ones((1, far.sum())), rtol=1e-10)
    # saturated theta_e of the unperturbed columns (the inversion's own formula)
    t, qs = ic.invert_T_for_thetae(0.02, 320.0, pc)
    Tc = t - c.T0
    es = c.ES0 * np.exp(c.A_SAT * Tc / (c.B_SAT + Tc))
    cpl = c.Cpd * 0.98 + c.Cpl * 0.02
    the = t * ((pc - es) / c.P0) ** (-c.Rd / cpl) * np.exp(c.Lv * qs / (cpl * t))
    # the reference brackets T in [250, 350] K: above ~7 km the root lies below 250 K and the
    # bisection returns the bracket's end, as moist.jl:275-289 does
    inside = t > 250.0 + 1e-6
    assert inside.sum() >= 30
    np.testing.assert_allclose(the[inside], 320.0, atol=1e-6)
    # the bubble: theta scaled by 1 + theta_pert cos^2 / 300 at its centre
    th = s.rho_theta[inner] / s.rho[inner]
    assert th[:, ~far].max() > th[:, far].max() or th[:, ~far].max() > 0


def test_create_rejects_bad_members_before_touching_the_device():
    """moist_create validates every member first (no HIP call on this path):
    halo > nx and a member whose field takes 2^32 bytes or more return -2
    with a message."""
    L = core.lib()
    p = make_params(np.float64, Nx=16, Nz=8, Lx=1600.0, H=800.0, dz0=100.0, z_fact=1.0, ns=6, tout_int=1.0)

    def create(mutate):
        c = p.to_c()
        mutate(c)
        arr = (ctypes.c_void_p * 1)(ctypes.addressof(c))
        h = ctypes.c_void_p()
        rc = L.moist_create(arr, 1, core.FP64, ctypes.byref(h))
        msg = L.moist_last_error(h).decode() if h.value else ""
        if h.value:
            L.moist_destroy(h)
        return rc, msg

    rc, msg = create(lambda c: setattr(c, "halo", 17))
    assert rc == -2 and "invalid member params" in msg
    # 2^26 columns x (8 + 1) levels x 8 bytes = 4.8e9 bytes >= 2^32
    rc, msg = create(lambda c: setattr(c, "nx", 1 << 26))
    assert rc == -2 and "member too large" in msg
