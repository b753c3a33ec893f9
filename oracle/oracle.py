"""ctypes front end of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; it is the checker, never the thing measured or shipped.  The
functions mirror the reference's routines one to one (moist.jl line ranges
in moist_oracle.cpp) and take host State/Params objects of moist2d_amd.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ORDER_SNAPSHOT = 0
ORDER_SERIAL = 1

NAN_FIELDS = ("rho", "rho_u", "rho_w", "rho_theta", "rho_qv", "rho_qc", "rho_qr", "d_rho_u_phys",
              "d_rho_theta_phys", "d_rho_qv_phys", "d_rho_qc_phys", "d_rho_qr_phys")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        for sfx, ct in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
            P = ctypes.POINTER(ct)
            for name, res, args in (
                ("f2h", ct, [ct, ct, ct, ct]),
                ("koren", ct, [ct]),
                ("sat_vp", None, [ct, P, P]),
                ("sat_q", None, [ct, ct, P, P]),
                ("thomas", None, [ctypes.c_int, P, P, P, P]),
                ("implicit_correction", None, [ctypes.c_void_p, P, P, P, P, ct]),
                ("flux_x", ct, [ctypes.c_void_p, P, P, P, ctypes.c_int, ctypes.c_int]),
                ("flux_z", ct, [ctypes.c_void_p, P, P, P, ctypes.c_int, ctypes.c_int]),
                ("exchange_halo", None, [ctypes.c_void_p, P]),
                ("cloud_microphysics", None, [ctypes.c_void_p] + [P] * 9),
                ("rain_sedimentation", None, [ctypes.c_void_p, P, P, P, P]),
                ("surface_fluxes", None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
                ("gray_radiation", None, [ctypes.c_void_p, P, ctypes.c_void_p, P]),
                ("compute_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ct, ctypes.c_int, ctypes.c_int]),
                ("rk3_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
            ):
                f = getattr(_lib, f"orc_{name}_{sfx}")
                f.restype = res
                f.argtypes = args
        V = ctypes.c_void_p
        for name, res, args in (
            ("compute_step_adj", ctypes.c_int, [V, V, V, V, ctypes.c_double, ctypes.c_int, V, V, V, V]),
            ("physics_fwd", None, [V, V, V, V, V]),
            ("physics_adj", None, [V, V, V, V, V, V, V, V]),
            ("rk3_step_adj", ctypes.c_int, [V, V, V, ctypes.c_double, V, V, V]),
            ("adjoint_run", ctypes.c_int, [V, V, ctypes.c_int, ctypes.c_int, V, V, V, V]),
            ("output_history", None, [V, V, V, V]),
            ("slow_tendencies", ctypes.c_int, [V, V, V, V]),
            ("jvp_run", ctypes.c_int, [V, V, V, V, ctypes.c_int, V, V, V]),
        ):
            f = getattr(_lib, f"orc_{name}_f64")
            f.restype = res
            f.argtypes = args
        _lib.orc_set_threads.restype = ctypes.c_int
        _lib.orc_set_threads.argtypes = [ctypes.c_int]
    return _lib


def _sfx(T):
    T = np.dtype(T)
    if T == np.float64:
        return "f64", ctypes.c_double
    if T == np.float32:
        return "f32", ctypes.c_float
    raise TypeError(T)


def _ptr(a, ct):
    assert a.flags["F_CONTIGUOUS"] or a.ndim == 1
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _fn(name, T):
    sfx, ct = _sfx(T)
    return getattr(lib(), f"orc_{name}_{sfx}"), ct


def set_threads(n: int) -> int:
    return lib().orc_set_threads(int(n))


# ---- scalar helpers ---------------------------------------------------------
def f2h(a, b, c, d, T=np.float64):
    f, ct = _fn("f2h", T)
    return np.dtype(T).type(f(a, b, c, d))


def koren_limiter(r, T=np.float64):
    f, ct = _fn("koren", T)
    return np.dtype(T).type(f(r))


def saturation_vapor_pressure(Tk, T=np.float64):
    f, ct = _fn("sat_vp", T)
    es, des = ct(), ct()
    f(Tk, ctypes.byref(es), ctypes.byref(des))
    return np.dtype(T).type(es.value), np.dtype(T).type(des.value)


def saturation_specific_humidity(p, es, T=np.float64):
    f, ct = _fn("sat_q", T)
    q, dq = ct(), ct()
    f(p, es, ctypes.byref(q), ctypes.byref(dq))
    return np.dtype(T).type(q.value), np.dtype(T).type(dq.value)


def solve_tridiagonal(a, b, c, d):
    """solve_tridiagonal!(a, b, c, d): modifies c and d in place."""
    f, ct = _fn("thomas", d.dtype)
    f(len(b), _ptr(a, ct), _ptr(b, ct), _ptr(c, ct), _ptr(d, ct))


def implicit_correction(rho_w, drho_w, theta, rt2pres, dt, p):
    f, ct = _fn("implicit_correction", rho_w.dtype)
    cp = p.to_c()
    f(ctypes.byref(cp), _ptr(rho_w, ct), _ptr(drho_w, ct), _ptr(theta, ct), _ptr(rt2pres, ct), dt)


def compute_limited_flux_x(scalar_rho, rho, rho_u, k, i, p):
    f, ct = _fn("flux_x", rho.dtype)
    cp = p.to_c()
    return f(ctypes.byref(cp), _ptr(scalar_rho, ct), _ptr(rho, ct), _ptr(rho_u, ct), k, i)


def compute_limited_flux_z(scalar_rho, rho, rho_w, k, i, p):
    f, ct = _fn("flux_z", rho.dtype)
    cp = p.to_c()
    return f(ctypes.byref(cp), _ptr(scalar_rho, ct), _ptr(rho, ct), _ptr(rho_w, ct), k, i)


def exchange_halo(A, p):
    f, ct = _fn("exchange_halo", A.dtype)
    cp = p.to_c()
    f(ctypes.byref(cp), _ptr(A, ct))


def cloud_microphysics(d_qv, d_qc, d_qr, d_th, rho, rho_theta, rho_qv, rho_qc, rho_qr, p):
    f, ct = _fn("cloud_microphysics", rho.dtype)
    cp = p.to_c()
    f(ctypes.byref(cp), *[_ptr(a, ct) for a in (d_qv, d_qc, d_qr, d_th, rho, rho_theta, rho_qv, rho_qc, rho_qr)])


def rain_sedimentation(d_sed, d_prec, rho_qr, rho, p):
    f, ct = _fn("rain_sedimentation", rho.dtype)
    cp = p.to_c()
    f(ctypes.byref(cp), *[_ptr(a, ct) for a in (d_sed, d_prec, rho_qr, rho)])


def _ptr_array(arrs, ct):
    arr = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return arr


def surface_fluxes(out, s, p):
    """out = (d_rho_u_sf, d_rho_theta_sf, d_rho_qv_sf, mflux, shflux, lhflux)."""
    f, ct = _fn("surface_fluxes", s.rho.dtype)
    cp = p.to_c()
    o = _ptr_array(out, ct)
    ss = _ptr_array([getattr(s, n) for n in ("rho", "rho_u", "rho_w", "rho_theta", "rho_qv", "rho_qc", "rho_qr")], ct)
    f(ctypes.byref(cp), o, ss)


def gray_radiation(d_rd, s, p):
    f, ct = _fn("gray_radiation", s.rho.dtype)
    cp = p.to_c()
    ss = _ptr_array([getattr(s, n) for n in ("rho", "rho_u", "rho_w", "rho_theta", "rho_qv", "rho_qc", "rho_qr")], ct)
    f(ctypes.byref(cp), _ptr(d_rd, ct), ss, _ptr(s.theta_ref, ct))


_PROG = ("rho", "rho_u", "rho_w", "rho_theta", "rho_qv", "rho_qc", "rho_qr")


def compute_step(phys, s, s0, p, dt, ns, order=ORDER_SNAPSHOT):
    """compute_step!(phys..., s, s0, p, dt, ns); returns the NaN code (0 = none)."""
    f, ct = _fn("compute_step", s.rho.dtype)
    cp = p.to_c()
    ph = _ptr_array(list(phys), ct)
    a = _ptr_array([getattr(s, n) for n in _PROG], ct)
    b = _ptr_array([getattr(s0, n) for n in _PROG], ct)
    return f(ctypes.byref(cp), ph, a, b, dt, ns, order)


def rk3_step(s, p, nsteps=1, order=ORDER_SNAPSHOT):
    """nsteps x rk3_step!(s, p); returns the NaN code of the failing stage (0 = none)."""
    f, ct = _fn("rk3_step", s.rho.dtype)
    cp = p.to_c()
    arrs = s.arrays()
    for a in arrs:
        assert a.dtype == s.rho.dtype and (a.flags["F_CONTIGUOUS"])
    return f(ctypes.byref(cp), _ptr_array(arrs, ct), nsteps, order)


# ---- adjoint (FP64) ---------------------------------------------------------
def _f64(arrs):
    for a in arrs:
        assert a.dtype == np.float64 and (a.flags["F_CONTIGUOUS"] or a.ndim == 1), a.dtype
    return _ptr_array(list(arrs), ctypes.c_double)


def compute_step_adj(phys, s, s0, p, dt, ns, lam_out):
    """Adjoint of compute_step! at (phys, s, s0): returns (lam_s[7], lam_s0[7], lam_phys[6])
    for the adjoint `lam_out` (7 fields) of the stage output."""
    shape = s.rho.shape
    ls = [np.zeros(shape, order="F") for _ in range(7)]
    ls0 = [np.zeros(shape, order="F") for _ in range(7)]
    lph = [np.zeros(shape, order="F") for _ in range(6)]
    cp = p.to_c()
    rc = lib().orc_compute_step_adj_f64(ctypes.byref(cp), _f64(phys), _f64([getattr(s, n) for n in _PROG]),
                                        _f64([getattr(s0, n) for n in _PROG]), dt, ns, _f64(lam_out), _f64(ls),
                                        _f64(ls0), _f64(lph))
    if rc:
        raise RuntimeError(f"compute_step_adj failed ({rc})")
    return ls, ls0, lph


def slow_tendencies(phys, s, p):
    """The slow tendencies of compute_step! at (phys, s) (Float64):
    dict of R_rho, R_rho_u, R_rho_w, R_rho_theta and the face velocities
    u_face, w_face, as the step itself computes them."""
    shape = s.rho.shape
    out = [np.zeros(shape, order="F") for _ in range(6)]
    cp = p.to_c()
    rc = lib().orc_slow_tendencies_f64(ctypes.byref(cp), _f64(phys), _f64([getattr(s, n) for n in _PROG]),
                                       _f64(out))
    if rc:
        raise RuntimeError(f"slow_tendencies: NaN code {rc}")
    return dict(zip(("R_rho", "R_rho_u", "R_rho_w", "R_rho_theta", "u_face", "w_face"), out))


def _vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def physics_fwd(s, p):
    """The six combined tendencies of rk3_step! and d_prec (surface fluxes read
    s's halo columns; radiation relaxes to s.theta_ref)."""
    shape = s.rho.shape
    ph = [np.zeros(shape, order="F") for _ in range(6)]
    dprec = np.zeros(shape[1])
    cp = p.to_c()
    tr = np.ascontiguousarray(s.theta_ref, dtype=np.float64)
    lib().orc_physics_fwd_f64(ctypes.byref(cp), _f64([getattr(s, n) for n in _PROG]), _f64(ph), _vp(dprec), _vp(tr))
    return ph, dprec


def physics_adj(s, p, lam_phys, lam_dprec):
    """Adjoint of physics_fwd: returns (lam_s[7], lam_K[3], lam_stale[2]); lam_stale
    holds the adjoints of the two halo values the surface fluxes read (rho_u
    left of the first column, rho right of the last)."""
    shape = s.rho.shape
    ls = [np.zeros(shape, order="F") for _ in range(7)]
    lK = np.zeros(3)
    st = np.zeros(2)
    cp = p.to_c()
    lam_dprec = np.ascontiguousarray(lam_dprec, dtype=np.float64)
    tr = np.ascontiguousarray(s.theta_ref, dtype=np.float64)
    lib().orc_physics_adj_f64(ctypes.byref(cp), _f64([getattr(s, n) for n in _PROG]), _f64(lam_phys),
                              _vp(lam_dprec), _f64(ls), _vp(lK), _vp(tr), _vp(st))
    return ls, lK, st


def rk3_step_adj(s, p, lam, lam_dprec=0.0, stale=None):
    """Adjoint of one rk3_step! at s (its halo columns as the forward step saw
    them): `lam` (7 fields, adjoint of the step output) is replaced by the
    adjoint of the step input; returns lam_K[3].  `stale` (2 values, in/out,
    surface fluxes on): on entry the adjoints of the next step's two halo
    reads, which are this step's stage-3 input at the ground; on exit those of
    this step's own halo reads."""
    lK = np.zeros(3)
    cp = p.to_c()
    tr = np.ascontiguousarray(s.theta_ref, dtype=np.float64)
    st = np.zeros(2) if stale is None else stale
    assert st.dtype == np.float64 and st.flags["C_CONTIGUOUS"] and st.size == 2
    rc = lib().orc_rk3_step_adj_f64(ctypes.byref(cp), _f64([getattr(s, n) for n in _PROG]), _f64(lam),
                                    float(lam_dprec), _vp(lK), _vp(tr), _vp(st))
    if rc:
        raise RuntimeError(f"rk3_step_adj failed ({rc})")
    return lK


def adjoint_run(s, p, nsteps, theta_target, ckpt=50):
    """Loss J = dx sum(prec) + 1/2 sum((theta - theta_target)^2) after nsteps from s
    (s is advanced in place) and its gradient: (J, grad[7] w.r.t. the initial
    prognostic fields, grad_K[3] w.r.t. K_auto, Qc_threshold, K_accr)."""
    shape = s.rho.shape
    grad = [np.zeros(shape, order="F") for _ in range(7)]
    gK = np.zeros(3)
    J = ctypes.c_double()
    cp = p.to_c()
    arrs = s.arrays()
    tt = np.asfortranarray(theta_target, dtype=np.float64)
    rc = lib().orc_adjoint_run_f64(ctypes.byref(cp), _f64(arrs), nsteps, ckpt, tt.ctypes.data_as(ctypes.c_void_p),
                                   ctypes.byref(J), _f64(grad), gK.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise RuntimeError(f"adjoint_run failed ({rc})")
    return J.value, grad, gK


def jvp_run(s, p, nsteps, theta_target, direction, dK=(0.0, 0.0, 0.0)):
    """Forward-mode derivative of adjoint_run's loss: (J, dJ), dJ the
    derivative of J along `direction` ({field name: array in the State's
    layout} for any of the 7 prognostic fields) and dK (tangents of K_auto,
    Qc_threshold, K_accr), from nsteps rk3_step!s run on dual numbers
    (branches follow the primal trajectory, as ForwardDiff's do).  s is not
    modified."""
    shape = s.rho.shape
    dirs = [np.asfortranarray(direction.get(n, np.zeros(shape)), dtype=np.float64) for n in _PROG]
    J, dJ = ctypes.c_double(), ctypes.c_double()
    cp = p.to_c()
    arrs = [np.array(a, dtype=np.float64, order="F") for a in s.arrays()]
    tt = np.asfortranarray(theta_target, dtype=np.float64)
    dk = np.asarray(dK, dtype=np.float64)
    rc = lib().orc_jvp_run_f64(ctypes.byref(cp), _f64(arrs), _f64(dirs), dk.ctypes.data_as(ctypes.c_void_p), nsteps,
                               tt.ctypes.data_as(ctypes.c_void_p), ctypes.byref(J), ctypes.byref(dJ))
    if rc:
        raise RuntimeError(f"jvp_run failed ({rc})")
    return J.value, dJ.value


HIST_FIELDS = ("rho", "rhou", "rhow", "u", "w", "theta", "qv", "qc", "qr", "RH", "temp", "pres", "thetae")


def output_history(s, p):
    """output_history's derived fields ((Nz, Nx), x fastest) and stats
    (Etot sum, w min/max, qc/qr max g/kg, RH min/max) of a Float64 State."""
    out = [np.zeros((p.Nz, p.Nx)) for _ in HIST_FIELDS]
    st = np.zeros(7)
    cp = p.to_c()
    lib().orc_output_history_f64(ctypes.byref(cp), _f64([getattr(s, n) for n in _PROG]),
                                 _ptr_array(out, ctypes.c_double), st.ctypes.data_as(ctypes.c_void_p))
    return dict(zip(HIST_FIELDS, out)), st
