"""Params{T} and set_constants! of the reference, mirrored on the host.

`Params(T, **kwargs)` follows the reference's keyword struct and its
constructor (moist.jl:70-222): the same field names and defaults, dt derived
from CFL 0.6 and c_s = 340 m/s (moist.jl:187-195), the stretched z grid
(moist.jl:203-216) built in the requested precision, and the same error
when the stretched grid does not reach H.  K_auto, Qc_threshold and K_accr,
globals in the reference (moist.jl:46-48), are fields here so that ensembles
and the adjoint can vary them.

`to_c()` flattens a Params into the `moist_params` struct of
include/moist2d.h (values widened to double, exact for Float32).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field, fields, replace
from typing import Any

import numpy as np


class moist_params(ctypes.Structure):
    """ctypes image of `moist_params` (include/moist2d.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("nz", ctypes.c_int32), ("halo", ctypes.c_int32), ("ns", ctypes.c_int32),
        ("do_mp", ctypes.c_int32), ("do_auto", ctypes.c_int32), ("do_accr", ctypes.c_int32),
        ("do_revap", ctypes.c_int32), ("do_sed", ctypes.c_int32), ("do_rd", ctypes.c_int32),
        ("do_sf", ctypes.c_int32), ("reserved0", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dt", ctypes.c_double), ("H", ctypes.c_double),
        ("nu", ctypes.c_double), ("kappa", ctypes.c_double), ("gamma_d", ctypes.c_double),
        ("beta_offcenter", ctypes.c_double), ("zf_sponge", ctypes.c_double),
        ("Cd", ctypes.c_double), ("Ch", ctypes.c_double), ("Ce", ctypes.c_double),
        ("uabs_min", ctypes.c_double), ("temp_surf", ctypes.c_double),
        ("k_abs_qv", ctypes.c_double), ("k_abs_qc", ctypes.c_double), ("k_abs_qr", ctypes.c_double),
        ("surf_emissivity", ctypes.c_double), ("tau_cooling", ctypes.c_double),
        ("z_tropopose", ctypes.c_double),
        ("K_auto", ctypes.c_double), ("Qc_threshold", ctypes.c_double), ("K_accr", ctypes.c_double),
        ("z", ctypes.POINTER(ctypes.c_double)), ("zf", ctypes.POINTER(ctypes.c_double)),
        ("dz", ctypes.POINTER(ctypes.c_double)),
    ]


class Constants:
    """set_constants!(T) (moist.jl:26-60): the physical constants in type T."""

    def __init__(self, T=np.float32):
        T = np.dtype(T).type
        self.T = T
        self.GRAV = T(9.81)
        self.Rd = T(287.04)
        self.Rv = T(461.5)
        self.Cpd = T(1004.64)
        self.Cpv = T(1846.0)
        self.Cpl = T(4218.0)
        self.Cvd = T(self.Cpd - self.Rd)
        self.P0 = T(1e5)
        self.Lv = T(2.501e6)
        self.Ka = T(0.4)
        self.SIGMA = T(5.67e-8)
        self.EPSvap = T(self.Rd / self.Rv)
        self.EPSTvap = T(self.Rv / self.Rd - T(1.0))
        self.K_auto = T(1e-3)
        self.Qc_threshold = T(1e-3)
        self.K_accr = T(2.2)
        self.K_evap = T(1e-2)
        self.A_term = T(36.3)
        self.B_term = T(0.1346)
        self.T0 = T(273.15)
        self.ES0 = T(610.78)
        self.A_SAT = T(17.67)
        self.B_SAT = T(243.5)
        self.EPS = T(1e-12)


# Keyword fields of Params{T} with their defaults (moist.jl:70-178).  Entries
# that are None are derived from other fields when not given.
_DEFAULTS: dict[str, Any] = dict(
    Lx=160e3, H=25e3, Nx=640, Nz=250, halo=2, ia=None, ka=None, is_=None, ie=None, ks=None, ke=None,
    dx=None, dz0=50.0, z_fact=1.02, nu=1.0, Pr=0.7, kappa=None, gamma_d=0.1, beta_offcenter=1.0,
    zf_sponge=0.2, do_mp=True, do_auto=True, do_accr=True, do_revap=True, do_sed=True, do_rd=True,
    do_sf=True, Cd=1.2e-3, Ch=0.8e-3, Ce=1.0e-3, uabs_min=5.0, temp_surf=300.0, k_abs_qv=0.15,
    k_abs_qc=20.0, k_abs_qr=10.0, surf_emissivity=1.0, tau_cooling=864000.0, u0=0.0, theta0=300.0,
    bndlayer_depth=1000.0, z_tropopose=15e3, dtheta_dz_trop=4.0e-3, dtheta_dz_stra=18.3e-3,
    theta_pert=2.0, rel_hum_surf=90.0, rel_hum_decay=4000.0, t_end=86400.0 * 2.0, tout_int=300.0,
    ns=3, dt=0.0, N_steps=0, Nint_steps=0,
    K_auto=1e-3, Qc_threshold=1e-3, K_accr=2.2,
)
_INT_FIELDS = {"Nx", "Nz", "halo", "ia", "ka", "is_", "ie", "ks", "ke", "ns", "N_steps", "Nint_steps"}
_BOOL_FIELDS = {"do_mp", "do_auto", "do_accr", "do_revap", "do_sed", "do_rd", "do_sf"}
_F64_FIELDS = {"t_end", "tout_int"}  # ::Float64 in the reference


@dataclass
class Params:
    """Params{T} (moist.jl:70-178) plus its derived grid (moist.jl:180-222).

    Julia's 1-based `is/ie/ks/ke` are kept as in the reference (1-based) so
    that formulas transcribe literally; `is_` stands for `is` (a Python
    keyword).  Arrays `x, zf, dz, z` are numpy arrays of dtype T.
    """

    T: Any
    values: dict = field(default_factory=dict)
    x: np.ndarray = None
    zf: np.ndarray = None
    dz: np.ndarray = None
    z: np.ndarray = None
    constants: Constants = None

    def __getattr__(self, name):
        if name == "is":
            name = "is_"
        v = self.__dict__.get("values", {})
        if name in v:
            return v[name]
        raise AttributeError(name)

    # --- flattening for the C ABI -------------------------------------
    def to_c(self) -> moist_params:
        p = moist_params()
        for n in ("halo", "ns"):
            setattr(p, n, int(self.values[n]))
        p.nx = int(self.values["Nx"])
        p.nz = int(self.values["Nz"])
        for n in _BOOL_FIELDS:
            setattr(p, n, 1 if self.values[n] else 0)
        for n in ("dx", "dt", "H", "nu", "kappa", "gamma_d", "beta_offcenter", "zf_sponge", "Cd", "Ch", "Ce",
                  "uabs_min", "temp_surf", "k_abs_qv", "k_abs_qc", "k_abs_qr", "surf_emissivity",
                  "tau_cooling", "z_tropopose", "K_auto", "Qc_threshold", "K_accr"):
            setattr(p, n, float(self.values[n]))
        self._c_arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in (self.z, self.zf, self.dz)]
        dp = ctypes.POINTER(ctypes.c_double)
        p.z, p.zf, p.dz = (a.ctypes.data_as(dp) for a in self._c_arrays)
        p._keep = self._c_arrays  # keep the arrays alive with the struct
        return p

    def replace(self, **kw) -> "Params":
        """Like Parameters.jl's `reconstruct`: rebuild with changed keywords."""
        base = {k: v for k, v in self.values.items() if k not in ("dt", "N_steps", "Nint_steps")}
        base.update(kw)
        return make_params(self.T, **base)


def _jl_min(x, y):
    """Base.min for IEEE floats (NaN-propagating, -0.0 < +0.0)."""
    d = x - y
    if math.isnan(float(x)) or math.isnan(float(y)):
        return d
    return x if np.signbit(d) else y


def make_params(T=np.float32, grid=True, **kwargs) -> Params:
    """`Params(T; kwargs...)` (moist.jl:180-222).

    grid=False is the reference's inner constructor `Params{T}(; kwargs...)`:
    no dt derivation and no grid (the vectors are zero-filled of length Nz).
    """
    T = np.dtype(T).type
    unknown = set(kwargs) - set(_DEFAULTS) - {"is"}
    if unknown:
        raise TypeError(f"unknown Params keywords: {sorted(unknown)}")
    if "is" in kwargs:
        kwargs["is_"] = kwargs.pop("is")
    v: dict[str, Any] = {}
    # Field defaults in declaration order; derived defaults see earlier fields.
    for name, default in _DEFAULTS.items():
        if name in kwargs:
            val = kwargs[name]
        elif default is not None:
            val = default
        else:
            val = None
        if val is None:
            if name == "ia":
                val = v["Nx"] + 2 * v["halo"]
            elif name == "ka":
                val = v["Nz"]
            elif name == "is_":
                val = 1 + v["halo"]
            elif name == "ie":
                val = v["Nx"] + v["halo"]
            elif name == "ks":
                val = 1
            elif name == "ke":
                val = v["Nz"]
            elif name == "dx":
                val = T(v["Lx"]) / T(v["Nx"])
            elif name == "kappa":
                val = T(v["nu"]) / T(v["Pr"])
        if name in _INT_FIELDS:
            val = int(val)
        elif name in _BOOL_FIELDS:
            val = bool(val)
        elif name in _F64_FIELDS:
            val = float(val)
        else:
            val = T(val)
        v[name] = val

    if not grid:
        c = Constants(T)
        c.K_auto, c.Qc_threshold, c.K_accr = T(v["K_auto"]), T(v["Qc_threshold"]), T(v["K_accr"])
        zero = np.zeros(v["ka"], dtype=T)
        return Params(T=T, values=v, x=np.zeros(v["ia"], dtype=T), zf=zero.copy(), dz=zero.copy(), z=zero.copy(),
                      constants=c)

    # Time step (moist.jl:187-195): Float64 arithmetic, then dt = T(...)
    CFL, cs = 0.6, 340.0
    dt_raw = CFL * float(v["dx"]) * v["ns"] / cs
    Nint = math.ceil(v["tout_int"] / dt_raw)
    dt = T(v["tout_int"] / Nint)
    N_steps = math.ceil(v["t_end"] / float(dt))
    if "dt" not in kwargs:
        v["dt"] = dt
    v["N_steps"] = N_steps
    v["Nint_steps"] = Nint

    # Axis (moist.jl:203-216), in T
    ia, ka, is_, ks, ke = v["ia"], v["ka"], v["is_"], v["ks"], v["ke"]
    x = np.empty(ia, dtype=T)
    for i in range(1, ia + 1):
        x[i - 1] = (T(i - is_) + T(0.5)) * v["dx"]
    zf = np.empty(ka, dtype=T)
    dz = np.empty(ka, dtype=T)
    z = np.empty(ka, dtype=T)
    zf[ks - 1] = v["dz0"]
    dz[ks - 1] = v["dz0"]
    z[ks - 1] = zf[ks - 1] * T(0.5)
    for k in range(ks + 1, ke + 1):
        a = dz[k - 2] * v["z_fact"]
        b = (v["H"] - zf[k - 2]) / T(ke - k + 1)
        dz[k - 1] = _jl_min(T(a), T(b))
        zf[k - 1] = zf[k - 2] + dz[k - 1]
        z[k - 1] = (zf[k - 2] + zf[k - 1]) * T(0.5)
    if zf[ke - 1] < v["H"]:
        raise ValueError(f"z_fact is too small: {v['z_fact']}, {zf[ke - 1]} < {v['H']}")
    c = Constants(T)
    c.K_auto, c.Qc_threshold, c.K_accr = T(v["K_auto"]), T(v["Qc_threshold"]), T(v["K_accr"])
    return Params(T=T, values=v, x=x, zf=zf, dz=dz, z=z, constants=c)
