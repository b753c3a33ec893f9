"""History output, restart input and the run driver (moist.jl:1686-1963).

`History` writes the reference's history variables (init_history,
moist.jl:1718-1795; output_history, moist.jl:1798-1889) in NetCDF classic
(64-bit offset) format with the same dimensions ("x", "z", "time"), names,
units and layout: data are written transposed (`rho_int'`, moist.jl:1862),
so x is fastest on disk, which is the device layout.  The derived fields
(u, w, theta, q, RH, temperature, pressure in hPa, theta_e) and the energy
budget come from the device (`Engine.history`, moist_history in the C ABI).
The reference writes NetCDF-4 through NCDatasets; this image has no HDF5
library, and classic format carries the same variables for any NetCDF reader.

`read_restart` fills a State from the last time slice of a history file as
run_sim's restart branch does (moist.jl:1908-1919), except that rho_qr is
read from "qr" (the reference reads "qc" there).

`run_sim` mirrors the reference's driver loop (moist.jl:1922-1963) on the
device-resident Engine: initial state, history at t = 0 and every
Nint_steps steps, NaN check of rho, rho_theta, rho_qv at each output.
"""
from __future__ import annotations

import logging

import numpy as np
from scipy.io import netcdf_file

from . import core
from .ic import init_state
from .state import PROGNOSTIC, State, allocate_state, exchange_halo

log = logging.getLogger("moist2d_amd")

# (name, units, dims) in init_history's order; 2-D fields are ("time", "z", "x") on disk
VARS_2D = [
    ("rho", "kg/m^3"), ("rhou", "kg/m^2/s"), ("rhow", "kg/m^2/s"), ("u", "m/s"), ("w", "m/s"), ("theta", "K"),
    ("qv", "kg/kg"), ("qc", "kg/kg"), ("qr", "kg/kg"), ("RH", "%"), ("temp", "K"), ("pres", "hPa"),
    ("thetae", "K"),
]
VARS_PHYS_2D = [
    ("d_rho_theta_mp", "kg K/m^3/s"), ("d_rho_qv_mp", "kg/m^3/s"), ("d_rho_qc_mp", "kg/m^3/s"),
    ("d_rho_qr_mp", "kg/m^3/s"), ("d_rho_qr_sed", "kg/m^3/s"), ("d_rho_theta_rd", "kg K/m^3/s"),
]
VARS_1D = [
    ("prec", "kg/m^2"), ("mflux", "kg/m/s^2"), ("shflux", "W/m^2"), ("lhflux", "W/m^2"),
    ("d_rho_u_sf", "kg/m^2/s^2"), ("d_rho_theta_sf", "kg K/m^3/s"), ("d_rho_qv_sf", "kg/kg/m^3/s"),
]


class History:
    """A history file (init_history, moist.jl:1718-1795)."""

    def __init__(self, path, p, dtype=None):
        self.p = p
        T = np.dtype(dtype or p.T)
        self.T = T
        self.nc = netcdf_file(path, "w", version=2)
        nc = self.nc
        nc.createDimension("time", None)  # the record dimension is defined first in this format
        nc.createDimension("x", p.Nx)
        nc.createDimension("z", p.Nz)
        is0 = p.is_ - 1
        for name, arr in (("x", p.x[is0:is0 + p.Nx]), ("z", p.z), ("zf", p.zf)):
            v = nc.createVariable(name, T, ("x",) if name == "x" else ("z",))
            v[:] = np.asarray(arr, dtype=T)
            v.units = "m"
        t = nc.createVariable("time", T, ("time",))
        t.units = "s"
        for name, units in VARS_2D + VARS_PHYS_2D:
            nc.createVariable(name, T, ("time", "z", "x")).units = units
        for name, units in VARS_1D:
            nc.createVariable(name, T, ("time", "x")).units = units
        self.n = 0
        self.Etot0 = None

    def write(self, t, eng: core.Engine, s: State, member=0):
        """output_history (moist.jl:1798-1889): device-derived fields and budget,
        State physics diagnostics from `s` (downloaded beforehand).  Returns the
        stats line the reference logs."""
        fields, st = eng.history(member)
        return self.write_fields(t, fields, st, s)

    def write_fields(self, t, fields, st, s: State):
        """Append one record from derived `fields` (dict of (Nz, Nx) arrays), the
        stats dict and the State's physics diagnostics."""
        if self.Etot0 is None:
            self.Etot0 = st["Etot"]
        st = dict(st, Etot=st["Etot"] - self.Etot0)
        nc, p, n = self.nc, self.p, self.n
        nc.variables["time"][n] = t
        for name, _ in VARS_2D:
            nc.variables[name][n] = fields[name].astype(self.T)
        is0 = p.is_ - 1
        inner = np.s_[:, is0:is0 + p.Nx]
        for name, _ in VARS_PHYS_2D:
            nc.variables[name][n] = np.ascontiguousarray(getattr(s, name)[inner]).astype(self.T)
        for name, _ in VARS_1D:
            nc.variables[name][n] = getattr(s, name)[is0:is0 + p.Nx].astype(self.T)
        self.n += 1
        nc.flush()
        log.info("t=%.2f s | w(min,max)=(%.2f, %.2f) m/s | qcmax=%.2f g/kg | qrmax=%.2f g/kg | "
                 "RH(min,max)=(%.2f, %.2f) %% | Etot=%.2e", t, st["wmin"], st["wmax"], st["qcmax"], st["qrmax"],
                 st["RHmin"], st["RHmax"], st["Etot"])
        return st

    def close(self):
        self.nc.close()


def read_restart(path, p) -> State:
    """State from the last time slice of a history file (moist.jl:1908-1919)."""
    s = allocate_state(p)
    with netcdf_file(path, "r", mmap=False) as nc:
        v = nc.variables
        is0 = p.is_ - 1
        inner = np.s_[:, is0:is0 + p.Nx]
        rho = np.asarray(v["rho"][-1], dtype=np.float64)
        s.rho[inner] = rho
        s.rho_u[inner] = v["rhou"][-1]
        s.rho_w[inner] = v["rhow"][-1]
        s.rho_theta[inner] = np.asarray(v["theta"][-1], dtype=np.float64) * rho
        s.rho_qv[inner] = np.asarray(v["qv"][-1], dtype=np.float64) * rho
        s.rho_qc[inner] = np.asarray(v["qc"][-1], dtype=np.float64) * rho
        s.rho_qr[inner] = np.asarray(v["qr"][-1], dtype=np.float64) * rho
        s.prec[is0:is0 + p.Nx] = v["prec"][-1]
    for n in PROGNOSTIC:
        exchange_halo(getattr(s, n), p)
    return s


def run_sim(p, path="history.nc", precision=None, restart=None, state=None, seed=0, graph=True):
    """run_sim (moist.jl:1922-1963) on the device: returns the final State."""
    log.info("Configuration: %d x %d | dt=%s | dx=%s", p.Nx, p.Nz, p.dt, p.dx)
    if restart:
        s = read_restart(restart, p)
    elif state is not None:
        s = state
    else:
        s = init_state(allocate_state(p), p, seed=seed)
        s.prec[...] = 0
    for n in PROGNOSTIC:
        exchange_halo(getattr(s, n), p)
    prec = precision if precision is not None else (core.FP64 if s.rho.dtype == np.float64 else core.MIXED)
    eng = core.Engine(p, prec)
    hist = History(path, p)
    try:
        eng.upload(s)
        eng.use_graph(graph)
        hist.write(0.0, eng, s)
        nint = max(1, int(p.Nint_steps))
        done = 0
        while done < p.N_steps:
            n = min(nint, p.N_steps - done)
            eng.step(n)
            done += n
            if done % nint == 0:
                eng.download(s)
                for f in ("rho", "rho_theta", "rho_qv"):
                    if np.isnan(getattr(s, f)).any():
                        log.error("NaN detected at step %d, time %s. Aborting.", done, done * float(p.dt))
                        return s
                hist.write(float(done * float(p.dt)), eng, s)
        eng.download(s)
    finally:
        hist.close()
        eng.close()
    return s
