"""ctypes binding of libmoist2d_hip.so and the drop-in step routines.

`rk3_step(s, p)` and `compute_step(d_rho_phys, ..., s, s0, p, dt, ns)` keep
the signatures and in-place semantics of the reference's `rk3_step!`
(moist.jl:1589) and `compute_step!` (moist.jl:1039-1040), and raise the
reference's `error("NaN detected in <field>")` as RuntimeError with the same
text.  `Engine` is the device-resident form (no per-step copies): an ensemble
of one or more domains that advances in place on the GPU.

There is no CPU fallback: loading fails with an error when the HIP library
is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .state import PROGNOSTIC, State

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmoist2d_hip.so")

FP64 = 0
MIXED = 1

NAN_FIELDS = ("rho", "rho_u", "rho_w", "rho_theta", "rho_qv", "rho_qc", "rho_qr", "d_rho_u_phys",
              "d_rho_theta_phys", "d_rho_qv_phys", "d_rho_qc_phys", "d_rho_qr_phys")

DIAG_NAMES = ("dry_mass", "total_water", "rain", "mass", "nan")

_lib = None


def lib():
    """Load libmoist2d_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library {LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip = ctypes.c_void_p, ctypes.c_int
        sig = {
            "moist_create": (ip, [vp, ip, ip, ctypes.POINTER(vp)]),
            "moist_upload": (ip, [vp, ip, vp, vp, vp, ip]),
            "moist_download": (ip, [vp, ip, vp, vp, ip]),
            "moist_download_phys": (ip, [vp, ip, vp, vp, ip]),
            "moist_rk3_step": (ip, [vp, ip]),
            "moist_compute_step": (ip, [vp, ip, vp, vp, ctypes.c_double, ip, ip]),
            "moist_diagnostics": (ip, [vp, ip, ctypes.POINTER(ctypes.c_double)]),
            "moist_diagnostics_noalloc": (ip, [vp, ip, ctypes.POINTER(ctypes.c_double)]),
            "moist_use_graph": (ip, [vp, ip]),
            "moist_set_option": (ip, [vp, ip, ip]),
            "moist_time_steps": (ip, [vp, ip, ctypes.POINTER(ctypes.c_double)]),
            "moist_kernel_events": (ip, [vp, ip]),
            "moist_kernel_stats": (ip, [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]),
            "moist_adjoint_run": (ip, [vp, ip, ip, vp, ip, ctypes.POINTER(ctypes.c_double)]),
            "moist_adjoint_grad": (ip, [vp, ip, vp, ctypes.POINTER(ctypes.c_double), ip]),
            "moist_adjoint_times": (ip, [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
            "moist_autotune": (ip, [vp, ip, ctypes.POINTER(ip), ctypes.POINTER(ip), ctypes.POINTER(ctypes.c_double)]),
            "moist_set_acoustic_shape": (ip, [vp, ip, ip]),
            "moist_history": (ip, [vp, ip, ctypes.POINTER(ctypes.c_double), vp, ip]),
            "moist_device_bytes": (ip, [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
            "moist_last_error": (ctypes.c_char_p, [vp]),
            "moist_destroy": (None, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptrs(arrs):
    return (ctypes.c_void_p * len(arrs))(*[None if a is None else a.ctypes.data for a in arrs])


def _raise_nan(code):
    raise RuntimeError(f"NaN detected in {NAN_FIELDS[code - 1]}")


class Engine:
    """Device-resident ensemble of domains (one moist_handle)."""

    def __init__(self, params, precision=FP64):
        if not isinstance(params, (list, tuple)):
            params = [params]
        self.params = list(params)
        self.precision = precision
        self._cp = [p.to_c() for p in self.params]
        arr = (ctypes.c_void_p * len(self._cp))(*[ctypes.addressof(c) for c in self._cp])
        h = ctypes.c_void_p()
        rc = lib().moist_create(arr, len(self._cp), precision, ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = lib().moist_last_error(h).decode() if h.value else ""
            if h.value:
                lib().moist_destroy(h)
            self._h = ctypes.c_void_p()
            raise RuntimeError(f"moist_create failed ({rc}): {msg}")

    def _check(self, rc, what):
        if rc < 0:
            raise RuntimeError(f"{what} failed ({rc}): {lib().moist_last_error(self._h).decode()}")
        if rc > 0:
            _raise_nan(rc)
        return rc

    def close(self):
        if self._h and self._h.value:
            lib().moist_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, s: State, member=0):
        eb = s.rho.dtype.itemsize
        fl = [np.asfortranarray(getattr(s, n)) for n in PROGNOSTIC]
        self._keep = fl
        self._check(lib().moist_upload(self._h, member, _ptrs(fl), s.prec.ctypes.data, s.theta_ref.ctypes.data, eb),
                    "moist_upload")

    def download(self, s: State, member=0, phys=True):
        eb = s.rho.dtype.itemsize
        fl = [getattr(s, n) for n in PROGNOSTIC]
        for a in fl:
            assert a.flags["F_CONTIGUOUS"]
        self._check(lib().moist_download(self._h, member, _ptrs(fl), s.prec.ctypes.data, eb), "moist_download")
        if phys:
            o2 = [s.d_rho_theta_mp, s.d_rho_qv_mp, s.d_rho_qc_mp, s.d_rho_qr_mp, s.d_rho_qr_sed, s.d_rho_theta_rd]
            o1 = [s.d_rho_u_sf, s.d_rho_theta_sf, s.d_rho_qv_sf, s.mflux, s.shflux, s.lhflux]
            self._check(lib().moist_download_phys(self._h, member, _ptrs(o2), _ptrs(o1), eb), "moist_download_phys")

    def step(self, nsteps=1):
        return self._check(lib().moist_rk3_step(self._h, nsteps), "moist_rk3_step")

    def compute_step(self, phys, s0: State, dt, ns, member=0):
        eb = s0.rho.dtype.itemsize
        ph = [np.asfortranarray(a) for a in phys]
        s0f = [np.asfortranarray(getattr(s0, n)) for n in PROGNOSTIC]
        return self._check(lib().moist_compute_step(self._h, member, _ptrs(ph), _ptrs(s0f), float(dt), int(ns), eb),
                           "moist_compute_step")

    def diagnostics(self, member=0) -> dict:
        out = (ctypes.c_double * 5)()
        self._check(lib().moist_diagnostics(self._h, member, out), "moist_diagnostics")
        return dict(zip(DIAG_NAMES, list(out)))

    def diagnostics_noalloc(self, member=0) -> dict:
        """diagnostics() through moist_diagnostics_noalloc (static device scratch,
        no allocation)."""
        out = (ctypes.c_double * 5)()
        self._check(lib().moist_diagnostics_noalloc(self._h, member, out), "moist_diagnostics_noalloc")
        return dict(zip(DIAG_NAMES, list(out)))

    def use_graph(self, on=True):
        self._check(lib().moist_use_graph(self._h, 1 if on else 0), "moist_use_graph")

    def store_phys_diag(self, on=True):
        self._check(lib().moist_set_option(self._h, 0, 1 if on else 0), "moist_set_option")

    def adjoint_records(self, mode):
        """Forward records of the reverse sweep: "auto", "step", "segment",
        "pipelined" (whole segments, one segment's recording pass beside the
        next one's reverse sweep) or "step_pipelined" (per-step records, one
        step's recording pass beside the reverse of the step after it)."""
        v = {"auto": 0, "step": 1, "segment": 2, "pipelined": 3, "step_pipelined": 4}[mode]
        self._check(lib().moist_set_option(self._h, 1, v), "moist_set_option")

    def adjoint_column_kernel(self, on=True):
        """Vertical part of each reverse acoustic substep as one kernel per column
        (default) or as three kernels (same arithmetic)."""
        self._check(lib().moist_set_option(self._h, 2, 1 if on else 0), "moist_set_option")

    def kernel_events(self, on=True):
        self._check(lib().moist_kernel_events(self._h, 1 if on else 0), "moist_kernel_events")

    def time_steps(self, nsteps):
        ms = ctypes.c_double()
        self._check(lib().moist_time_steps(self._h, nsteps, ctypes.byref(ms)), "moist_time_steps")
        return ms.value

    def adjoint(self, nsteps, theta_target, ckpt=50):
        """Loss and reverse sweep over `nsteps` large steps from the uploaded
        state.  The trajectory and the forward records are in the handle's form
        (FP64, or mixed with FP32 records); the adjoint arithmetic is FP64.
        theta_target: one (ka, ia) array per member.
        Returns J per member; gradients come from `adjoint_grad`."""
        tt = [np.asfortranarray(t, dtype=np.float64) for t in theta_target]
        self._keep_tt = tt
        J = (ctypes.c_double * len(self.params))()
        self._check(lib().moist_adjoint_run(self._h, nsteps, ckpt, _ptrs(tt), 8, J), "moist_adjoint_run")
        return list(J)

    def adjoint_grad(self, member=0):
        """(dJ/d initial prognostic fields (7 arrays (ka, ia), interior columns),
        dJ/d (K_auto, Qc_threshold, K_accr))."""
        p = self.params[member]
        g = [np.zeros((p.Nz, p.Nx + 2 * p.halo), order="F") for _ in range(7)]
        gk = (ctypes.c_double * 3)()
        self._check(lib().moist_adjoint_grad(self._h, member, _ptrs(g), gk, 8), "moist_adjoint_grad")
        return g, np.array(list(gk))

    def adjoint_times(self):
        """(forward ms, reverse ms) of the last adjoint run."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(lib().moist_adjoint_times(self._h, ctypes.byref(a), ctypes.byref(b)), "moist_adjoint_times")
        return a.value, b.value

    def autotune(self, reps=20):
        """Time every acoustic-kernel shape on this handle's domains and keep the
        fastest; returns (levels per chunk, threads per workgroup, us per launch)."""
        L, NT, us = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._check(lib().moist_autotune(self._h, reps, ctypes.byref(L), ctypes.byref(NT), ctypes.byref(us)),
                    "moist_autotune")
        return L.value, NT.value, us.value

    def set_acoustic_shape(self, levels, threads):
        self._check(lib().moist_set_acoustic_shape(self._h, levels, threads), "moist_set_acoustic_shape")

    HIST_FIELDS = ("rho", "rhou", "rhow", "u", "w", "theta", "qv", "qc", "qr", "RH", "temp", "pres", "thetae")
    HIST_STATS = ("Etot", "wmin", "wmax", "qcmax", "qrmax", "RHmin", "RHmax")

    def history(self, member=0, fields=True):
        """output_history's derived fields ((Nz, Nx) arrays, x fastest, dict keyed
        by the history variable names) and stats (dict), computed on the device."""
        p = self.params[member]
        st = (ctypes.c_double * 7)()
        out = {n: np.zeros((p.Nz, p.Nx)) for n in self.HIST_FIELDS} if fields else None
        ptrs = _ptrs([out[n] for n in self.HIST_FIELDS]) if fields else None
        self._check(lib().moist_history(self._h, member, st, ptrs, 8), "moist_history")
        return out, dict(zip(self.HIST_STATS, list(st)))

    def device_bytes(self):
        """(state bytes, adjoint bytes) of device memory held by the handle."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(lib().moist_device_bytes(self._h, ctypes.byref(a), ctypes.byref(b)), "moist_device_bytes")
        return a.value, b.value

    def kernel_stats(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(lib().moist_kernel_stats(self._h, ctypes.byref(ms), ctypes.byref(n)), "moist_kernel_stats")
        return ms.value, n.value


def _precision_for(s: State, precision):
    if precision is not None:
        return precision
    return FP64 if s.rho.dtype == np.float64 else MIXED


# Handles of the host-State entry points, kept between calls: the reference
# calls rk3_step! once per step (moist.jl:1943-1944), so each call re-uploads
# the State into a cached handle instead of creating one.  Keyed on the
# Params values, the z grid and the precision; a few are kept, least
# recently used first out.
_CACHE: dict = {}
_CACHE_MAX = 4


def _params_key(p, precision):
    vals = tuple(sorted((k, repr(v)) for k, v in p.values.items()))
    grid = tuple(np.asarray(a, dtype=np.float64).tobytes() for a in (p.z, p.zf, p.dz))
    return (np.dtype(p.T).str, vals, grid, precision)


def _cached_engine(p, precision):
    key = _params_key(p, precision)
    eng = _CACHE.pop(key, None)
    if eng is None:
        while len(_CACHE) >= _CACHE_MAX:
            _CACHE.pop(next(iter(_CACHE))).close()
        eng = Engine(p, precision)
    _CACHE[key] = eng  # most recently used last
    return eng


def release_cached() -> None:
    """Free the device memory of the handles kept by rk3_step / compute_step."""
    while _CACHE:
        _CACHE.popitem()[1].close()


def rk3_step(s: State, p, precision=None, nsteps=1):
    """rk3_step!(s, p) (moist.jl:1589-1683) on the GPU, in place on host State `s`.

    Float64 states run the FP64 form, Float32 states the mixed form, unless
    `precision` says otherwise.  `nsteps` > 1 advances several large steps
    with one upload and one download.  The device handle is created on the
    first call for these Params and reused by later calls.
    """
    eng = _cached_engine(p, _precision_for(s, precision))
    eng.upload(s)
    rc = eng.step(nsteps)
    eng.download(s)
    return rc


def compute_step(d_rho_phys, d_rho_u_phys, d_rho_theta_phys, d_rho_qv_phys, d_rho_qc_phys, d_rho_qr_phys, s: State,
                 s0: State, p, dt, ns, precision=None):
    """compute_step!(...) (moist.jl:1039-1580): one RK stage, in place on `s`
    (the device handle is cached as for rk3_step)."""
    eng = _cached_engine(p, _precision_for(s, precision))
    eng.upload(s)
    eng.compute_step((d_rho_phys, d_rho_u_phys, d_rho_theta_phys, d_rho_qv_phys, d_rho_qc_phys, d_rho_qr_phys),
                     s0, dt, ns)
    eng.download(s, phys=False)
