"""State{A,B,C} and allocate_state of the reference, mirrored on the host.

Field names, order and shapes follow moist.jl:294-345: 2-D fields are
(ka, ia) arrays in Julia's column-major order (numpy order='F', z fastest),
1-D x fields have length ia, `theta_ref` has length ka.  Indexing is 0-based
here: `s.rho[k, i]` is the reference's `s.rho[k+1, i+1]`.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass

import numpy as np

PROGNOSTIC = ("rho", "rho_u", "rho_w", "rho_theta", "rho_qv", "rho_qc", "rho_qr")
FIELDS_2D = PROGNOSTIC + ("d_rho_theta_mp", "d_rho_qv_mp", "d_rho_qc_mp", "d_rho_qr_mp", "d_rho_qr_sed",
                          "d_rho_theta_rd")
FIELDS_X = ("prec", "mflux", "shflux", "lhflux", "d_rho_u_sf", "d_rho_theta_sf", "d_rho_qv_sf")
# Field order of the reference's State constructor (moist.jl:295-316)
STATE_ORDER = ("rho", "rho_u", "rho_w", "rho_theta", "rho_qv", "rho_qc", "rho_qr", "prec", "mflux", "shflux",
               "lhflux", "theta_ref", "d_rho_theta_mp", "d_rho_qv_mp", "d_rho_qc_mp", "d_rho_qr_mp",
               "d_rho_qr_sed", "d_rho_u_sf", "d_rho_theta_sf", "d_rho_qv_sf", "d_rho_theta_rd")


@dataclass
class State:
    rho: np.ndarray
    rho_u: np.ndarray
    rho_w: np.ndarray
    rho_theta: np.ndarray
    rho_qv: np.ndarray
    rho_qc: np.ndarray
    rho_qr: np.ndarray
    prec: np.ndarray
    mflux: np.ndarray
    shflux: np.ndarray
    lhflux: np.ndarray
    theta_ref: np.ndarray
    d_rho_theta_mp: np.ndarray
    d_rho_qv_mp: np.ndarray
    d_rho_qc_mp: np.ndarray
    d_rho_qr_mp: np.ndarray
    d_rho_qr_sed: np.ndarray
    d_rho_u_sf: np.ndarray
    d_rho_theta_sf: np.ndarray
    d_rho_qv_sf: np.ndarray
    d_rho_theta_rd: np.ndarray

    @property
    def dtype(self):
        return self.rho.dtype

    def copy(self) -> "State":
        """deepcopy(s) (moist.jl:1647)."""
        return copy.deepcopy(self)

    def arrays(self):
        return [getattr(self, n) for n in STATE_ORDER]


def allocate_state(p) -> State:
    """allocate_state(p) (moist.jl:319-345): zero-filled arrays of type T."""
    T = p.T
    ia, ka = p.ia, p.ka
    kw = {}
    for n in STATE_ORDER:
        if n in FIELDS_2D:
            kw[n] = np.zeros((ka, ia), dtype=T, order="F")
        elif n == "theta_ref":
            kw[n] = np.zeros(ka, dtype=T)
        else:
            kw[n] = np.zeros(ia, dtype=T)
    return State(**kw)


def exchange_halo(A: np.ndarray, p) -> None:
    """exchange_halo!(A, p) (moist.jl:233-241): periodic x halo copy."""
    h, is0, ie0 = p.halo, p.is_ - 1, p.ie - 1
    A[:, 0:h] = A[:, ie0 - h + 1:ie0 + 1]
    A[:, ie0 + 1:ie0 + 1 + h] = A[:, is0:is0 + h]
