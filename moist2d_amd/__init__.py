"""moist2d_amd — MI355X-native time-stepping core of the 2-D moist C-grid model.

Host-side mirror of the reference's Julia API (Params, State, allocate_state,
rk3_step!, compute_step!) over the C ABI of libmoist2d_hip.so (include/moist2d.h).
The device path is the only compute path: `moist2d_amd.core` raises if the
HIP library is missing.
"""
from .params import Constants, Params, make_params, moist_params  # noqa: F401
from .state import State, allocate_state, exchange_halo  # noqa: F401

__all__ = ["Constants", "Params", "make_params", "moist_params", "State", "allocate_state", "exchange_halo"]
