import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


PROG = ("rho", "rho_u", "rho_w", "rho_theta", "rho_qv", "rho_qc", "rho_qr")


def interior(p, a):
    is0 = p.is_ - 1
    return a[:, is0:is0 + p.Nx]


def rel_l2(a, b):
    """||a - b|| / ||b||; 0 when both are zero."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    nd = np.linalg.norm(a - b)
    if nb == 0.0:
        return 0.0 if nd == 0.0 else np.inf
    return nd / nb


def field_errors(p, s, ref):
    return {n: rel_l2(interior(p, getattr(s, n)), interior(p, getattr(ref, n))) for n in PROG}
