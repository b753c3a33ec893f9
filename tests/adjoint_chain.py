"""The loss gradient along a given trajectory, as the chain of the oracle's
one-step adjoints (rk3_step_adj): lambda_N from the loss at the final state,
then lambda_n = (d step / d x_n)^T lambda_n+1 for n = N-1 .. 0, with the
Kessler-parameter adjoints summed over the steps.  Along the oracle's own
trajectory this is the oracle's adjoint_run; along a trajectory taken from
the device it checks a reverse sweep that linearises about those states.
"""
import numpy as np

import oracle as O

PROG7 = ("rho", "rho_u", "rho_w", "rho_theta", "rho_qv", "rho_qc", "rho_qr")


def chain_adjoint(p, states, tt):
    """states: x_0 .. x_N (Float64 States with halos); tt: theta target (ka, ia).
    Returns (J, grad[7], gK[3])."""
    h = p.halo
    xN = states[-1]
    inner = np.s_[:, h:h + p.Nx]
    th = xN.rho_theta / xN.rho
    d = np.zeros(xN.rho.shape)
    d[inner] = (th - tt)[inner]
    J = p.dx * float(np.sum(xN.prec[h:h + p.Nx])) + 0.5 * float(np.sum(d[inner] ** 2))
    lam = [np.zeros(xN.rho.shape, order="F") for _ in range(7)]
    lam[0][...] = -d * th / xN.rho
    lam[3][...] = d / xN.rho
    gK = np.zeros(3)
    stale = np.zeros(2)  # adjoints of the surface fluxes' halo reads, carried one step back
    for x in reversed(states[:-1]):
        gK += O.rk3_step_adj(x.copy(), p, lam, lam_dprec=p.dx * p.dt, stale=stale)  # prec += d_prec * dt
    if p.do_sf:
        # the first step's halo reads are periodic copies of the initial state
        lam[1][0, p.ie - 1] += stale[0]
        lam[0][0, p.is_ - 1] += stale[1]
    return J, lam, gK
