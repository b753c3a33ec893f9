"""Initial conditions: the reference's init_state! and the benchmark workloads.

The reference's IC code (moist.jl:351-450) draws its noise from per-thread
Xoshiro generators (`rngs[Threads.threadid()]`, moist.jl:370-375), so its
noise depends on the thread schedule.  Here every random number is a pure
function of (seed, member, stream, i, k): a SplitMix64 hash chain, uniform
doubles from the top 53 bits and Box-Muller for Gaussians.  The same inputs
give the same state on any machine and at any thread count.

Workloads (all in Float64, periodic x, rigid lid, z_fact = 1, dz0 = dz):
  bubble_128x64     128 x 64, dx = dz = 156.25 m, dt = 1 s, ns = 6
  coldpool_8192x256 8192 x 256, dx = dz = 100 m, dt = 0.5 s, ns = 8
  ensemble_256      256 domains, nx in {256, 320, ..., 2048}, nz in {64, 128},
                    dx = dz = 125 m, dt = 1 s, ns = 6
  adjoint_512x128   512 x 128, dx = dz = 78.125 m, dt = 0.5 s, ns = 6
  ensemble_adj_64   64 domains of 1024 x 128, dx = dz = 100 m, dt = 1 s, ns = 6
The base state is theta(z) = 300 exp(N^2 z / g), N = 0.01 1/s, p_sfc = 1000 hPa,
in the discrete hydrostatic balance of the model's vertical momentum equation
(p_k - p_{k-1} + g (z_k - z_{k-1}) (rho_{k-1} + rho_k) / 2 = 0, the balance of
moist.jl:1271-1274), with qv = RH * qsat(T, p) solved jointly with rho.
Perturbations of theta are applied at fixed rho*theta (pressure unchanged).
"""
from __future__ import annotations

import math

import numpy as np

from .params import Constants, make_params
from .state import PROGNOSTIC, State, allocate_state, exchange_halo

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def rng_uniform(seed, member, stream, i, k):
    """Uniform doubles in [0, 1) as a pure function of the five integers."""
    h = _splitmix64(np.uint64(seed))
    h = _splitmix64(h ^ np.asarray(member, dtype=np.uint64))
    h = _splitmix64(h ^ np.asarray(stream, dtype=np.uint64))
    h = _splitmix64(h ^ np.asarray(i, dtype=np.uint64))
    h = _splitmix64(h ^ np.asarray(k, dtype=np.uint64))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def rng_normal(seed, member, stream, i, k):
    """Standard normal via Box-Muller from streams `stream` and `stream+1`."""
    u1 = rng_uniform(seed, member, stream, i, k)
    u2 = rng_uniform(seed, member, stream + 1, i, k)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


# ---------------------------------------------------------------------------
# thermodynamics in float64 (same formulas as moist.jl:254-273)
# ---------------------------------------------------------------------------
_C = Constants(np.float64)


def _qsat(Tk, p):
    Tc = Tk - _C.T0
    es = _C.ES0 * np.exp(_C.A_SAT * Tc / (_C.B_SAT + Tc))
    tmp = p - (1.0 - _C.EPSvap) * es
    return np.where(tmp > _C.EPS, _C.EPSvap * es / np.where(tmp > _C.EPS, tmp, 1.0), 0.0)


def _eos_rho(p, theta, qv):
    """Density from pressure, theta and vapour (qc = qr = 0), inverting
    pres = P0 (r rho theta / P0)^(cp/cv) (moist.jl:1058)."""
    qd = 1.0 - qv
    r = _C.Rd * qd + _C.Rv * qv
    cp = _C.Cpd * qd + _C.Cpv * qv
    cv = cp - r
    rho_theta = _C.P0 / r * (p / _C.P0) ** (cv / cp)
    return rho_theta / theta


def _temperature(p, theta, qv):
    qd = 1.0 - qv
    r = _C.Rd * qd + _C.Rv * qv
    cp = _C.Cpd * qd + _C.Cpv * qv
    return theta * (p / _C.P0) ** (r / cp)


def hydrostatic_moist_columns(z, theta, rh, p_sfc=1e5, iters=12):
    """Discretely balanced columns: theta (nz, ncol) and rh (nz, ncol) given,
    returns (rho, qv, p), each (nz, ncol), float64.  Level 1 is tied to the
    surface pressure by p_1 = p_sfc - g z_1 rho_1."""
    nz = theta.shape[0]
    rho = np.empty_like(theta)
    qv = np.empty_like(theta)
    p = np.empty_like(theta)
    g = float(_C.GRAV)
    for k in range(nz):
        th = theta[k]
        if k == 0:
            pk = np.full_like(th, p_sfc)
            q = np.zeros_like(th)
        else:
            pk = p[k - 1].copy()
            q = qv[k - 1].copy()
        for _ in range(iters):
            Tk = _temperature(pk, th, q)
            q = rh[k] * _qsat(Tk, pk)
            rk = _eos_rho(pk, th, q)
            if k == 0:
                pk = p_sfc - g * z[0] * rk
            else:
                pk = p[k - 1] - g * (z[k] - z[k - 1]) * 0.5 * (rho[k - 1] + rk)
        Tk = _temperature(pk, th, q)
        q = rh[k] * _qsat(Tk, pk)
        rho[k] = _eos_rho(pk, th, q)
        qv[k] = q
        p[k] = pk
    return rho, qv, p


def _rh_profile(z, rh_low, z_low, rh_top=0.20, z_top=8000.0):
    z = np.asarray(z, dtype=np.float64)
    lin = rh_low + (rh_top - rh_low) * (z - z_low) / (z_top - z_low)
    return np.where(z <= z_low, rh_low, np.where(z >= z_top, rh_top, lin))


def _cos2_bump(x, z, xc, zc, xr, zr, amp):
    r = np.sqrt(((x - xc) / xr) ** 2 + ((z - zc) / zr) ** 2)
    c = np.cos(np.pi * r * 0.5)
    return np.where(r < 1.0, amp * c * c, 0.0)


def _fill_state(p, theta_base, theta_pert, rh, qv_pert=None) -> State:
    """Build a State from the environment theta_base(z), the perturbation
    theta_pert(k, i) (interior) and rh(z): hydrostatic environment, then theta
    perturbed at fixed rho*theta."""
    T = p.T
    nz, nx = p.Nz, p.Nx
    z = p.z.astype(np.float64)
    rho_e, qv_e, _ = hydrostatic_moist_columns(z, theta_base[:, None], rh[:, None])
    rho_e, qv_e = rho_e[:, 0], qv_e[:, 0]
    rho_theta = (rho_e * theta_base)[:, None] * np.ones((1, nx))
    theta = theta_base[:, None] + theta_pert
    rho = rho_theta / theta
    qv = np.broadcast_to(qv_e[:, None], (nz, nx)).copy()
    if qv_pert is not None:
        qv = np.maximum(qv + qv_pert, 0.0)
    s = allocate_state(p)
    is0 = p.is_ - 1
    sl = np.s_[:, is0:is0 + nx]
    s.rho[sl] = rho
    s.rho_theta[sl] = rho_theta
    s.rho_qv[sl] = rho * qv
    s.theta_ref[:] = theta_base.astype(T)
    for n in PROGNOSTIC:
        exchange_halo(getattr(s, n), p)
    return s


def _grid(p):
    is0 = p.is_ - 1
    x = p.x[is0:is0 + p.Nx].astype(np.float64)
    z = p.z.astype(np.float64)
    X, Z = np.meshgrid(x, z)  # (nz, nx)
    I, K = np.meshgrid(np.arange(p.Nx, dtype=np.uint64), np.arange(p.Nz, dtype=np.uint64))
    return X, Z, I, K


def theta_profile(z, theta0=300.0, N=0.01):
    return theta0 * np.exp(N * N * np.asarray(z, dtype=np.float64) / float(_C.GRAV))


def _workload_params(T, nx, nz, d, dt, ns, steps, **kw):
    kw = {"do_sf": False, "do_rd": False, **kw}
    return make_params(T, Nx=nx, Nz=nz, Lx=nx * d, H=nz * d, dz0=d, z_fact=1.0, ns=ns, tout_int=dt,
                       t_end=steps * dt, **kw)


def bubble(T=np.float64, nx=128, nz=64, d=156.25, dt=1.0, ns=6, steps=600, amp=2.0, radius=2000.0,
           xc=None, zc=2000.0, noise=0.01, seed=0, member=0, rh_low=0.90, z_low=2000.0, rh_top=0.20, **kw):
    """Moist thermal bubble (the 128x64 workload by default)."""
    p = _workload_params(T, nx, nz, d, dt, ns, steps, **kw)
    X, Z, I, K = _grid(p)
    xc = 0.5 * nx * d if xc is None else xc
    th_p = _cos2_bump(X, Z, xc, zc, radius, radius, amp)
    th_p = th_p + noise * (2.0 * rng_uniform(seed, member, 0, I, K) - 1.0)
    z = p.z.astype(np.float64)
    s = _fill_state(p, theta_profile(z), th_p, _rh_profile(z, rh_low, z_low, rh_top=rh_top))
    return p, s


def coldpool(T=np.float64, nx=8192, nz=256, d=100.0, dt=0.5, ns=8, steps=1200, seed=1, **kw):
    """Cold pool in a moist boundary layer (the 8192x256 workload)."""
    p = _workload_params(T, nx, nz, d, dt, ns, steps, **kw)
    X, Z, I, K = _grid(p)
    Lx = nx * d
    th_p = _cos2_bump(X, Z, 0.5 * Lx, 0.0, 5000.0, 2000.0, -4.0)
    low = Z < 1000.0
    th_p = th_p + np.where(low, 0.1 * rng_normal(seed, 0, 0, I, K), 0.0)
    qv_p = np.where(low, 1e-4 * rng_normal(seed, 0, 2, I, K), 0.0)
    z = p.z.astype(np.float64)
    s = _fill_state(p, theta_profile(z), th_p, _rh_profile(z, 0.95, 1500.0), qv_pert=qv_p)
    return p, s


def ensemble(T=np.float64, members=256, seed=2, d=125.0, dt=1.0, ns=6, steps=600, nz_choices=(64, 128),
             nx_choices=None, **kw):
    """Ensemble of bubbles on domains of uneven size (the 256-member workload)."""
    if nx_choices is None:
        nx_choices = list(range(256, 2048 + 1, 64))
    out = []
    for m in range(members):
        u = [float(rng_uniform(seed, m, 1000 + j, 0, 0)) for j in range(7)]
        nx = nx_choices[min(int(u[0] * len(nx_choices)), len(nx_choices) - 1)]
        nz = nz_choices[min(int(u[1] * len(nz_choices)), len(nz_choices) - 1)]
        Lx = nx * d
        out.append(bubble(T, nx=nx, nz=nz, d=d, dt=dt, ns=ns, steps=steps, amp=1.0 + 2.0 * u[2],
                          radius=1500.0 + 1500.0 * u[3], xc=(0.3 + 0.4 * u[4]) * Lx, zc=1500.0 + 1000.0 * u[5],
                          seed=seed, member=m, **kw))
    return out


def adjoint_bubble(T=np.float64, nx=512, nz=128, d=78.125, dt=0.5, ns=6, steps=600, seed=3, **kw):
    """Bubble of the 512x128 adjoint workload."""
    return bubble(T, nx=nx, nz=nz, d=d, dt=dt, ns=ns, steps=steps, seed=seed, **kw)


def batched_adjoint(T=np.float64, members=64, seed=4, steps=300):
    """Members of the 64 x 1024x128 batched adjoint workload: bubbles at
    dx = dz = 100 m, dt = 1 s, ns = 6, with seeded amplitude U[1, 3] K,
    centre x U[0.3, 0.7] Lx, z U[1.5, 2.5] km and K_auto ~ U[5e-4, 2e-3];
    returns [(params, state, theta_target)]."""
    out = []
    for m in range(members):
        u = [float(rng_uniform(seed, m, 2000 + j, 0, 0)) for j in range(5)]
        p, s = bubble(T, nx=1024, nz=128, d=100.0, dt=1.0, ns=6, steps=steps, amp=1.0 + 2.0 * u[0],
                      xc=(0.3 + 0.4 * u[1]) * 1024 * 100.0, zc=1500.0 + 1000.0 * u[2], seed=seed, member=m,
                      K_auto=5e-4 + 1.5e-3 * u[3])
        out.append((p, s, theta_target(p, seed=seed, member=m)))
    return out


def theta_target(p, seed=3, sigma=0.5, member=0):
    """Target theta of the adjoint workload's loss: the base profile plus a
    seeded Gaussian field of standard deviation `sigma` (K), in the (ka, ia)
    layout (halo columns are periodic copies)."""
    X, Z, I, K = _grid(p)
    th = theta_profile(Z) + sigma * rng_normal(seed, member, 7, I, K)
    out = np.zeros((p.Nz, p.Nx + 2 * p.halo), order="F")
    h = p.halo
    out[:, h:h + p.Nx] = th
    out[:, :h] = th[:, p.Nx - h:]
    out[:, h + p.Nx:] = th[:, :h]
    return out


def invert_T_for_thetae(qt, theta_e, p_k):
    """Temperature (and saturation qv) whose saturated equivalent potential
    temperature at pressure p_k is theta_e: 30 bisection steps on [250, 350] K
    (invert_T_for_thetae, moist.jl:275-289).  Vectorised over p_k."""
    p_k = np.asarray(p_k, dtype=np.float64)
    lo = np.full(p_k.shape, 250.0)
    hi = np.full(p_k.shape, 350.0)
    qsat = np.zeros(p_k.shape)
    qd = 1.0 - qt
    cp = _C.Cpd * qd + _C.Cpl * qt
    for _ in range(30):
        t = (lo + hi) / 2
        Tc = t - _C.T0
        es = _C.ES0 * np.exp(_C.A_SAT * Tc / (_C.B_SAT + Tc))
        qsat = _qsat(t, p_k)
        th = t * ((p_k - es) / _C.P0) ** (-_C.Rd / cp) * np.exp(_C.Lv * qsat / (cp * t))
        above = th > theta_e
        hi = np.where(above, t, hi)
        lo = np.where(above, lo, t)
    return (lo + hi) / 2, qsat


def init_state_bf2002(s: State, p, qt=0.02, theta_e=320.0) -> State:
    """The saturated neutral atmosphere of Bryan and Fritsch (2002) with a
    cos^2 warm bubble: init_state_BF2002!(s, p) (moist.jl:453-551), in Float64.

    qv + qc = qt everywhere and the saturated equivalent potential temperature
    is theta_e; the column pressure is iterated to the model's discrete
    hydrostatic balance (moist.jl:489-514, relaxation 0.7, tolerance 1e-10).
    The reference's first-level update reads an index `k` that is not defined
    at that point (moist.jl:494, the function is commented out of run_sim);
    here that layer is the trapezoid from the surface to z[ks],
    p_ks = P0 - g (rho_s + rho_ks) z[ks] / 2, which is what its factor
    0.25 * (z[k] - z[k-1]) equals on a uniform grid."""
    T = p.T
    c = _C
    is0, nx, nz = p.is_ - 1, p.Nx, p.Nz
    z = p.z.astype(np.float64)
    qd = 1.0 - qt
    P0, G = c.P0, c.GRAV
    # first guess (moist.jl:465-484)
    t_s, qv = invert_T_for_thetae(qt, theta_e, np.array(P0))
    theta = float(t_s)
    qv = float(qv)
    r = c.Rd * qd + c.Rv * qv
    cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qt - qv)
    Pi = np.zeros(nz)
    pc = np.zeros(nz)
    Pi[0] = 1.0 - G * z[0] / (cp * theta)
    pc[0] = P0 * Pi[0] ** (cp / r)
    t_k, qv = invert_T_for_thetae(qt, theta_e, pc[0])
    theta = float(t_k) / Pi[0]
    qv = float(qv)
    for k in range(1, nz):
        r = c.Rd * qd + c.Rv * qv
        cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qt - qv)
        Pi[k] = Pi[k - 1] - G * (z[k] - z[k - 1]) / (cp * theta)
        pc[k] = P0 * Pi[k] ** (cp / r)
        t_k, qv = invert_T_for_thetae(qt, theta_e, pc[k])
        theta = float(t_k) / Pi[k]
        qv = float(qv)
    # hydrostatic iteration (moist.jl:487-514)
    omega = 0.7
    t_s = float(t_s)
    for _ in range(100):
        max_diff = 0.0
        t_k, qv = invert_T_for_thetae(qt, theta_e, pc[0])
        r = c.Rd * qd + c.Rv * float(qv)
        rho_s = P0 / (r * t_s)
        rho_k = pc[0] / (r * float(t_k))
        p_new = P0 - G * (rho_s + rho_k) * z[0] * 0.5
        d = omega * (p_new - pc[0])
        pc[0] += d
        max_diff = max(max_diff, abs(d / p_new))
        for k in range(1, nz):
            t_m, qv_m = invert_T_for_thetae(qt, theta_e, pc[k - 1])
            rho_m = pc[k - 1] / ((c.Rd * qd + c.Rv * float(qv_m)) * float(t_m))
            t_k, qv = invert_T_for_thetae(qt, theta_e, pc[k])
            rho_k = pc[k] / ((c.Rd * qd + c.Rv * float(qv)) * float(t_k))
            p_new = pc[k - 1] - G * (rho_m + rho_k) * 0.5 * (z[k] - z[k - 1])
            d = omega * (p_new - pc[k])
            pc[k] += d
            max_diff = max(max_diff, abs(d / p_new))
        if max_diff < 1e-10:
            break
    # fields (moist.jl:518-548)
    t_k, qv = invert_T_for_thetae(qt, theta_e, pc)
    r = c.Rd * qd + c.Rv * qv
    cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qt - qv)
    theta = t_k * (P0 / pc) ** (r / cp)
    x = p.x[is0:is0 + nx].astype(np.float64)
    xc, zc, xr, zr = float(p.Lx) / 2, 2e3, 2e3, 2e3
    X, Z = np.meshgrid(x, z)
    dist = np.sqrt(((X - xc) / xr) ** 2 + ((Z - zc) / zr) ** 2)
    tdiff = float(p.theta_pert) * np.cos(np.pi * dist * 0.5) ** 2 / 300.0
    th2 = theta[:, None] * np.where(dist < 1.0, tdiff + 1.0, 1.0)
    qs = _qsat(t_k, pc)
    qv2 = np.minimum(qt, qs)
    qc2 = qt - qv2
    r2 = c.Rd * qd + c.Rv * qv2
    rho = (pc / (r2 * t_k))[:, None] * np.ones((1, nx))
    sl = np.s_[:, is0:is0 + nx]
    s.rho[sl] = rho.astype(T)
    s.rho_u[sl] = (rho * float(p.u0)).astype(T)
    s.rho_w[...] = 0
    s.rho_theta[sl] = (rho * th2).astype(T)
    s.rho_qv[sl] = (rho * qv2[:, None]).astype(T)
    s.rho_qc[sl] = (rho * qc2[:, None]).astype(T)
    s.rho_qr[...] = 0
    for n in PROGNOSTIC:
        exchange_halo(getattr(s, n), p)
    s._bf2002_pressure = pc
    return s


def init_state(s: State, p, seed=0) -> State:
    """The reference's init_state!(s, p, rngs) (moist.jl:351-450) with the
    portable generator in place of the per-thread Xoshiro streams.  Float64
    arithmetic for the column Newton iteration, results stored in T."""
    T = p.T
    c = Constants(np.float64)
    is0, nx, nz = p.is_ - 1, p.Nx, p.Nz
    z = p.z.astype(np.float64)
    th_ref = np.where(z < p.bndlayer_depth, float(p.theta0),
                      np.where(z < p.z_tropopose, float(p.theta0) + float(p.dtheta_dz_trop) * (z - float(p.bndlayer_depth)),
                               float(p.theta0) + float(p.dtheta_dz_trop) * (float(p.z_tropopose) - float(p.bndlayer_depth))
                               + float(p.dtheta_dz_stra) * (z - float(p.z_tropopose))))
    s.theta_ref[:] = th_ref.astype(T)
    I, K = np.meshgrid(np.arange(nx, dtype=np.uint64), np.arange(nz, dtype=np.uint64))
    theta_init = th_ref.astype(T)[:, None].astype(np.float64) + float(p.theta_pert) * (
        rng_uniform(seed, 0, 0, I, K) * 2.0 - 1.0)
    RH_surf = float(p.rel_hum_surf) / 100.0
    rho = np.zeros((nz, nx))
    qv = np.zeros((nz, nx))
    Pi = np.ones(nx)
    p_km = None
    for k in range(nz):
        if k == 0:
            th = (float(p.theta0) + theta_init[0]) * 0.5
            qv[k] = 0.0
            zm = 0.0
        else:
            th = (theta_init[k - 1] + theta_init[k]) * 0.5
            qv[k] = qv[k - 1]
            zm = z[k - 1]
        qd = 1.0 - qv[k]
        r = c.Rd * qd + c.Rv * qv[k]
        cp = c.Cpd * qd + c.Cpv * qv[k]
        Pi = Pi - c.GRAV * (z[k] - zm) / (cp * th)
        pk = c.P0 * Pi ** (cp / r)
        Tk = th_ref[k] * Pi
        rho[k] = pk / (r * Tk)
        for _ in range(10):
            qs = _qsat(Tk, pk)
            qv_new = np.maximum(0.0, RH_surf * math.exp(-z[k] / float(p.rel_hum_decay)) * qs)
            qv[k] += 0.8 * (qv_new - qv[k])
            qd = 1.0 - qv[k]
            r = c.Rd * qd + c.Rv * qv[k]
            cp = c.Cpd * qd + c.Cpv * qv[k]
            cv = cp - r
            Pi = (rho[k] * r * theta_init[k] / c.P0) ** (r / cv)
            pk = c.P0 * Pi ** (cp / r)
            Tk = th_ref[k] * Pi
            if k > 0:
                l_ = pk - p_km + c.GRAV * (z[k] - z[k - 1]) * 0.5 * (rho[k - 1] + rho[k])
                dldr = cp / cv * pk / rho[k] + c.GRAV * (z[k] - z[k - 1]) * 0.5
                rho[k] -= l_ / dldr
        p_km = pk
    sl = np.s_[:, is0:is0 + nx]
    s.rho_w[...] = 0
    s.rho_qc[...] = 0
    s.rho_qr[...] = 0
    s.rho[sl] = rho.astype(T)
    s.rho_u[sl] = (s.rho[sl] * T(p.u0)).astype(T)
    s.rho_theta[sl] = (s.rho[sl].astype(np.float64) * theta_init).astype(T)
    s.rho_qv[sl] = (s.rho[sl].astype(np.float64) * qv).astype(T)
    for n in PROGNOSTIC:
        exchange_halo(getattr(s, n), p)
    return s
