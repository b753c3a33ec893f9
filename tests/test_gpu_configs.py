"""GPU parity at the workloads' own sizes, against the oracle.

* 8192x256 cold pool (dx = dz = 100 m, dt = 0.5 s, ns = 8): FP64 form after
  10 large steps, per-field rel-L2 <= 1e-11 (the drop-in bar); the mixed
  form against the same oracle run, with its error stated and bounded.
* 512x128 adjoint bubble (dx = dz = 78.125 m, dt = 0.5 s, ns = 6): the
  checkpointed loss gradient over 4 steps (checkpoints every 2) in the FP64
  form against the oracle's adjoint_run, and in the mixed form against the
  chain of the oracle's one-step adjoints along the device's own mixed
  trajectory.
* 64 x 1024x128 batched adjoint: the whole 64-member handle, 3 steps;
  members 0 and 63 against the oracle as above, in both forms.

Reference: rk3_step! moist.jl:1589-1683; the adjoint has no counterpart
in the reference (its loss and controls are those of the adjoint workloads).
"""
import numpy as np
import pytest

import oracle as O
from adjoint_chain import chain_adjoint
from conftest import field_errors
from moist2d_amd import core, ic

pytestmark = pytest.mark.gpu

PROG7 = ("rho", "rho_u", "rho_w", "rho_theta", "rho_qv", "rho_qc", "rho_qr")


def _print(tag, errs):
    print(tag, " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)


def _rel(a, b):
    n = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / n if n > 0 else d


@pytest.fixture(scope="module")
def coldpool_ref():
    print("coldpool 8192x256: IC and 10 oracle steps", flush=True)
    p, s = ic.coldpool(np.float64)
    ref = s.copy()
    assert O.rk3_step(ref, p, 10) == 0
    return p, s, ref


def _gpu_steps(p, s, steps, precision, graph=True):
    eng = core.Engine(p, precision)
    eng.use_graph(graph)
    eng.upload(s)
    eng.step(steps)
    out = s.copy()
    eng.download(out)
    eng.close()
    return out


def test_coldpool_8192x256_fp64_10_steps(coldpool_ref):
    p, s, ref = coldpool_ref
    out = _gpu_steps(p, s, 10, core.FP64)
    errs = field_errors(p, out, ref)
    _print("coldpool 8192x256 fp64, 10 steps vs oracle", errs)
    for n, e in errs.items():
        assert e <= 1e-11, (n, e)
    h = p.halo
    np.testing.assert_allclose(out.prec[h:h + p.Nx], ref.prec[h:h + p.Nx], rtol=1e-11, atol=1e-300)


def test_coldpool_8192x256_mixed_10_steps(coldpool_ref):
    """Stated error of the mixed form at the workload's size after 10 steps
    (5 s).  Measured on MI355X: rho, rho_theta ~1e-9; rho_qv ~1e-7; rho_u,
    rho_w ~1e-5; rho_qc, rho_qr (small fields early in the run) ~1e-4.
    Bars: 1e-6 (rho, rho_theta), 1e-5 (rho_qv), 1e-4 (rho_u, rho_w), 2e-3
    (rho_qc, rho_qr), and total water + rain to 1e-7 relative."""
    p, s, ref = coldpool_ref
    out = _gpu_steps(p, s, 10, core.MIXED)
    errs = field_errors(p, out, ref)
    _print("coldpool 8192x256 mixed, 10 steps vs FP64 oracle", errs)
    bars = dict(rho=1e-6, rho_theta=1e-6, rho_qv=1e-5, rho_u=1e-4, rho_w=1e-4, rho_qc=2e-3, rho_qr=2e-3)
    for n, e in errs.items():
        assert e <= bars[n], (n, e)
    h = p.halo
    sl = np.s_[:, h:h + p.Nx]

    def water(st):
        return float(((st.rho_qv + st.rho_qc + st.rho_qr)[sl] * p.dz[:, None]).sum() + st.prec[h:h + p.Nx].sum())

    assert abs(water(out) - water(ref)) <= 1e-7 * water(ref)


def _trajectory(eng, members, states, nsteps):
    """x_0 .. x_nsteps of the listed members on one handle, as FP64 States."""
    xs = {m: [] for m in members}
    for n in range(nsteps + 1):
        for m in members:
            a = states[m].copy()
            eng.download(a, member=m, phys=False)
            xs[m].append(a)
        if n < nsteps:
            eng.step(1)
    return xs


def _check_fp64(tag, p, s, tt, g, gK, J, nsteps, ckpt):
    Jo, go, gKo = O.adjoint_run(s.copy(), p, nsteps, tt, ckpt=ckpt)
    h = p.halo
    errs = {n: _rel(g[f][:, h:h + p.Nx], go[f][:, h:h + p.Nx]) for f, n in enumerate(PROG7)}
    _print(f"{tag}: J {J:.12e} vs oracle {Jo:.12e}; grad rel-L2", errs)
    assert abs(J - Jo) <= 1e-12 * abs(Jo), (J, Jo)
    for n, e in errs.items():
        assert e <= 1e-8, (tag, n, e)
    for j in range(3):
        assert abs(gK[j] - gKo[j]) <= 1e-8 * abs(gKo[j]) + 1e-14, (tag, j, gK[j], gKo[j])


def _check_mixed(tag, p, traj, tt, g, gK, J):
    """Bars of the mixed adjoint against the one-step adjoint chain along its
    own trajectory (FP32 records): all fields 1e-3, each field 5e-3."""
    Jc, gc, gKc = chain_adjoint(p, traj, tt)
    h = p.halo
    errs = {n: _rel(g[f][:, h:h + p.Nx], gc[f][:, h:h + p.Nx]) for f, n in enumerate(PROG7)}
    tot = np.sqrt(sum(np.linalg.norm(g[f][:, h:h + p.Nx] - gc[f][:, h:h + p.Nx]) ** 2 for f in range(7)) /
                  sum(np.linalg.norm(gc[f][:, h:h + p.Nx]) ** 2 for f in range(7)))
    _print(f"{tag}: J {J:.12e} vs chain {Jc:.12e}; all fields {tot:.1e}; rel-L2", errs)
    assert abs(J - Jc) <= 1e-12 * abs(Jc), (J, Jc)
    assert tot <= 1e-3, tot
    for n, e in errs.items():
        assert e <= 5e-3, (tag, n, e)
    for j in range(3):
        assert abs(gK[j] - gKc[j]) <= 1e-3 * abs(gKc[j]) + 1e-12, (tag, j, gK[j], gKc[j])


def test_adjoint_bubble_512x128_fp64():
    p, s = ic.adjoint_bubble(np.float64)
    tt = ic.theta_target(p, seed=3)
    print("adjoint 512x128 fp64, 4 steps ckpt 2", flush=True)
    eng = core.Engine(p, core.FP64)
    eng.upload(s)
    J = eng.adjoint(4, [tt], ckpt=2)[0]
    g, gK = eng.adjoint_grad(0)
    eng.close()
    _check_fp64("adjoint 512x128 fp64", p, s, tt, g, gK, J, 4, 2)


def test_adjoint_bubble_512x128_mixed():
    p, s = ic.adjoint_bubble(np.float64)
    tt = ic.theta_target(p, seed=3)
    print("adjoint 512x128 mixed, 4 steps ckpt 2", flush=True)
    eng = core.Engine(p, core.MIXED)
    eng.upload(s)
    traj = _trajectory(eng, [0], [s], 4)[0]
    eng.upload(s)
    J = eng.adjoint(4, [tt], ckpt=2)[0]
    g, gK = eng.adjoint_grad(0)
    eng.close()
    _check_mixed("adjoint 512x128 mixed", p, traj, tt, g, gK, J)


@pytest.fixture(scope="module")
def batched_members():
    print("batched adjoint: 64 member ICs", flush=True)
    return ic.batched_adjoint(np.float64, members=64, seed=4)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_batched_adjoint_64x1024x128(batched_members, form):
    mem = batched_members
    nsteps, ckpt = 3, 2
    check = (0, 63)
    print(f"batched adjoint 64x1024x128 {form}, {nsteps} steps ckpt {ckpt}", flush=True)
    eng = core.Engine([p for p, _, _ in mem], core.FP64 if form == "fp64" else core.MIXED)
    for m, (p, s, _) in enumerate(mem):
        eng.upload(s, member=m)
    traj = None
    if form == "mixed":
        traj = _trajectory(eng, check, [s for _, s, _ in mem], nsteps)
        for m, (p, s, _) in enumerate(mem):
            eng.upload(s, member=m)
    J = eng.adjoint(nsteps, [t for _, _, t in mem], ckpt=ckpt)
    grads = {m: eng.adjoint_grad(m) for m in check}
    eng.close()
    for m in check:
        p, s, tt = mem[m]
        g, gK = grads[m]
        if form == "fp64":
            _check_fp64(f"batched fp64 member {m}", p, s, tt, g, gK, J[m], nsteps, ckpt)
        else:
            _check_mixed(f"batched mixed member {m}", p, traj[m], tt, g, gK, J[m])
