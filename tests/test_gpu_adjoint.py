"""GPU reverse mode (moist_adjoint_run / moist_adjoint_grad) against the
oracle's adjoint (itself pinned by central finite differences and by the
oracle's forward mode in test_adjoint_oracle.py), against the oracle's
forward-mode derivative (dual numbers, the ForwardDiff counterpart) and
against central finite differences of the GPU's own forward step.

FP64 form: bars are the loss to 1e-12 relative; per-field gradient rel-L2
<= 1e-8 (the oracle and the GPU run the same FP64 operations in a different
order, and the reverse sweep amplifies their 1e-13 trajectory differences by
at most the condition of the step); parameter gradients to 1e-8.  Both ways
of keeping the forward records (whole segment, per step) are covered.

Mixed form: the trajectory advances in the mixed form and each step's
adjoint runs in FP64 from the mixed step's own FP32 records.  The reference
is the chain of the oracle's one-step adjoints along the device's own
mixed-form trajectory (downloaded step by step): the two differ by the FP32
rounding of the records (bar 1e-4 per field).  Its distance from the
FP64-form gradient is the trajectory's precision error, checked separately.
"""
import numpy as np
import pytest

import oracle as O
from adjoint_chain import chain_adjoint
from conftest import PROG
from moist2d_amd import core, ic

pytestmark = pytest.mark.gpu

PROG7 = ("rho", "rho_u", "rho_w", "rho_theta", "rho_qv", "rho_qc", "rho_qr")


def _wrap(p, a):
    h = p.halo
    a[:, :h] = a[:, p.Nx:p.Nx + h]
    a[:, h + p.Nx:] = a[:, h:2 * h]
    return a


def _case(nx=16, nz=12, seed=2, **kw):
    rng = np.random.default_rng(seed)
    p, s = ic.bubble(np.float64, nx=nx, nz=nz, d=200.0, **kw)
    s.rho_u[...] = 3.0 * s.rho
    O.rk3_step(s, p, 3)
    s.rho_qc[...] += 1.5e-3 * s.rho * (1 + 0.2 * rng.random(s.rho.shape))
    s.rho_qr[...] += 1e-4 * s.rho * (1 + 0.2 * rng.random(s.rho.shape))
    for n in PROG7:
        _wrap(p, getattr(s, n))
    tt = np.asfortranarray(s.rho_theta / s.rho + rng.normal(0, 0.5, size=s.rho.shape))
    return p, s, tt


def _rel(a, b):
    d = np.linalg.norm(a - b)
    n = np.linalg.norm(b)
    return d / n if n > 0 else d


def _compare(tag, p, s, tt, nsteps, ckpt, tol=1e-8, records="auto"):
    Jo, go, gKo = O.adjoint_run(s.copy(), p, nsteps, tt, ckpt=ckpt)
    eng = core.Engine(p, core.FP64)
    eng.adjoint_records(records)
    eng.upload(s)
    J = eng.adjoint(nsteps, [tt], ckpt=ckpt)[0]
    g, gK = eng.adjoint_grad(0)
    fwd, rev = eng.adjoint_times()
    eng.close()
    h = p.halo
    errs = {n: _rel(g[f][:, h:h + p.Nx], go[f][:, h:h + p.Nx]) for f, n in enumerate(PROG7)}
    kerr = [abs(gK[j] - gKo[j]) / max(abs(gKo[j]), 1e-300) for j in range(3)]
    print(f"{tag}: J {J:.12e} vs {Jo:.12e}; grad rel-L2 " + " ".join(f"{k}={v:.1e}" for k, v in errs.items())
          + " dK " + " ".join(f"{e:.1e}" for e in kerr) + f"; fwd {fwd:.2f} ms rev {rev:.2f} ms", flush=True)
    assert abs(J - Jo) <= 1e-12 * abs(Jo), (J, Jo)
    for n, e in errs.items():
        assert e <= tol, (tag, n, e)
    for j in range(3):
        assert kerr[j] <= tol or abs(gK[j] - gKo[j]) <= 1e-12, (tag, j, gK[j], gKo[j])
    return g, gK


def test_adjoint_one_step_dynamics_only():
    p, s, tt = _case(do_mp=False)
    _compare("dynamics 1 step", p, s, tt, 1, 1)


def test_adjoint_one_step_with_physics():
    p, s, tt = _case()
    _compare("physics 1 step", p, s, tt, 1, 1)


@pytest.mark.parametrize("records", ["segment", "step"])
def test_adjoint_checkpointed_run(records):
    p, s, tt = _case()
    _compare(f"12 steps ckpt 5, {records} records", p, s, tt, 12, 5, records=records)


def _trajectory(p, s, nsteps, precision):
    """x_0 .. x_nsteps of the device's forward in `precision`, as FP64 States."""
    eng = core.Engine(p, precision)
    eng.upload(s)
    xs = []
    for n in range(nsteps + 1):
        a = s.copy()
        eng.download(a, phys=False)
        xs.append(a)
        if n < nsteps:
            eng.step(1)
    eng.close()
    return xs


def _device_adjoint(p, s, tt, nsteps, ckpt, precision, records="auto"):
    eng = core.Engine(p, precision)
    eng.adjoint_records(records)
    eng.upload(s)
    J = eng.adjoint(nsteps, [tt], ckpt=ckpt)[0]
    g, gK = eng.adjoint_grad(0)
    eng.close()
    return J, g, gK


@pytest.mark.parametrize("records", ["segment", "step"])
@pytest.mark.parametrize("do_mp", [False, True])
def test_mixed_adjoint_is_the_step_adjoint_chain_along_its_trajectory(records, do_mp):
    """FP32 records: each step is linearised at the mixed step's own stage
    values (stage inputs, theta, rt2pres, tendencies, acoustic
    perturbations), which differ from the FP64 step's by the mixed form's
    one-step error.  Measured: all fields together rel-L2 1.4e-6 (no
    physics) and 9e-5 (physics); per field at most 1.1e-3 (rho_qv without
    physics, a small component).  Bars: 1e-3 all fields, 5e-3 per field."""
    p, s, tt = _case(do_mp=do_mp)
    nsteps, ckpt = 5, 2
    Jc, gc, gKc = chain_adjoint(p, _trajectory(p, s, nsteps, core.MIXED), tt)
    J, g, gK = _device_adjoint(p, s, tt, nsteps, ckpt, core.MIXED, records)
    h = p.halo
    errs = {n: _rel(g[f][:, h:h + p.Nx], gc[f][:, h:h + p.Nx]) for f, n in enumerate(PROG7)}
    kerr = [abs(gK[j] - gKc[j]) / max(abs(gKc[j]), 1e-300) for j in range(3)]
    print(f"mixed {records} do_mp={do_mp}: J {J:.12e} vs chain {Jc:.12e}; rel-L2 "
          + " ".join(f"{k}={v:.1e}" for k, v in errs.items()) + " dK " + " ".join(f"{e:.1e}" for e in kerr),
          flush=True)
    tot = np.sqrt(sum(np.linalg.norm(g[f][:, h:h + p.Nx] - gc[f][:, h:h + p.Nx]) ** 2 for f in range(7)) /
                  sum(np.linalg.norm(gc[f][:, h:h + p.Nx]) ** 2 for f in range(7)))
    print(f"   all fields rel-L2 {tot:.1e}", flush=True)
    assert abs(J - Jc) <= 1e-12 * abs(Jc), (J, Jc)
    assert tot <= 1e-3, tot
    for n, e in errs.items():
        assert e <= 5e-3, (n, e)
    for j in range(3):
        assert kerr[j] <= 1e-3 or abs(gK[j] - gKc[j]) <= 1e-12, (j, gK[j], gKc[j])


def test_mixed_adjoint_record_modes_agree():
    """Whole-segment and per-step records replay the same mixed steps: the
    gradients are bitwise equal."""
    p, s, tt = _case()
    a = _device_adjoint(p, s, tt, 7, 3, core.MIXED, "segment")
    b = _device_adjoint(p, s, tt, 7, 3, core.MIXED, "step")
    assert a[0] == b[0]
    for f in range(7):
        np.testing.assert_array_equal(a[1][f], b[1][f])
    np.testing.assert_array_equal(a[2], b[2])


@pytest.mark.parametrize("case", ["bubble", "sf_rd"])
@pytest.mark.parametrize("precision", [core.FP64, core.MIXED])
def test_pipelined_segments_match_segment_records(precision, case):
    """Pipelined segments (one segment's recording pass on a second stream
    beside the next segment's reverse sweep, the segment buffers in two
    halves) run the same kernels on the same values as whole-segment records:
    J and the gradients are bitwise equal.  7 steps, checkpoints every 2: four
    segments, so both halves are written twice; with surface fluxes on, the
    stale halo pairs of the segments' steps are in the halves too."""
    p, s, tt = _case() if case == "bubble" else _case_sf_rd()
    a = _device_adjoint(p, s, tt, 7, 2, precision, "segment")
    b = _device_adjoint(p, s, tt, 7, 2, precision, "pipelined")
    assert a[0] == b[0]
    for f in range(7):
        np.testing.assert_array_equal(a[1][f], b[1][f], err_msg=PROG7[f])
    np.testing.assert_array_equal(a[2], b[2])


@pytest.mark.parametrize("case", ["bubble", "sf_rd"])
@pytest.mark.parametrize("precision", [core.FP64, core.MIXED])
def test_pipelined_steps_match_step_records(precision, case):
    """Per-step records with each step's recording pass on a second stream
    beside the reverse of the step after it (two record blocks used in turn)
    replay the same values as per-step records: J and the gradients are
    bitwise equal.  7 steps, checkpoints every 3: segments of 3, 3 and 1
    steps, so both blocks are reused within a segment and across segments."""
    p, s, tt = _case() if case == "bubble" else _case_sf_rd()
    a = _device_adjoint(p, s, tt, 7, 3, precision, "step")
    b = _device_adjoint(p, s, tt, 7, 3, precision, "step_pipelined")
    assert a[0] == b[0]
    for f in range(7):
        np.testing.assert_array_equal(a[1][f], b[1][f], err_msg=PROG7[f])
    np.testing.assert_array_equal(a[2], b[2])


def test_adjoint_buffers_that_do_not_fit_fail_cleanly_twice():
    """An adjoint run whose checkpoints cannot fit in device memory (100,000
    steps of the 512x128 bubble with a checkpoint every step: about 370 GB)
    returns an error before any step runs, and so does the same call again:
    the failed allocation leaves no half-built buffers behind for the second
    call to find.  The handle then runs a small adjoint."""
    p, s = ic.bubble(np.float64, nx=512, nz=128)
    tt = np.asfortranarray(s.rho_theta / s.rho)
    eng = core.Engine(p, core.FP64)
    eng.upload(s)
    for _ in range(2):
        with pytest.raises(RuntimeError, match="moist_adjoint_run"):
            eng.adjoint(100000, [tt], ckpt=1)
    J = eng.adjoint(2, [tt], ckpt=1)[0]
    assert np.isfinite(J)
    eng.close()


def test_mixed_adjoint_close_to_fp64_adjoint():
    """The mixed-form gradient differs from the FP64 form's by what the
    trajectory's FP32 deviations change.  Measured on this case: the loss
    2.5e-6 relative, per-field gradient rel-L2 <= 3e-4; bars 1e-5 and 1e-3
    (the mixed form's forward bar is 1e-4 state rel-L2; a gradient, being a
    derivative along that trajectory, carries a few times its error)."""
    p, s, tt = _case()
    nsteps, ckpt = 12, 5
    J64, g64, gK64 = _device_adjoint(p, s, tt, nsteps, ckpt, core.FP64)
    Jm, gm, gKm = _device_adjoint(p, s, tt, nsteps, ckpt, core.MIXED)
    h = p.halo
    errs = {n: _rel(gm[f][:, h:h + p.Nx], g64[f][:, h:h + p.Nx]) for f, n in enumerate(PROG7)}
    kerr = [abs(gKm[j] - gK64[j]) / max(abs(gK64[j]), 1e-300) for j in range(3)]
    print(f"mixed vs fp64: J rel {abs(Jm - J64) / abs(J64):.1e}; rel-L2 "
          + " ".join(f"{k}={v:.1e}" for k, v in errs.items()) + " dK " + " ".join(f"{e:.1e}" for e in kerr),
          flush=True)
    assert abs(Jm - J64) <= 1e-5 * abs(J64)
    for n, e in errs.items():
        assert e <= 1e-3, (n, e)


def test_adjoint_bubble_64x32():
    p, s = ic.bubble(np.float64, nx=64, nz=32, d=312.5, dt=1.0)
    O.rk3_step(s, p, 30)  # let cloud and rain form
    rng = np.random.default_rng(9)
    tt = np.asfortranarray(s.rho_theta / s.rho + rng.normal(0, 0.5, size=s.rho.shape))
    _compare("bubble 64x32, 20 steps", p, s, tt, 20, 6, tol=1e-7)


def test_adjoint_equals_forward_mode_derivative_64x32():
    """The device's FP64 adjoint against forward-mode differentiation of the
    oracle (dual numbers through the restated rk3_step!, branches following
    the primal trajectory, as ForwardDiff.jl through moist.jl) on the 64x32
    bubble with cloud and rain, 20 steps, checkpoints every 6: directions of
    theta', qv and rho_u and the three Kessler parameters, to 1e-8 (the two
    trajectories differ at roundoff)."""
    p, s = ic.bubble(np.float64, nx=64, nz=32, d=312.5, dt=1.0)
    O.rk3_step(s, p, 30)
    rng = np.random.default_rng(12)
    # cloud above the autoconversion threshold and rain everywhere, so that
    # the three Kessler parameters act
    s.rho_qc[...] += 1.5e-3 * s.rho * (1 + 0.2 * rng.random(s.rho.shape))
    s.rho_qr[...] += 1e-4 * s.rho * (1 + 0.2 * rng.random(s.rho.shape))
    for n in PROG7:
        _wrap(p, getattr(s, n))
    tt = np.asfortranarray(s.rho_theta / s.rho + rng.normal(0, 0.5, size=s.rho.shape))
    nsteps = 20
    eng = core.Engine(p, core.FP64)
    eng.upload(s)
    J = eng.adjoint(nsteps, [tt], ckpt=6)[0]
    g, gK = eng.adjoint_grad(0)
    eng.close()
    h = p.halo
    for f, n, sc in ((3, "rho_theta", 0.1), (4, "rho_qv", 1e-4), (1, "rho_u", 0.5)):
        v = np.asfortranarray(_wrap(p, rng.normal(size=s.rho.shape)) * sc * s.rho)
        Jf, dJ = O.jvp_run(s, p, nsteps, tt, {n: v})
        ad = float(np.sum(g[f][:, h:h + p.Nx] * v[:, h:h + p.Nx]))
        print(f"{n}: adjoint {ad:.12e} forward mode {dJ:.12e} rel {abs(ad - dJ) / abs(dJ):.1e}", flush=True)
        assert abs(Jf - J) <= 1e-12 * abs(J)
        assert abs(ad - dJ) <= 1e-8 * abs(dJ), (n, ad, dJ)
    for j in range(3):
        dk = [0.0, 0.0, 0.0]
        dk[j] = 1.0
        _, dJ = O.jvp_run(s, p, nsteps, tt, {}, dK=dk)
        print(f"K{j}: adjoint {gK[j]:.12e} forward mode {dJ:.12e}", flush=True)
        assert dJ != 0.0 and abs(gK[j] - dJ) <= 1e-8 * abs(dJ), (j, gK[j], dJ)


def test_adjoint_ensemble_members_match_single_runs():
    cases = [_case(nx=16, nz=12, seed=2), _case(nx=24, nz=10, seed=5)]
    eng = core.Engine([c[0] for c in cases], core.FP64)
    for m, (p, s, tt) in enumerate(cases):
        eng.upload(s, member=m)
    J = eng.adjoint(6, [c[2] for c in cases], ckpt=4)
    for m, (p, s, tt) in enumerate(cases):
        Jo, go, gKo = O.adjoint_run(s.copy(), p, 6, tt, ckpt=4)
        g, gK = eng.adjoint_grad(m)
        h = p.halo
        assert abs(J[m] - Jo) <= 1e-12 * abs(Jo)
        for f, n in enumerate(PROG7):
            e = _rel(g[f][:, h:h + p.Nx], go[f][:, h:h + p.Nx])
            assert e <= 1e-8, (m, n, e)
        np.testing.assert_allclose(gK, gKo, rtol=1e-8, atol=1e-12)
    eng.close()


def test_adjoint_matches_fd_of_gpu_forward():
    """Central FD (h = 1e-6, smooth directions) of the GPU's own FP64 forward."""
    p, s, tt = _case()
    nsteps = 8
    eng = core.Engine(p, core.FP64)
    eng.upload(s)
    eng.adjoint(nsteps, [tt], ckpt=3)
    g, _ = eng.adjoint_grad(0)
    eng.close()
    h = p.halo
    rng = np.random.default_rng(11)

    def loss(s_):
        e = core.Engine(p, core.FP64)
        e.upload(s_)
        e.step(nsteps)
        a = s_.copy()
        e.download(a)
        e.close()
        th = (a.rho_theta / a.rho)[:, h:h + p.Nx]
        return p.dx * a.prec[h:h + p.Nx].sum() + 0.5 * np.sum((th - tt[:, h:h + p.Nx]) ** 2)

    x = np.arange(p.Nx) / p.Nx
    z = (np.arange(p.Nz) + 0.5) / p.Nz
    errs = []
    for f, n, sc in ((3, "rho_theta", 0.01), (4, "rho_qv", 1e-5), (3, "rho_theta", 0.01), (4, "rho_qv", 1e-5)):
        v = np.zeros((p.Nz, p.Nx))
        for m in range(1, 4):
            for k in range(3):
                a, b = rng.normal(size=2)
                v += np.outer(np.sin(np.pi * k * z + 3 * rng.random()),
                              a * np.cos(2 * np.pi * m * x) + b * np.sin(2 * np.pi * m * x))
        full = np.zeros(s.rho.shape)
        full[:, h:h + p.Nx] = v * sc
        full = _wrap(p, full) * s.rho
        eps = 1e-6
        sp, sm = s.copy(), s.copy()
        getattr(sp, n)[...] += eps * full
        getattr(sm, n)[...] -= eps * full
        fd = (loss(sp) - loss(sm)) / (2 * eps)
        ad = float(np.sum(g[f][:, h:h + p.Nx] * full[:, h:h + p.Nx]))
        errs.append(abs(fd - ad) / abs(ad))
    print("GPU FD rel errors", " ".join(f"{e:.1e}" for e in errs), flush=True)
    assert np.median(errs) <= 1e-5 and max(errs) <= 1e-4, errs


@pytest.mark.parametrize("kw", [dict(beta_offcenter=0.6), dict(ns=8), dict(ns=1), dict(gamma_d=0.0)])
def test_adjoint_edge_cases(kw):
    """Off-centering (fact = (1 - beta)/(1 + beta) != 0 in the w-p right-hand side and its
    adjoint), unequal stage substeps, one substep, no divergence damping."""
    p, s, tt = _case(**kw)
    _compare(f"edge {kw}", p, s, tt, 4, 2)


@pytest.mark.parametrize("nz", [100, 160])
def test_adjoint_two_wave_column_kernel(nz):
    """Columns above 64 levels take two waves in the column kernel (100
    levels: one level per thread; 160: two).  Its solve then runs in 128
    chunks, the lower wave's aggregate composed into the upper wave's scans
    through LDS (and the upper wave's into the lower one's, top down), where
    the three-kernel form runs 64 chunks in one wave: the scans' products
    group differently, so the two forms agree to rounding, not bitwise.  FP64
    form against the oracle's adjoint (bar 1e-8, as every adjoint parity
    case) and against the three-kernel form (bar 1e-11)."""
    p, s, tt = _case(nz=nz)
    g, gK = _compare(f"two-wave column kernel nz={nz}", p, s, tt, 3, 2)
    eng = core.Engine(p, core.FP64)
    eng.adjoint_column_kernel(False)
    eng.upload(s)
    eng.adjoint(3, [tt], ckpt=2)
    g3, gK3 = eng.adjoint_grad(0)
    eng.close()
    h = p.halo
    for f, n in enumerate(PROG7):
        e = _rel(g[f][:, h:h + p.Nx], g3[f][:, h:h + p.Nx])
        assert e <= 1e-11, (n, e)


@pytest.mark.parametrize("precision", [core.FP64, core.MIXED])
def test_adjoint_column_kernel_matches_three_kernels(precision):
    """The vertical part of each reverse acoustic substep as one kernel per
    column (the default) against the three-kernel form: the same arithmetic
    in the same order (one column per wave in both), so the gradients are
    bitwise equal."""
    p, s, tt = _case()
    out = []
    for col in (True, False):
        eng = core.Engine(p, precision)
        eng.adjoint_column_kernel(col)
        eng.upload(s)
        J = eng.adjoint(7, [tt], ckpt=3)[0]
        g, gK = eng.adjoint_grad(0)
        eng.close()
        out.append((J, g, gK))
    assert out[0][0] == out[1][0]
    for f in range(7):
        np.testing.assert_array_equal(out[0][1][f], out[1][1][f], err_msg=PROG7[f])
    np.testing.assert_array_equal(out[0][2], out[1][2])


def _case_sf_rd(seed=7):
    """The reference's default switches: surface fluxes and gray radiation on,
    with wind, cloud and rain, spun up 3 oracle steps so the halo columns
    (which the surface fluxes read at the domain's ends) hold what a step
    leaves there."""
    rng = np.random.default_rng(seed)
    p, s = ic.bubble(np.float64, nx=16, nz=12, d=200.0, do_sf=True, do_rd=True, temp_surf=302.0)
    s.theta_ref[:] = (s.rho_theta / s.rho)[:, p.halo]
    s.rho_u[...] = 6.0 * s.rho
    s.rho_qc[...] += 1.5e-3 * s.rho * (1 + 0.2 * rng.random(s.rho.shape))
    s.rho_qr[...] += 1e-4 * s.rho * (1 + 0.2 * rng.random(s.rho.shape))
    for n in PROG7:
        _wrap(p, getattr(s, n))
    assert O.rk3_step(s, p, 3) == 0
    tt = np.asfortranarray(s.rho_theta / s.rho + rng.normal(0, 0.5, size=s.rho.shape))
    return p, s, tt


@pytest.mark.parametrize("records", ["segment", "step"])
def test_adjoint_with_surface_fluxes_and_radiation(records):
    """FP64 form against the oracle's adjoint_run with both switches on,
    including the adjoints of the surface fluxes' stale halo reads carried
    one step back."""
    p, s, tt = _case_sf_rd()
    _compare(f"sf+rd 9 steps ckpt 4, {records} records", p, s, tt, 9, 4, records=records)


def test_mixed_adjoint_with_surface_fluxes_and_radiation():
    """Mixed form with both switches on against the one-step adjoint chain
    along its own trajectory (bars as for the other mixed-form chain tests)."""
    p, s, tt = _case_sf_rd(8)
    nsteps, ckpt = 5, 2
    traj = _trajectory(p, s, nsteps, core.MIXED)
    h = p.halo
    # the first step's halo reads are the uploaded State's halo columns
    for n in ("rho", "rho_u"):
        getattr(traj[0], n)[:, :h] = getattr(s, n)[:, :h]
        getattr(traj[0], n)[:, h + p.Nx:] = getattr(s, n)[:, h + p.Nx:]
    Jc, gc, gKc = chain_adjoint(p, traj, tt)
    J, g, gK = _device_adjoint(p, s, tt, nsteps, ckpt, core.MIXED)
    errs = {n: _rel(g[f][:, h:h + p.Nx], gc[f][:, h:h + p.Nx]) for f, n in enumerate(PROG7)}
    tot = np.sqrt(sum(np.linalg.norm(g[f][:, h:h + p.Nx] - gc[f][:, h:h + p.Nx]) ** 2 for f in range(7)) /
                  sum(np.linalg.norm(gc[f][:, h:h + p.Nx]) ** 2 for f in range(7)))
    print(f"mixed sf+rd: J {J:.12e} vs chain {Jc:.12e}; all fields {tot:.1e}; rel-L2 "
          + " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
    assert abs(J - Jc) <= 1e-12 * abs(Jc), (J, Jc)
    assert tot <= 1e-3, tot
    for n, e in errs.items():
        assert e <= 5e-3, (n, e)


def test_two_handles_in_sequence_with_graph_replay():
    """The sequence bench.py's adjoint leg runs in one process: an FP64-form
    adjoint whose forward steps replay a captured graph, that handle
    destroyed, then a mixed-form handle's adjoint, also with graph replay.
    Each handle captures its own graph against its own pools.  The FP64 run
    matches the oracle's loss; the mixed run equals, bit for bit, the same
    adjoint on a handle that launches directly."""
    p, s, tt = _case(nx=32, nz=16)
    Jo, _, _ = O.adjoint_run(s.copy(), p, 4, tt, ckpt=2)
    print("FP64 adjoint with graph replay, 32x16, 4 steps", flush=True)
    eng = core.Engine(p, core.FP64)
    eng.use_graph(True)
    eng.upload(s)
    J64 = eng.adjoint(4, [tt], ckpt=2)[0]
    eng.close()
    assert abs(J64 - Jo) <= 1e-12 * abs(Jo)
    out = []
    for graph in (True, False):
        print(f"mixed adjoint, graph replay {graph}", flush=True)
        eng = core.Engine(p, core.MIXED)
        eng.use_graph(graph)
        eng.upload(s)
        J = eng.adjoint(4, [tt], ckpt=2)[0]
        g, gK = eng.adjoint_grad(0)
        eng.close()
        out.append((J, g, gK))
    assert out[0][0] == out[1][0]
    for f in range(7):
        np.testing.assert_array_equal(out[0][1][f], out[1][1][f], err_msg=PROG7[f])
    np.testing.assert_array_equal(out[0][2], out[1][2])
    assert np.isfinite(out[0][0]) and all(np.all(np.isfinite(a)) for a in out[0][1])
