"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

FP64 form: per-field relative L2 error <= 1e-11 after 10 large steps of the
128x64 bubble (snapshot scalar order), the bar of the drop-in.  Mixed form:
error against the FP64 oracle stated per field (<= 1e-4 after 10 steps).
"""
import numpy as np
import pytest

import oracle as O
from conftest import PROG, field_errors, interior, rel_l2
from moist2d_amd import core, ic
from moist2d_amd.state import allocate_state

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-11


def _print(tag, errs):
    print(tag, " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)


def test_fp64_bubble_10_steps():
    p, s = ic.bubble(np.float64)
    ref = s.copy()
    assert O.rk3_step(ref, p, 10) == 0
    eng = core.Engine(p, core.FP64)
    eng.upload(s)
    eng.step(10)
    out = s.copy()
    eng.download(out)
    errs = field_errors(p, out, ref)
    _print("fp64 bubble 10 steps", errs)
    for n, e in errs.items():
        assert e <= TOL_FP64, (n, e)
    np.testing.assert_allclose(out.prec, ref.prec, rtol=1e-11, atol=1e-300)


def _run_gpu(p, s, steps, precision=core.FP64, graph=False):
    eng = core.Engine(p, precision)
    eng.use_graph(graph)
    eng.upload(s)
    eng.step(steps)
    out = s.copy()
    eng.download(out)
    d = eng.diagnostics()
    eng.close()
    return out, d


def test_mixed_bubble_10_steps_vs_fp64_oracle():
    p, s = ic.bubble(np.float64)
    ref = s.copy()
    assert O.rk3_step(ref, p, 10) == 0
    out, _ = _run_gpu(p, s, 10, core.MIXED)
    errs = field_errors(p, out, ref)
    _print("mixed bubble 10 steps vs FP64 oracle", errs)
    # stated error of the mixed form against the FP64 oracle
    for n in ("rho", "rho_theta", "rho_qv"):
        assert errs[n] <= 1e-6, (n, errs[n])
    for n in ("rho_u", "rho_w"):
        assert errs[n] <= 1e-4, (n, errs[n])


def test_fp64_compute_step_entry():
    """compute_step! with non-zero tendencies, s != s0, ns = 3, in a uniform
    5 m/s wind.  (Two steps from rest |rho_u| <= 6e-3 and the oracle's own
    rho_u moves by 4e-10 under a 1-ulp input perturbation; with the wind that
    sensitivity is 4e-15, so the 1e-11 bar measures the kernels.)"""
    p, s = ic.bubble(np.float64, nx=64, nz=32, d=156.25)
    s.rho_u[...] = 5.0 * s.rho
    rng = np.random.default_rng(5)
    s0 = s.copy()
    O.rk3_step(s, p, 2)  # s moved away from s0
    phys = [np.asfortranarray(rng.normal(0, sc, size=s.rho.shape)) for sc in (1e-6, 1e-4, 1e-4, 1e-7, 1e-7, 1e-7)]
    ref = s.copy()
    assert O.compute_step(phys, ref, s0, p, 0.5, 3) == 0
    eng = core.Engine(p, core.FP64)
    eng.upload(s)
    eng.compute_step(phys, s0, 0.5, 3)
    out = s.copy()
    eng.download(out, phys=False)
    eng.close()
    errs = field_errors(p, out, ref)
    _print("fp64 compute_step", errs)
    for n, e in errs.items():
        assert e <= TOL_FP64, (n, e)
    # halo columns: the reference leaves periodic copies of the stage input
    for n in ("rho", "rho_u", "rho_w", "rho_qv"):
        np.testing.assert_array_equal(getattr(out, n)[:, :p.halo], getattr(ref, n)[:, :p.halo])


def test_mixed_compute_step_entry():
    """The mixed form of moist_compute_step (core.compute_step's default for
    Float32 states): explicit dt, s0 given apart from s, non-zero tendencies,
    ns = 3, against the FP64 oracle.  Bar: FP32 rounding, i.e. rel-L2 <=
    1e-5 per field, or 30 times what rounding s, s0 and the tendencies to
    FP32 once changes in the oracle's own result."""
    p, s = ic.bubble(np.float64, nx=64, nz=32, d=156.25)
    s.rho_u[...] = 5.0 * s.rho
    rng = np.random.default_rng(5)
    s0 = s.copy()
    O.rk3_step(s, p, 2)
    phys = [np.asfortranarray(rng.normal(0, sc, size=s.rho.shape)) for sc in (1e-6, 1e-4, 1e-4, 1e-7, 1e-7, 1e-7)]
    ref = s.copy()
    assert O.compute_step(phys, ref, s0, p, 0.5, 3) == 0

    def r32(st):
        a = st.copy()
        for n in PROG:
            getattr(a, n)[...] = getattr(st, n).astype(np.float32).astype(np.float64)
        return a

    pert = r32(s)
    phys32 = [np.asfortranarray(x.astype(np.float32).astype(np.float64)) for x in phys]
    assert O.compute_step(phys32, pert, r32(s0), p, 0.5, 3) == 0
    sens = field_errors(p, pert, ref)
    eng = core.Engine(p, core.MIXED)
    eng.upload(s)
    eng.compute_step(phys, s0, 0.5, 3)
    out = s.copy()
    eng.download(out, phys=False)
    eng.close()
    errs = field_errors(p, out, ref)
    _print("mixed compute_step", errs)
    _print("   oracle change from FP32-rounded inputs", sens)
    for n, e in errs.items():
        assert e <= max(1e-5, 30.0 * sens[n]), (n, e, sens[n])


def test_fp64_default_physics_surface_and_radiation():
    """Default Params switches (surface fluxes and gray radiation on) with the
    reference's init_state! profile."""
    from moist2d_amd.params import make_params
    p = make_params(np.float64, Nx=32, Nz=40, Lx=32 * 500.0, H=20e3, dz0=200.0, z_fact=1.05, ns=3, tout_int=2.0)
    s = ic.init_state(allocate_state(p), p, seed=7)
    ref = s.copy()
    assert O.rk3_step(ref, p, 3) == 0
    out, _ = _run_gpu(p, s, 3)
    errs = field_errors(p, out, ref)
    _print("fp64 sf+rd 3 steps", errs)
    for n, e in errs.items():
        assert e <= TOL_FP64, (n, e)
    for n in ("mflux", "shflux", "lhflux", "d_rho_u_sf", "d_rho_theta_sf", "d_rho_qv_sf"):
        a, b = getattr(out, n)[p.halo:p.halo + p.Nx], getattr(ref, n)[p.halo:p.halo + p.Nx]
        assert rel_l2(a, b) <= 1e-11, n
    for n in ("d_rho_theta_mp", "d_rho_qv_mp", "d_rho_qc_mp", "d_rho_qr_mp", "d_rho_qr_sed", "d_rho_theta_rd"):
        e = rel_l2(interior(p, getattr(out, n)), interior(p, getattr(ref, n)))
        assert e <= 1e-10, (n, e)


def test_ensemble_members_match_single_runs():
    members = ic.ensemble(np.float64, members=3, nx_choices=[64, 128, 192], nz_choices=(16, 24), steps=5)
    eng = core.Engine([p for p, _ in members], core.FP64)
    for m, (p, s) in enumerate(members):
        eng.upload(s, member=m)
    eng.step(3)
    for m, (p, s) in enumerate(members):
        ref = s.copy()
        assert O.rk3_step(ref, p, 3) == 0
        out = s.copy()
        eng.download(out, member=m)
        errs = field_errors(p, out, ref)
        _print(f"ensemble member {m} ({p.Nx}x{p.Nz})", errs)
        for n, e in errs.items():
            assert e <= TOL_FP64, (m, n, e)
    eng.close()


def test_mixed_ensemble_of_odd_and_even_columns_matches_single_runs():
    """Mixed form (level-pair work fields, members padded to an even nz): an
    ensemble of three members with nz 17, 13, 24 and nx 64, 70, 128 in one
    handle gives each member the same state, bit for bit, as a handle of its
    own, with graph replay and the default tile shape."""
    members = [ic.bubble(np.float64, nx=nx, nz=nz, d=250.0, seed=m)
               for m, (nx, nz) in enumerate([(64, 17), (70, 13), (128, 24)])]
    eng = core.Engine([p for p, _ in members], core.MIXED)
    eng.use_graph(True)
    for m, (p, s) in enumerate(members):
        eng.upload(s, member=m)
    eng.step(4)
    for m, (p, s) in enumerate(members):
        out = s.copy()
        eng.download(out, member=m)
        one = core.Engine(p, core.MIXED)
        one.use_graph(True)
        one.upload(s)
        one.step(4)
        ref = s.copy()
        one.download(ref)
        one.close()
        for n in PROG:
            np.testing.assert_array_equal(interior(p, getattr(out, n)), interior(p, getattr(ref, n)), err_msg=n)
        assert np.all(np.isfinite(interior(p, out.rho_w)))
    eng.close()


def test_ensemble_workload_members_match_oracle():
    """Members 0, 1 and 255 of the 256-member workload (seed 2) in one handle
    with all 256 domains, graph replay, 10 steps, FP64 form against the oracle.
    (After 2 steps from rest the oracle's own rho_u moves by 4e-11 when rho_theta
    is perturbed by 1 ulp; after 10 steps by 2e-12, below the 1e-11 bar.)"""
    members = ic.ensemble(np.float64, members=256, steps=10)
    eng = core.Engine([p for p, _ in members], core.FP64)
    for m, (p, s) in enumerate(members):
        eng.upload(s, member=m)
    eng.use_graph(True)
    eng.step(10)
    for m in (0, 1, 255):
        p, s = members[m]
        ref = s.copy()
        assert O.rk3_step(ref, p, 10) == 0
        out = s.copy()
        eng.download(out, member=m)
        errs = field_errors(p, out, ref)
        _print(f"ensemble-256 member {m} ({p.Nx}x{p.Nz})", errs)
        for n, e in errs.items():
            assert e <= TOL_FP64, (m, n, e)
    eng.close()


@pytest.mark.parametrize("precision", [core.FP64, core.MIXED])
def test_graph_replay_matches_direct_and_is_deterministic(precision):
    p, s = ic.bubble(np.float64, nx=128, nz=32)
    a, _ = _run_gpu(p, s, 4, precision, graph=False)
    b, _ = _run_gpu(p, s, 4, precision, graph=True)
    c, _ = _run_gpu(p, s, 4, precision, graph=True)
    for n in PROG:
        np.testing.assert_array_equal(getattr(a, n), getattr(b, n), err_msg=n)
        np.testing.assert_array_equal(getattr(b, n), getattr(c, n), err_msg=n)


def test_nan_is_reported_like_the_reference():
    p, s = ic.bubble(np.float64, nx=64, nz=16)
    s.rho[5, p.halo + 3] = np.nan
    with pytest.raises(RuntimeError, match="NaN detected in rho$"):
        core.rk3_step(s, p)


def test_dropin_step_reuses_its_handle():
    """The reference's driver calls rk3_step!(s, p) once per step
    (moist.jl:1943-1944): three such calls on host States reuse one cached
    device handle and equal three oracle steps; a call that hits a NaN
    leaves the handle usable for the next clean State."""
    core.release_cached()
    p, s = ic.bubble(np.float64, nx=64, nz=32, d=156.25)
    ref = s.copy()
    assert O.rk3_step(ref, p, 3) == 0
    out = s.copy()
    engines = []
    for _ in range(3):
        core.rk3_step(out, p)
        engines.append(next(iter(core._CACHE.values())))
    assert len(core._CACHE) == 1 and engines[0] is engines[1] is engines[2]
    errs = field_errors(p, out, ref)
    _print("drop-in rk3_step x3", errs)
    for n, e in errs.items():
        assert e <= TOL_FP64, (n, e)
    bad = s.copy()
    bad.rho_qr[4, p.halo + 5] = np.nan
    with pytest.raises(RuntimeError, match="NaN detected in "):
        core.rk3_step(bad, p)
    again = s.copy()
    core.rk3_step(again, p, nsteps=3)
    for n in PROG:
        np.testing.assert_array_equal(getattr(again, n), getattr(out, n), err_msg=n)
    core.release_cached()


@pytest.mark.parametrize("precision", [core.FP64, core.MIXED])
@pytest.mark.parametrize("field", ["rho", "rho_theta", "rho_qc"])
def test_nan_report_names_the_first_field_in_both_forms(precision, field):
    """The state fields' NaN checks run in k_scalar (FP64 form) or in the last
    acoustic substep of the stage (mixed form); either way the first field in
    the reference's order is named.  A NaN in rho_theta or rho_qc reaches the
    pressure (through cp/cv and R), then rho_u' and the acoustic rho' update
    within the first stage, so rho is the first field named, as in the
    reference's dataflow."""
    p, s = ic.bubble(np.float64, nx=64, nz=16)
    getattr(s, field)[5, p.halo + 3] = np.nan
    eng = core.Engine(p, precision)
    eng.upload(s)
    with pytest.raises(RuntimeError, match="NaN detected in rho$"):
        eng.step(1)
    eng.close()


def test_conservation_coldpool_1024x256():
    p, s = ic.coldpool(np.float64, nx=1024, nz=256)
    for prec, tol in ((core.FP64, 1e-13), (core.MIXED, 1e-6)):
        eng = core.Engine(p, prec)
        eng.upload(s)
        d0 = eng.diagnostics()
        eng.step(40)
        d1 = eng.diagnostics()
        eng.close()
        w0 = d0["total_water"] + d0["rain"]
        dw = (d1["total_water"] + d1["rain"] - w0) / w0
        m0 = d0["mass"] + d0["rain"]
        dm = (d1["mass"] + d1["rain"] - m0) / m0
        dd = (d1["dry_mass"] - d0["dry_mass"]) / d0["dry_mass"]
        print(f"conservation prec={prec}: water+rain {dw:.2e} mass+rain {dm:.2e} dry {dd:.2e}", flush=True)
        assert abs(dw) <= tol and abs(dm) <= tol and abs(dd) <= tol


@pytest.mark.parametrize("levels,threads", [(2, 256), (2, 512), (2, 1024), (4, 256), (4, 512), (4, 1024),
                                            (2, -512), (2, -1024), (4, -512), (4, -1024)])
def test_fp64_every_acoustic_shape(levels, threads):
    """Every shape the autotuner may pick keeps the FP64 parity bar."""
    p, s = ic.bubble(np.float64)
    ref = s.copy()
    assert O.rk3_step(ref, p, 10) == 0
    eng = core.Engine(p, core.FP64)
    eng.set_acoustic_shape(levels, threads)
    eng.upload(s)
    eng.step(10)
    out = s.copy()
    eng.download(out)
    eng.close()
    errs = field_errors(p, out, ref)
    _print(f"fp64 shape L={levels} NT={threads}", errs)
    for n, e in errs.items():
        assert e <= TOL_FP64, (n, e)


def test_autotune_keeps_parity():
    p, s = ic.bubble(np.float64)
    ref = s.copy()
    assert O.rk3_step(ref, p, 10) == 0
    for prec, tol in ((core.FP64, TOL_FP64), (core.MIXED, 1e-4)):
        eng = core.Engine(p, prec)
        eng.upload(s)
        L, NT, us = eng.autotune(reps=5)
        assert L in (2, 4) and NT in (256, 512, 1024, -512, -1024) and us > 0
        eng.upload(s)
        eng.use_graph(True)
        eng.step(10)
        out = s.copy()
        eng.download(out)
        eng.close()
        errs = field_errors(p, out, ref)
        _print(f"autotuned prec={prec} L={L} NT={NT} ({us:.1f} us)", errs)
        for n, e in errs.items():
            assert e <= tol, (prec, n, e)


@pytest.mark.parametrize("kw", [
    dict(beta_offcenter=0.6),           # off-centering terms active (zero at the default beta = 1)
    dict(ns=1),                         # one substep in every stage
    dict(ns=8),                         # 3, 4, 8 substeps: unequal stage dt_s
    dict(gamma_d=0.0),                  # no divergence damping
    dict(nz=3),                         # two unknowns in the w-p system
    dict(nx=3),                         # halo of 2 on 3 columns (periodic wrap of the stencils)
    dict(nx=70, nz=17),                 # ragged tiles in x and z
])
def test_fp64_edge_cases(kw):
    nx, nz = kw.pop("nx", 48), kw.pop("nz", 24)
    p, s = ic.bubble(np.float64, nx=nx, nz=nz, d=250.0, **kw)
    s.rho_u[...] = 4.0 * s.rho  # a mean wind exercises both upwind directions' branches
    ref = s.copy()
    assert O.rk3_step(ref, p, 6) == 0
    # the oracle's own change under a 1-ulp perturbation of rho_theta: on tiny
    # columns (nz = 3) rho_w moves by 2e-11, so the bar is the larger of 1e-11
    # and twice that sensitivity
    pert = s.copy()
    pert.rho_theta[...] = np.nextafter(pert.rho_theta, np.inf)
    assert O.rk3_step(pert, p, 6) == 0
    sens = field_errors(p, pert, ref)
    out, _ = _run_gpu(p, s, 6)
    errs = field_errors(p, out, ref)
    _print(f"fp64 edge {kw} {nx}x{nz}", errs)
    _print("   oracle 1-ulp sensitivity", sens)
    for n, e in errs.items():
        assert e <= max(TOL_FP64, 2.0 * sens[n]), (n, e, sens[n])


MIXED_EDGE_CASES = [
    dict(beta_offcenter=0.6),  # off-centering terms active
    dict(ns=1),                # one substep a stage: the first substep is also the last (the final update)
    dict(ns=8),                # 3, 4, 8 substeps
    dict(nz=3),                # one level pair and a padded top level
    dict(nx=3),                # periodic wrap on 3 columns
    dict(nx=70, nz=17),        # ragged tiles, odd nz (padded level pair)
]


@pytest.mark.parametrize("kw", MIXED_EDGE_CASES)
def test_mixed_edge_cases(kw):
    """The mixed form's kernels (level-pair work fields, the last substep's
    final update) on the FP64 edge cases, against the FP64 oracle after 6
    steps, and bit-identical between direct launches and graph replay.  The
    bar is FP32 rounding: rel-L2 <= 1e-4 per field, or 30 times what rounding
    the initial state once to FP32 changes in the oracle's own result (the
    mixed form rounds at every step; on a 3-level column the oracle's rho_w
    moves 2e-11 under a 1-ulp change of rho_theta, so FP32 rounding there is
    felt at the 1e-3 level)."""
    kw = dict(kw)
    nx, nz = kw.pop("nx", 48), kw.pop("nz", 24)
    p, s = ic.bubble(np.float64, nx=nx, nz=nz, d=250.0, **kw)
    s.rho_u[...] = 4.0 * s.rho
    ref = s.copy()
    assert O.rk3_step(ref, p, 6) == 0
    r32 = s.copy()
    for n in PROG:
        getattr(r32, n)[...] = getattr(s, n).astype(np.float32).astype(np.float64)
    assert O.rk3_step(r32, p, 6) == 0
    sens = field_errors(p, r32, ref)
    out, _ = _run_gpu(p, s, 6, core.MIXED)
    out_g, _ = _run_gpu(p, s, 6, core.MIXED, graph=True)
    errs = field_errors(p, out, ref)
    _print(f"mixed edge {kw} {nx}x{nz}", errs)
    _print("   oracle change from FP32-rounded initial state", sens)
    for n in PROG:
        np.testing.assert_array_equal(getattr(out, n), getattr(out_g, n), err_msg=n)
    for n, e in errs.items():
        assert e <= max(1e-4, 30.0 * sens[n]), (n, e, sens[n])
