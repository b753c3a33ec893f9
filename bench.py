#!/usr/bin/env python3
"""Benchmark: cell-steps/s of the 8192x256 cold-pool run on one MI355X.

Prints one JSON line.  `value` is the mixed form (FP32 prognostic deviations,
FP64 tridiagonal/base state/accumulators); the FP64 form of the same run is
timed beside it (`fp64_form`), and the mixed form's error against the FP64
form after 600 s of model time (1200 large steps, whatever K is) is stated
(`mixed_vs_fp64`; the FP64 form itself matches the CPU oracle to <= 1e-11 per
field in the parity tests).  Every timed step stores the State diagnostics
rk3_step! writes (moist.jl:1593-1630).  `dropin` times the host-State call
core.rk3_step(s, p), one step per call as the reference's driver makes it.

Timing: W warm-up steps, then exactly K large steps between two HIP events on
the handle's stream with the state resident in HBM (the C ABI synchronizes the
stream before returning).  The roofline object is for the acoustic column
kernel, timed live with a HIP event pair around each of its launches in a
separate pass of K steps.  `cpu_baseline` times the CPU oracle (FP64,
snapshot order, OpenMP over columns) on a bounded sample of the same
workload.

Beside the headline line: the 256-member ensemble (`ensemble`), the 512x128
adjoint workload (`adjoint`: reverse-sweep time / forward time, central-FD
check of the gradient) and the 64-member batched adjoint (`batched_adjoint`:
gradients/s, device memory).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from moist2d_amd import core, ic  # noqa: E402
from moist2d_amd.state import PROGNOSTIC  # noqa: E402

METRIC = "cell-steps/sec on 1 GPU (FP32 vs FP64 form, stated err) + adjoint/forward time"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of the acoustic substep per cell, in field passes, as the
# reference's dataflow needs them (moist.jl:1360-1459): reads (14) rho_theta'
# (n-1), rho_theta' (n-2), rt2pres, R_rho_u, rho_u', mom_x, R_rho, R_rho_theta,
# theta, rho', R_rho_w, rho_w', rho_w of the stage input, mom_z; writes (7)
# rho_u', mom_x, rho', rho_theta' (n), the saved rho_theta' (old), rho_w',
# mom_z.  (Neighbour columns and levels of the stencils are the same fields,
# not extra passes.  The kernel here rotates three rho_theta' buffers instead
# of copying `old`, so it moves 20 of these 21 passes.)  The roofline object
# is for the plain substeps; in the mixed form the last substep of each stage
# is a second variant of the kernel (FIN = true), which also does the stage's
# final update of rho, rho_u, rho_theta, rho_w, and is not among the timed
# launches.
ACOUSTIC_PASSES = 21
# Field passes of one large step: physics 10, per stage slow tendencies 23 and
# scalar transport + final update 26, per acoustic substep 21.
def step_passes(ns):
    n_sub = max(1, -(-ns // 3)) + max(1, -(-ns // 2)) + ns
    return 10 + 3 * (23 + 26) + ACOUSTIC_PASSES * n_sub


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_acoustic_pairs_coldpool_8192x256_mixed.json")


def measured_traffic(kernel, nx, nz):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of the same workload (None if absent or another shape)."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except OSError:
        return None
    if t.get("kernel") != kernel or t.get("nx") != nx or t.get("nz") != nz:
        return None
    return t["traffic_bytes"]


AUTOTUNE = True


FIXED_SHAPE = None  # (levels, threads) from --shape: no autotuning


def tune(eng):
    """Pick the acoustic kernel's tile shape for this handle (None when disabled);
    a shape fixed with --shape is set on the handle as it is."""
    if FIXED_SHAPE:
        if eng is not None:
            eng.set_acoustic_shape(FIXED_SHAPE[0], FIXED_SHAPE[1])
        return {"levels_per_chunk": FIXED_SHAPE[0], "threads": FIXED_SHAPE[1], "us_per_launch": None}
    if not AUTOTUNE:
        return None
    L, NT, us = eng.autotune(reps=10)
    return {"levels_per_chunk": L, "threads": NT, "us_per_launch": us}


def run_form(p, s, precision, steps, warmup, graph=True):
    """Warm up on a throwaway engine, then time `steps` large steps from t = 0."""
    shape = None
    if FIXED_SHAPE:
        shape = tune(None)
    if warmup:
        eng = core.Engine(p, precision)
        eng.upload(s)
        shape = shape or tune(eng)
        if shape:
            eng.set_acoustic_shape(shape["levels_per_chunk"], shape["threads"])
        eng.use_graph(graph)
        eng.time_steps(warmup)
        eng.close()
    eng = core.Engine(p, precision)
    eng.upload(s)
    if shape:
        eng.set_acoustic_shape(shape["levels_per_chunk"], shape["threads"])
    eng.shape = shape
    eng.use_graph(graph)
    t0 = time.perf_counter()
    ms = eng.time_steps(steps)
    wall = time.perf_counter() - t0
    return eng, ms, wall


def kernel_pass(p, s, precision, steps, shape=None):
    eng = core.Engine(p, precision)
    eng.upload(s)
    if shape:
        eng.set_acoustic_shape(shape["levels_per_chunk"], shape["threads"])
    eng.kernel_events(True)
    eng.time_steps(1)
    eng.time_steps(steps)
    ms, n = eng.kernel_stats()
    eng.close()
    return ms, n


def ensemble_run(steps, members=256, seed=2):
    """The batched workload: 256 bubbles on domains of uneven nx (256..2048) and
    nz (64, 128), all advanced by the same launches (one handle)."""
    mem = ic.ensemble(np.float64, members=members, seed=seed)
    params = [p for p, _ in mem]
    cells = sum(p.Nx * p.Nz for p in params)
    out = {"workload": f"ensemble_{members}", "members": members, "cells": cells,
           "columns": sum(p.Nx for p in params), "dt": float(params[0].dt), "ns": params[0].ns, "steps": steps}
    for name, prec in (("mixed", core.MIXED), ("fp64", core.FP64)):
        eng = core.Engine(params, prec)
        for m, (_, s) in enumerate(mem):
            eng.upload(s, member=m)
        shape = tune(eng)
        eng.use_graph(True)
        eng.time_steps(2)
        ms = eng.time_steps(steps)
        d = [eng.diagnostics(m) for m in (0, members - 1)]
        eng.close()
        out[name] = {"cell_steps_per_s": cells * steps / (ms * 1e-3), "ms_per_step": ms / steps, "acoustic_shape": shape,
                     "first_last_member_total_water": [x["total_water"] for x in d]}
    return out


def _smooth_direction(p, rng):
    """Seeded random field of low wavenumbers (x modes 1-3, z modes 0-2), (ka, ia) layout."""
    x = np.arange(p.Nx) / p.Nx
    z = (np.arange(p.Nz) + 0.5) / p.Nz
    v = np.zeros((p.Nz, p.Nx))
    for m in range(1, 4):
        for n in range(3):
            a, b = rng.normal(size=2)
            v += np.outer(np.sin(np.pi * n * z + 3 * rng.random()),
                          a * np.cos(2 * np.pi * m * x) + b * np.sin(2 * np.pi * m * x))
    h = p.halo
    out = np.zeros((p.Nz, p.Nx + 2 * h), order="F")
    out[:, h:h + p.Nx] = v
    out[:, :h] = v[:, p.Nx - h:]
    out[:, h + p.Nx:] = v[:, :h]
    return out


def _loss_np(p, a, tt):
    h = p.halo
    th = (a.rho_theta / a.rho)[:, h:h + p.Nx]
    return float(p.dx) * a.prec[h:h + p.Nx].sum() + 0.5 * np.sum((th - tt[:, h:h + p.Nx]) ** 2)


def _adjoint_form(p, s, tt, precision, steps, ckpt):
    """Forward timing and one checkpointed adjoint run in one form."""
    eng = core.Engine(p, precision)
    eng.use_graph(True)
    eng.upload(s)
    shape = tune(eng)
    eng.upload(s)
    eng.time_steps(2)
    eng.upload(s)
    ms_plain = eng.time_steps(steps)
    eng.upload(s)
    J = eng.adjoint(steps, [tt], ckpt=ckpt)[0]
    ms_fwd, ms_rev = eng.adjoint_times()
    g, gK = eng.adjoint_grad(0)
    st_b, ad_b = eng.device_bytes()
    res = {"acoustic_shape": shape, "J": J, "dJ_dK": [float(x) for x in gK], "forward_ms": ms_plain,
           "forward_with_checkpoints_ms": ms_fwd, "reverse_ms": ms_rev, "adjoint_forward_ratio": ms_rev / ms_plain,
           "device_bytes": st_b + ad_b}
    return eng, res, g, gK


def adjoint_bench(steps=600, ckpt=50, fd_dirs=32, cpu_steps=2):
    """The 512x128 adjoint workload: loss J = dx sum(prec) + 1/2 |theta -
    theta_target|^2 after 600 steps, gradient w.r.t. the initial theta' and qv
    (at fixed density) and K_auto, Qc_threshold, K_accr; checkpoints every 50
    steps; FP64 form, and the mixed form (mixed-form trajectory, FP64 step
    adjoints along it).  Central FD of the GPU's own FP64 forward on 32
    seeded directions checks both gradients: smooth fields of unit RMS scaled
    by the RMS of the initial theta' (theta - base profile) or qv, steps
    h = 1e-4 and 1e-6 relative (an FP32 trajectory cannot resolve such steps,
    so the mixed gradient is checked against the FP64 forward's differences)."""
    p, s = ic.adjoint_bubble(np.float64, steps=steps)
    tt = ic.theta_target(p)
    eng, out, g, gK = _adjoint_form(p, s, tt, core.FP64, steps, ckpt)
    rng = np.random.default_rng(100)
    hh = p.halo
    th0 = (s.rho_theta / s.rho)[:, hh:hh + p.Nx] - ic.theta_profile(p.z)[:, None]
    qv0 = (s.rho_qv / s.rho)[:, hh:hh + p.Nx]
    rms = {3: float(np.sqrt(np.mean(th0 ** 2))), 4: float(np.sqrt(np.mean(qv0 ** 2)))}
    # central FD at two relative steps: h = 1e-4, where a sweep of h (1e-2 .. 1e-7)
    # puts the FD's own error lowest (truncation below, roundoff of the
    # 600-step losses above), and h = 1e-6
    hs = (1e-4, 1e-6)
    errs = {h: [] for h in hs}
    fds = []
    ads = []
    for d in range(fd_dirs):
        f, n = (3, "rho_theta") if d % 2 == 0 else (4, "rho_qv")
        v = _smooth_direction(p, rng)
        v *= rms[f] / np.sqrt(np.mean(v[:, hh:hh + p.Nx] ** 2))
        v = v * s.rho
        ad = float(np.sum(g[f][:, hh:hh + p.Nx] * v[:, hh:hh + p.Nx]))
        ads.append(ad)
        fd_h = {}
        for eps in hs:
            vals = []
            for sgn in (1.0, -1.0):
                sp = s.copy()
                getattr(sp, n)[...] += sgn * eps * v
                eng.upload(sp)
                eng.step(steps)
                a = s.copy()
                eng.download(a, phys=False)
                vals.append(_loss_np(p, a, tt))
            fd_h[eps] = (vals[0] - vals[1]) / (2 * eps)
            errs[eps].append(abs(fd_h[eps] - ad) / abs(ad))
        fds.append((f, v, fd_h))
    eng.close()
    out = dict({"workload": f"adjoint_bubble_{p.Nx}x{p.Nz}", "form": "fp64", "steps": steps, "ckpt": ckpt,
                "dt": float(p.dt), "ns": p.ns}, **out)

    def fd_summary(e, h):
        return {"directions": fd_dirs, "h": h, "median_rel_err": float(np.median(e)), "max_rel_err": float(max(e)),
                "above_1e-5": int(sum(x > 1e-5 for x in e))}

    out["fd_check"] = dict(fd_summary(errs[hs[0]], hs[0]), controls="theta' and qv at fixed rho",
                           direction_scale={"theta_K": rms[3], "qv": rms[4]})
    out["fd_check_small_h"] = fd_summary(errs[hs[1]], hs[1])
    # per direction (seeded, reproducible): the adjoint's directional derivative and both FDs
    out["fd_check"]["per_direction"] = {"adjoint": ads, "fd_h1e-4": [fd[hs[0]] for _, _, fd in fds],
                                        "fd_h1e-6": [fd[hs[1]] for _, _, fd in fds]}
    # mixed form
    eng, mx, gm, gKm = _adjoint_form(p, s, tt, core.MIXED, steps, ckpt)
    eng.close()
    errs_m = {h: [] for h in hs}
    ads_m = []
    for f, v, fd_h in fds:
        ad = float(np.sum(gm[f][:, hh:hh + p.Nx] * v[:, hh:hh + p.Nx]))
        ads_m.append(ad)
        for h in hs:
            errs_m[h].append(abs(fd_h[h] - ad) / abs(ad))
    mx["form"] = "mixed"
    mx["vs_fp64"] = {"J_rel": abs(mx["J"] - out["J"]) / abs(out["J"]),
                     "grad_theta_rel_l2": rel_l2(gm[3][:, hh:hh + p.Nx], g[3][:, hh:hh + p.Nx]),
                     "grad_qv_rel_l2": rel_l2(gm[4][:, hh:hh + p.Nx], g[4][:, hh:hh + p.Nx]),
                     "dJ_dK_rel": [abs(a - b) / max(abs(b), 1e-300) for a, b in zip(gKm, gK)]}
    mx["fd_check_vs_fp64_forward"] = fd_summary(errs_m[hs[0]], hs[0])
    mx["fd_check_vs_fp64_forward_small_h"] = fd_summary(errs_m[hs[1]], hs[1])
    mx["fd_check_vs_fp64_forward"]["per_direction_adjoint"] = ads_m
    out["mixed"] = mx
    if cpu_steps:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # the checker, timed here as the CPU baseline only
        threads = O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        out["cpu_oracle_threads"] = threads
        t0 = time.perf_counter()
        O.adjoint_run(s.copy(), p, cpu_steps, tt, ckpt=cpu_steps)
        cpu = (time.perf_counter() - t0) / cpu_steps
        out["cpu_oracle_ms_per_step"] = cpu * 1e3
        out["gpu_ms_per_step"] = (out["forward_with_checkpoints_ms"] + out["reverse_ms"]) / steps
        out["gpu_ms_per_step_mixed"] = (mx["forward_with_checkpoints_ms"] + mx["reverse_ms"]) / steps
    return out


def batched_adjoint_bench(members=64, steps=300, ckpt=50, seed=4, forms=("mixed", "fp64")):
    """64 bubbles at 1024x128 (dx = dz = 100 m, dt = 1 s, ns = 6) with seeded
    amplitude / position and K_auto ~ U[5e-4, 2e-3]: forward + checkpointed
    reverse sweep of every member's loss in one batched launch sequence, in
    the mixed form (FP32 bulk, FP64 tridiagonal and step adjoints) and in the
    FP64 form."""
    mem = ic.batched_adjoint(np.float64, members=members, seed=seed, steps=steps)
    cells = sum(p.Nx * p.Nz for p, _, _ in mem)
    res = {}
    for form in forms:
        eng = core.Engine([p for p, _, _ in mem], core.MIXED if form == "mixed" else core.FP64)
        eng.use_graph(True)
        for m, (p, s, _) in enumerate(mem):
            eng.upload(s, member=m)
        shape = tune(eng)
        for m, (p, s, _) in enumerate(mem):
            eng.upload(s, member=m)
        eng.time_steps(2)
        for m, (p, s, _) in enumerate(mem):
            eng.upload(s, member=m)
        ms_plain = eng.time_steps(steps)
        for m, (p, s, _) in enumerate(mem):
            eng.upload(s, member=m)
        J = eng.adjoint(steps, [t for _, _, t in mem], ckpt=ckpt)
        ms_fwd, ms_rev = eng.adjoint_times()
        _, gK0 = eng.adjoint_grad(0)
        st_b, ad_b = eng.device_bytes()
        eng.close()
        res[form] = {"workload": f"batched_adjoint_{members}x1024x128", "form": form, "members": members,
                     "cells": cells, "acoustic_shape": shape, "steps": steps, "ckpt": ckpt, "forward_ms": ms_plain,
                     "forward_with_checkpoints_ms": ms_fwd, "reverse_ms": ms_rev,
                     "adjoint_forward_ratio": ms_rev / ms_plain,
                     "gradients_per_s": members / ((ms_fwd + ms_rev) * 1e-3), "device_bytes": st_b + ad_b,
                     "device_GiB": (st_b + ad_b) / 2 ** 30, "J_member0": J[0],
                     "dJ_dK_member0": [float(x) for x in gK0]}
    out = dict(res[forms[0]])
    for f in forms[1:]:
        out[f] = res[f]
    return out


def cpu_baseline(p, s, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker, timed here as the CPU baseline only

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = O.set_threads(threads)
    ref = s.copy()
    O.rk3_step(ref, p, 1)  # warm-up
    t0 = time.perf_counter()
    rc = O.rk3_step(ref, p, steps)
    dt = time.perf_counter() - t0
    cells = p.Nx * p.Nz
    return {"value": cells * steps / dt, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"{p.Nx}x{p.Nz} cold pool, {steps} large steps after 1 warm-up step, FP64, snapshot order, "
                      f"OpenMP over columns ({dt:.1f} s)", "rc": rc}


def copy_calibration(nbytes=256 << 20):
    """One device-to-device copy of known size (the runtime's copy kernel), so
    that rocprofv3's FETCH_SIZE / WRITE_SIZE readings can be scaled to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def _run_steps(p, s, precision, steps, shape=None):
    """State of the device form after `steps` large steps from s (graph replay)."""
    eng = core.Engine(p, precision)
    eng.upload(s)
    if shape:
        eng.set_acoustic_shape(shape["levels_per_chunk"], shape["threads"])
    eng.use_graph(True)
    eng.time_steps(steps)
    out = s.copy()
    eng.download(out, phys=False)
    diag = eng.diagnostics()
    eng.close()
    return out, diag


def accuracy_leg(p, s, steps, shape_m=None, shape_d=None):
    """Stated error of the mixed form: per-field rel-L2 against the FP64 form
    after `steps` large steps from t = 0 (1200 x 0.5 s = 600 s), the drift of
    total water + rain between them, and the FP32 sensitivity floor (the FP64
    form from the initial state rounded once to FP32)."""
    out_m, diag_m = _run_steps(p, s, core.MIXED, steps, shape_m)
    out_d, diag_d = _run_steps(p, s, core.FP64, steps, shape_d)
    is0 = p.is_ - 1
    sl = np.s_[:, is0:is0 + p.Nx]
    err = {n: rel_l2(getattr(out_m, n)[sl], getattr(out_d, n)[sl]) for n in PROGNOSTIC}
    w0 = diag_d["total_water"] + diag_d["rain"]
    err["total_water_plus_rain_rel_diff"] = (diag_m["total_water"] + diag_m["rain"] - w0) / w0 if w0 else 0.0
    err["sim_seconds"] = steps * float(p.dt)
    s32 = s.copy()
    for n in PROGNOSTIC:
        getattr(s32, n)[...] = getattr(s, n).astype(np.float32).astype(np.float64)
    out_f, _ = _run_steps(p, s32, core.FP64, steps, shape_d)
    floor = {n: rel_l2(getattr(out_f, n)[sl], getattr(out_d, n)[sl]) for n in PROGNOSTIC}
    floor["what"] = "FP64 form from the FP32-rounded initial state vs FP64 form, same time"
    return err, floor


def fp32_step_floor(p, s, steps, out_d=None, deviations=False):
    """The FP64 form with the prognostic state rounded to FP32 after every
    large step (host round trip per step) against the plain FP64 form: the
    error a scheme that stores its prognostics in FP32 carries with FP64
    arithmetic everywhere else.  deviations=True stores rho and rho_theta as
    the mixed form does, as FP32 deviations from their horizontal means at
    t = 0 (an FP64 profile per level)."""
    if out_d is None:
        out_d, _ = _run_steps(p, s, core.FP64, steps)
    is0 = p.is_ - 1
    sl = np.s_[:, is0:is0 + p.Nx]
    base = {n: getattr(s, n)[sl].mean(axis=1, keepdims=True) if deviations and n in ("rho", "rho_theta") else 0.0
            for n in PROGNOSTIC}

    def rnd(st):
        for n in PROGNOSTIC:
            a = getattr(st, n)
            a[...] = (a - base[n]).astype(np.float32).astype(np.float64) + base[n]

    eng = core.Engine(p, core.FP64)
    a = s.copy()
    rnd(a)
    t0 = time.perf_counter()
    for j in range(steps):
        eng.upload(a)
        eng.step(1)
        eng.download(a, phys=False)
        rnd(a)
        if (j + 1) % 200 == 0:
            print(f"fp32 step floor: {j + 1}/{steps} steps, {time.perf_counter() - t0:.0f} s", file=sys.stderr,
                  flush=True)
    eng.close()
    res = {n: rel_l2(getattr(a, n)[sl], getattr(out_d, n)[sl]) for n in PROGNOSTIC}
    res["what"] = ("FP64 form with the state rounded to FP32 after every large step (rho, rho_theta as deviations "
                   "from their t = 0 level means) vs FP64 form" if deviations else
                   "FP64 form with the state rounded to FP32 after every large step vs FP64 form")
    res["sim_seconds"] = steps * float(p.dt)
    return res


def dropin_leg(p, s, calls=5):
    """The literal drop-in path: core.rk3_step(s, p) on a host State, one large
    step per call as the reference's driver calls rk3_step! (moist.jl:1943-1944):
    upload, step, download of the State and its diagnostics over PCIe, into a
    handle cached after the first call."""
    core.release_cached()
    a = s.copy()
    t0 = time.perf_counter()
    core.rk3_step(a, p, precision=core.MIXED)
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(calls):
        core.rk3_step(a, p, precision=core.MIXED)
    per = (time.perf_counter() - t0) / calls
    core.release_cached()
    return {"form": "mixed", "first_call_ms": first * 1e3, "ms_per_step": per * 1e3, "calls": calls,
            "what": "host State in, host State out each call (PCIe-inclusive), handle cached"}


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1200, help="large steps timed (1200 x 0.5 s = 600 s)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=8192)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=64, help="large steps of the CPU baseline (about 15 s on 16 host threads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=30, help="0 skips the 256-member ensemble")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint workloads")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default acoustic tile shape")
    ap.add_argument("--shape", type=int, nargs=2, metavar=("LEVELS", "THREADS"),
                    help="fix the acoustic tile shape of every form (no autotuning), e.g. for profiling")
    ap.add_argument("--accuracy-steps", type=int, default=1200,
                    help="large steps of the mixed-vs-FP64 accuracy leg (1200 x 0.5 s = 600 s)")
    ap.add_argument("--fp32-step-floor", action="store_true",
                    help="also run the FP64 form with the state rounded to FP32 after every step "
                         "(host round trip per step, about a minute)")
    ap.add_argument("--no-dropin", action="store_true", help="skip timing the host-State drop-in call")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device copy (counter calibration under rocprofv3)")
    ap.add_argument("--batched-profile", type=int, metavar="STEPS", default=0,
                    help="only run the 64-member batched adjoint over STEPS steps (checkpoints every 50, mixed "
                         "form; both forms without --no-fp64) and print its times; for a kernel trace")
    ap.add_argument("--adjoint-profile", type=int, metavar="STEPS", default=0,
                    help="only run the 512x128 adjoint workload over STEPS steps (checkpoints every 50 steps, "
                         "no FD check; both forms, or the mixed one with --no-fp64) and print its times; "
                         "for a kernel trace of the reverse sweep")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    global AUTOTUNE, FIXED_SHAPE
    AUTOTUNE = not args.no_autotune
    FIXED_SHAPE = tuple(args.shape) if args.shape else None
    if args.batched_profile > 0:
        forms = ("mixed",) if args.no_fp64 else ("mixed", "fp64")
        print(json.dumps(batched_adjoint_bench(steps=args.batched_profile, forms=forms)), flush=True)
        return
    if args.adjoint_profile > 0:
        n = args.adjoint_profile
        p, s = ic.adjoint_bubble(np.float64, steps=n)
        tt = ic.theta_target(p)
        res = {}
        forms = (("mixed", core.MIXED),) if args.no_fp64 else (("fp64", core.FP64), ("mixed", core.MIXED))
        for name, prec in forms:
            eng, out, _, _ = _adjoint_form(p, s, tt, prec, n, 50)
            eng.close()
            res[name] = {k: out[k] for k in ("forward_ms", "reverse_ms", "adjoint_forward_ratio", "acoustic_shape")}
        print(json.dumps({"workload": "adjoint_bubble_512x128", "steps": n, "ckpt": 50, **res}), flush=True)
        return
    if args.copy_calibration:
        copy_calibration()
    p, s = ic.coldpool(np.float64, nx=args.nx, nz=args.nz)
    cells = p.Nx * p.Nz
    K, W = args.steps, args.warmup

    eng_m, ms_m, wall_m = run_form(p, s, core.MIXED, K, W)
    mixed = {"cell_steps_per_s": cells * K / (ms_m * 1e-3), "ms_per_step": ms_m / K, "wall_s": wall_m,
             "acoustic_shape": eng_m.shape}
    eng_m.close()

    fp64 = None
    err = None
    floor = None
    if not args.no_fp64:
        eng_d, ms_d, wall_d = run_form(p, s, core.FP64, K, W)
        fp64 = {"value": cells * K / (ms_d * 1e-3), "ms_per_step": ms_d / K, "wall_s": wall_d,
                "acoustic_shape": eng_d.shape}
        eng_d.close()
        mixed["speedup_vs_fp64"] = mixed["cell_steps_per_s"] / fp64["value"]
        err, floor = accuracy_leg(p, s, args.accuracy_steps, mixed["acoustic_shape"], fp64["acoustic_shape"])
    if args.fp32_step_floor:
        floor = dict(floor or {})
        out_d, _ = _run_steps(p, s, core.FP64, args.accuracy_steps)
        floor["per_step"] = fp32_step_floor(p, s, args.accuracy_steps, out_d)
        floor["per_step_deviations"] = fp32_step_floor(p, s, args.accuracy_steps, out_d, deviations=True)
    dropin = None if args.no_dropin else dropin_leg(p, s)

    # dominant kernel: acoustic column kernel, timed live with HIP events
    kms, nlaunch = kernel_pass(p, s, core.MIXED, min(K, 40), mixed["acoustic_shape"])
    alg_bytes = ACOUSTIC_PASSES * cells * 4
    achieved = alg_bytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    sh = mixed["acoustic_shape"] or {"levels_per_chunk": 4, "threads": 1024}
    nt = sh["threads"]
    # the mixed form's acoustic kernel (work fields in level pairs)
    kname = f"mo::k_acoustic_pairs<{sh['levels_per_chunk']}, {abs(nt)}, {8 if nt < 0 else 1}, false>"
    roof = {"kernel": kname, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kname, p.Nx, p.Nz),
            "avg_launch_ms": kms, "launches": nlaunch, "alg_bytes_per_launch": alg_bytes}
    step_bytes = step_passes(p.ns) * cells * 4
    step_gbs = step_bytes / (mixed["ms_per_step"] * 1e-3) / 1e9
    roof["whole_step"] = {"field_passes": step_passes(p.ns), "alg_bytes": step_bytes, "achieved": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS}

    ens = ensemble_run(args.ensemble_steps) if args.ensemble_steps > 0 else None
    adj = badj = None
    if not args.no_adjoint:
        adj = adjoint_bench(cpu_steps=0 if args.no_cpu else 2)
        badj = batched_adjoint_bench()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(p, s, args.cpu_steps)

    line = {
        "metric": METRIC,
        "value": mixed["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 tridiagonal, base state, accumulators)",
        "data": "synthetic",
        "config": {"workload": f"coldpool_{p.Nx}x{p.Nz}", "nx": p.Nx, "nz": p.Nz, "dx": float(p.dx),
                   "dt": float(p.dt), "ns": p.ns, "form": "mixed", "hipgraph": True, "phys_diag": True},
        "fp64_form": fp64,
        "mixed_vs_fp64": err,
        "fp64_sensitivity_floor": floor,
        "mixed_wall_s": mixed["wall_s"],
        "dropin": dropin,
        "acoustic_shape": mixed["acoustic_shape"],
        "roofline": roof,
        "cpu_baseline": cpu,
        "ensemble": ens,
        "adjoint_forward_ratio": adj["adjoint_forward_ratio"] if adj else None,
        "adjoint": adj,
        "batched_adjoint": badj,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
