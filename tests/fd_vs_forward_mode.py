"""The 600-step 512x128 adjoint workload's 32 finite-difference directions,
checked against forward-mode differentiation of the oracle (CPU only; about
20 minutes on 8 threads, so a script rather than a collected test).

bench.py's adjoint leg checks the device gradient with central finite
differences of the device's own FP64 forward along 32 seeded directions and
records, per direction, the adjoint's directional derivative and both
differences (`fd_check.per_direction` of the bench record).  This script
rebuilds the same directions (same seeds and scaling), runs the oracle's
forward mode (rk3_step! on dual numbers: the ForwardDiff counterpart) along
each over the 600 steps, runs the oracle's own checkpointed adjoint once, and
prints how far each of them (oracle adjoint, device adjoint, FD at both steps)
is from the forward-mode derivative.

    python tests/fd_vs_forward_mode.py [bench record .json]
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402

import bench  # noqa: E402
from moist2d_amd import ic  # noqa: E402


def main():
    rec_path = sys.argv[1] if len(sys.argv) > 1 else os.path.join(ROOT, "profiles",
                                                                    "bench_coldpool_8192x256_with_adjoint.json")
    rec = json.load(open(rec_path))["adjoint"]
    per = rec["fd_check"]["per_direction"]
    O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    steps, ckpt, ndir = rec["steps"], rec["ckpt"], len(per["adjoint"])
    p, s = ic.adjoint_bubble(np.float64, steps=steps)
    tt = ic.theta_target(p)
    hh = p.halo
    # the directions of bench.adjoint_bench, in its order
    rng = np.random.default_rng(100)
    th0 = (s.rho_theta / s.rho)[:, hh:hh + p.Nx] - ic.theta_profile(p.z)[:, None]
    qv0 = (s.rho_qv / s.rho)[:, hh:hh + p.Nx]
    rms = {3: float(np.sqrt(np.mean(th0 ** 2))), 4: float(np.sqrt(np.mean(qv0 ** 2)))}
    dirs = []
    for d in range(ndir):
        f, n = (3, "rho_theta") if d % 2 == 0 else (4, "rho_qv")
        v = bench._smooth_direction(p, rng)
        v *= rms[f] / np.sqrt(np.mean(v[:, hh:hh + p.Nx] ** 2))
        dirs.append((f, n, v * s.rho))
    t0 = time.perf_counter()
    _, g, _ = O.adjoint_run(s.copy(), p, steps, tt, ckpt=ckpt)
    print(f"oracle adjoint: {time.perf_counter() - t0:.0f} s", flush=True)
    rel = {"oracle_adjoint": [], "device_adjoint": [], "fd_h1e-4": [], "fd_h1e-6": []}
    for d, (f, n, v) in enumerate(dirs):
        t0 = time.perf_counter()
        _, dj = O.jvp_run(s, p, steps, tt, {n: v})
        ad_o = float(np.sum(g[f][:, hh:hh + p.Nx] * v[:, hh:hh + p.Nx]))
        vals = {"oracle_adjoint": ad_o, "device_adjoint": per["adjoint"][d], "fd_h1e-4": per["fd_h1e-4"][d],
                "fd_h1e-6": per["fd_h1e-6"][d]}
        for k, x in vals.items():
            rel[k].append(abs(x - dj) / abs(dj))
        print(f"direction {d:2d} ({n}): forward mode {dj:.12e}; rel diff " +
              " ".join(f"{k} {rel[k][-1]:.1e}" for k in rel) + f" ({time.perf_counter() - t0:.0f} s)", flush=True)
    print(json.dumps({k: {"max": float(max(v)), "median": float(np.median(v)), "above_1e-5": int(sum(x > 1e-5 for x in v))}
                      for k, v in rel.items()}), flush=True)


if __name__ == "__main__":
    main()
