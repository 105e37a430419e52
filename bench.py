"""Benchmark: DIVA solves/s to a 1e-4 relative FP64 nonlinear residual on one MI355X.

Workload: the marine channel, 401 x 41 cells at 2 km (bed -150 - 0.5 x[km]
with 500 m Gaussian side walls, seeded thickness profile 1500 -> 300 m,
A = 2e-17, Weertman m = 3, beta_sq = 1e4), DIVA, cold start.  One step is
one complete solve_DIVA from the uploaded state (device-resident inputs; the
state reset is inside the timed region) until ||A(u)u - b(u)|| / ||b(u)||
<= 1e-4 in FP64.

Prints one JSON line (see the measurement section of DESIGN.md for the
roofline and baseline definitions).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "imau-ice_amd"))

HBM_PEAK_GBS = 8000.0
# one fine-level smoother sweep (5 colour launches of k_sor_color): read u, v,
# N_a, N_b, beta_eff_cx, beta_eff_cy and the two right-hand sides, write u, v
# (10 FP64 words per cell) plus the two row-type bytes
SWEEP_BYTES_PER_CELL_FP64 = 10 * 8 + 2


def workload():
    from imau_vel import config, geometry
    g, cfg, st = geometry.channel(nx=401, ny=41, dx=2e3, seed=0)
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-4, DIVA_visc_it_nit=500,
                      DIVA_choice_matrix_solver="PETSc", check_interval=8)
    st.DIVA_PETSc_rtol = 0.1      # inexact Picard: each linear solve reduces its residual 10x
    st.DIVA_PETSc_abstol = 0.0
    return g, cfg, st


def cpu_baseline(g, cfg, st):
    """The CPU oracle (FP64, exact banded-LU linear solves standing in for
    PETSc, 1 thread) on one full solve of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    s = st.copy()
    t0 = time.perf_counter()
    O.solve("diva", g, cfg, s, O.LIN_BANDED_LU)
    dt = time.perf_counter() - t0
    return s, dict(value=1.0 / dt, unit="solves/s", cores=1, kind="port",
                   sample="1 full DIVA solve of the 401x41 channel to 1e-4 FP64 residual "
                          "(%d Picard iterations, %.1f s)" % (s.stats["picard_iterations"], dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(ensemble members would be sharded across GPUs by size, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from imau_vel import solver

    g, cfg, st = workload()
    st0 = st.copy()                       # the uploaded initial state (download() writes into st)
    ctx = solver.Context()
    prob = ctx.problem([g], cfg, [st])
    for _ in range(args.warmup):
        prob.reset_state()
        prob.solve()
    # timed region: K complete solves, inputs resident in HBM
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prob.reset_state()
        prob.solve()
    dt = time.perf_counter() - t0
    out = prob.download()[0]
    stats = dict(out.stats)
    # kernel timing pass over the same solves (HIP events on the library stream)
    ctx.kernel_timing(True)
    for _ in range(args.steps):
        prob.reset_state()
        prob.solve()
    n_sweeps, sweep_ms = ctx.kernel_timing(False)
    avg_us = 1e3 * sweep_ms / max(n_sweeps, 1)
    cells = g.nx * g.ny
    bytes_per_launch = cells * SWEEP_BYTES_PER_CELL_FP64
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9
    res = dict(
        metric="DIVA solves/s at 1e-4 rel residual: learned+FP32 vs plain FP64 GPU",
        value=args.steps / dt, unit="solves/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=1e3 * dt / args.steps, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f64", data="synthetic",
        config=dict(workload="marine channel 401x41 at 2 km, DIVA cold start, plain FP64 GPU solve", nx=g.nx,
                    ny=g.ny, method="fp64", stop="FP64 nonlinear residual <= 1e-4"),
        picard_iterations=stats["picard_iterations"], sweeps_per_solve=stats["linear_sweeps"],
        final_residual=stats["nonlinear_residual"],
        roofline=dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                      traffic=None, kernel="k_sor_color (one 5-colour fine-level sweep)", avg_launch_us=avg_us, launches=n_sweeps,
                      bytes_per_launch=bytes_per_launch),
    )
    if not args.no_cpu_baseline:
        ref, cb = cpu_baseline(g, cfg, st0)
        res["cpu_baseline"] = cb
        u, ur = out["u_vav_cx"], ref["u_vav_cx"]
        res["err_vs_oracle"] = float(np.abs(u - ur).max() / np.abs(ur).max())
    prob.close()
    ctx.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
