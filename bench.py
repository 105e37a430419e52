"""Benchmark: DIVA solves/s to a 1e-4 relative FP64 nonlinear residual on one MI355X.

Default workload ("ensemble"): 64 DIVA solves of uneven sizes, n x n with
n ~ U[96, 512] (rounded to odd), 4 km grids, seeded land-terminating random
ice caps (Matern bed, parabolic surface with fractal noise), A and the beta_sq
scale log-uniform over one decade, Weertman m = 3.  All members are
device-resident in one batched problem; one step is one complete batched solve
from the uploaded cold state: the learned warm start (network features,
forward pass, velocities onto the staggered grids, column-local completion of
viscosity and basal stress), then the
Picard loop until every member has ||A(u)u - b(u)|| / ||b(u)|| <= 1e-4 in
FP64.  value = 64 x steps / seconds.  "--method mixed" is the same solver from
the cold start.

"--workload channel": the 401 x 41 marine channel at 2 km, one solve per step.

The line also carries:
  roofline_net  the network's forward pass over the ensemble: its multiply-add
                work / wall time around a drained stream, against the dense
                BF16 matrix-core peak;
  cold_start, learned_gain
                the same solver from the cold start, and the ratio;
  roofline      the dominant kernel, k_visc (3D viscosity, one launch per
                Picard iteration): algorithmic bytes / launch duration from
                HIP events on the library stream, over the members each
                launch processed;
  roofline_smoother
                the same for one fine-level smoother sweep (k_smooth_fused);
  cpu_baseline  the CPU oracle (FP64, banded-LU linear solves) on bounded
                samples: the smallest ensemble member on one thread (value),
                the dome and the marine channel on one thread and on every
                usable core, with the core count and CPU model;
  fp64_multigrid, plain_fp64
                the same workload with the FP64 multigrid method and with the
                plain FP64 method (BiCGStab with tile Gauss-Seidel
                preconditioning, no multigrid), timed with the headline's
                warm-up and step count;
  err_vs_oracle_at_1e-4
                the headline's and the plain method's velocities against the
                oracle's own solves of the three smallest members.
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
MFMA_BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 matrix-core peak
# one fine-level smoother sweep (one k_smooth_fused launch): read u, v,
# N_a, N_b, beta_eff_cx, beta_eff_cy and the two right-hand sides, write u, v
# (10 words per cell, 4 bytes each in the mixed mode, 8 in FP64) plus the two
# row-type bytes
SWEEP_WORDS_PER_CELL = 10
SWEEP_BYTES_EXTRA = 2


def VISC_BYTES_PER_CELL(nz, word):
    """k_strain + k_visc per a-grid cell: k_strain reads u, v and writes the
    strain invariant (8 B each); k_visc reads the previous viscosity and the
    rate-factor term B (nz words each), writes the new viscosity (nz words;
    4 bytes each with FP32 viscosity, else 8), reads the strain invariant, H,
    taub_cx, taub_cy (8 B each) and the ice mask (4 B) and writes visc_int,
    N_a, F2 (8 B each)."""
    return 3 * nz * word + 3 * 8 + 4 * 8 + 4 + 3 * 8


# HBM bytes per processed cell measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on the ensemble workload (FETCH_SIZE scaled by 2.0, WRITE_SIZE by
# 1.0: the factors a 1 GiB device copy reads; profiles/ensemble_learned_kernel_stats.md),
# for the headline's FP32 viscosity; other precisions are not measured.
# Re-measure after changing either kernel.
VISC_TRAFFIC_BYTES_PER_CELL = {"mixed_visc": 309.7}
SWEEP_TRAFFIC_BYTES_PER_CELL = {"mixed_visc": 45.2}
METHODS = {
    "learned": dict(precision="mixed_visc", mg_levels=0, picard_relax_max=0.0, ksp_type=1, anderson_m=3, learned=True,
                    text="learned warm start (BF16 conv net on the matrix cores, viscosity passes), then GCR in FP64 "
                         "(operator, residuals, update) right-preconditioned by an FP32 aggregation-multigrid V-cycle; "
                         "FP32 3D viscosity; Anderson-accelerated Picard loop (3 differences); cold solve as fallback"),
    "mixed": dict(precision="mixed_visc", mg_levels=0, picard_relax_max=0.0, ksp_type=1, anderson_m=3,
                  text="GCR in FP64 (operator, residuals, update) right-preconditioned by an FP32 aggregation-"
                       "multigrid V-cycle; FP32 3D viscosity; Anderson-accelerated Picard loop (3 differences)"),
    "mixed64visc": dict(precision="mixed", mg_levels=0, picard_relax_max=0.0, ksp_type=1, anderson_m=3,
                        text="as mixed, with the viscosity in FP64"),
    "adaptive": dict(precision="mixed", mg_levels=0, picard_relax_max=1.0, ksp_type=1, anderson_m=0,
                     text="GCR in FP64 right-preconditioned by an FP32 aggregation-multigrid V-cycle; Picard loop "
                          "with adaptive relaxation (0.4 raised x1.25 per residual decrease up to 1.0)"),
    "fp64": dict(precision="fp64", mg_levels=0, picard_relax_max=0.0, ksp_type=0, anderson_m=0,
                 text="FP64 BiCGStab preconditioned by an FP64 aggregation-multigrid V-cycle; reference Picard "
                      "loop (fixed relaxation 0.4)"),
    "plain": dict(precision="fp64", mg_levels=1, picard_relax_max=0.0, ksp_type=0, anderson_m=0,
                  text="FP64 BiCGStab preconditioned by tile Gauss-Seidel (no multigrid); reference Picard loop "
                       "(fixed relaxation 0.4)"),
}


def use_method(cfg, meth, relax_max=None):
    from imau_vel import config
    config.set_config(cfg, precision=meth["precision"], mg_levels=meth["mg_levels"], ksp_type=meth["ksp_type"],
                      anderson_m=meth["anderson_m"], mg_rebuild=meth.get("mg_rebuild", 1),
                      picard_relax_max=meth["picard_relax_max"] if relax_max is None else relax_max)


def workload(name, lin_rtol=0.3, seed=0):
    from imau_vel import config, geometry
    if name == "channel":
        g, cfg, st = geometry.channel(nx=401, ny=41, dx=2e3, seed=0)
        members = [(g, st)]
        desc = "marine channel 401x41 at 2 km, DIVA cold start"
    else:
        cfg, members = geometry.ensemble(64, 96, 512, seed=seed)
        desc = "ensemble of 64 DIVA solves, n x n with n ~ U[96,512] at 4 km (seed %d), cold start" % seed
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-4, DIVA_visc_it_nit=500,
                      DIVA_choice_matrix_solver="PETSc", check_interval=8)
    for _, st in members:
        st.DIVA_PETSc_rtol = lin_rtol  # inexact Picard: each linear solve reduces its residual by this factor
        st.DIVA_PETSc_abstol = 0.0
        st.DIVA_SOR_nit = 5000        # Krylov iteration cap per linear solve
    return cfg, members, desc


def cpu_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return dict(nproc=os.cpu_count(), usable=usable, model=model)


def _oracle_concurrent(jobs):
    """Run oracle DIVA solves [(grid, cfg, state)] concurrently, one thread each
    (the oracle's C call releases the interpreter lock); returns the wall time
    of the whole set and of each solve."""
    import threading
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    errs, each = [], [0.0] * len(jobs)

    def run(q, g, c, st):
        t = time.perf_counter()
        try:
            O.solve("diva", g, c, st, O.LIN_BANDED_LU)
        except Exception as e:  # reported below
            errs.append(e)
        each[q] = time.perf_counter() - t
    threads = [threading.Thread(target=run, args=(q,) + tuple(job)) for q, job in enumerate(jobs)]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errs:
        raise errs[0]
    return time.perf_counter() - t0, each


def cpu_baseline(cfg, members, n_err=3):
    """The CPU oracle (FP64, exact banded-LU linear solves standing in for
    PETSc, one thread per solve):
      * the headline sample: the n_err smallest ensemble members solved to the
        1e-4 residual at once (one thread each); `value` is the smallest
        member's rate, and the solutions are kept for the error of the GPU
        results;
      * the dome (29 x 29) and the marine channel (401 x 41): one solve on one
        thread, and one solve per usable core concurrently (the reference's
        parallelism is MPI over columns; independent solves on every core give
        the machine's throughput)."""
    from imau_vel import config, geometry
    info = cpu_info()
    order = sorted(range(len(members)), key=lambda q: members[q][0].nx * members[q][0].ny)[:n_err]
    refs = [members[k][1].copy() for k in order]
    _, each = _oracle_concurrent([(members[k][0], cfg, r) for k, r in zip(order, refs)])
    k0, r0, dt = order[0], refs[0], each[0]
    g0 = members[k0][0]
    cb = dict(value=1.0 / dt, unit="solves/s", cores=1, kind="port", cpu=info)
    cb["sample"] = ("member %d (%dx%d, the smallest; larger members cost more) solved to the 1e-4 FP64 residual on "
                    "one thread (beside two other members on their own threads): %d Picard iterations, %.1f s" %
                    (k0, g0.nx, g0.ny, r0.stats["picard_iterations"], dt))
    # the reference configurations: 1 thread and every usable core
    ncores = max(1, min(16, info["usable"]))
    for name in ("dome", "channel"):
        if name == "dome":
            g, c2, st = geometry.dome()
        else:
            g, c2, st = geometry.channel()
        config.set_config(c2, stop_rule=1, nonlinear_rtol=1e-4, DIVA_visc_it_nit=500)
        t1, _ = _oracle_concurrent([(g, c2, st.copy())])
        tn, _ = _oracle_concurrent([(g, c2, st.copy()) for _ in range(ncores)])
        cb[name] = dict(grid="%dx%d" % (g.nx, g.ny), one_thread=1.0 / t1, all_cores=ncores / tn, cores=ncores,
                        unit="solves/s", stop="FP64 residual <= 1e-4, reference Picard relaxation, banded LU")
    return order, refs, cb


def copy_probe(nbytes):
    """Five copies of nbytes between two device buffers: the reference reading
    for the FETCH_SIZE / WRITE_SIZE counters (each copy reads and writes nbytes)."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps(dict(copy_probe_bytes=nbytes, copies=5)))


def timed_solves(prob, cfg, steps, warmup, ws=None):
    """steps complete solves from the uploaded state (with the learned warm
    start when ws is given); each solve returns only after its stream has
    drained, so wall time brackets the GPU work."""
    def step():
        prob.reset_state()
        if ws is not None:
            ws.solve(cfg)
        else:
            prob.solve(cfg)
    for _ in range(warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("ensemble", "channel"), default="ensemble")
    ap.add_argument("--seed", type=int, default=0, help="seed of the ensemble")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--method", choices=tuple(METHODS), default="learned")
    ap.add_argument("--warm-iters", type=int, default=0,
                    help="viscosity passes completing the learned warm start (warm_visc_iters)")
    ap.add_argument("--warm-column", type=int, default=9,
                    help="column-local completion iterations before those passes (warm_column_iters)")
    ap.add_argument("--warm-relax", type=float, default=0.7,
                    help="relaxation of the unmixed Picard steps of a warm-started solve")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 comparison legs")
    ap.add_argument("--no-cold", action="store_true", help="skip the cold-start leg of the learned method")
    ap.add_argument("--relax-max", type=float, default=None, help="override the headline's picard_relax_max")
    ap.add_argument("--ksp", choices=("bicgstab", "gcr"), default=None, help="override the headline's Krylov method")
    ap.add_argument("--mg-rebuild", type=int, default=None, help="override the headline's mg_rebuild")
    ap.add_argument("--anderson-m", type=int, default=None, help="override the headline's Anderson depth")
    ap.add_argument("--relax", type=float, default=None,
                    help="DIVA_visc_it_relax of the headline (its unmixed Picard steps)")
    ap.add_argument("--lin-rtol", type=float, default=0.3,
                    help="relative tolerance of each linear solve (inexact Picard; DIVA_PETSc_rtol)")
    ap.add_argument("--visc-tile", default="auto",
                    help="tile height of the DIVA 3D viscosity kernel: auto (timed on the problem before the timed "
                         "region) or 0, 4, 8, 16, 32 (0: tiled kernel, else row-marching)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run 5 plain device-to-device copies of BYTES each (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(ensemble members would be sharded across GPUs by size)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    meth = dict(METHODS[args.method])
    if meth.get("learned"):
        from imau_vel import learned   # imports torch before the library initialises HIP
    from imau_vel import config, solver

    cfg, members, desc = workload(args.workload, args.lin_rtol, args.seed)
    grids = [g for g, _ in members]
    states = [st.copy() for _, st in members]    # download() writes into these
    nsolve = len(members)
    ctx = solver.Context()
    prob = ctx.problem(grids, cfg, states)
    ws = (learned.WarmStartSolver(ctx, prob, warm_iters=args.warm_iters, column_iters=args.warm_column,
                                  relax=args.warm_relax)
          if meth.get("learned") else None)

    # headline method
    if args.mg_rebuild:
        meth["mg_rebuild"] = args.mg_rebuild
    if args.anderson_m is not None:
        meth["anderson_m"] = args.anderson_m
    if args.ksp:
        meth["ksp_type"] = 1 if args.ksp == "gcr" else 0
        meth["text"] = meth["text"].replace("GCR", "BiCGStab") if args.ksp == "bicgstab" else meth["text"]
    use_method(cfg, meth, args.relax_max)
    if args.relax is not None:
        cfg.DIVA_visc_it_relax = args.relax
    if args.visc_tile == "auto":
        visc_tile = prob.autotune(cfg)   # set-up, outside the timed region
    else:
        visc_tile = int(args.visc_tile)
        prob.set_tile_rows(visc_tile)
    dt = timed_solves(prob, cfg, args.steps, args.warmup, ws)
    t_dl = time.perf_counter()
    out = prob.download()
    t_dl = time.perf_counter() - t_dl
    dl_bytes = sum(a.nbytes for o in out for a in o.f.values())
    its = [o.stats["picard_iterations"] for o in out]
    apps = [o.stats["linear_sweeps"] for o in out]
    res = [o.stats["nonlinear_residual"] for o in out]
    # kernel timing passes over the same solves (HIP events on the library stream):
    # the 3D viscosity kernel, then one fine-level smoother sweep
    timed = {}
    for what in ("visc", "smoother"):
        ctx.kernel_timing(what)
        timed_solves(prob, cfg, args.steps, 0, ws)
        timed[what] = ctx.kernel_timing(False)
    cells = sum(g.nx * g.ny for g in grids)
    word = 4 if meth["precision"] in ("mixed", "mixed_visc") else 8
    nz = cfg.nz

    def roof(what, kernel, bytes_per_cell, traffic_per_cell):
        n, ms, cells_t = timed[what]
        avg_us = 1e3 * ms / max(n, 1)
        per_launch = bytes_per_cell * cells_t / max(n, 1)
        achieved = bytes_per_cell * cells_t / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        traffic = traffic_per_cell * cells_t / max(n, 1) if traffic_per_cell else None
        return dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=traffic, kernel=kernel, avg_launch_us=avg_us, launches=n, cells_processed=cells_t,
                    bytes_per_launch=per_launch, bytes_per_cell=bytes_per_cell)

    measured = (args.workload == "ensemble" and args.seed == 0 and args.method in ("learned", "mixed", "mixed64visc") and args.ksp is None and args.relax_max is None
                and args.lin_rtol == 0.3 and args.mg_rebuild is None and args.anderson_m is None
                and args.relax is None)
    # the traffic figures were measured with the tiled kernel (height 0)
    visc_traffic = VISC_TRAFFIC_BYTES_PER_CELL.get(meth["precision"]) if measured and visc_tile == 0 else None
    sweep_traffic = SWEEP_TRAFFIC_BYTES_PER_CELL.get(meth["precision"]) if measured else None
    vk = "k_visc" if visc_tile == 0 else "k_visc_march (%d rows)" % visc_tile
    roof_visc = roof("visc", "k_strain + %s (strain rates, 3D viscosity, N, F2: once per Picard iteration)" % vk,
                     VISC_BYTES_PER_CELL(nz, 4 if meth["precision"] == "mixed_visc" else 8), visc_traffic)
    roof_sweep = roof("smoother", "k_smooth_fused (one fine-level smoother sweep)" if meth["mg_levels"] == 0
                      else "k_prec_tile (one tile Gauss-Seidel application)",
                      SWEEP_WORDS_PER_CELL * word + SWEEP_BYTES_EXTRA, sweep_traffic if meth["mg_levels"] == 0 else None)

    dtype = {"mixed": "f32 preconditioner / f64 residual", "mixed_visc": "f32 viscosity and preconditioner / f64 "
             "residual and update"}.get(meth["precision"], "f64")
    if ws is not None:
        dtype = "bf16 network / " + dtype
    result = dict(
        metric="DIVA solves/s at 1e-4 rel residual: learned+FP32 vs plain FP64 GPU",
        value=nsolve * args.steps / dt, unit="solves/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=1e3 * dt / args.steps, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype=dtype, data="synthetic",
        config=dict(workload=desc, members=nsolve, cells=cells, method=meth["text"],
                    stop="FP64 nonlinear residual <= 1e-4 for every member", learned_warm_start=ws is not None,
                    visc_tile_rows=visc_tile),
        picard_iterations=dict(min=min(its), max=max(its), mean=float(np.mean(its))),
        preconditioner_applications=dict(min=min(apps), max=max(apps)),
        max_final_residual=max(res),
        roofline=roof_visc,
        roofline_smoother=roof_sweep,
    )
    if ws is not None:
        # the network: one forward pass over the whole ensemble (features, 10
        # convolutions, pooling, upsampling), host-timed around a drained stream
        import torch
        flops = learned.forward_flops(ws.io.plan)
        for _ in range(2):
            ws.net.forward(ws.io.plan, ws.io.features(cfg))
        ws.net.sync()
        reps = 10
        t0 = time.perf_counter()
        for _ in range(reps):
            ws.net.forward(ws.io.plan, ws.io.features(cfg))
        ws.net.sync()
        t_fwd = (time.perf_counter() - t0) / reps
        result["roofline_net"] = dict(bound="mfma", achieved=flops / t_fwd / 1e12, peak=MFMA_BF16_PEAK_TFLOPS,
                                      unit="TFLOP/s", frac=flops / t_fwd / 1e12 / MFMA_BF16_PEAK_TFLOPS, traffic=None,
                                      kernel="features + forward pass of the warm-start network (k_conv x10, "
                                             "pool, upsample), one launch sequence over the ensemble",
                                      ms_per_forward=1e3 * t_fwd, flop_per_forward=flops, pixels=ws.io.plan.npix[0])
        result["warm_start"] = dict(network=os.path.relpath(learned.WEIGHTS, os.path.dirname(__file__)),
                                    warm_visc_iters=args.warm_iters, warm_column_iters=args.warm_column,
                                    warm_relax=args.warm_relax,
                                    fallbacks=ws.fallbacks,
                                    fallback="cold solve of the whole ensemble when any member's FP64 residual "
                                             "stays above 1e-4")
    result["download"] = dict(ms=1e3 * t_dl, bytes=dl_bytes, note="every field of every member into the host "
                              "IceState arrays after the timed solves (outside the timed region)")
    if ws is not None and not args.no_cold:
        # the same solver from the cold start
        dtc = timed_solves(prob, cfg, args.steps, 0)
        itc, rc = prob.stats()
        result["cold_start"] = dict(value=nsolve * args.steps / dtc, unit="solves/s",
                                    picard_iterations=dict(min=min(itc), max=max(itc), mean=float(np.mean(itc))),
                                    max_final_residual=max(rc))
        result["learned_gain"] = result["value"] / result["cold_start"]["value"]
    if not args.no_compare:
        for key in ("fp64", "plain"):
            if key == args.method:
                continue
            m2 = METHODS[key]
            use_method(cfg, m2)
            # the same warm-up and step count as the headline
            dtp = timed_solves(prob, cfg, args.steps, args.warmup)
            outp = prob.download()
            name = "plain_fp64" if key == "plain" else "fp64_multigrid"
            result[name] = dict(value=nsolve * args.steps / dtp, unit="solves/s", method=m2["text"],
                                steps=args.steps, warmup=args.warmup,
                                picard_max=max(o.stats["picard_iterations"] for o in outp),
                                max_final_residual=max(o.stats["nonlinear_residual"] for o in outp))
        if "plain_fp64" in result:
            result["speedup_vs_plain_fp64"] = result["value"] / result["plain_fp64"]["value"]
        use_method(cfg, meth, args.relax_max)
    if not args.no_cpu_baseline:
        order, refs, cb = cpu_baseline(cfg, members)
        result["cpu_baseline"] = cb
        # the headline's solves of those members against the oracle's (both
        # stopped at 1e-4), and the plain FP64 method's
        errs = {}
        for key, m2 in (("headline", meth), ("plain_fp64", METHODS["plain"])):
            use_method(cfg, m2)
            prob.reset_state()
            if key == "headline" and ws is not None:
                ws.solve(cfg)
            else:
                prob.solve(cfg)
            out = prob.download()
            errs[key] = {"member %d (%dx%d)" % (k, grids[k].nx, grids[k].ny):
                         float(max(np.abs(out[k][f] - r[f]).max() / np.abs(r[f]).max() for f in ("u_vav_cx", "v_vav_cy")))
                         for k, r in zip(order, refs)}
        use_method(cfg, meth, args.relax_max)
        result["err_vs_oracle_at_1e-4"] = dict(errs, note="max relative velocity difference to the oracle's own solve; "
                                                            "both stop at the 1e-4 residual, so this is the spread of "
                                                            "1e-4-accurate solutions, not the fixed-point error (that is "
                                                            "tested at 1e-6 / 1e-5 at residual 1e-8)")
    prob.close()
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
