"""Train the warm-start network of the DIVA solve on the GPU and write its weights.

    python -m imau_vel.train_warmstart --samples 2000 --epochs 20 --out imau_vel/warmstart_net.npz

1. Training data: `samples` seeded 256 x 256 random ice caps at 4 km (rate
   factor and friction scale log-uniform over one decade each, as the
   ensemble), solved to an FP64 nonlinear residual of 1e-4 by the library's
   batched solver; features and targets are cut from the solved problems on the
   device (imau_net_features / imau_net_targets).
2. Training: Adam on the ice-weighted squared error of the asinh-scaled a-grid
   velocities, through the library's own forward and backward kernels.
3. Evaluation on `holdout` further seeds: Picard iterations and wall time of
   the cold solve against the learned warm start.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch  # noqa: F401  (before libimauvel.so: see imau_vel.learned)

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    __package__ = "imau_vel"

from imau_vel import config, geometry, learned, solver  # noqa: E402


def solver_config(lin_rtol=0.3):
    """The headline solver: FP32 multigrid V-cycle inside FP64 GCR, FP32 3D
    viscosity, Anderson-accelerated Picard loop, stop at FP64 residual 1e-4."""
    cfg = geometry.synthetic_config()
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-4, DIVA_visc_it_nit=300, DIVA_choice_matrix_solver="PETSc",
                      check_interval=8, precision="mixed_visc", mg_levels=0, ksp_type="gcr", anderson_m=3)
    return cfg


def prepare(states, lin_rtol=0.3):
    for s in states:
        s.DIVA_PETSc_rtol = lin_rtol
        s.DIVA_PETSc_abstol = 0.0
        s.DIVA_SOR_nit = 5000


def holdout_members(n_members, n, seed0):
    out = []
    for k in range(n_members):
        g, st = learned._unpack(learned._sample_geometry((seed0 + k, n)))
        out.append((g, st))
    return out


def evaluate(ctx, net, cfg, members, warm_iters=(2, 3), log=print):
    """Cold against warm-started solves of the same members (one batched problem)."""
    import torch
    grids = [g for g, _ in members]
    states = [st.copy() for _, st in members]
    prepare(states)
    prob = solver.Problem(ctx, grids, cfg, states)
    io = learned.ProblemIO(net, prob)
    res = {}
    for label, wv in [("cold", 0)] + [("learned_w%d" % w, w) for w in warm_iters]:
        config.set_config(cfg, warm_visc_iters=wv)
        times = []
        for rep in range(2):
            prob.reset_state()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if wv:
                io.warm_start(cfg)
            prob.solve(cfg)
            times.append(time.perf_counter() - t0)
        out = prob.download(outputs=set())
        its = [o.stats["picard_iterations"] for o in out]
        r = [o.stats["nonlinear_residual"] for o in out]
        res[label] = dict(picard_mean=float(np.mean(its)), picard_max=int(max(its)), max_residual=float(max(r)),
                          seconds=float(min(times)), solves_per_s=len(members) / min(times))
        log("%s: %s" % (label, json.dumps(res[label])))
    config.set_config(cfg, warm_visc_iters=0)
    net.release(io.plan)
    io.plan.close()
    prob.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--holdout", type=int, default=64)
    ap.add_argument("--seed0", type=int, default=10_000, help="first training seed")
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", default=learned.WEIGHTS)
    ap.add_argument("--eval-only", metavar="WEIGHTS", default=None,
                    help="only evaluate these weights on the held-out seeds")
    ap.add_argument("--warm-iters", default="3,5,8", help="completion passes to evaluate (comma separated)")
    args = ap.parse_args()
    warm = tuple(int(w) for w in args.warm_iters.split(","))

    def log(msg):
        print(msg, flush=True)

    cfg = solver_config()
    ctx = solver.Context()
    if args.eval_only:
        net = learned.Net.load(ctx, args.eval_only)
        hold = holdout_members(args.holdout, args.n, seed0=args.seed0 + 1_000_000)
        evaluate(ctx, net, cfg, hold, warm_iters=warm, log=log)
        ctx.close()
        return
    t0 = time.time()
    X, T, W = learned.training_data(ctx, cfg, args.samples, n=args.n, batch=64, seed0=args.seed0, log=log,
                                    workers=args.workers, prepare=prepare)
    log("training data ready: %d samples (%.1f s)" % (args.samples, time.time() - t0))
    net = learned.Net(ctx, seed=0)
    hist = learned.train(net, X, T, W, n=args.n, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
    del X, T, W
    hold = holdout_members(args.holdout, args.n, seed0=args.seed0 + 1_000_000)
    ev = evaluate(ctx, net, cfg, hold, warm_iters=warm, log=log)
    net.save(args.out, loss_history=np.array(hist), samples=args.samples, epochs=args.epochs, n=args.n,
             seed0=args.seed0, evaluation=json.dumps(ev))
    log("weights written to %s (%.1f s total)" % (args.out, time.time() - t0))
    ctx.close()


if __name__ == "__main__":
    main()
