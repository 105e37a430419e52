"""The measured workloads at their stated sizes, on the GPU.

* The 64-member ensemble (n x n, n ~ U[96, 512] at 4 km, seed 0): the
  headline method (learned warm start, FP32 viscosity, FP32 multigrid inside
  FP64 GCR, Anderson(3)) brings every member to the 1e-4 FP64 residual, an
  independent FP64 evaluation of that residual agrees, the cold start of the
  same solver reaches the same answer, and on the smallest member the plain
  FP64 GPU solve and the headline solve meet the oracle's fixed point (1e-6
  and the FP32-viscosity tolerance 1e-5).
* The learned warm start on the 64 held-out 256 x 256 seeds: every member
  converges, no fallback, and the learned start needs at most 0.8x the cold
  start's Picard iterations.
"""
import os
import threading

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch  # noqa: F401  (its HIP runtime initialises before the library's)
    from imau_vel import learned, solver, train_warmstart
    ctx = solver.Context()
    yield learned, solver, train_warmstart, ctx
    ctx.close()


def headline_config():
    from imau_vel import config, geometry
    cfg = geometry.synthetic_config()
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-4, DIVA_visc_it_nit=500, DIVA_choice_matrix_solver="PETSc",
                      check_interval=8, precision="mixed_visc", mg_levels=0, ksp_type="gcr", anderson_m=3)
    return cfg


def linear_settings(states, rtol=0.3):
    for s in states:
        s.DIVA_PETSc_rtol, s.DIVA_PETSc_abstol, s.DIVA_SOR_nit = rtol, 0.0, 5000


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.timeout(900)
def test_ensemble_64_members(env):
    learned, solver, tw, ctx = env
    from imau_vel import config, geometry
    if not os.path.exists(learned.WEIGHTS):
        pytest.skip("no trained weights")
    _, members = geometry.ensemble(64, 96, 512, seed=0)
    cfg = headline_config()
    grids = [g for g, _ in members]
    k = min(range(len(members)), key=lambda q: grids[q].nx * grids[q].ny)
    print("ensemble: 64 members, cells", sum(g.nx * g.ny for g in grids), "smallest", k, grids[k].nx, flush=True)

    # the oracle's fixed point of the smallest member, in a thread (the
    # oracle's C call releases the interpreter lock) while the GPU works
    ref = members[k][1].copy()
    ocfg = headline_config()
    config.set_config(ocfg, precision="fp64", anderson_m=0, ksp_type="bicgstab", nonlinear_rtol=1e-8,
                      DIVA_visc_it_nit=3000)
    worker = threading.Thread(target=O.solve, args=("diva", grids[k], ocfg, ref, O.LIN_BANDED_LU))
    worker.start()

    states = [st.copy() for _, st in members]
    linear_settings(states)
    prob = solver.Problem(ctx, grids, cfg, states)
    ws = learned.WarmStartSolver(ctx, prob)
    its_w, res_w = ws.solve(cfg)
    print("learned: picard mean %.2f max %d, max residual %.3e, fallbacks %d" %
          (np.mean(its_w), max(its_w), max(res_w), ws.fallbacks), flush=True)
    assert max(res_w) <= 1e-4 and ws.fallbacks == 0
    learned_out = [(s["u_vav_cx"].copy(), s["v_vav_cy"].copy()) for s in prob.download()]
    # an independent FP64 evaluation of each member's residual at the result
    fcfg = headline_config()
    config.set_config(fcfg, precision="fp64")
    worst = 0.0
    for g, s in zip(grids, prob.states):
        r = ctx.nonlinear_residual(g, fcfg, s)
        worst = max(worst, r)
    print("independent FP64 residual, worst member %.3e" % worst, flush=True)
    assert worst <= 1.05e-4

    # the cold start of the same solver reaches the same solution
    prob.reset_state()
    prob.solve(cfg)
    its_c, res_c = prob.stats()
    assert max(res_c) <= 1e-4
    diff = max(max(relerr(s["u_vav_cx"], u), relerr(s["v_vav_cy"], v))
               for s, (u, v) in zip(prob.download(), learned_out))
    print("cold: picard mean %.2f; learned vs cold max rel difference %.3e" % (np.mean(its_c), diff), flush=True)
    assert diff <= 1e-2
    ws.close()
    prob.close()

    # the smallest member against the oracle's fixed point (residual 1e-8)
    gk, stk = members[k]
    plain = stk.copy()
    linear_settings([plain], rtol=0.1)
    pcfg = headline_config()
    config.set_config(pcfg, precision="fp64", anderson_m=0, ksp_type="bicgstab", nonlinear_rtol=1e-8,
                      DIVA_visc_it_nit=3000)
    ctx.solve_diva(gk, pcfg, plain)
    fast = stk.copy()
    linear_settings([fast], rtol=0.1)
    config.set_config(cfg, nonlinear_rtol=1e-8, DIVA_visc_it_nit=3000)
    ctx.solve_diva(gk, cfg, fast)
    worker.join()
    e_plain = max(relerr(plain[f], ref[f]) for f in ("u_vav_cx", "v_vav_cy", "taub_cx", "u_3D_cx"))
    e_fast = max(relerr(fast[f], ref[f]) for f in ("u_vav_cx", "v_vav_cy", "taub_cx", "u_3D_cx"))
    print("member %d vs oracle (residual %.2e): plain FP64 %.3e, headline precision %.3e" %
          (k, ref.stats["nonlinear_residual"], e_plain, e_fast), flush=True)
    assert ref.stats["nonlinear_residual"] <= 1e-8
    assert e_plain < 1e-6 and e_fast < 1e-5


@pytest.mark.timeout(600)
def test_learned_gain_on_64_heldout_seeds(env):
    """The learned configuration's evaluation set: 64 held-out 256 x 256 seeds
    (never in the training set, which starts at seed 10 000)."""
    learned, solver, tw, ctx = env
    if not os.path.exists(learned.WEIGHTS):
        pytest.skip("no trained weights")
    cfg = tw.solver_config()
    mem = [learned._unpack(learned._sample_geometry((4_000_000 + s, 256))) for s in range(64)]
    states = [st.copy() for _, st in mem]
    tw.prepare(states)
    prob = solver.Problem(ctx, [g for g, _ in mem], cfg, states)
    prob.solve(cfg)
    its_c, res_c = prob.stats()
    ws = learned.WarmStartSolver(ctx, prob)
    prob.reset_state()
    its_w, res_w = ws.solve(cfg)
    print("64 held-out seeds at 256^2: cold %.2f, learned %.2f Picard iterations; max residual %.2e / %.2e" %
          (np.mean(its_c), np.mean(its_w), max(res_c), max(res_w)), flush=True)
    assert max(res_c) <= 1e-4 and max(res_w) <= 1e-4
    assert ws.fallbacks == 0
    assert np.mean(its_w) <= 0.8 * np.mean(its_c)
    ws.close()
    prob.close()
