"""GPU tests of the learned warm start (include/imau_net.h, imau_vel.learned):
the network's problem I/O against the numpy restatement, what the trained
network gains over the cold start of the same solver on held-out seeds, that
both starts reach the same fixed point, and the cold fallback (forced from the
host with an iteration cap the warm-started solve cannot meet)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import net_oracle as NO  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch  # noqa: F401  (before the library: see imau_vel.learned)
    from imau_vel import learned, solver, train_warmstart
    ctx = solver.Context()
    yield learned, solver, train_warmstart, ctx
    ctx.close()


def members(learned, seeds, n):
    return [learned._unpack(learned._sample_geometry((s, n))) for s in seeds]


def problem(solver, train_warmstart, ctx, cfg, mem):
    states = [st.copy() for _, st in mem]
    train_warmstart.prepare(states)
    return solver.Problem(ctx, [g for g, _ in mem], cfg, states), states


def unbits(t):
    u = t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


def test_features_targets_and_warm_mapping_match_oracle(env):
    learned, solver, tw, ctx = env
    import torch
    cfg = tw.solver_config()
    mem = members(learned, [3_000_001, 3_000_002], 97) + members(learned, [3_000_003], 64)
    prob, states = problem(solver, tw, ctx, cfg, mem)
    prob.solve(cfg)
    states = prob.download()          # solved velocities for the targets
    net = learned.Net(ctx, seed=0)
    io = learned.ProblemIO(net, prob)
    x = unbits(io.features(cfg))
    t = torch.empty(io.plan.npix[0] * 2, dtype=torch.float32, device="cuda")
    w = torch.empty(io.plan.npix[0], dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()          # the library writes on its own stream
    io.targets(t, w)
    net.sync()
    t, w = t.cpu().numpy().astype(np.float64), w.cpu().numpy().astype(np.float64)
    k = 0
    for (g, _), st in zip(mem, states):
        n = g.nx * g.ny
        ref = NO.features(st["Hi_a"], st["Hs_a"], st["Hb_a"], st["dHs_dx_a"], st["dHs_dy_a"], st["beta_sq_a"],
                          st["A_flow_vav_a"]).reshape(-1)
        got = x[k * 16:(k + n) * 16]
        ref_b = NO.bf16(ref)
        assert np.all(np.abs(got - ref_b) <= 2.0 ** -7 * np.abs(ref_b) + 1e-30)
        tr, wr = NO.targets(st["u_vav_cx"], st["v_vav_cy"], st["Hi_a"])
        assert np.allclose(t[2 * k:2 * (k + n)], tr.reshape(-1), rtol=1e-6, atol=1e-6)
        assert np.array_equal(w[k:k + n], wr.reshape(-1))
        k += n
    # warm mapping of a known network output
    rng = np.random.default_rng(0)
    o = rng.normal(scale=3.0, size=io.plan.npix[0] * 2).astype(np.float32)
    L = learned._bind(ctx.L)
    ctx._check(L.imau_net_warm_start(ctx.h, prob.h, __import__("ctypes").byref(cfg), io.plan.h,
                                     learned._ptr(torch.from_numpy(o).to("cuda"))))
    out = prob.download()
    k = 0
    for (g, _), st in zip(mem, out):
        n = g.nx * g.ny
        u, v = NO.warm_velocities(o[2 * k:2 * (k + n)].astype(np.float64).reshape(g.nx, g.ny, 2))
        assert np.allclose(st["u_vav_cx"], u, rtol=1e-6, atol=1e-9)
        assert np.allclose(st["v_vav_cy"], v, rtol=1e-6, atol=1e-9)
        k += n
    net.release(io.plan)
    io.plan.close()
    prob.close()


def test_learned_warm_start_cuts_picard_iterations(env):
    """Held-out seeds (never in the training set) of three sizes: every member
    converges to the 1e-4 FP64 residual, with fewer Picard iterations than the
    cold start of the same solver and no fallback."""
    learned, solver, tw, ctx = env
    if not os.path.exists(learned.WEIGHTS):
        pytest.skip("no trained weights")
    cfg = tw.solver_config()
    mem = (members(learned, range(4_000_000, 4_000_012), 256) + members(learned, range(4_100_000, 4_100_004), 161)
           + members(learned, range(4_200_000, 4_200_004), 353))
    prob, _ = problem(solver, tw, ctx, cfg, mem)
    prob.solve(cfg)
    its_cold, res_cold = prob.stats()
    # both completions of the warm state: column-local (the default) and staggered passes
    for kw in (dict(), dict(warm_iters=6, column_iters=0)):
        ws = learned.WarmStartSolver(ctx, prob, **kw)
        prob.reset_state()
        its_w, res_w = ws.solve(cfg)
        print(kw, "cold", np.mean(its_cold), its_cold, "learned", np.mean(its_w), its_w)
        assert max(res_cold) <= 1e-4 and max(res_w) <= 1e-4
        assert ws.fallbacks == 0
        assert np.mean(its_w) <= 0.8 * np.mean(its_cold)
        ws.close()
    prob.close()


def test_warm_and_cold_starts_reach_the_same_fixed_point(env):
    learned, solver, tw, ctx = env
    if not os.path.exists(learned.WEIGHTS):
        pytest.skip("no trained weights")
    cfg = tw.solver_config()
    cfg.nonlinear_rtol = 1e-8
    mem = members(learned, range(4_300_000, 4_300_003), 193)
    prob, _ = problem(solver, tw, ctx, cfg, mem)
    prob.solve(cfg)
    cold = [(o["u_vav_cx"].copy(), o["v_vav_cy"].copy()) for o in prob.download()]
    ws = learned.WarmStartSolver(ctx, prob)
    prob.reset_state()
    its, res = ws.solve(cfg)
    assert max(res) <= 1e-8 and ws.fallbacks == 0
    for (uc, vc), o in zip(cold, prob.download()):
        assert np.abs(o["u_vav_cx"] - uc).max() <= 1e-5 * np.abs(uc).max()
        assert np.abs(o["v_vav_cy"] - vc).max() <= 1e-5 * np.abs(vc).max()
    ws.close()
    prob.close()


def test_fallback_to_the_cold_solve(env):
    """An iteration cap of 2 keeps the warm-started residual above 1e-4: the
    whole problem is solved again from the cold state, and the result equals a
    direct cold solve bit for bit.  That cold solve cannot meet the target
    within the cap either, so the solver reports IMAU_ERR_STALLED."""
    learned, solver, tw, ctx = env
    cfg = tw.solver_config()
    cfg.DIVA_visc_it_nit = 2
    mem = members(learned, range(4_400_000, 4_400_002), 96)
    prob, _ = problem(solver, tw, ctx, cfg, mem)
    prob.solve(cfg)
    direct = [(o["u_vav_cx"].copy(), o["v_vav_cy"].copy()) for o in prob.download()]
    net = learned.Net(ctx, seed=0)          # untrained: any warm start will do
    ws = learned.WarmStartSolver(ctx, prob, net=net, warm_iters=2, stall_window=0)
    prob.reset_state()
    with pytest.raises(solver.ImauError) as e:
        ws.solve(cfg)
    assert e.value.code == -5
    assert ws.fallbacks == 1
    for (uc, vc), o in zip(direct, prob.download()):
        assert np.array_equal(o["u_vav_cx"], uc) and np.array_equal(o["v_vav_cy"], vc)
    ws.close()
    prob.close()


def test_stalled_warm_start_restarts_cold_inside_the_loop(env):
    """The 401 x 41 marine channel (2 km, floating ice) lies outside the
    network's training distribution: from the learned start the FP64 residual
    stalls near 0.3 for tens of Picard iterations (93 iterations in all
    without the restart).  The stall detector restarts the member from the
    cold state after stall_window iterations, so the learned solve costs at
    most stall_window + 1 iterations more than the cold solve, and reaches
    the same target."""
    from imau_vel import geometry
    learned, solver, tw, ctx = env
    if not os.path.exists(learned.WEIGHTS):
        pytest.skip("no trained weights")
    g, _, st = geometry.channel()
    cfg = tw.solver_config()
    for c in ("u", "v"):
        setattr(cfg, "DIVA_boundary_BC_%s_west" % c, 1)
    cfg.DIVA_boundary_BC_v_south = cfg.DIVA_boundary_BC_v_north = 1
    cfg.DIVA_visc_it_nit = 500
    states = [st.copy()]
    tw.prepare(states)
    prob = solver.Problem(ctx, [g], cfg, states)
    prob.solve(cfg)
    its_c, res_c = prob.stats()
    ws = learned.WarmStartSolver(ctx, prob)
    prob.reset_state()
    its_w, res_w = ws.solve(cfg)
    print("channel: cold %d iterations, learned %d with %d restart(s)" % (its_c[0], its_w[0], ws.restarts))
    assert res_c[0] <= 1e-4 and res_w[0] <= 1e-4
    assert ws.restarts == 1 and ws.fallbacks == 0
    assert its_w[0] <= its_c[0] + ws.stall_window + 1
    ws.close()
    prob.close()
