"""The entry point that allocates nothing, imau_vel_problem_residual_noalloc,
against the existing imau_vel_nonlinear_residual, through the C ABI: two
calls in a row on problems of different sizes (the second smaller), each
checked member by member.  Both evaluate the same kernels in FP64 with the
same per-member reduction order, so they agree to the last bits (asserted to
1e-13 relative)."""
import numpy as np
import pytest

from imau_vel import config, geometry

pytestmark = pytest.mark.gpu


def _check(ctx, prob, grids, cfg, label):
    import torch
    n = len(grids)
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    prob.residual_noalloc(out.data_ptr(), cfg)
    torch.cuda.synchronize()          # the library writes on its own stream
    got = out.cpu().numpy()
    states = prob.download()
    ref = np.array([ctx.nonlinear_residual(g, cfg, s) for g, s in zip(grids, states)])
    print(label, "no-alloc", got, "reference entry point", ref, flush=True)
    assert np.all(np.isfinite(got)) and np.all(ref > 0)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)
    return got


def test_noalloc_residual_two_calls_second_smaller():
    from imau_vel import solver
    ctx = solver.Context()
    try:
        # call 1: two members, after a few Picard iterations (a non-trivial state)
        cases = [geometry.random_geometry(65, 65, seed=7), geometry.dome()]
        cfg = cases[0][1]
        config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-3, DIVA_visc_it_nit=6)
        grids = [c[0] for c in cases]
        prob = solver.Problem(ctx, grids, cfg, [c[2].copy() for c in cases])
        prob.solve(cfg)
        r1 = _check(ctx, prob, grids, cfg, "call 1 (65x65 + 29x29)")
        # the state is unchanged by the evaluation: a second evaluation repeats it
        r1b = _check(ctx, prob, grids, cfg, "call 1 repeated")
        np.testing.assert_array_equal(r1, r1b)
        prob.close()
        # call 2: a smaller problem (one 41 x 37 member, FP32 viscosity state)
        g2, cfg2, st2 = geometry.random_geometry(41, 37, seed=4)
        config.set_config(cfg2, stop_rule=1, nonlinear_rtol=1e-2, DIVA_visc_it_nit=4, precision="mixed_visc",
                          ksp_type="gcr")
        prob2 = solver.Problem(ctx, [g2], cfg2, [st2.copy()])
        prob2.solve(cfg2)
        _check(ctx, prob2, [g2], cfg2, "call 2 (41x37)")
        prob2.close()
    finally:
        ctx.close()
