"""The multigrid V-cycle replayed from HIP graphs (mg_graph = 1) gives the
same solves, bit for bit, as the V-cycle launched kernel by kernel
(mg_graph = 0): a batch of a dome and a ragged cap, and the marine channel,
in both precisions and with both Krylov methods.  Each problem is solved
several times in a row (the graphs are captured in the first V-cycles of a
solve and replayed for the rest, with the solver's own work and host
synchronisations between the replays), and every solve is compared."""
import numpy as np
import pytest

from imau_vel import config, geometry, solver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = solver.Context()
    yield c
    c.close()


def members(kind):
    if kind == "batch":
        g1, cfg, s1 = geometry.dome()
        g2, _, s2 = geometry.random_geometry(71, 75, seed=5, cfg=cfg)
        return [g1, g2], cfg, [s1, s2]
    g, cfg, s = geometry.channel(nx=121, ny=41, dx=6e3)
    return [g], cfg, [s]


@pytest.mark.parametrize("kind", ["batch", "channel"])
@pytest.mark.parametrize("precision,ksp", [("mixed_visc", "gcr"), ("fp64", "bicgstab")])
def test_graph_vcycle_equals_kernel_launches(ctx, kind, precision, ksp):
    runs = {}
    for graph in (0, 1):
        grids, cfg, states = members(kind)
        config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-8, DIVA_visc_it_nit=200, precision=precision,
                          DIVA_choice_matrix_solver="PETSc", mg_levels=0, ksp_type=ksp, mg_graph=graph)
        for s in states:
            s.DIVA_PETSc_rtol, s.DIVA_PETSc_abstol = 0.3, 0.0
        prob = solver.Problem(ctx, grids, cfg, states)
        out = []
        try:
            for rep in range(3):
                prob.reset_state()
                prob.solve(cfg)
                out.append((prob.stats(), [m.copy() for m in prob.download()]))
        finally:
            prob.close()
        runs[graph] = out
    for rep in range(3):
        (s0, m0), (s1, m1) = runs[0][rep], runs[1][rep]
        assert s0 == s1, (rep, s0, s1)
        for a, b in zip(m0, m1):
            for k in ("u_vav_cx", "v_vav_cy", "visc_eff_3D_a", "w_3D_a"):
                np.testing.assert_array_equal(a[k], b[k], err_msg="%s rep %d" % (k, rep))
    its = runs[1][0][0][0][0]
    assert its > 2
