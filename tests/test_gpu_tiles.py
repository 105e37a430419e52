"""Tile heights of the DIVA 3D viscosity kernel (nz = 15): the tiled kernel
(height 0) and the row-marching kernel (4 to 32 rows per wavefront) compute
the same viscosity, so a solve ends in the same state bit for bit whichever
height is set or the autotuner picks; heights outside the candidates are
rejected."""
import numpy as np
import pytest

from imau_vel import config, geometry, solver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = solver.Context()
    yield c
    c.close()


def two_members():
    """The dome and a ragged random cap (71 x 75: partial tiles along both
    axes, whatever the height)."""
    g1, cfg, s1 = geometry.dome()
    g2, _, s2 = geometry.random_geometry(71, 75, seed=5, cfg=cfg)
    return [g1, g2], cfg, [s1, s2]


@pytest.mark.parametrize("precision", ["fp64", "mixed_visc"])
def test_every_tile_height_gives_the_same_solve(ctx, precision):
    grids, cfg, states = two_members()
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-6, DIVA_visc_it_nit=40, precision=precision)
    assert cfg.nz == 15
    prob = solver.Problem(ctx, grids, cfg, states)
    out = {}
    try:
        for rows in (0, 4, 8, 16, 32):
            prob.set_tile_rows(rows)
            prob.reset_state()
            prob.solve(cfg)
            out[rows] = (prob.stats(), [m.copy() for m in prob.download()])
        chosen = prob.autotune(cfg)
        assert chosen in (0, 4, 8, 16, 32)
        prob.reset_state()
        prob.solve(cfg)
        out["auto"] = (prob.stats(), [m.copy() for m in prob.download()])
    finally:
        prob.close()
    its0, st0 = out[0]
    assert its0[0][0] > 2
    for key, (its, st) in out.items():
        assert its == its0, (key, its, its0)
        for m in range(len(grids)):
            for f in ("u_vav_cx", "v_vav_cy", "visc_eff_3D_a", "N_a", "F2_a"):
                np.testing.assert_array_equal(st[m][f], st0[m][f], err_msg="%s %s member %d" % (key, f, m))


def test_tile_height_outside_the_candidates_is_rejected(ctx):
    g, cfg, st = geometry.dome()
    prob = solver.Problem(ctx, [g], cfg, [st])
    try:
        with pytest.raises(solver.ImauError):
            prob.set_tile_rows(12)
    finally:
        prob.close()
