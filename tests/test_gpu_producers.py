"""GPU parity of the input producers through the C ABI against the oracle:
update_general_ice_model_data (src/general_ice_model_data_module.f90:30-346)
with and without the ocean flood fill (:868-986), and calc_ice_rheology
(src/thermodynamics_module.f90:641-753) for every choice.  Masks bit for bit;
thicknesses, slopes and flow factors within 1e-14 relative (the same
arithmetic without contracted multiply-adds)."""
import numpy as np
import pytest

import _oracle as O
import test_producer_oracle as T
from imau_vel import config
from imau_vel.state import PRODUCER_FIELDS, ProducerState

pytestmark = pytest.mark.gpu

MASKS = [n for n, (_, _, dt) in PRODUCER_FIELDS.items() if dt == np.int32]


@pytest.fixture(scope="module")
def ctx():
    from imau_vel import solver
    c = solver.Context()
    yield c
    c.close()


def close(a, b, rtol=1e-14):
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale) <= rtol


def with_rates_and_A(g, ps, seed):
    rng = np.random.default_rng(seed)
    ps["dHi_dt_a"] = rng.standard_normal((g.ny, g.nx))
    ps["dHb_dt_a"] = 1e-3 * rng.standard_normal((g.ny, g.nx))
    ps["mask_land_a"] = (rng.uniform(size=(g.ny, g.nx)) < 0.5).astype(np.int32)
    ps["A_flow_3D_a"] = 10.0 ** rng.uniform(-17, -16, (ps.nz, g.ny, g.nx))
    ps["A_flow_vav_a"] = 10.0 ** rng.uniform(-17, -16, (g.ny, g.nx))
    return ps


@pytest.mark.parametrize("flood", [1, 0])
@pytest.mark.parametrize("seed", [0, 1])
def test_update_general_ice_model_data_matches_oracle(ctx, seed, flood):
    g, ps = T.marine_case(nx=70 + seed, ny=45, seed=seed)
    with_rates_and_A(g, ps, seed)
    c = config.producer_config(do_ocean_floodfill=flood)
    gpu, ref = ps.copy(), ps.copy()
    ctx.update_general_ice_model_data(g, c, gpu)
    O.producers_update(g, c, ref)
    for name, (kind, is3d, dt) in PRODUCER_FIELDS.items():
        if name in ("Ti_a", "Hi_a", "Hb_a", "SL_a", "dHi_dt_a", "dHb_dt_a", "A_flow_3D_a", "A_flow_vav_a"):
            continue
        if dt == np.int32:
            np.testing.assert_array_equal(gpu[name], ref[name], err_msg=name)
        else:
            assert close(gpu[name], ref[name]), name
    assert np.any(ref["mask_gl_a"]) and np.any(ref["mask_cf_a"])


def test_flood_fill_through_a_long_winding_channel(ctx):
    """A one-cell serpentine channel of floating cells from the edge through
    the whole domain: the fill needs many rounds across tiles; the result
    equals the oracle's stack walk."""
    from imau_vel.state import Grid
    nx, ny = 150, 90
    g = Grid(nx, ny, 1e3)
    ps = ProducerState(g, 15)
    Hb = np.full((ny, nx), 50.0)
    for r in range(2, ny - 2, 4):                       # rows of channel, joined alternately at the ends
        Hb[r, 1:nx - 1] = -300.0
        c = nx - 2 if (r // 4) % 2 == 0 else 1
        Hb[r:r + 5, c] = -300.0
    Hb[2, 0] = -300.0                                   # the channel's mouth on the west edge
    ps["Hb_a"] = Hb
    gpu, ref = ps.copy(), ps.copy()
    c = config.producer_config()
    ctx.update_general_ice_model_data(g, c, gpu)
    O.producers_update(g, c, ref)
    np.testing.assert_array_equal(gpu["mask_ocean_a"], ref["mask_ocean_a"])
    assert ref["mask_ocean_a"].sum() > 1000


def test_null_outputs_are_skipped(ctx):
    g, ps = T.marine_case(seed=3)
    omit = ("TAF_a", "mask_a", "dHs_dx_cy", "dHi_dt_a", "A_flow_3D_a", "A_flow_3D_cx")
    gpu, ref = ps.copy(), ps.copy()
    gpu["TAF_a"] = -7.0
    c = config.producer_config()
    ctx.update_general_ice_model_data(g, c, gpu, omit=omit)
    O.producers_update(g, c, ref, omit=omit)
    assert np.all(gpu["TAF_a"] == -7.0)
    np.testing.assert_array_equal(gpu["mask_ocean_a"], ref["mask_ocean_a"])
    assert close(gpu["dHs_dy_cy"], ref["dHs_dy_cy"])


@pytest.mark.parametrize("rheo,margin,time", [("Huybrechts1992", "infinite_slab", 0.0), ("Huybrechts1992", "BC", 0.0),
                                              ("uniform", "infinite_slab", 0.0), ("MISMIP_mod", "BC", 20000.0)])
def test_calc_ice_rheology_matches_oracle(ctx, rheo, margin, time):
    g, ps = T.marine_case(nx=53, ny=41, seed=4)
    O.producers_update(g, config.producer_config(), ps)
    rng = np.random.default_rng(5)
    ps["Ti_a"] = rng.uniform(235.0, 273.0, (ps.nz, g.ny, g.nx))
    c = config.producer_config(choice_ice_rheology=rheo, choice_ice_margin=margin, m_enh_sheet=2.5, m_enh_shelf=0.7)
    gpu, ref = ps.copy(), ps.copy()
    ctx.calc_ice_rheology(g, c, gpu, time)
    O.calc_ice_rheology(g, c, ref, time)
    for name in ("A_flow_3D_a", "A_flow_vav_a", "A_flow_3D_cx", "A_flow_3D_cy", "A_flow_vav_cx", "A_flow_vav_cy"):
        assert close(gpu[name], ref[name], 1e-13), name
