"""Pins the oracle's input producers -- update_general_ice_model_data with the
full masks (src/general_ice_model_data_module.f90:30-346), the ocean flood
fill (:868-986) and calc_ice_rheology (src/thermodynamics_module.f90:641-753)
-- to facts they do not restate: connected components from scipy.ndimage for
the flood fill, morphological erosion / dilation for the transition masks,
the flotation closed forms, exact slopes of a plane, and the continuity and
textbook magnitude of the Huybrechts (1992) Arrhenius law."""
import numpy as np
import pytest
from scipy import ndimage

import _oracle as O
from imau_vel import config
from imau_vel.state import Grid, ProducerState

RHO_I, RHO_W = 910.0, 1028.0
CROSS = ndimage.generate_binary_structure(2, 1)
SQUARE = ndimage.generate_binary_structure(2, 2)


def marine_case(nx=40, ny=33, seed=0):
    """Ice sheet on a bed that dips below sea level towards the domain edge,
    with an enclosed sub-glacial basin (a floating pocket that is not ocean
    under the flood fill), a floating shelf, open water and an ice-free
    island."""
    rng = np.random.default_rng(seed)
    g = Grid(nx, ny, 8e3)
    ps = ProducerState(g, 15)
    X, Y = np.meshgrid(np.linspace(-1, 1, nx), np.linspace(-1, 1, ny))
    r = np.hypot(X, Y)
    Hb = 600.0 - 1400.0 * r ** 2 + 40.0 * rng.standard_normal((ny, nx))
    basin = np.hypot(X - 0.15, Y + 0.1) < 0.18
    Hb = np.where(basin, -900.0, Hb)                     # deep enclosed basin
    H = np.clip(2200.0 * (1.0 - r ** 2), 0.0, None)
    H = np.where(basin, 300.0, H)                        # thin ice floats over it
    H = np.where(r > 0.95, 0.0, H)
    island = np.hypot(X - 0.85, Y - 0.85) < 0.12          # ice-free land in the open sea (a coast)
    Hb = np.where(island, 300.0, Hb)
    H = np.where(island, 0.0, H)
    ps["Hi_a"] = H
    ps["Hb_a"] = Hb
    ps["SL_a"] = 0.0
    return g, ps


def floating(ps):
    return ps["Hi_a"] < (ps["SL_a"] - ps["Hb_a"]) * RHO_W / RHO_I


def edge_components(mask):
    """Cells of mask connected (8 neighbours) to the domain edge."""
    lab, n = ndimage.label(mask, structure=SQUARE)
    edge = set(np.unique(np.concatenate([lab[0], lab[-1], lab[:, 0], lab[:, -1]]))) - {0}
    return np.isin(lab, list(edge))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_flood_fill_is_the_edge_connected_floating_set(seed):
    g, ps = marine_case(seed=seed)
    O.producers_update(g, config.producer_config(do_ocean_floodfill=1), ps)
    expect = edge_components(floating(ps))
    np.testing.assert_array_equal(ps["mask_ocean_a"] == 1, expect)
    np.testing.assert_array_equal(ps["mask_land_a"], 1 - ps["mask_ocean_a"])
    fl = floating(ps)
    assert np.any(fl & ~expect)          # the enclosed basin floats but is no ocean
    O.producers_update(g, config.producer_config(do_ocean_floodfill=0), ps)
    np.testing.assert_array_equal(ps["mask_ocean_a"] == 1, fl)


def test_flood_fill_crosses_diagonal_links():
    """The fill visits the 8 neighbours: a floating pocket touching the sea
    only at a corner is ocean; one separated by a land cell is not."""
    g = Grid(12, 10, 1e3)
    ps = ProducerState(g, 15)
    Hb = np.full((10, 12), 100.0)
    Hb[:, :3] = -500.0                   # sea along the west edge
    Hb[4, 3] = -500.0                    # ... reaching to (4, 3)
    Hb[5, 4] = -500.0                    # diagonal neighbour of (4, 3)
    Hb[2, 6] = -500.0                    # isolated pocket
    ps["Hb_a"] = Hb
    O.producers_update(g, config.producer_config(), ps)
    assert ps["mask_ocean_a"][5, 4] == 1
    assert ps["mask_ocean_a"][2, 6] == 0


def test_transition_masks_by_morphology():
    g, ps = marine_case(seed=3)
    O.producers_update(g, config.producer_config(), ps)
    ice = ps["mask_ice_a"] == 1
    ocean = ps["mask_ocean_a"] == 1
    land = ps["mask_land_a"] == 1
    sheet, shelf = ps["mask_sheet_a"] == 1, ps["mask_shelf_a"] == 1
    np.testing.assert_array_equal(ice, ps["Hi_a"] > 0)
    np.testing.assert_array_equal(sheet, ice & land)
    np.testing.assert_array_equal(shelf, ice & ocean)
    inner = np.zeros_like(ice)
    inner[1:-1, 1:-1] = True
    dil = lambda m: ndimage.binary_dilation(m, structure=CROSS)
    expect = {
        "mask_margin_a": ice & ~ndimage.binary_erosion(ice, structure=CROSS, border_value=1),
        "mask_coast_a": land & ~ice & dil(ocean),
        "mask_gl_a": sheet & dil(shelf),
        "mask_cf_a": ice & dil(ocean & ~ice),
    }
    for name, m in expect.items():
        np.testing.assert_array_equal(ps[name] == 1, m & inner, err_msg=name)
        assert np.any(m & inner), name                   # the case exercises every mask


def test_total_mask_precedence():
    """mask_a: land/ocean, then coast, then sheet/shelf, then margin, grounding
    line and calving front, each overriding the previous (:282-346)."""
    g, ps = marine_case(seed=4)
    O.producers_update(g, config.producer_config(), ps)
    m = ps["mask_a"]
    gl, cf, margin = ps["mask_gl_a"] == 1, ps["mask_cf_a"] == 1, ps["mask_margin_a"] == 1
    assert np.all(m[cf] == 8)
    assert np.all(m[gl & ~cf] == 7)
    assert np.all(m[margin & ~gl & ~cf] == 6)
    plain = ~(gl | cf | margin)
    assert np.all(m[plain & (ps["mask_sheet_a"] == 1)] == 3)
    assert np.all(m[plain & (ps["mask_shelf_a"] == 1)] == 4)
    assert np.all(m[plain & (ps["mask_coast_a"] == 1)] == 5)
    rest = plain & (ps["mask_ice_a"] == 0) & (ps["mask_coast_a"] == 0)
    np.testing.assert_array_equal(m[rest], np.where(ps["mask_land_a"][rest] == 1, 0, 1))
    assert np.any(gl & margin) or np.any(cf & margin)


def test_flotation_closed_forms_and_planar_slopes():
    g, ps = marine_case(seed=5)
    Hi, Hb = ps["Hi_a"].copy(), ps["Hb_a"].copy()
    O.producers_update(g, config.producer_config(), ps)
    fl = floating(ps)
    np.testing.assert_allclose(ps["Hs_a"], np.where(fl, Hi * (1 - RHO_I / RHO_W), Hi + Hb), rtol=1e-14, atol=1e-9)
    np.testing.assert_allclose(ps["TAF_a"], Hi - np.maximum(0.0, -Hb * RHO_W / RHO_I), rtol=1e-14, atol=1e-9)
    # a tilted grounded plane: every surface slope is exact on every grid
    ps2 = ProducerState(g, 15)
    X, Y = np.meshgrid(g.x, g.y)
    ps2["Hb_a"] = 500.0
    ps2["Hi_a"] = 1000.0 + 2e-3 * X - 5e-4 * Y
    O.producers_update(g, config.producer_config(), ps2)
    for f, v in (("dHs_dx_a", 2e-3), ("dHs_dy_a", -5e-4), ("dHs_dx_cx", 2e-3), ("dHs_dy_cx", -5e-4),
                 ("dHs_dx_cy", 2e-3), ("dHs_dy_cy", -5e-4), ("dHi_dx_a", 2e-3), ("dHi_dy_a", -5e-4)):
        np.testing.assert_allclose(ps2[f], v, rtol=1e-9, err_msg=f)
    np.testing.assert_allclose(ps2["Hi_cx"], 1000.0 + 2e-3 * (X[:, :-1] + g.dx / 2) - 5e-4 * Y[:, :-1], rtol=1e-14)
    np.testing.assert_allclose(ps2["Hi_b"], 1000.0 + 2e-3 * (X[:-1, :-1] + g.dx / 2) - 5e-4 * (Y[:-1, :-1] + g.dx / 2),
                               rtol=1e-14)


def test_surface_rate_uses_the_entry_land_mask():
    g, ps = marine_case(seed=6)
    ps["dHi_dt_a"] = 0.7
    ps["dHb_dt_a"] = 0.01
    entry = np.zeros((g.ny, g.nx), dtype=np.int32)
    entry[:, : g.nx // 2] = 1
    ps["mask_land_a"] = entry
    O.producers_update(g, config.producer_config(), ps)
    np.testing.assert_allclose(ps["dHs_dt_a"], np.where(entry == 1, 0.71, 0.7 * (1 - RHO_I / RHO_W)), rtol=1e-14)


def test_huybrechts_rheology_is_continuous_and_of_textbook_magnitude():
    """The two Arrhenius branches meet at 263.15 K (within 1%: the published
    constants are rounded), A rises with temperature, and A(-10 C) is within a
    factor 1.5 of the textbook 3.5e-25 s^-1 Pa^-3 (Paterson 1994)."""
    g = Grid(6, 5, 1e3)
    ps = ProducerState(g, 15)
    ps["Hi_a"] = 1000.0
    ps["Hb_a"] = 100.0
    O.producers_update(g, config.producer_config(), ps)
    T = np.linspace(230.0, 273.0, 15)
    ps["Ti_a"] = T[:, None, None]
    O.calc_ice_rheology(g, config.producer_config(), ps)
    A = ps["A_flow_3D_a"][:, 2, 2]
    assert np.all(np.diff(A) > 0)
    Tc = np.full((15, 5, 6), 263.15)
    Tc[:, :, :3] = 263.15 - 1e-9                        # the low-temperature branch
    ps["Ti_a"] = Tc
    O.calc_ice_rheology(g, config.producer_config(), ps)
    lo, hi = ps["A_flow_3D_a"][0, 2, 1], ps["A_flow_3D_a"][0, 2, 4]
    assert lo != hi and abs(lo / hi - 1.0) < 0.01
    ps["Ti_a"] = 263.15
    O.calc_ice_rheology(g, config.producer_config(), ps)
    a_sec = ps["A_flow_3D_a"][0, 2, 2] / (365.25 * 24 * 3600)
    assert 3.5e-25 / 1.5 < a_sec < 3.5e-25 * 1.5


def test_rheology_choices_enhancement_and_maps():
    g, ps = marine_case(seed=7)
    O.producers_update(g, config.producer_config(), ps)
    ps["Ti_a"] = 255.0
    c = config.producer_config(m_enh_sheet=3.0, m_enh_shelf=0.5)
    O.calc_ice_rheology(g, c, ps)
    A = ps["A_flow_3D_a"]
    base = 1.14e-5 * np.exp(-6.0e4 / (8.314 * 255.0))
    free = 1.14e-5 * np.exp(-6.0e4 / (8.314 * 263.15))      # infinite slab: ice-free cells at 263.15 K
    expect = np.where(ps["mask_sheet_a"] == 1, 3.0 * base, np.where(ps["mask_shelf_a"] == 1, 0.5 * base, free))
    np.testing.assert_allclose(A, np.broadcast_to(expect, A.shape), rtol=1e-13)
    np.testing.assert_allclose(ps["A_flow_vav_a"], expect, rtol=1e-13)       # uniform columns
    np.testing.assert_allclose(ps["A_flow_3D_cx"], 0.5 * (A[:, :, :-1] + A[:, :, 1:]), rtol=1e-14)
    np.testing.assert_allclose(ps["A_flow_vav_cy"], 0.5 * (expect[:-1] + expect[1:]), rtol=1e-13)
    O.calc_ice_rheology(g, config.producer_config(choice_ice_margin="BC"), ps)
    assert np.all(ps["A_flow_3D_a"][:, ps["mask_ice_a"] == 0] == 0.0)
    for t, a in ((0.0, 1e-16), (20000.0, 1e-17), (40000.0, 1e-16), (60000.0, 1e-16)):
        O.calc_ice_rheology(g, config.producer_config(choice_ice_rheology="MISMIP_mod"), ps, time=t)
        np.testing.assert_allclose(ps["A_flow_3D_a"][:, ps["mask_ice_a"] == 0], a, rtol=0)
    O.calc_ice_rheology(g, config.producer_config(choice_ice_rheology="uniform", uniform_flow_factor=4e-17), ps)
    np.testing.assert_allclose(ps["A_flow_vav_a"][ps["mask_ice_a"] == 0], 4e-17, rtol=1e-14)
