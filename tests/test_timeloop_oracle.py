"""Pins the oracle's 'direct' time loop (run_ice_dynamics_direct,
src/ice_dynamics_module.f90:63-242, with update_ice_thickness :435-648) and
its building blocks to facts they do not restate: mass conservation over many
steps, the timers of the loop, the SIA critical time step against the
closed-form SIA column, the floating-margin fraction of hand-built fronts and
the unconnected-shelf removal against connected components."""
import numpy as np
import pytest
from scipy import ndimage

import _oracle as O
from imau_vel import config, geometry
from imau_vel.state import Grid, IceState, ProducerState

CROSS = ndimage.generate_binary_structure(2, 1)


def sia_dome(nx=21, ny=19):
    g, cfg, st = geometry.dome(nx=nx, ny=ny, dx=60e3)
    config.set_config(cfg, choice_ice_dynamics="SIA")
    return g, cfg, st


def loop_cfgs(dyn="SIA", **ts):
    tcfg = config.thickness_config(ice_thickness_west_BC="none", ice_thickness_east_BC="none",
                                   ice_thickness_south_BC="none", ice_thickness_north_BC="none",
                                   choice_ice_dynamics=dyn)
    pcfg = config.producer_config(do_ocean_floodfill=0)
    ts.setdefault("start_time_of_run", -1000.0)
    ts.setdefault("end_time_of_run", 1e6)
    scfg = config.timestep_config(**ts)
    return tcfg, pcfg, scfg


def test_fixed_geometry_steps_at_dt_max():
    """choice_ice_dynamics = 'none': no solves, the thickness is unchanged,
    and the loop takes dt_max steps (the last one shortened to t_end)."""
    g, cfg, st = sia_dome()
    config.set_config(cfg, choice_ice_dynamics="none")
    tcfg, pcfg, scfg = loop_cfgs("none", dt_max=7.0)
    ts = config.timestep_state(0.0)
    H0 = st["Hi_a"].copy()
    H = O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, st, 0.5, 0.0, ts, 50.0)
    assert ts.nsteps == 8 and ts.time == 50.0 and ts.dt == pytest.approx(1.0)
    np.testing.assert_array_equal(H, H0)


def test_sia_loop_conserves_mass_and_advances_its_timer():
    """SIA dome with a uniform surface mass balance, no thickness boundary
    rule and no floating ice: the ice volume grows by exactly SMB x area x
    elapsed time (every face flux leaves one cell and enters its neighbour),
    and each step is the SIA critical step clipped to [dt_min, dt_max]."""
    g, cfg, st = sia_dome()
    tcfg, pcfg, scfg = loop_cfgs(dt_max=5.0, dt_min=0.01)
    ts = config.timestep_state(0.0)
    smb = 0.3
    V0 = st["Hi_a"].sum()
    H = O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, st, smb, 0.0, ts, 40.0)
    assert ts.time == 40.0 and ts.nsteps >= 8
    assert np.all(H >= 0.0)
    dV = H.sum() - V0
    assert dV == pytest.approx(smb * g.nx * g.ny * 40.0, rel=1e-10)
    assert 0.01 <= ts.dt_crit_SIA <= 5.0
    assert ts.t_next_SIA >= ts.time


def test_critical_timestep_sia_closed_form():
    """A sheet-covered planar slab with the SIA column u(zeta) = 2 A (rho g H s)^3 H (1 - zeta^4) / 4:
    D = u_vav H / s with u_vav = 4/5 of the surface speed, dt = 0.9 dx^2 / (6 D);
    without a sheet the 3D advective bound 0.9 dx / max(|u| + |v|) is left."""
    nz = 41
    zeta = np.linspace(0.0, 1.0, nz)
    g = Grid(8, 7, 20e3)
    st = IceState(g, nz)
    H, s, A = 2000.0, 3e-3, 1e-16
    us = 2 * A * (910 * 9.81 * H * s) ** 3 * H / 4
    col = us * (1 - zeta ** 4)
    st["u_3D_SIA_cx"] = col[:, None, None]
    st["u_3D_cx"] = col[:, None, None]
    st["Hi_cx"] = H
    st["Hi_cy"] = H
    st["dHs_dx_cx"] = -s
    st["dHs_dy_cy"] = 1.0
    st["mask_sheet_a"] = 1
    D = 0.8 * us * H / s
    dt = O.critical_timestep_sia(g, zeta, 1e9, st)
    assert dt == pytest.approx(0.9 * g.dx ** 2 / (6 * D), rel=2e-3)
    st["mask_sheet_a"] = 0
    u3 = np.full(st["u_3D_cx"].shape, 10.0)
    u3[5, 3, 4] = -500.0
    st["u_3D_cx"] = u3
    st["v_3D_cy"] = 20.0
    assert O.critical_timestep_sia(g, zeta, 1e9, st) == pytest.approx(0.9 * g.dx / 520.0, rel=1e-14)
    assert O.critical_timestep_sia(g, zeta, 7.0, st) == pytest.approx(0.9 * 7.0, rel=1e-14)


def test_floating_margin_fraction_of_fronts():
    """A shelf calving-front cell takes H / (thickest non-front ice
    neighbour), or 1 when that neighbour is thinner or absent; other ice cells
    1, ice-free 0; the first and last rows are not written."""
    g = Grid(7, 6, 1e3)
    H = np.zeros((6, 7))
    ice = np.zeros((6, 7), dtype=np.int32)
    cf = np.zeros_like(ice)
    shelf = np.zeros_like(ice)
    H[2, 1:4] = [400.0, 300.0, 120.0]
    ice[2, 1:4] = 1
    shelf[2, 1:4] = 1
    cf[2, 3] = 1                          # front with a thicker interior neighbour
    H[4, 5] = 80.0                        # isolated front cell: no non-front neighbour
    ice[4, 5] = shelf[4, 5] = cf[4, 5] = 1
    H[3, 1], ice[3, 1], shelf[3, 1], cf[3, 1] = 900.0, 1, 1, 1   # front thicker than its neighbours
    fmf = np.full((6, 7), -1.0, order="F")
    O.floating_margin_fraction(g, H, ice, cf, shelf, fmf)
    assert fmf[2, 3] == pytest.approx(120.0 / 300.0, rel=1e-15)
    assert fmf[4, 5] == 1.0 and fmf[3, 1] == 1.0 and fmf[2, 1] == 1.0 and fmf[2, 2] == 1.0
    assert fmf[1, 0] == 0.0
    assert np.all(fmf[0] == -1.0) and np.all(fmf[-1] == -1.0)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_unconnected_shelves_are_the_unseeded_components(seed):
    """Shelf cells keep their ice if their 4-connected shelf component holds
    an interior shelf cell with a 4-neighbour sheet cell (scipy.ndimage.label)."""
    rng = np.random.default_rng(seed)
    ny, nx = 30, 34
    g = Grid(nx, ny, 1e3)
    sheet = (ndimage.gaussian_filter(rng.standard_normal((ny, nx)), 3) > 0.15).astype(np.int32)
    shelf = ((rng.uniform(size=(ny, nx)) < 0.55) & (sheet == 0)).astype(np.int32)
    H = np.where((sheet == 1) | (shelf == 1), rng.uniform(50, 500, (ny, nx)), 0.0)
    H = np.asfortranarray(H)
    lab, n = ndimage.label(shelf, structure=CROSS)
    inner = np.zeros((ny, nx), bool)
    inner[1:-1, 1:-1] = True
    seed_cells = (shelf == 1) & inner & ndimage.binary_dilation(sheet == 1, structure=CROSS)
    keep = np.isin(lab, np.unique(lab[seed_cells]))
    expect = np.where((shelf == 1) & ~keep, 0.0, H)
    O.remove_unconnected_shelves(g, H, shelf, sheet)
    np.testing.assert_array_equal(H, expect)
    assert np.any((shelf == 1) & ~keep) and np.any((shelf == 1) & keep)


def test_final_geometry_is_consistent():
    """After the loop the derived fields are those of the final thickness
    (the last update_general_ice_model_data)."""
    g, cfg, st = sia_dome()
    tcfg, pcfg, scfg = loop_cfgs(dt_max=5.0)
    ts = config.timestep_state(0.0)
    geo = ProducerState(g, cfg.nz)
    H = O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, st, 0.1, 0.0, ts, 20.0, geo=geo)
    ref = ProducerState(g, cfg.nz)
    ref["Hi_a"], ref["Hb_a"], ref["SL_a"] = H, st["Hb_a"], st["SL_a"]
    O.producers_update(g, pcfg, ref)
    for f in ("Hs_a", "mask_ice_a", "mask_sheet_a", "Hi_cx", "dHs_dx_cx", "dHs_dy_cy"):
        np.testing.assert_array_equal(geo[f], ref[f], err_msg=f)
