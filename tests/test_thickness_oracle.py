"""Pins the oracle's ice-thickness update (calc_dHi_dt,
src/ice_thickness_module.f90:25-909) and critical time step
(calc_critical_timestep_adv, src/ice_dynamics_module.f90:726-762) to facts it
does not restate: mass conservation of both schemes, the upwind closed form
in one dimension, the steady state of uniform flow, the explicit limit of the
semi-implicit scheme, the boundary rules, and the CFL formula."""
import numpy as np
import pytest

import _oracle as O
from imau_vel import config
from imau_vel.state import Grid, ThicknessState


def thickness_case(nx=24, ny=18, dx=10e3, seed=0, ice_everywhere=True, ocean=False):
    """Random smooth thickness, velocities and mass balance; no flow through
    the outer faces (the reference's fluxes stop at the domain edge)."""
    rng = np.random.default_rng(seed)
    g = Grid(nx, ny, dx)
    ts = ThicknessState(g)
    X, Y = np.meshgrid(np.linspace(0, 1, nx), np.linspace(0, 1, ny))
    H = 800.0 + 300.0 * np.sin(2.5 * X + rng.uniform(0, 3)) * np.cos(1.7 * Y + rng.uniform(0, 3))
    if not ice_everywhere:
        H = np.where((X - 0.5) ** 2 + (Y - 0.5) ** 2 < 0.12, H, 0.0)
    ts["Hi_a"] = H
    ts["Hb_a"] = 200.0
    ts["SL_a"] = 0.0
    ts["u_vav_cx"] = 150.0 * np.sin(3.0 * X[:, :-1] + 1.0) + 60.0 * rng.standard_normal((ny, nx - 1))
    ts["v_vav_cy"] = 120.0 * np.cos(2.0 * Y[:-1, :] + 0.5) + 60.0 * rng.standard_normal((ny - 1, nx))
    ts["SMB_year"] = rng.uniform(-0.5, 1.0, (ny, nx))
    ts["BMB"] = rng.uniform(-0.2, 0.0, (ny, nx))
    ts["float_margin_frac_a"] = 1.0
    ts["mask_ice_a"] = (H > 0).astype(np.int32)
    if ocean:
        ts["Hb_a"] = -2000.0
        ts["mask_shelf_a"] = (H > 0).astype(np.int32)
        ts["mask_ocean_a"] = 1
    return g, ts


def tcfg(**kw):
    kw.setdefault("ice_thickness_west_BC", "none")
    kw.setdefault("ice_thickness_east_BC", "none")
    kw.setdefault("ice_thickness_south_BC", "none")
    kw.setdefault("ice_thickness_north_BC", "none")
    return config.thickness_config(**kw)


@pytest.mark.parametrize("method", ["explicit", "semi-implicit"])
def test_mass_conservation(method):
    """Every face flux leaves one cell and enters its neighbour with the same
    value (explicit: the rescaled upwind volume; semi-implicit: fs times the
    new and 1 - fs times the old upwind flux), so without clipping the ice
    volume changes by the mass balance alone."""
    g, ts = thickness_case(seed=1)
    dt = 0.5
    O.thickness_update(g, tcfg(choice_ice_integration_method=method), ts, dt)
    mb = (ts["SMB_year"] + ts["BMB"]).sum() * dt
    dv = (ts["Hi_tplusdt_a"] - ts["Hi_a"]).sum()
    assert ts["Hi_tplusdt_a"].min() > 0.0
    assert abs(dv - mb) <= 1e-9 * np.abs(ts["Hi_a"]).sum() * 1e-3 + 1e-9 * abs(mb)
    np.testing.assert_allclose(ts["dHi_dt_a"], (ts["Hi_tplusdt_a"] - ts["Hi_a"]) / dt, rtol=1e-12, atol=1e-12)


def test_explicit_upwind_closed_form():
    """Uniform eastward flow u over a thickness profile H(x): the explicit
    scheme is first-order upwind, dH/dt = -u (H_i - H_(i-1)) / dx + MB, and the
    explicit fluxes are u H_upwind dx dt."""
    g, ts = thickness_case(seed=2)
    u = 120.0
    ts["u_vav_cx"] = u
    ts["v_vav_cy"] = 0.0
    dt = 0.2
    O.thickness_update(g, tcfg(), ts, dt)
    H = ts["Hi_a"]
    expect = -u * (H[:, 1:-1] - H[:, :-2]) / g.dx + ts["SMB_year"][:, 1:-1] + ts["BMB"][:, 1:-1]
    np.testing.assert_allclose(ts["dHi_dt_a"][:, 1:-1], expect, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(ts["Qx_cx"], u * H[:, :-1] * g.dx * dt, rtol=1e-14)


def test_uniform_flow_is_steady_and_semiimplicit_tends_to_explicit():
    """Uniform thickness carried by uniform flow stays uniform in the interior
    (both schemes, no mass balance); and as dt -> 0 the semi-implicit step
    approaches the explicit one to first order in dt (a wrong sign or weight of
    an implicit term would leave an O(1) difference in dH/dt)."""
    g, ts = thickness_case(seed=3)
    ts["SMB_year"] = 0.0
    ts["BMB"] = 0.0
    base = ts.copy()
    ts["Hi_a"] = 900.0
    ts["u_vav_cx"] = 80.0
    ts["v_vav_cy"] = -50.0
    s = ts.copy()
    O.thickness_update(g, tcfg(choice_ice_integration_method="explicit"), s, 2.0)
    np.testing.assert_allclose(s["dHi_dt_a"][1:-1, 1:-1], 0.0, atol=1e-9)
    # implicitly the thinning of the inflow edge (no flux enters the domain)
    # reaches downstream, decaying geometrically: steady ten cells away
    s = ts.copy()
    O.thickness_update(g, tcfg(choice_ice_integration_method="semi-implicit"), s, 2.0)
    np.testing.assert_allclose(s["dHi_dt_a"][1:g.ny - 10, 10:-1], 0.0, atol=1e-9)
    assert abs(s["dHi_dt_a"][1, 0]) > 1e-3
    diffs = []
    for dt in (1e-2, 1e-3):
        a, b = base.copy(), base.copy()
        O.thickness_update(g, tcfg(choice_ice_integration_method="explicit"), a, dt)
        O.thickness_update(g, tcfg(choice_ice_integration_method="semi-implicit"), b, dt)
        diffs.append(np.abs(a["dHi_dt_a"] - b["dHi_dt_a"]).max() / np.abs(a["dHi_dt_a"]).max())
    assert diffs[1] < 0.15 * diffs[0] and diffs[1] < 1e-2, diffs


def test_explicit_outflux_rescaled_to_available_ice():
    """A thin cell drained by strong outflow: its outfluxes are scaled so that
    it empties exactly (the thickness reaches zero, not below), and what its
    neighbours receive is the scaled volume (mass still conserved)."""
    g, ts = thickness_case(seed=4)
    ts["SMB_year"] = 0.0
    ts["BMB"] = 0.0
    H = ts["Hi_a"].copy()
    H[8, 10] = 5.0
    ts["Hi_a"] = H
    u = ts["u_vav_cx"].copy()
    u[8, 10], u[8, 9] = 3000.0, -3000.0
    ts["u_vav_cx"] = u
    v = ts["v_vav_cy"].copy()
    v[7, 10] = v[8, 10] = 0.0          # nothing flows in through the other faces
    ts["v_vav_cy"] = v
    dt = 5.0                           # the outflow would carry 3x the cell's ice
    O.thickness_update(g, tcfg(), ts, dt)
    assert ts["Hi_tplusdt_a"][8, 10] == pytest.approx(0.0, abs=1e-9)
    assert abs((ts["Hi_tplusdt_a"] - ts["Hi_a"]).sum()) <= 1e-9 * ts["Hi_a"].sum()


def test_boundary_rules():
    """'zero' clears one row (two under DIVA), 'fixed' keeps the old
    thickness, 'periodic' copies the opposite interior row of the old field,
    'infinite' keeps the surface elevation of the inner neighbour."""
    g, ts = thickness_case(seed=5)
    c = tcfg(ice_thickness_west_BC="zero", ice_thickness_east_BC="fixed", ice_thickness_south_BC="periodic",
             ice_thickness_north_BC="infinite", choice_ice_dynamics="DIVA")
    O.thickness_update(g, c, ts, 0.5)
    Ht, H = ts["Hi_tplusdt_a"], ts["Hi_a"]
    assert np.all(Ht[1:-1, 0] == 0.0) and np.all(Ht[1:-1, 1] == 0.0)      # west, two columns under DIVA
    np.testing.assert_array_equal(Ht[1:-1, -1], H[1:-1, -1])              # east fixed
    np.testing.assert_array_equal(Ht[0, 2:-1], H[-2, 2:-1])               # south periodic from the old field
    hs_in = Ht[-2, :] + np.maximum(ts["SL_a"][-2, :] - 910.0 / 1028.0 * Ht[-2, :], ts["Hb_a"][-2, :])
    np.testing.assert_allclose(Ht[-1, :] + ts["Hb_a"][-1, :], hs_in, rtol=1e-13)   # north: grounded, same surface
    c1 = tcfg(ice_thickness_west_BC="zero", choice_ice_dynamics="SIA/SSA")
    ts2 = thickness_case(seed=5)[1]
    O.thickness_update(g, c1, ts2, 0.5)
    assert np.all(ts2["Hi_tplusdt_a"][:, 0] == 0.0) and np.all(ts2["Hi_tplusdt_a"][1:-1, 1] > 0.0)


def test_critical_timestep_closed_form():
    g = Grid(12, 9, 5e3)
    u = np.full((9, 11), 200.0)
    v = np.full((8, 12), -100.0)
    u[4, 6] = 1700.0                   # the fastest interior point sets the step
    dt = O.critical_timestep_adv(g, u, v, 10.0)
    assert dt == pytest.approx(0.9 * 5e3 / (1700.0 + 100.0), rel=1e-14)
    assert O.critical_timestep_adv(g, u * 0 + 1e-3, v * 0, 10.0) == 10.0     # capped at dt_max
