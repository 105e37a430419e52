"""Pins the oracle's 'direct' time loop (run_ice_dynamics_direct,
src/ice_dynamics_module.f90:63-242, with update_ice_thickness :435-648) and
its building blocks to facts they do not restate: mass conservation over many
steps, the timers of the loop, the SIA critical time step against the
closed-form SIA column, the floating-margin fraction of hand-built fronts and
the unconnected-shelf removal against connected components, the threshold
calving of hand-built shelves, the fixed-geometry switches against
morphological dilation, and the ice removals."""
import numpy as np
import pytest
from scipy import ndimage

import _oracle as O
from imau_vel import config, geometry
from imau_vel.state import Grid, IceState, ProducerState, RefGeo

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
    ts.setdefault("choice_calving_law", "none")
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


def shelf_bay(profile, nx=14, ny=7):
    """Land walls in rows 0 and ny-1 and columns 0-2 (bed +100 m), a sheet of
    500 m on the land between the walls, a shelf over a 1000 m deep bay in
    columns 3 .. 3+len(profile)-1 with the given thickness per column, open
    ocean beyond; the shelf's only calving front faces the open ocean."""
    g = Grid(nx, ny, 10e3)
    cfg = geometry.synthetic_config()
    config.set_config(cfg, choice_ice_dynamics="none")
    st = IceState(g, cfg.nz)
    Hb = np.full((ny, nx), -1000.0)
    Hb[:, :3] = 100.0
    Hb[0, :] = Hb[-1, :] = 100.0
    H = np.zeros((ny, nx))
    H[1:-1, :3] = 500.0
    H[1:-1, 3:3 + len(profile)] = profile
    st["Hb_a"], st["Hi_a"] = Hb, H
    st["A_flow_3D_a"] = 1e-16
    st["A_flow_vav_a"] = 1e-16
    return g, cfg, st


def one_update(g, cfg, st, ref=None, **ts):
    """One step of the loop with fixed dynamics (choice_ice_dynamics = 'none'):
    the thickness only changes through update_ice_thickness."""
    tcfg, pcfg, _ = loop_cfgs("none")
    ts.setdefault("choice_calving_law", "none")
    scfg = config.timestep_config(start_time_of_run=-1000.0, dt_max=5.0, **ts)
    tstate = config.timestep_state(0.0)
    H = O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, st, 0.0, 0.0, tstate, 5.0, ref=ref)
    assert tstate.nsteps == 1
    return H


@pytest.mark.parametrize("thr, profile, calved", [
    (200.0, [150.0] * 8, [9, 10]),                       # thin uniform shelf: one front column per pass, two passes
    (200.0, [250.0] * 8, []),                            # thicker than the threshold: nothing
    (200.0, [300 - 25 * k for k in range(8)], [9, 10]),  # thinning shelf: front 125 (eff. 150), then 150 (eff. 175)
    (170.0, [300 - 25 * k for k in range(8)], [10]),     # ... the second front's effective 175 m survives 170 m
])
def test_threshold_calving_of_a_bay_shelf(thr, profile, calved):
    """threshold_thickness_calving applied twice per update
    (src/ice_dynamics_module.f90:574-588): a front cell goes when its effective
    thickness (the thickest non-front ice neighbour, or its own when that is
    thinner) is below the threshold; the second pass sees the new front."""
    g, cfg, st = shelf_bay(profile)
    H0 = st["Hi_a"].copy()
    H = one_update(g, cfg, st, choice_calving_law="threshold_thickness", calving_threshold_thickness=thr)
    expect = H0.copy()
    expect[:, calved] = 0.0
    np.testing.assert_array_equal(H, expect)
    Hn = one_update(g, cfg, st, choice_calving_law="none")
    np.testing.assert_array_equal(Hn, H0)


@pytest.mark.parametrize("seed", [0, 1])
def test_fixed_geometry_switches_are_dilations(seed):
    """fixed_shelf_geometry keeps H on the shelf and on sheet cells with a
    shelf cell in their 3x3 neighbourhood, fixed_sheet_geometry the converse,
    fixed_grounding_line both grounding-line sets
    (src/ice_dynamics_module.f90:460-559): morphological dilations."""
    rng = np.random.default_rng(seed)
    ny, nx = 23, 27
    g = Grid(nx, ny, 1e3)
    sheet = (ndimage.gaussian_filter(rng.standard_normal((ny, nx)), 2) > 0.0).astype(np.int32)
    shelf = ((rng.uniform(size=(ny, nx)) < 0.5) & (sheet == 0)).astype(np.int32)
    Hi = rng.uniform(10, 100, (ny, nx))
    Ht0 = rng.uniform(200, 300, (ny, nx))
    box = np.ones((3, 3), bool)
    gl_sheet = (sheet == 1) & ndimage.binary_dilation(shelf == 1, structure=box)
    gl_shelf = (shelf == 1) & ndimage.binary_dilation(sheet == 1, structure=box)
    for flags, fixed in (({"fixed_shelf_geometry": 1}, (shelf == 1) | gl_sheet),
                         ({"fixed_sheet_geometry": 1}, (sheet == 1) | gl_shelf),
                         ({"fixed_grounding_line": 1}, gl_sheet | gl_shelf)):
        sc = config.timestep_config(**flags)
        Ht = np.asfortranarray(Ht0.copy())
        O.fix_geometry(g, sc, Hi, sheet, shelf, Ht)
        np.testing.assert_array_equal(Ht, np.where(fixed, Hi, Ht0), err_msg=str(flags))
        assert fixed.any() and (~fixed).any()


def test_ice_removals():
    """mask_noice, do_remove_shelves, remove_shelves_larger_than_PD and
    continental_shelf_calving (src/ice_dynamics_module.f90:590-635): each
    removes exactly the cells its rule names, and nothing else changes."""
    g, cfg, st = shelf_bay([400.0] * 8)
    H0 = st["Hi_a"].copy()
    noice = np.zeros_like(H0, dtype=np.int32)
    noice[2:4, 1] = 1
    H = one_update(g, cfg, st.copy(), ref=RefGeo(g, mask_noice_a=noice))
    np.testing.assert_array_equal(H, np.where(noice == 1, 0.0, H0))
    H = one_update(g, cfg, st.copy(), do_remove_shelves=1)
    np.testing.assert_array_equal(H, np.where(st["Hb_a"] < 0, 0.0, H0))
    PD_Hi = H0.copy()
    PD_Hi[:, 7:] = 0.0                                    # present-day front at column 7
    H = one_update(g, cfg, st.copy(), ref=RefGeo(g, Hi_PD=PD_Hi, Hb_PD=st["Hb_a"]), remove_shelves_larger_than_PD=1)
    expect = H0.copy()
    expect[:, 7:] = 0.0
    np.testing.assert_array_equal(H, expect)
    GIA_Hb = np.where(np.arange(g.nx)[None, :] >= 9, -2500.0, -1000.0) * np.ones((g.ny, 1))
    GIA_Hb[:, :3] = 100.0
    H = one_update(g, cfg, st.copy(), ref=RefGeo(g, Hi_GIAeq=np.zeros_like(H0), Hb_GIAeq=GIA_Hb),
                   continental_shelf_calving=1, continental_shelf_min_height=-2000.0)
    expect = H0.copy()
    expect[:, 9:] = 0.0
    np.testing.assert_array_equal(H, expect)
    with pytest.raises(RuntimeError):
        one_update(g, cfg, st.copy(), remove_shelves_larger_than_PD=1)


# ---------------------------------------------------------------- predictor-corrector ('pc')
def _pc_run(dt, pc, t_end=100.0, smb=0.0, **ts):
    """The SIA dome from velocities solved at its geometry, steps fixed at dt
    (dt_min = dt_max = dt; a huge pc_epsilon keeps the controller off)."""
    from imau_vel.state import PcState
    g, cfg, st = sia_dome()
    O.solve("sia", g, cfg, st)
    ts.setdefault("pc_epsilon", 1e30)
    tcfg, pcfg, scfg = loop_cfgs(dt_max=dt, dt_min=dt, **ts)
    tst = config.timestep_state(0.0)
    p = PcState(g, scfg, time=0.0) if pc else None
    H = O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, st, smb, 0.0, tst, t_end, pc=p)
    return H, tst, p, g


def test_pc_is_second_order_in_time():
    """run_ice_dynamics_pc (src/ice_dynamics_module.f90:243-432) integrates
    dH/dt = F(H, u(H)) with a trapezoidal corrector: halving the step divides
    the error (against a run at an eighth of the smallest step) by about 4,
    where the 'direct' loop's forward step divides it by about 2.  A dropped
    or misweighted corrector term falls back to first order.  The estimate
    pc_eta ~ |H_corr - H_pred| / dt of calc_pc_truncation_error (:763-816)
    falls by about 4 as well, which the AB2 predictor's weights
    (1 + zeta/2, -zeta/2) give and a forward-Euler predictor would not (2)."""
    ratios = {}
    for pc in (False, True):
        ref = _pc_run(0.625, pc)[0]
        m = ref > 1500.0                      # the interior, away from the moving margin
        errs, etas = [], []
        for dt in (10.0, 5.0, 2.5):
            H, tst, p, _ = _pc_run(dt, pc)
            assert tst.nsteps == round(100.0 / dt)
            errs.append(np.abs(H - ref)[m].max())
            if pc:
                etas.append(p.s.pc_eta)
        ratios[pc] = [errs[k] / errs[k + 1] for k in range(2)]
        if pc:
            assert all(etas[k] / etas[k + 1] > 3.5 for k in range(2)), etas
    assert all(r > 3.5 for r in ratios[True]), ratios
    assert all(1.7 < r < 2.6 for r in ratios[False]), ratios


def test_pc_conserves_mass_and_keeps_its_state():
    """With a uniform SMB and no outflow the corrected rate (the mean of two
    conservative flux divergences) adds exactly SMB x area x time; the state
    handed back continues the run: two calls of 50 yr equal one of 100 yr."""
    smb = 0.3
    g, cfg, st = sia_dome()
    V0 = st["Hi_a"].sum()
    H, tst, p, g = _pc_run(5.0, True, smb=smb)
    assert H.sum() - V0 == pytest.approx(smb * g.nx * g.ny * 100.0, rel=1e-10)
    from imau_vel.state import PcState
    g, cfg, st = sia_dome()
    O.solve("sia", g, cfg, st)
    tcfg, pcfg, scfg = loop_cfgs(dt_max=5.0, dt_min=5.0, pc_epsilon=1e30)
    tst2, p2 = config.timestep_state(0.0), PcState(g, scfg, time=0.0)
    st["Hi_a"] = O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, st, smb, 0.0, tst2, 50.0, pc=p2)
    H2 = O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, st, smb, 0.0, tst2, 100.0, pc=p2)
    np.testing.assert_array_equal(H2, H)
    assert p2.s.pc_eta == p.s.pc_eta and tst2.nsteps == 10


def test_pc_step_controller_bounds():
    """The step of each velocity update (:316-323): within [dt_min, dt_max],
    within [0.5, 2] times the previous one and below the advective limit of
    the velocities it starts from; with a small pc_epsilon the truncation
    error, not the advective limit, sets it, and the error estimate stays
    near pc_epsilon.  Each call below runs just past one update."""
    from imau_vel.state import PcState
    g, cfg, st = sia_dome()
    O.solve("sia", g, cfg, st)
    tcfg, pcfg, scfg = loop_cfgs(dt_max=50.0, dt_min=0.01, pc_epsilon=2e-4)
    tst, p = config.timestep_state(0.0), PcState(g, scfg, time=0.0)
    seq, etas = [p.s.dt_crit_ice], []
    for k in range(25):
        adv = O.critical_timestep_adv(g, st["u_vav_cx"], st["v_vav_cy"], scfg.dt_max)
        st["Hi_a"] = O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, st, 0.0, 0.0, tst, p.s.t_next + 1e-6, pc=p)
        seq.append(p.s.dt_crit_ice)
        etas.append(p.s.pc_eta)
        assert scfg.dt_min <= seq[-1] <= scfg.dt_max
        assert 0.5 * seq[-2] * (1 - 1e-12) <= seq[-1] <= 2.0 * seq[-2] * (1 + 1e-12)
        assert seq[-1] <= max(adv, scfg.dt_min) * (1 + 1e-12)
    assert max(seq) < 0.5 * O.critical_timestep_adv(g, st["u_vav_cx"], st["v_vav_cy"], scfg.dt_max)
    assert 0.2 * scfg.pc_epsilon < np.median(etas[5:]) < 5.0 * scfg.pc_epsilon, etas


def test_pc_none_dynamics_keeps_the_geometry():
    g, cfg, st = sia_dome()
    config.set_config(cfg, choice_ice_dynamics="none")
    from imau_vel.state import PcState
    tcfg, pcfg, scfg = loop_cfgs("none", dt_max=7.0)
    tst, p = config.timestep_state(0.0), PcState(g, scfg, time=0.0)
    H0 = st["Hi_a"].copy()
    H = O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, st, 0.5, 0.0, tst, 50.0, pc=p)
    assert tst.nsteps == 8 and tst.time == 50.0
    np.testing.assert_array_equal(H, H0)


def test_pc_rejects_a_timer_behind_the_clock():
    """The update timer must not lie before the current time (the loop would
    step backwards); the state's rate fields must be given."""
    from imau_vel.state import PcState
    g, cfg, st = sia_dome()
    tcfg, pcfg, scfg = loop_cfgs(dt_max=5.0)
    p = PcState(g, scfg, time=-1.0)
    with pytest.raises(RuntimeError, match="before the current time"):
        O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, st, 0.0, 0.0, config.timestep_state(0.0), 10.0, pc=p)
