"""GPU parity of the 'direct' time loop (run_ice_dynamics_direct,
src/ice_dynamics_module.f90:63-242, with update_ice_thickness :435-648)
against the oracle's loop: the same number of steps and timers, the final
thickness, masks and velocities, with threshold calving (the reference
default) and with the fixed-geometry and ice-removal options.  The velocity solves are converged tightly
(FP64 residual 1e-10, Krylov 1e-12) so the two loops follow the same path."""
import numpy as np
import pytest

import _oracle as O
from imau_vel import config, geometry
from imau_vel.state import ProducerState, RefGeo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from imau_vel import solver
    c = solver.Context()
    yield c
    c.close()


def tight(cfg, st):
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-10, DIVA_visc_it_nit=400)
    st.DIVA_PETSc_rtol = 1e-12
    st.DIVA_PETSc_abstol = 0.0


def case(name):
    if name == "sia_dome":
        g, cfg, st = geometry.dome(nx=21, ny=19, dx=60e3)
        config.set_config(cfg, choice_ice_dynamics="SIA")
        return g, cfg, st, 0.3, 40.0, "none"
    if name == "hybrid_dome":
        g, cfg, st = geometry.dome(nx=21, ny=19, dx=60e3)
        config.set_config(cfg, choice_ice_dynamics="SIA/SSA")
        tight(cfg, st)
        return g, cfg, st, 0.2, 30.0, "none"
    g, cfg, st = geometry.channel(nx=61, ny=21, dx=12e3)      # marine: floating front at the east edge
    config.set_config(cfg, choice_ice_dynamics="SIA/SSA" if name == "hybrid_channel" else "SSA")
    tight(cfg, st)
    # 'zero' thickness boundaries leave H + dt (-H / dt) = +-1 ulp of ice where H was removed, and the
    # sign of that residue (a mask bit) follows the last bits of dt: where a calving front meets such a
    # boundary row, 1e-13 differences in dt flip the front's mask.  That case keeps 'infinite' ones.
    return g, cfg, st, 0.1, 20.0, "infinite" if name == "fixed_gl_pd_channel" else "zero"


def removal_inputs(g, st, name):
    """update_ice_thickness options of the channel cases and their inputs."""
    if name == "calving_channel":          # a threshold that the thinner part of the shelf front falls below
        return {"calving_threshold_thickness": 450.0}, None
    if name == "fixed_gl_pd_channel":      # grounding line held, a present-day front, a no-ice strip before it
        PD = st["Hi_a"].copy()
        PD[:, -4:] = 0.0
        noice = np.zeros(PD.shape, dtype=np.int32)
        noice[:, -6:-4] = 1
        return ({"fixed_grounding_line": 1, "remove_shelves_larger_than_PD": 1},
                RefGeo(g, mask_noice_a=noice, Hi_PD=PD, Hb_PD=st["Hb_a"]))
    if name == "shelf_removal_channel":    # all floating ice removed; the sheet held; a continental-shelf edge
        GIA = np.where(g.x[None, :] > g.x[-8], -3000.0, st["Hb_a"])
        return ({"do_remove_shelves": 1, "fixed_sheet_geometry": 1, "continental_shelf_calving": 1},
                RefGeo(g, Hi_GIAeq=np.zeros((g.ny, g.nx)), Hb_GIAeq=GIA))
    return {}, None


@pytest.mark.parametrize("name", ["sia_dome", "hybrid_dome", "ssa_channel", "hybrid_channel", "calving_channel",
                                  "fixed_gl_pd_channel", "shelf_removal_channel"])
def test_time_loop_matches_oracle(ctx, name):
    g, cfg, st, smb, t_end, bc = case(name)
    tcfg = config.thickness_config(choice_ice_dynamics=cfg.choice_ice_dynamics, ice_thickness_west_BC=bc,
                                   ice_thickness_east_BC=bc, ice_thickness_south_BC=bc, ice_thickness_north_BC=bc)
    pcfg = config.producer_config()
    opts, rgeo = removal_inputs(g, st, name)
    scfg = config.timestep_config(start_time_of_run=-1000.0, dt_max=5.0, **opts)
    gpu, ref = st.copy(), st.copy()
    ts_g, ts_r = config.timestep_state(0.0), config.timestep_state(0.0)
    geo_g, geo_r = ProducerState(g, cfg.nz), ProducerState(g, cfg.nz)
    Hg = ctx.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, gpu, smb, 0.0, ts_g, t_end, geo=geo_g, ref=rgeo)
    Hr = O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, ref, smb, 0.0, ts_r, t_end, geo=geo_r, ref=rgeo)
    if opts:  # the options bite: more ice-free cells than the same loop without them
        plain = config.timestep_config(start_time_of_run=-1000.0, dt_max=5.0)
        Hp = O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, plain, st.copy(), smb, 0.0, config.timestep_state(0.0),
                                       t_end)
        assert (Hr == 0).sum() > (Hp == 0).sum(), ((Hr == 0).sum(), (Hp == 0).sum())
    assert ts_g.nsteps == ts_r.nsteps and ts_g.nsteps >= 4, (ts_g.nsteps, ts_r.nsteps)
    assert ts_g.time == ts_r.time == t_end
    for f in ("t_next_SIA", "t_next_SSA", "dt_crit_SIA", "dt_crit_SSA", "dt"):
        assert getattr(ts_g, f) == pytest.approx(getattr(ts_r, f), rel=1e-7, abs=1e-12), f
    scale = np.abs(Hr).max()
    assert np.abs(Hg - Hr).max() <= 1e-8 * scale, np.abs(Hg - Hr).max() / scale
    for m in ("mask_ice_a", "mask_sheet_a", "mask_shelf_a", "mask_cf_a", "mask_a"):
        np.testing.assert_array_equal(geo_g[m], geo_r[m], err_msg=m)
    for f in ("float_margin_frac_a", "Hi_eff_cf_a"):
        np.testing.assert_allclose(geo_g[f][1:-1], geo_r[f][1:-1], rtol=1e-8, atol=1e-8 * scale, err_msg=f)
    if gpu.stats.get("picard_iterations", 0) > 0:   # the last solve converged (a capped loop is no parity case)
        assert gpu.stats["nonlinear_residual"] <= 1e-9, gpu.stats
    u = np.abs(ref["u_vav_cx"]).max()
    assert np.abs(gpu["u_vav_cx"] - ref["u_vav_cx"]).max() <= 1e-6 * u


@pytest.mark.parametrize("name", ["pc_sia_dome", "pc_diva_dome", "pc_ssa_channel", "pc_hybrid_dome"])
def test_pc_time_loop_matches_oracle(ctx, name):
    """The predictor-corrector loop (run_ice_dynamics_pc, :243-432) against
    the oracle's: the DIVA dome (the loop the reference runs DIVA with), the
    SSA channel with its floating front and threshold calving, and the SIA
    dome with the truncation-error step controller active.  The DIVA and SSA
    cases keep the controller off (a large pc_epsilon): its step follows
    pc_eta^0.4, a difference of two thickness estimates that rounding in the
    solves moves by more than the 1e-8 thickness bar."""
    from imau_vel.state import PcState
    if name == "pc_sia_dome":
        g, cfg, st, smb, t_end, bc = case("sia_dome")
        eps = 2e-4
    elif name == "pc_hybrid_dome":
        g, cfg, st, smb, t_end, bc = case("hybrid_dome")
        eps = 1e30
    elif name == "pc_diva_dome":
        g, cfg, st = geometry.dome(nx=21, ny=19, dx=60e3)
        tight(cfg, st)
        smb, t_end, bc, eps = 0.2, 30.0, "none", 1e30
    else:
        g, cfg, st, smb, t_end, bc = case("ssa_channel")
        eps = 1e30
    tcfg = config.thickness_config(choice_ice_dynamics=cfg.choice_ice_dynamics, ice_thickness_west_BC=bc,
                                   ice_thickness_east_BC=bc, ice_thickness_south_BC=bc, ice_thickness_north_BC=bc)
    pcfg = config.producer_config()
    scfg = config.timestep_config(start_time_of_run=-1000.0, dt_max=5.0, pc_epsilon=eps)
    gpu, ref = st.copy(), st.copy()
    ts_g, ts_r = config.timestep_state(0.0), config.timestep_state(0.0)
    pc_g, pc_r = PcState(g, scfg, time=0.0), PcState(g, scfg, time=0.0)
    geo_g, geo_r = ProducerState(g, cfg.nz), ProducerState(g, cfg.nz)
    Hg = ctx.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, gpu, smb, 0.0, ts_g, t_end, geo=geo_g, pc=pc_g)
    Hr = O.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, ref, smb, 0.0, ts_r, t_end, geo=geo_r, pc=pc_r)
    print(name, "steps", ts_g.nsteps, ts_r.nsteps, "dt_crit_ice", pc_g.s.dt_crit_ice, pc_r.s.dt_crit_ice,
          "pc_eta", pc_g.s.pc_eta, pc_r.s.pc_eta)
    assert ts_g.nsteps == ts_r.nsteps and ts_g.nsteps >= 4, (ts_g.nsteps, ts_r.nsteps)
    assert ts_g.time == ts_r.time == t_end
    for f in ("dt_crit_ice", "dt_crit_ice_prev", "pc_zeta", "t_next", "dt"):
        assert getattr(pc_g.s, f) == pytest.approx(getattr(pc_r.s, f), rel=1e-7, abs=1e-12), f
    assert pc_g.s.pc_eta == pytest.approx(pc_r.s.pc_eta, rel=1e-3)
    scale = np.abs(Hr).max()
    assert np.abs(Hg - Hr).max() <= 1e-8 * scale, np.abs(Hg - Hr).max() / scale
    rate = max(np.abs(pc_r.dHi_dt_a).max(), 1e-30)
    assert np.abs(pc_g.dHi_dt_a - pc_r.dHi_dt_a).max() <= 1e-6 * rate
    assert np.abs(pc_g.dHidt_Hn_un - pc_r.dHidt_Hn_un).max() <= 1e-6 * max(np.abs(pc_r.dHidt_Hn_un).max(), 1e-30)
    for m in ("mask_ice_a", "mask_sheet_a", "mask_shelf_a", "mask_cf_a", "mask_a"):
        np.testing.assert_array_equal(geo_g[m], geo_r[m], err_msg=m)
    if gpu.stats.get("picard_iterations", 0) > 0:
        assert gpu.stats["nonlinear_residual"] <= 1e-9, gpu.stats
    uf = "u_vav_cx"
    u = np.abs(ref[uf]).max()
    assert np.abs(gpu[uf] - ref[uf]).max() <= 1e-6 * u


def test_pc_rejects_bad_state(ctx):
    """The library rejects, before any device work, an update timer behind the
    clock and a non-positive controller state (IMAU_ERR_CONFIG), as the oracle
    does for the timer."""
    from imau_vel.state import PcState
    g, cfg, st, smb, t_end, bc = case("sia_dome")
    tcfg = config.thickness_config(choice_ice_dynamics=cfg.choice_ice_dynamics)
    pcfg = config.producer_config()
    scfg = config.timestep_config(start_time_of_run=-1000.0, dt_max=5.0)
    p = PcState(g, scfg, time=-1.0)
    with pytest.raises(RuntimeError, match="before the current time"):
        ctx.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, st.copy(), smb, 0.0, config.timestep_state(0.0), 5.0,
                                    pc=p)
    p = PcState(g, scfg, time=0.0)
    p.s.pc_eta = 0.0
    with pytest.raises(RuntimeError, match="must be positive"):
        ctx.run_ice_dynamics_direct(g, cfg, tcfg, pcfg, scfg, st.copy(), smb, 0.0, config.timestep_state(0.0), 5.0,
                                    pc=p)
