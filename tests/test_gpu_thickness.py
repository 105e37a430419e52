"""GPU parity of the ice-thickness update (calc_dHi_dt,
src/ice_thickness_module.f90:25-909) and of the advective critical time step
(src/ice_dynamics_module.f90:726-762) against the oracle, through the C ABI:

* explicit scheme: thickness change, new thickness and the rescaled face
  fluxes within 1e-12 relative (the same arithmetic, one thread per cell);
* semi-implicit scheme: the GPU's BiCGStab solved to a relative residual of
  1e-14 against the oracle's exact banded LU, within 1e-9;
* both conserve mass on the GPU as the oracle does (pinned there);
* the velocities of a device-resident problem give the same update as the
  same velocities handed over from the host, bit for bit."""
import numpy as np
import pytest

import _oracle as O
import test_thickness_oracle as T
from imau_vel import config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from imau_vel import solver
    c = solver.Context()
    yield c
    c.close()


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def shelf_case(seed=7):
    """A floating ice shelf in open ocean: calving-front cells (mask_cf) with
    partly filled cells (float_margin_frac < 1), so both schemes take their
    calving-front branches."""
    g, ts = T.thickness_case(seed=seed, ice_everywhere=False, ocean=True)
    H = ts["Hi_a"]
    ice = H > 0
    front = ice & ~(np.roll(ice, 1, 0) & np.roll(ice, -1, 0) & np.roll(ice, 1, 1) & np.roll(ice, -1, 1))
    ts["mask_cf_a"] = front.astype(np.int32)
    rng = np.random.default_rng(seed)
    ts["float_margin_frac_a"] = np.where(front, rng.uniform(0.3, 1.0, H.shape), 1.0)
    return g, ts


CASES = {"grounded": lambda: T.thickness_case(seed=11), "cap": lambda: T.thickness_case(seed=12, ice_everywhere=False),
         "shelf": shelf_case}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("method", ["explicit", "semi-implicit"])
def test_thickness_matches_oracle(ctx, name, method):
    g, ts = CASES[name]()
    c = config.thickness_config(choice_ice_integration_method=method, ice_thickness_west_BC="zero",
                                ice_thickness_east_BC="infinite", ice_thickness_south_BC="periodic",
                                ice_thickness_north_BC="fixed", dHi_choice_matrix_solver="PETSc",
                                dHi_PETSc_rtol=1e-14, dHi_PETSc_abstol=0.0)
    gpu, ref = ts.copy(), ts.copy()
    ctx.thickness_update(g, c, gpu, 2.0)
    O.thickness_update(g, c, ref, 2.0)
    tol = 1e-12 if method == "explicit" else 1e-9
    fields = ("dHi_dt_a", "Hi_tplusdt_a") + (("Qx_cx", "Qy_cy") if method == "explicit" else ())
    for f in fields:
        e = relerr(gpu[f], ref[f])
        assert e < tol, (name, method, f, e)


@pytest.mark.parametrize("method", ["explicit", "semi-implicit"])
def test_gpu_conserves_mass(ctx, method):
    g, ts = T.thickness_case(seed=13)
    c = T.tcfg(choice_ice_integration_method=method, dHi_choice_matrix_solver="PETSc", dHi_PETSc_rtol=1e-14,
               dHi_PETSc_abstol=0.0)
    ctx.thickness_update(g, c, ts, 0.5)
    mb = (ts["SMB_year"] + ts["BMB"]).sum() * 0.5
    dv = (ts["Hi_tplusdt_a"] - ts["Hi_a"]).sum()
    assert abs(dv - mb) <= 1e-9 * ts["Hi_a"].sum()


def test_sor_stop_rule_reaches_its_tolerance(ctx):
    """dHi_choice_matrix_solver = 'SOR': the solve stops once the largest
    update is below dHi_SOR_tol (m); with a tight tolerance the result is the
    exact solution."""
    g, ts = T.thickness_case(seed=14)
    c = T.tcfg(choice_ice_integration_method="semi-implicit", dHi_choice_matrix_solver="SOR", dHi_SOR_tol=1e-10,
               dHi_SOR_nit=5000)
    gpu, ref = ts.copy(), ts.copy()
    ctx.thickness_update(g, c, gpu, 2.0)
    O.thickness_update(g, c, ref, 2.0)
    assert relerr(gpu["Hi_tplusdt_a"], ref["Hi_tplusdt_a"]) < 1e-9


def test_critical_timestep(ctx):
    g, ts = T.thickness_case(seed=15)
    for u, v in ((ts["u_vav_cx"], ts["v_vav_cy"]), (ts["u_vav_cx"] * 1e-6, ts["v_vav_cy"] * 1e-6)):
        a = ctx.critical_timestep_adv(g, u, v, 10.0)
        b = O.critical_timestep_adv(g, u, v, 10.0)
        assert a == b, (a, b)


def test_problem_velocities_stay_on_the_device(ctx):
    """A DIVA solve of the dome, then the thickness update of its member from
    the problem (velocities read on the device) equals the update from the
    downloaded velocities bit for bit, and the oracle's within 1e-12."""
    from imau_vel import geometry, solver
    from imau_vel.state import ThicknessState
    g, cfg, st = geometry.dome()
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-6, DIVA_visc_it_nit=300)
    prob = solver.Problem(ctx, [g], cfg, [st])
    prob.solve(cfg)
    s = prob.download()[0]
    ts = ThicknessState(g)
    for f in ("Hi_a", "Hb_a", "SL_a", "mask_ice_a", "mask_shelf_a", "mask_ocean_a"):
        ts[f] = s[f]
    ts["u_vav_cx"], ts["v_vav_cy"] = s["u_vav_cx"], s["v_vav_cy"]
    ts["SMB_year"] = 0.3
    ts["float_margin_frac_a"] = 1.0
    c = config.thickness_config()
    a, b, ref = ts.copy(), ts.copy(), ts.copy()
    a["u_vav_cx"] = 0.0          # ignored: the problem's own velocities are used
    a["v_vav_cy"] = 0.0
    prob.thickness(0, c, a, 5.0)
    ctx.thickness_update(g, c, b, 5.0)
    O.thickness_update(g, c, ref, 5.0)
    prob.close()
    np.testing.assert_array_equal(a["Hi_tplusdt_a"], b["Hi_tplusdt_a"])
    assert relerr(b["Hi_tplusdt_a"], ref["Hi_tplusdt_a"]) < 1e-12
    assert np.abs(a["dHi_dt_a"]).max() > 0.0
