"""Pins the CPU oracle (the checker) to facts it does not restate itself.

The reference ships no test suite or golden vectors for this path, and its
Fortran cannot be built here (it needs MPI, NetCDF and PETSc).  Each oracle
function is therefore checked against something independent of its own
formula: closed forms, physical limits, brute force, a library routine
(scipy's sparse direct solver), symmetries and invariants.
"""
import ctypes
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import _oracle as O
from imau_vel import config, geometry
from imau_vel.state import Grid, IceState

ZETA = config.DEFAULT_ZETA


# ---------------------------------------------------------------- vertical integrals
def test_vertical_integrals_exact_for_linear_profiles():
    # the trapezoid rule is exact for f linear in zeta on the non-uniform grid
    z = np.array(ZETA)
    f = 3.0 - 2.0 * z
    assert O.vertical_integrate(z, f) == pytest.approx(3.0 - 1.0, abs=1e-15)
    assert O.vertical_average(z, f) == pytest.approx(2.0, abs=1e-15)
    # bottom-to-zeta integral is negative: int_1^zeta f = -(F(1) - F(zeta))
    out = O.integrate_bottom_to_zeta(z, f)
    expect = -((3.0 * 1 - 1.0) - (3.0 * z - z ** 2))
    np.testing.assert_allclose(out, expect, atol=1e-14)


def test_vertical_integral_converges_to_closed_form():
    # int_0^1 zeta^4 = 1/5 ; the trapezoid error decreases as dz^2
    errs = []
    for n in (8, 16, 32):
        z = np.linspace(0.0, 1.0, n)
        errs.append(abs(O.vertical_integrate(z, z ** 4) - 0.2))
    assert errs[1] < errs[0] / 3.5 and errs[2] < errs[1] / 3.5


# ---------------------------------------------------------------- grounded fractions
def test_fraction_above_zero_trivial_cases():
    assert O.fraction_above_zero(1.0, 2.0, 3.0, 4.0) == 1.0
    assert O.fraction_above_zero(-1.0, -2.0, -3.0, -4.0) == 0.0


@pytest.mark.parametrize("a,b", [(1.0, 0.3), (2.0, -1.0), (0.5, 1.7)])
def test_fraction_above_zero_half_plane(a, b):
    # f(x,y) = a*(x - x0) with corners in x only: exact fraction is the area right of x0
    x0 = 0.37
    f = lambda x, y: a * (x - x0) + 0.0 * y + 0.0 * b
    phi = O.fraction_above_zero(f(0, 1), f(1, 1), f(0, 0), f(1, 0))
    assert phi == pytest.approx(1.0 - x0, abs=2e-3)


def _mc_fraction(nw, ne, sw, se, n=400):
    xs = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xs, xs)
    f = sw * (1 - X) * (1 - Y) + se * X * (1 - Y) + nw * (1 - X) * Y + ne * X * Y
    return float((f > 0).mean())


def test_fraction_above_zero_brute_force():
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(300):
        c = rng.uniform(-1.0, 1.0, 4)
        c = np.where(np.abs(c) < 0.05, np.sign(c) * 0.05 + 0.0, c)
        phi = O.fraction_above_zero(*c)
        worst = max(worst, abs(phi - _mc_fraction(*c)))
    assert worst < 5e-3, worst


# ---------------------------------------------------------------- sliding laws
def _cfg(**kw):
    return config.default_config(**kw)


@pytest.mark.parametrize("u", [10.0, 200.0, 3000.0])
def test_weertman_stress_velocity_relation(u):
    # tau_b = beta |u| = beta_sq |u|^(1/m)  (power law)
    c = _cfg(choice_sliding_law="Weertman")
    beta = O.sliding_beta(c, u, 0.0, beta_sq=1e4)
    assert beta * u == pytest.approx(1e4 * u ** (1.0 / 3.0), rel=1e-7)  # delta_v = 1e-3 regularisation


def test_coulomb_laws_plastic_limits():
    c = _cfg(choice_sliding_law="Coulomb")
    for u in (1.0, 50.0, 900.0):
        # |u| is regularised as sqrt(delta_v^2 + u^2), delta_v = 1e-3
        assert O.sliding_beta(c, u, 0.0, tauc=7e4) * u == pytest.approx(7e4, rel=1e-6)
    c = _cfg(choice_sliding_law="Coulomb_regularised")
    # tau = tauc (|u|/u_t)^q ; at |u| = u_t the stress is tauc
    assert O.sliding_beta(c, 100.0, 0.0, tauc=5e4) * 100.0 == pytest.approx(5e4, rel=1e-9)
    assert O.sliding_beta(c, 0.0, 400.0, tauc=5e4) * 400.0 == pytest.approx(5e4 * 4 ** 0.3, rel=1e-9)


def test_tsai_and_schoof_limits():
    N, a2, b2 = 2e6, 0.5, 1e4
    for law in ("Tsai2015", "Schoof2005"):
        c = _cfg(choice_sliding_law=law)
        slow = 1e-2
        assert O.sliding_beta(c, slow, 0.0, beta_sq=b2, alpha_sq=a2, neff=N) * slow == pytest.approx(
            b2 * slow ** (1 / 3), rel=1e-2)
        fast = 1e9
        assert O.sliding_beta(c, fast, 0.0, beta_sq=b2, alpha_sq=a2, neff=N) * fast == pytest.approx(a2 * N, rel=1e-2)


def test_zoet_iverson_limit():
    c = _cfg(choice_sliding_law="Zoet-Iverson")
    N, phi = 1e6, 20.0
    u = 1e9
    assert O.sliding_beta(c, u, 0.0, phi=phi, neff=N) * u == pytest.approx(N * math.tan(math.radians(phi)), rel=1e-6)


# ---------------------------------------------------------------- producers
def test_numpy_producer_matches_oracle_producer():
    import ctypes
    g, cfg, st = geometry.channel(nx=41, ny=13, dx=20e3, seed=1)
    L = O.lib()
    out = {k: np.zeros_like(st[k]) for k in ("Hs_a", "TAF_a", "mask_land_a", "mask_ocean_a", "mask_ice_a",
                                             "mask_sheet_a", "mask_shelf_a", "Hi_cx", "Hi_cy", "Hi_b",
                                             "A_flow_3D_cx", "A_flow_3D_cy", "A_flow_vav_a", "dHi_dx_a", "dHi_dy_a",
                                             "dHs_dx_a", "dHs_dy_a", "dHs_dx_cx", "dHs_dy_cx", "dHs_dx_cy", "dHs_dy_cy")}
    p = O._ptr
    L.orc_update_general_ice_model_data(ctypes.byref(g.struct()), ctypes.byref(cfg), p(st["Hi_a"]), p(st["Hb_a"]),
                                        p(st["SL_a"]), p(st["A_flow_3D_a"]),
                                        *[p(out[k]) for k in out])
    for k, v in out.items():
        np.testing.assert_array_equal(v, st[k], err_msg=k)


def test_producer_slopes_exact_on_planar_surface():
    g = Grid(9, 7, 1e3)
    cfg = geometry.synthetic_config()
    st = IceState(g, cfg.nz)
    X, Y = np.meshgrid(g.x, g.y)
    st["Hb_a"] = 100.0 + 0.002 * X - 0.003 * Y
    st["Hi_a"] = 500.0
    st["A_flow_3D_a"] = 1e-16
    geometry.update_general_ice_model_data(st, ZETA)
    for k, val in (("dHs_dx_a", 0.002), ("dHs_dy_a", -0.003), ("dHs_dx_cx", 0.002), ("dHs_dy_cx", -0.003),
                   ("dHs_dx_cy", 0.002), ("dHs_dy_cy", -0.003)):
        np.testing.assert_allclose(st[k], val, rtol=1e-9, err_msg=k)


# ---------------------------------------------------------------- linearised system
def _coeff_state(seed=0):
    g, cfg, st = geometry.channel(nx=21, ny=11, dx=8e3, seed=seed)
    c2 = geometry.synthetic_config()
    config.set_config(c2, DIVA_visc_it_nit=3, DIVA_visc_it_norm_dUV_tol=0.0)
    O.solve("diva", g, c2, st, O.LIN_BANDED_LU)  # fills N_a, N_b, beta_eff, taud
    return g, cfg, st


def test_free_rows_form_a_symmetric_block():
    """The interior (free) rows of the staggered DIVA operator are the
    Hessian of a convex functional: their mutual block is symmetric."""
    g, cfg, st = _coeff_state()
    ptr, idx, val, b, x = O.assemble_csr(g, cfg, st, st["u_vav_cx"], st["v_vav_cy"])
    m = len(b)
    A = sp.csr_matrix((val, idx, ptr), shape=(m, m)).toarray()
    nrows = np.diff(ptr)
    free = np.where(nrows == 9)[0]
    Af = A[np.ix_(free, free)]
    assert np.abs(Af - Af.T).max() <= 1e-12 * np.abs(Af).max()


def test_constant_mode_of_free_rows():
    """A uniform velocity field only feels the friction term: row sums of the
    u (v) coefficients equal -beta_eff dx^2, and the cross terms cancel."""
    g, cfg, st = _coeff_state(1)
    u1 = np.ones_like(st["u_vav_cx"])
    v0 = np.zeros_like(st["v_vav_cy"])
    ptr, idx, val, b, x = O.assemble_csr(g, cfg, st, u1, v0)
    m = len(b)
    A = sp.csr_matrix((val, idx, ptr), shape=(m, m))
    isu = np.zeros(m, bool)
    n = 0
    nx, ny = g.nx, g.ny
    bcx = np.zeros(m)
    for i in range(nx):
        for j in range(ny):
            if i < nx - 1:
                isu[n] = True
                bcx[n] = st["beta_eff_cx"][j, i]
                n += 1
            if j < ny - 1:
                n += 1
    free = np.diff(ptr) == 9
    ones_u = isu.astype(float)
    r = A @ ones_u
    dx2 = g.dx ** 2
    np.testing.assert_allclose(r[free & isu], -bcx[free & isu] * dx2, rtol=1e-9, atol=1e-3)
    np.testing.assert_allclose(r[free & ~isu], 0.0, atol=1e-3 * np.abs(A).max())


def test_banded_lu_and_sor_match_scipy_direct_solve():
    g, cfg, st = _coeff_state(2)
    ptr, idx, val, b, x0 = O.assemble_csr(g, cfg, st, st["u_vav_cx"], st["v_vav_cy"])
    m = len(b)
    A = sp.csr_matrix((val, idx, ptr), shape=(m, m)).tocsc()
    D = sp.diags(1.0 / np.abs(A).max(axis=1).toarray().ravel())
    xs = spl.spsolve((D @ A).tocsc(), D @ b)
    rc, xl = O.banded_lu(ptr, idx, val, b)
    assert rc == 0
    scale = np.abs(xs).max()
    assert np.abs(xl - xs).max() <= 1e-10 * scale
    rc, xo, its = O.sor(ptr, idx, val, b, x0, 200000, 1e-10, 1.3)
    assert rc == 0
    assert np.abs(xo - xs).max() <= 1e-6 * scale


# ---------------------------------------------------------------- solves against closed forms
def test_diva_uniform_slab_closed_form():
    """Uniform grounded slab with Weertman sliding.  With no membrane stresses
    tau_b = tau_d = rho g H |s'|; the shear-only viscosity gives
    F2 = int H/eta zeta^2 = 2 A H tau_b^2 / 5 (Lipscomb et al. 2019 Eq. 30,
    the SIA depth-averaged deformation 2 A tau^3 H / 5); and, the reference
    evaluating the sliding law with the depth-averaged velocity
    (src/ice_velocity_module.f90:320), u_vav solves
    tau = beta_sq u^(1/m - 1) (u - tau F2)."""
    from scipy.optimize import brentq
    nz = 32
    zeta = list(np.linspace(0.0, 1.0, nz))
    cfg = geometry.synthetic_config(zeta=zeta, DIVA_visc_it_nit=600, stop_rule=1, nonlinear_rtol=1e-12)
    g = Grid(8, 8, 10e3)
    st = IceState(g, nz)
    H, slope, A = 2000.0, 5e-3, 1e-16
    st["Hi_a"] = H
    st["Hb_a"] = -slope * np.meshgrid(g.x, g.y)[0]
    st["SL_a"] = -1e4
    st["A_flow_3D_a"] = A
    geometry.fill_uniform_roughness(st, cfg)
    geometry.update_general_ice_model_data(st, zeta)
    O.solve("diva", g, cfg, st, O.LIN_BANDED_LU)
    assert np.abs(st["u_vav_cx"]).max() < 0.5 * cfg.DIVA_vel_max   # velocity limiter inactive
    tau = 910.0 * 9.81 * H * slope
    F2 = 2.0 * A * H * tau ** 2 / 5.0
    assert F2 > 1.0 / (3.0 * 1e3)          # above the visc_eff_min floor of F2
    np.testing.assert_allclose(st["F2_a"], F2, rtol=3e-3)
    np.testing.assert_allclose(st["taub_cx"], tau, rtol=1e-8)
    F2n = float(st["F2_a"][3, 3])
    u_exp = brentq(lambda u: 1e4 * u ** (-2.0 / 3.0) * (u - tau * F2n) - tau, tau * F2n * 1.0001, 1e6)
    np.testing.assert_allclose(st["u_vav_cx"], u_exp, rtol=1e-8)
    np.testing.assert_allclose(st["u_base_cx"], u_exp - tau * F2n, rtol=1e-7)
    assert np.abs(st["v_vav_cy"]).max() < 1e-9 * u_exp


def test_ssa_schoof2006_ice_stream():
    """SSA with the Schoof (2006) till yield stress reproduces the closed-form
    ice-stream velocity (src/utilities_module.f90:836-877) up to the grid error."""
    L, mexp, tant, H, A = 100e3, 0.5, 3e-4, 2000.0, 1e-17
    ny, nx, dx = 121, 5, 5e3
    cfg = config.default_config(choice_ice_dynamics="SSA", choice_sliding_law="Coulomb",
                                choice_basal_roughness="parameterised", choice_param_basal_roughness="SSA_icestream",
                                choice_basal_hydrology="dry", include_basal_freezing=0, SSA_icestream_L=L,
                                SSA_icestream_m=mexp, DIVA_visc_it_nit=300, DIVA_visc_it_norm_dUV_tol=1e-9)
    g = Grid(nx, ny, dx, xmin=0.0, ymin=-(ny - 1) / 2 * dx)
    st = IceState(g, cfg.nz)
    st["Hi_a"] = H
    st["Hb_a"] = -tant * np.meshgrid(g.x, g.y)[0]
    st["SL_a"] = -1e5
    st["A_flow_3D_a"] = A
    geometry.update_general_ice_model_data(st, config.zeta_of(cfg))
    ua = np.array([O.lib() and _schoof(y, L, mexp, tant, H, A) for y in g.y])
    st["u_SSA_cx"] = 0.5 * ua[:, None]
    O.solve("ssa", g, cfg, st, O.LIN_BANDED_LU)
    u = st["u_SSA_cx"][:, 2]
    assert np.abs(u - ua).max() <= 0.03 * ua.max()


def _schoof(y, L, m, tant, h0, A):
    import ctypes
    U, tc = ctypes.c_double(), ctypes.c_double()
    O.lib().orc_ssa_schoof2006(ctypes.c_double(3.0), ctypes.c_double(L), ctypes.c_double(m), ctypes.c_double(tant),
                               ctypes.c_double(h0), ctypes.c_double(A), ctypes.c_double(y), ctypes.byref(U),
                               ctypes.byref(tc))
    return U.value


def test_sia_column_matches_closed_form():
    """SIA on a planar slab: u(zeta) = -2 A (rho g H)^3 |s'|^2 s' H (1 - zeta^4)/4."""
    nz = 32
    zeta = list(np.linspace(0.0, 1.0, nz))
    cfg = geometry.synthetic_config(choice_ice_dynamics="SIA", zeta=zeta)
    g = Grid(6, 6, 10e3)
    st = IceState(g, nz)
    H, slope, A = 1500.0, 2e-3, 1e-16
    st["Hi_a"] = H
    st["Hb_a"] = -slope * np.meshgrid(g.x, g.y)[0]
    st["SL_a"] = -1e4
    st["A_flow_3D_a"] = A
    geometry.update_general_ice_model_data(st, zeta)
    O.solve("sia", g, cfg, st)
    z = np.array(zeta)
    expect = 2 * A * (910 * 9.81 * H) ** 3 * slope ** 3 * H * (1 - z ** 4) / 4
    np.testing.assert_allclose(st["u_3D_SIA_cx"][:, 3, 2], expect, rtol=2e-3, atol=1e-9)


# ---------------------------------------------------------------- invariants
def test_no_grounded_ice_gives_zero_velocity():
    g, cfg, st = geometry.dome()
    st["Hi_a"] = 0.0
    geometry.update_general_ice_model_data(st, ZETA)
    st["u_vav_cx"] = 5.0
    O.solve("diva", g, cfg, st)
    assert np.all(st["u_vav_cx"] == 0.0) and np.all(st["v_vav_cy"] == 0.0)


def test_dome_fixed_point_is_transpose_symmetric():
    g, cfg, st = geometry.dome()
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-11, DIVA_visc_it_nit=600)
    O.solve("diva", g, cfg, st, O.LIN_BANDED_LU)
    u, v = st["u_vav_cx"], st["v_vav_cy"]
    assert np.abs(u - v.T).max() <= 1e-9 * np.abs(u).max()


def test_fixed_point_independent_of_relaxation_and_linear_solver():
    g, cfg, st = geometry.channel(nx=41, ny=11, dx=10e3, seed=5)
    a, b = st.copy(), st.copy()
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-11, DIVA_visc_it_nit=1500)
    O.solve("diva", g, cfg, a, O.LIN_BANDED_LU)
    c2 = geometry.synthetic_config(DIVA_boundary_BC_u_west="zero", DIVA_boundary_BC_v_west="zero",
                                   DIVA_boundary_BC_v_south="zero", DIVA_boundary_BC_v_north="zero",
                                   stop_rule=1, nonlinear_rtol=1e-11, DIVA_visc_it_nit=1500, DIVA_visc_it_relax=0.7)
    b.DIVA_SOR_tol = 1e-10
    b.DIVA_SOR_nit = 500000
    O.solve("diva", g, c2, b, O.LIN_SOR)
    assert np.abs(a["u_vav_cx"] - b["u_vav_cx"]).max() <= 1e-7 * np.abs(a["u_vav_cx"]).max()


# ---------------------------------------------------------------- 3D velocities, BC-mode rows, idealised laws
def _inclined_slab(nz=32, H=2000.0, slope=5e-3, A=1e-16, n=8):
    zeta = list(np.linspace(0.0, 1.0, nz))
    cfg = geometry.synthetic_config(zeta=zeta, DIVA_visc_it_nit=800, stop_rule=1, nonlinear_rtol=1e-12)
    g = Grid(n, n, 10e3)
    st = IceState(g, nz)
    st["Hi_a"] = H
    st["Hb_a"] = -slope * np.meshgrid(g.x, g.y)[0]
    st["SL_a"] = -1e4
    st["A_flow_3D_a"] = A
    geometry.fill_uniform_roughness(st, cfg)
    geometry.update_general_ice_model_data(st, zeta)
    O.solve("diva", g, cfg, st, O.LIN_BANDED_LU)
    return np.array(zeta), cfg, st


def test_diva_slab_velocity_profile_closed_form():
    """u(zeta) = u_b + tau_b F1(zeta) (Lipscomb et al. 2019 Eq. 29): on a uniform
    slab the shear-only viscosity 1/(2 A (tau zeta)^2) makes F1 the SIA
    profile, u(zeta) = u_b + 2 A tau^3 H (1 - zeta^4) / 4, and the surface
    velocity exceeds the depth average by the SIA ratio (5/4 of the
    deformation part)."""
    z, cfg, st = _inclined_slab()
    H, slope, A = 2000.0, 5e-3, 1e-16
    tau = 910.0 * 9.81 * H * slope
    ub = st["u_base_cx"][3, 3]
    expect = ub + 2.0 * A * tau ** 3 * H * (1.0 - z ** 4) / 4.0
    np.testing.assert_allclose(st["u_3D_cx"][:, 3, 3], expect, rtol=2e-3)
    # the viscosity column (away from the surface, where the floor of the
    # strain rate, DIVA_epsilon_sq_0, takes over)
    eta = st["visc_eff_3D_a"][:, 3, 3]
    k = z > 0.2
    np.testing.assert_allclose(eta[k], 1.0 / (2.0 * A * (tau * z[k]) ** 2), rtol=1e-6)
    # F1 at the surface is F2's counterpart without the zeta^2 weight
    F1s = st["F1_3D_a"][0, 3, 3]
    assert F1s == pytest.approx(2.0 * A * tau ** 2 * H / 4.0, rel=2e-3)
    deform = st["u_surf_a"][3, 3] - ub
    assert deform == pytest.approx(1.25 * (st["u_vav_cx"][3, 3] - ub), rel=3e-3)


def test_slab_vertical_velocity_follows_the_bed():
    """Mass conservation on an inclined slab of uniform thickness: the flow is
    parallel to the bed, so w(zeta) = -|s'| u(zeta) at every level.  The
    reference's calc_3D_vertical_velocities takes the zeta-derivative term
    from the a-grid velocities of the previous call (u_3D_a is refreshed only
    afterwards, src/ice_velocity_module.f90:479-492): a first call from zero
    a-grid velocities gives the basal kinematic value -|s'| u_b at every
    level, a second call from the converged state the full profile."""
    z, cfg, st = _inclined_slab(nz=15)
    slope = 5e-3
    ub = st["u_base_a"][2:-2, 2:-2]
    np.testing.assert_allclose(st["w_3D_a"][:, 2:-2, 2:-2], -slope * np.broadcast_to(ub, (15,) + ub.shape),
                               rtol=1e-9)
    O.solve("diva", Grid(8, 8, 10e3), cfg, st, O.LIN_BANDED_LU)
    u3, w3 = st["u_3D_a"], st["w_3D_a"]
    np.testing.assert_allclose(w3[:, 2:-2, 2:-2], -slope * u3[:, 2:-2, 2:-2], rtol=1e-9)
    assert abs(w3[0, 3, 3] - w3[-1, 3, 3]) > 0.05 * abs(w3[-1, 3, 3])   # the levels differ


def _floating_shelf(kind, A=1e-17, H=400.0, nx=40, ny=6, dx=5e3):
    """A uniform floating shelf in choice_ice_margin = 'BC', uniform in y: a
    wall (u = 0) to the west, a grounded strip along it (the reference
    returns zero velocities when no ice is grounded), the calving front two
    cells before the eastern edge."""
    dyn = "SSA" if kind == "ssa" else "DIVA"
    cfg = geometry.synthetic_config(choice_ice_dynamics=dyn, choice_ice_margin="BC", DIVA_boundary_BC_u_west="zero",
                                    DIVA_boundary_BC_v_south="zero", DIVA_boundary_BC_v_north="zero",
                                    stop_rule=1, nonlinear_rtol=1e-11, DIVA_visc_it_nit=2000)
    g = Grid(nx, ny, dx)
    st = IceState(g, cfg.nz)
    Hi = np.full((ny, nx), H)
    Hi[:, nx - 2:] = 0.0
    st["Hi_a"] = Hi
    Hb = np.full((ny, nx), -1000.0)
    Hb[:, :4] = -340.0            # just grounded: a small surface step at the grounding line
    st["Hb_a"] = Hb
    st["SL_a"] = 0.0
    st["A_flow_3D_a"] = A
    geometry.fill_uniform_roughness(st, cfg)
    geometry.update_general_ice_model_data(st, config.zeta_of(cfg))
    return g, cfg, st


@pytest.mark.parametrize("kind", ["diva", "ssa"])
def test_unconfined_shelf_spreads_at_weertman_rate(kind):
    """Weertman (1957): a freely floating shelf of uniform thickness in plane
    flow spreads at du/dx = A (rho_i g H (1 - rho_i/rho_w) / 4)^n.  This pins
    the calving-front stress rows of choice_ice_margin = 'BC'
    (src/ice_velocity_module.f90:2192-2341): the front row carries the whole
    load, the free rows keep the strain rate uniform, and a wrong factor or
    sign in the front row changes the rate."""
    A, H = 1e-17, 400.0
    g, cfg, st = _floating_shelf(kind, A=A, H=H)
    O.solve(kind, g, cfg, st, O.LIN_BANDED_LU)
    u = st["u_SSA_cx" if kind == "ssa" else "u_vav_cx"]
    rate = A * (910.0 * 9.81 * H * (1.0 - 910.0 / 1028.0) / 4.0) ** 3
    dudx = np.diff(u[2, 8: g.nx - 2]) / g.dx      # the floating part, away from the grounding line
    np.testing.assert_allclose(dudx, rate, rtol=1e-6)
    assert np.abs(st["v_SSA_cy" if kind == "ssa" else "v_vav_cy"]).max() <= 1e-9 * np.abs(u).max()
    # the front row is the stress condition, not the free row
    assert st["DIVA_isfront_inner"][2, g.nx - 3] == 1 and st["DIVA_isfront_outer"][2, g.nx - 2] == 1


@pytest.mark.parametrize("which", ["ISMIP_HOM_C", "ISMIP_HOM_D", "ISMIP_HOM_F"])
def test_idealised_friction_fields(which):
    """The ISMIP-HOM friction fields (Pattyn et al. 2008; reference
    src/basal_conditions_and_sliding_module.f90:1288-1432) evaluated by a
    solve at the grid's own coordinates: beta = 1000 + 1000 sin(2 pi x/L)
    sin(2 pi y/L) (C), 1000 + 1000 sin(2 pi x/L) (D), (A_vav 1000)^-1 (F)."""
    L = 80e3
    cfg = geometry.synthetic_config(choice_sliding_law="idealised", choice_idealised_sliding_law=which, ISMIP_HOM_L=L,
                                    DIVA_visc_it_nit=2)
    g, cfg, st = geometry.random_geometry(31, 27, seed=21, cfg=cfg)
    O.solve("diva", g, cfg, st, O.LIN_BANDED_LU)
    X, Y = np.meshgrid(g.x, g.y)
    if which == "ISMIP_HOM_C":
        beta = 1000.0 + 1000.0 * np.sin(2 * np.pi * X / L) * np.sin(2 * np.pi * Y / L)
    elif which == "ISMIP_HOM_D":
        beta = 1000.0 + 1000.0 * np.sin(2 * np.pi * X / L)
    else:
        beta = 1.0 / (st["A_flow_vav_a"] * 1000.0)
    np.testing.assert_allclose(st["beta_a"], beta, rtol=1e-12, atol=1e-9)


def _quad_mc(c, n=600):
    return _mc_fraction(*c, n=n)


@pytest.mark.parametrize("scenario", [1, 2, 3, 4])
def test_fraction_above_zero_every_scenario_and_rotation(scenario):
    """Each of the four CISM scenarios (one corner above zero, one below, two
    adjacent, two opposite) in each of its four rotations, against a brute-force
    count of the bilinear quad (the rotations of rotate_quad_until_match must
    not change the area)."""
    rng = np.random.default_rng(40 + scenario)
    sign = {1: [1, -1, -1, -1], 2: [-1, 1, 1, 1], 3: [1, 1, -1, -1], 4: [1, -1, 1, -1]}[scenario]
    worst = 0.0
    for _ in range(25):
        mag = rng.uniform(0.2, 1.0, 4) * 10.0 ** rng.uniform(-1, 2)
        base = np.array(sign) * mag           # NW, NE, SE, SW order around the quad
        for r in range(4):
            nw, ne, se, sw = np.roll(base, r)
            phi = O.fraction_above_zero(nw, ne, sw, se)
            worst = max(worst, abs(phi - _quad_mc((nw, ne, sw, se))))
    assert worst < 4e-3, worst


def test_fraction_above_zero_near_zero_corners_and_planar_quads():
    """The ftol perturbation (corners within 1e-4 of zero are moved to
    +-1e-4 x 1.1/1.21/1.13/0.97) and the |dd| < 1e-4 shift of the SW corner by
    0.1 (src/general_ice_model_data_module.f90:539-574, 608-680) change the
    fraction only by their own size: exact zeros, tiny corners and exactly
    planar quads (dd = 0) of realistic TAF magnitudes (metres to hundreds of
    metres) against a brute-force count."""
    rng = np.random.default_rng(9)
    worst = 0.0
    for _ in range(60):
        c = rng.uniform(-1.0, 1.0, 4) * 10.0 ** rng.uniform(1, 2)
        c[rng.integers(4)] = rng.choice([0.0, 1e-7, -1e-7, 5e-5, -5e-5])
        phi = O.fraction_above_zero(*c)
        worst = max(worst, abs(phi - _quad_mc(c)))
    for _ in range(60):
        a, b, d = rng.uniform(-1, 1, 3) * np.array([30.0, 80.0, 80.0])
        f = lambda x, y: a + b * x + d * y      # planar: NE + SW - NW - SE = 0
        c = (f(0, 1), f(1, 1), f(0, 0), f(1, 0))
        phi = O.fraction_above_zero(*c)
        worst = max(worst, abs(phi - _quad_mc(c)))
    assert worst < 4e-3, worst


def _periodic_slab(axis, shift, nx=23, ny=8):
    """Inclined slab with a thickness wave along `axis` ('x' or 'y'), uniform
    across it; periodic boundaries along the wave and 'infinite' ones across
    it.  The reference's periodic rows tie u(1) to u(nx-2) and u(nx-1) to u(2)
    (1-based), so u is periodic with nx - 3 cells along x and v with ny - 3
    cells along y (src/ice_velocity_module.f90:1932-2191); the wave has that
    period.  SSA: in DIVA the boundary cells' one-sided strain rates reach the
    interior rows through the staggered vertical shear, so only the SSA is
    exactly periodic."""
    along = ("west", "east") if axis == "x" else ("south", "north")
    across = ("south", "north") if axis == "x" else ("west", "east")
    bcs = {"DIVA_boundary_BC_%s_%s" % (c, s): "periodic" for c in ("u", "v") for s in along}
    bcs.update({"DIVA_boundary_BC_%s_%s" % (c, s): "infinite" for c in ("u", "v") for s in across})
    cfg = geometry.synthetic_config(choice_ice_dynamics="SSA", stop_rule=1, nonlinear_rtol=1e-12,
                                    DIVA_visc_it_nit=2000, **bcs)
    if axis == "y":
        nx, ny = ny, nx
    g = Grid(nx, ny, 5e3)
    st = IceState(g, cfg.nz)
    P = (nx if axis == "x" else ny) - 3
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    k = (I if axis == "x" else J) + shift
    st["Hi_a"] = 1000.0 + 150.0 * np.sin(2 * np.pi * k / P) + 60.0 * np.cos(6 * np.pi * k / P)
    X, Y = np.meshgrid(g.x, g.y)
    st["Hb_a"] = -4e-3 * (X if axis == "x" else Y)
    st["SL_a"] = -1e4
    st["A_flow_3D_a"] = 1e-16
    geometry.fill_uniform_roughness(st, cfg)
    geometry.update_general_ice_model_data(st, config.zeta_of(cfg))
    st.DIVA_SOR_tol, st.DIVA_SOR_nit = 1e-9, 200000      # x-periodic rows: the oracle solves with SOR
    O.solve("ssa", g, cfg, st, O.LIN_BANDED_LU)
    return P, st


@pytest.mark.parametrize("axis", ["x", "y"])
def test_periodic_rows_are_translation_invariant(axis):
    """The solution repeats with the period of the boundary rows (to the 1e-12
    nonlinear residual of the solves), and shifting the input wave by d cells
    shifts the solution by d cells (to the trace the boundary cells leave); a
    wrong opposite index breaks both.
    The corner rows, where a periodic side meets an 'infinite' one, take the
    'infinite' rule (zero before infinite before periodic,
    src/ice_velocity_module.f90:1954-1989): they copy the diagonal inner
    neighbour."""
    P, a = _periodic_slab(axis, 0)
    _, b = _periodic_slab(axis, 3)
    if axis == "x":
        u0, u3, w = a["u_SSA_cx"], b["u_SSA_cx"], a["v_SSA_cy"]
    else:     # transpose so the wave runs along the second index
        u0, u3, w = a["v_SSA_cy"].T, b["v_SSA_cy"].T, a["u_SSA_cx"]
    tol = 1e-7 * np.abs(u0).max()
    n = u0.shape[1]
    inner = slice(1, u0.shape[0] - 1)           # the rows across the wave that are not boundary rows
    np.testing.assert_allclose(u0[inner, P:], u0[inner, :n - P], rtol=0, atol=tol)
    # the boundary cells' one-sided strain rates and the corner rows leave a
    # trace of a few 1e-5 next to the periodic edges; a shift by one cell in
    # the opposite index would change the solution by about 10%
    loose = 1e-4 * np.abs(u0).max()
    np.testing.assert_allclose(u3[inner, : n - 3], u0[inner, 3:], rtol=0, atol=loose)
    assert np.abs(u0[inner] - u0[1]).max() <= loose                   # uniform across the wave
    np.testing.assert_allclose(u0[0, 1:-1], u0[1, 1:-1], rtol=0, atol=tol)   # 'infinite' across
    for (r, c), (ri, ci) in (((0, 0), (1, 1)), ((0, n - 1), (1, n - 2)), ((-1, 0), (-2, 1)), ((-1, n - 1), (-2, n - 2))):
        assert abs(u0[r, c] - u0[ri, ci]) <= tol, (r, c)               # corners
    assert abs(u0[1, 0] - u0[1, 1]) > 10 * tol                        # ... which periodic rows would not give
    assert np.abs(w).max() <= 1e-4 * np.abs(u0).max()
    assert np.ptp(u0) > 0.05 * np.abs(u0).max()                      # the wave matters


def test_martin2011_hydrology_and_roughness():
    """Martin et al. (2011, The Cryosphere 5, 727-740) as configured by the
    reference (src/basal_conditions_and_sliding_module.f90:90-294, defaults
    src/configuration_module.f90:362-377): pore water pressure 0.96 x the
    overburden times lambda, lambda = 1 for a bed at or below sea level + 0 m,
    falling linearly to 0 at 1000 m above it; till friction angle 5 deg at or
    below -1000 m bed elevation, 20 deg at or above 0 m, linear between.  The
    hydrology reads the bed relative to sea level, the roughness the bed
    itself: a sea level of 50 m tells them apart.  Values by hand."""
    g = Grid(8, 5, 1e3)
    cfg = geometry.synthetic_config()
    config.set_config(cfg, choice_basal_hydrology="Martin2011", choice_basal_roughness="parameterised",
                      choice_param_basal_roughness="Martin2011", choice_sliding_law="Coulomb_regularised")
    st = IceState(g, cfg.nz)
    Hb = np.array([-1500.0, -1000.0, -500.0, -250.0, 0.0, 300.0, 1050.0, 2000.0])
    st["Hb_a"] = Hb[None, :]
    st["Hi_a"] = 1000.0
    st["SL_a"] = 50.0
    io = st.io()
    phi = np.zeros((g.ny, g.nx), order="F")
    tauc = np.zeros((g.ny, g.nx), order="F")
    assert O.lib().orc_calc_basal_conditions(ctypes.byref(g.struct()), ctypes.byref(cfg), ctypes.byref(io),
                                             O._ptr(phi), O._ptr(tauc)) == 0
    po = 910.0 * 9.81 * 1000.0
    # Hb - SL = -1550, -1050, -550, -300, -50, 250, 1000, 1950 -> lambda 1, 1, 1, 1, 1, 0.75, 0, 0
    lam = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 0.75, 0.0, 0.0])
    np.testing.assert_allclose(st["overburden_pressure_a"], po, rtol=1e-15)
    np.testing.assert_allclose(st["pore_water_pressure_a"][2], 0.96 * po * lam, rtol=1e-14)
    np.testing.assert_allclose(st["Neff_a"][2] / po, [0.04] * 5 + [0.28, 1.0, 1.0], rtol=1e-13)
    np.testing.assert_allclose(phi[2], [5.0, 5.0, 12.5, 16.25, 20.0, 20.0, 20.0, 20.0], rtol=1e-15)
    assert np.all(phi == phi[2]) and np.all(st["Neff_a"] == st["Neff_a"][2])


def test_shelf_land_zero_rows_of_bc_mode():
    """choice_ice_margin = 'BC': a velocity between a shelf cell and an
    ice-free land cell is a zero row (src/ice_velocity_module.f90:1787-1796),
    in either order of the two cells.  An ice-free island in the middle of a
    y-uniform shelf: its four faces carry exactly zero velocity, the faces
    next to them do not, and the flow stays mirror-symmetric about the
    island's row (a rule applied to one orientation only would break both)."""
    g, cfg, st = _floating_shelf("ssa", ny=9)
    jc, ic = 4, 20
    Hb, Hi = st["Hb_a"].copy(), st["Hi_a"].copy()
    Hb[jc, ic], Hi[jc, ic] = 100.0, 0.0
    st["Hb_a"], st["Hi_a"] = Hb, Hi
    geometry.update_general_ice_model_data(st, config.zeta_of(cfg))
    O.solve("ssa", g, cfg, st, O.LIN_BANDED_LU)
    u, v = st["u_SSA_cx"], st["v_SSA_cy"]
    assert u[jc, ic - 1] == 0.0 and u[jc, ic] == 0.0          # shelf|land and land|shelf
    assert v[jc - 1, ic] == 0.0 and v[jc, ic] == 0.0
    scale = np.abs(u).max()
    assert abs(u[jc, ic - 2]) > 1e-3 * scale and abs(u[jc, ic + 1]) > 1e-3 * scale
    assert abs(u[jc - 1, ic - 1]) > 1e-3 * scale and abs(v[jc - 1, ic - 1]) > 1e-6 * scale
    np.testing.assert_allclose(u[1:jc], u[-2:jc:-1], rtol=0, atol=1e-7 * scale)      # mirror in y
    np.testing.assert_allclose(v[1:jc - 1], -v[-2:jc:-1], rtol=0, atol=1e-7 * scale)


def test_zero_boundary_wins_the_corner():
    """Domain-boundary corners take 'zero' when either side is 'zero', else
    'infinite' (src/ice_velocity_module.f90:1954-1989): with u_south = zero
    and u_east = infinite the south-east u corner is exactly zero, the
    north-east one (north and east infinite) copies its diagonal inner
    neighbour, and the east column between them copies its west neighbour."""
    g, cfg, st = _floating_shelf("ssa", ny=7)
    config.set_config(cfg, DIVA_boundary_BC_u_south="zero")
    O.solve("ssa", g, cfg, st, O.LIN_BANDED_LU)
    u = st["u_SSA_cx"]
    e = g.nx - 2                                  # the last u column
    assert u[0, e] == 0.0 and np.all(u[0] == 0.0)
    assert u[-1, e] == u[-2, e - 1]
    np.testing.assert_array_equal(u[1:-1, e], u[1:-1, e - 1])
    assert np.abs(u[-1]).max() > 0.0


# ---------------------------------------------------------------- Picard error estimate
def _visc_errs(u_new, u_old, chk):
    """2 |u_new - u_old| / |u_new + u_old + 1e-5| where |chk| > 1e-6, else 0."""
    return np.where(np.abs(chk) > 1e-6, 2.0 * np.abs(u_new - u_old) / np.abs(u_new + u_old + 1e-5), 0.0)


def test_visc_iter_errors_pair_consecutive_iterates():
    """estimate_visc_iter_UV_errors (src/ice_velocity_module.f90:1602-1652) runs
    after every Picard iteration that does not end the loop (:355): the
    estimate left by a run of K iterations compares iterate K-1 with K-2.
    The iterates come from independent runs of 1, 2 and 3 iterations of the
    same deterministic loop; a run of one iteration leaves the 1e5 set at the
    solve's start (:176-177)."""
    g, cfg, st = geometry.dome(nx=17, ny=15, dx=50e3)
    st["u_vav_cx"] = 1e-3 * np.sin(np.arange(st["u_vav_cx"].size)).reshape(st["u_vav_cx"].shape, order="F")
    runs = {}
    for K in (1, 2, 3):
        config.set_config(cfg, stop_rule=0, DIVA_visc_it_norm_dUV_tol=0.0, DIVA_visc_it_nit=K)
        runs[K] = O.solve("diva", g, cfg, st.copy())
        assert runs[K].stats["picard_iterations"] == K
    assert np.all(runs[1]["DIVA_err_cx"] == 1e5) and np.all(runs[1]["DIVA_err_cy"] == 1e5)
    iterate = {0: st, 1: runs[1], 2: runs[2]}
    for K in (2, 3):
        new, old = iterate[K - 1], iterate[K - 2]
        np.testing.assert_array_equal(runs[K]["u_cx_prev"], new["u_vav_cx"])
        for e, u in (("DIVA_err_cx", "u_vav_cx"), ("DIVA_err_cy", "v_vav_cy")):
            want = _visc_errs(new[u], old[u], new[u])
            np.testing.assert_allclose(runs[K][e], want, rtol=1e-14, atol=0.0, err_msg="%s K=%d" % (e, K))
    assert (runs[3]["DIVA_err_cx"] > 0.0).all() and (runs[3]["DIVA_err_cx"] < 1e5).all()


def test_visc_iter_errors_of_ssa_test_the_stale_vertical_average():
    """In solve_SSA the tolerance test reads ice%u_vav_cx, which the SSA loop
    does not update (:238, the estimate of u_SSA_cx): cells where the state's
    u_vav_cx is below 1e-6 get 0 whatever the SSA velocity does."""
    import branch_cases as B
    kind, g, cfg, st = B.build("ssa_dome")
    half = g.ny // 2
    st["u_vav_cx"] = 1.0
    st["u_vav_cx"][:half] = 0.0
    st["v_vav_cy"] = 1.0
    st["v_vav_cy"][:, : g.nx // 2] = 0.0
    runs = {}
    for K in (1, 2):
        config.set_config(cfg, stop_rule=0, DIVA_visc_it_norm_dUV_tol=0.0, DIVA_visc_it_nit=K)
        runs[K] = O.solve("ssa", g, cfg, st.copy())
    new = runs[1]
    for e, u, chk in (("DIVA_err_cx", "u_SSA_cx", st["u_vav_cx"]), ("DIVA_err_cy", "v_SSA_cy", st["v_vav_cy"])):
        np.testing.assert_allclose(runs[2][e], _visc_errs(new[u], st[u], chk), rtol=1e-14, atol=0.0, err_msg=e)
    assert np.all(runs[2]["DIVA_err_cx"][:half] == 0.0)
    assert (np.abs(new["u_SSA_cx"][:half]) > 1e-3).any()         # moving ice in the masked half
    assert (runs[2]["DIVA_err_cx"][half:] > 0.0).any()


def test_visc_iter_errors_under_the_residual_rule():
    """The residual stop rule ends the loop on the nonlinear residual, not on
    the reference's has_converged: every iteration, the last one included,
    leaves its estimate (final iterate against u_cx_prev)."""
    g, cfg, st = geometry.dome(nx=17, ny=15, dx=50e3)
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-6, DIVA_visc_it_nit=200)
    O.solve("diva", g, cfg, st)
    assert st.stats["picard_iterations"] > 2
    for e, u, p in (("DIVA_err_cx", "u_vav_cx", "u_cx_prev"), ("DIVA_err_cy", "v_vav_cy", "v_cy_prev")):
        np.testing.assert_allclose(st[e], _visc_errs(st[u], st[p], st[u]), rtol=1e-14, atol=0.0, err_msg=e)


# ---------------------------------------------------------------- ISMIP-HOM E
def _arolla_file(path, rows, extra=0, drop=0):
    x, Hb, Hs, sl = rows
    with open(path, "w") as f:
        for k in range(51 - drop):
            f.write("%.1f %.6f %.6f %d\n" % (x[k], Hb[k], Hs[k], sl[k]))
        for k in range(extra):
            f.write("9999.0 0.0 0.0 1\n")
    return path


def test_ismip_hom_e_file_and_geometry(tmp_path):
    """The Arolla-format file: 51 list-directed records (x, Hb, Hs, slides),
    later ones not read, fewer rejected with the reference's message; every
    row of the grid holds the file's column profile with Hi = Hs - Hb
    (src/reference_fields_module.f90:751-762)."""
    import branch_cases as B
    rows = B.arolla_like_rows()
    p = _arolla_file(tmp_path / "a.dat", rows, extra=3)
    x, Hb, Hs, sl = geometry.read_ismip_hom_e(p)
    np.testing.assert_allclose(Hb, rows[1], atol=1e-6)
    np.testing.assert_array_equal(sl, rows[3])
    g, cfg, st = geometry.ismip_hom_e(p, ny=6)
    assert (g.nx, g.ny, g.dx) == (51, 6, 100.0)
    for j in range(6):
        np.testing.assert_allclose(st["Hi_a"][j], rows[2] - rows[1], atol=2e-6)
    np.testing.assert_array_equal(st["ISMIP_HOM_E_slides"], rows[3])
    with pytest.raises(geometry.ArollaFileError, match="should be 51 lines"):
        geometry.read_ismip_hom_e(_arolla_file(tmp_path / "b.dat", rows, drop=1))


def test_ismip_hom_e_sliding_law():
    """calc_sliding_law_idealised_ISMIP_HOM_E (:1352-1399): record i of the
    file sets column i, beta = 0 where slides = 1 and 1e30 (clamped to
    DIVA_beta_max) elsewhere, on every row.  The solve equals, bit for bit,
    Weertman sliding with m = 1 (beta = beta_sq exactly) and beta_sq = 0 / 1e30
    on those columns, an independent law fed the intended pattern (a
    transposed or inverted pattern differs); with no slip column it equals
    the no-sliding law (beta_eff = 1/F2) to the Picard tolerance."""
    import branch_cases as B
    kind, g, cfg, st = B.build("ismip_e")
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-11, DIVA_visc_it_nit=500)
    e = O.solve("diva", g, cfg, st.copy())
    sl = st["ISMIP_HOM_E_slides"]
    assert 0 < sl.sum() < 51
    want = np.where(sl[None, :] == 1, 0.0, cfg.DIVA_beta_max) * np.ones((g.ny, 1))
    np.testing.assert_array_equal(e["beta_a"], want)
    cw = config.default_config()
    ctypes.memmove(ctypes.addressof(cw), ctypes.addressof(cfg), ctypes.sizeof(cfg))
    config.set_config(cw, choice_sliding_law="Weertman", slid_Weertman_m=1.0)
    w = st.copy()
    w["beta_sq_a"] = np.where(sl[None, :] == 1, 0.0, 1e30) * np.ones((g.ny, 1))
    O.solve("diva", g, cw, w)
    for k in ("u_vav_cx", "v_vav_cy", "beta_a", "u_base_a", "w_3D_a"):
        np.testing.assert_array_equal(e[k], w[k], err_msg=k)
    # no slip column: the no-sliding law
    st0 = st.copy()
    st0["ISMIP_HOM_E_slides"] = 0
    e0 = O.solve("diva", g, cfg, st0)
    cn = config.default_config()
    ctypes.memmove(ctypes.addressof(cn), ctypes.addressof(cfg), ctypes.sizeof(cfg))
    config.set_config(cn, choice_sliding_law="no_sliding")
    n0 = O.solve("diva", g, cn, st.copy())
    scale = np.abs(n0["u_vav_cx"]).max()
    assert np.abs(e0["u_vav_cx"] - n0["u_vav_cx"]).max() <= 1e-8 * scale
    # the slip columns speed the flow up there
    cols = np.where(sl == 1)[0][:-1]          # cx faces between two slip columns
    assert (np.abs(e["u_vav_cx"][:, cols]) > 1.5 * np.abs(n0["u_vav_cx"][:, cols])).all()


def test_ismip_hom_e_needs_51_columns():
    g, cfg, st = geometry.random_geometry(41, 37, seed=11, cfg=geometry.synthetic_config(
        choice_sliding_law="idealised", choice_idealised_sliding_law="ISMIP_HOM_E"))
    with pytest.raises(RuntimeError, match="51 lines"):
        O.solve("diva", g, cfg, st)
