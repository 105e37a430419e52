"""GPU parity of every branch of the velocity path against the CPU oracle.

Each case of branch_cases.py reaches one branch: a sliding law, a hydrology
or roughness choice, basal freezing, the BC-mode margin rows, the 'sans'
rows, periodic boundaries, the velocity limiter, solve_SSA, solve_SIA and
their SIA/SSA combination.  Through the C ABI the GPU solve is compared with
the oracle (exact banded-LU linear solves) on the same input:

* one Picard iteration from the same warm state: every coefficient field
  (viscosity, N, F2, beta, beta_eff, grounded fractions, hydrology,
  roughness) within 1e-10 relative and the SICOPOLIS masks exactly; the
  velocities after that one linear solve within 1e-8 (the GPU's Krylov
  solve runs to a relative linear residual of 1e-13);
* the converged fixed point (FP64 nonlinear residual 1e-11) within 1e-6
  relative, the plain-FP64 bar;
* where the Picard iteration has no fixed point to converge to (plastic
  beds that cycle, the velocity limiter), the same fixed number of Picard
  iterations of the reference loop within 1e-6;
* the reference Picard stop rule (relative velocity change below
  DIVA_visc_it_norm_dUV_tol) stops both after the same number of iterations.
"""
import ctypes

import numpy as np
import pytest

import _oracle as O
import branch_cases as B
from imau_vel import config

pytestmark = pytest.mark.gpu

COEF = ("visc_eff_3D_a", "N_a", "N_b", "F2_a", "beta_a", "beta_eff_a", "beta_eff_cx", "beta_eff_cy", "f_grnd_a",
        "f_grnd_cx", "f_grnd_cy", "f_grnd_b", "taudx_cx", "taudy_cy", "visc_eff_int_a", "Neff_a",
        "pore_water_pressure_a", "overburden_pressure_a", "phi_fric_a", "tauc_a", "beta_sq_a", "alpha_sq_a")
MASKS = ("DIVA_mask_combi", "DIVA_isfront_inner", "DIVA_isfront_outer")

# cases whose reference Picard loop converges to an FP64 residual of 1e-11
FIXED_POINT = ["no_sliding", "coulomb_saturated", "tsai_saturated", "schoof_mismip", "weertman_mismip",
               "zoet_iverson_dry", "ismip_c", "ismip_e", "ismip_f", "freezing", "bc_margin", "sans_random", "sans_channel",
               "periodic", "ssa_dome", "bc_margin_ssa"]
# cases that cycle (plastic beds, a zero-friction strip) or are held by the limiter
TRAJECTORY = ["coulomb_reg_martin", "zoet_iverson_martin", "ismip_d", "limiter", "ssa_channel", "ssa_icestream"]


@pytest.fixture(scope="module")
def ctx():
    from imau_vel import solver
    c = solver.Context()
    yield c
    c.close()


def relerr(a, b, scale=None):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    s = np.abs(b).max() if scale is None else scale
    return float(np.abs(a - b).max() / max(s, 1e-300))


# fields compared against the magnitude of the velocity field rather than
# their own: the transverse component where it vanishes by symmetry, and the
# basal velocity where a nearly frozen bed makes it a difference of close numbers
VEL_SCALED = ("v_SSA_cy", "v_vav_cy", "v_3D_cy", "u_SSA_cx", "u_vav_cx", "u_base_cx", "v_base_cy", "taub_cy")


def ice_mask(st, f):
    """Points of field f next to ice: in choice_ice_margin = 'BC' the rows of
    ice-free cells carry N = 0 and only the friction of a bare bed, so their
    velocities are set by couplings many orders of magnitude weaker than the
    ice's; a Krylov solve stopped on ||b - Ax|| <= rtol ||b|| (PETSc's rule,
    the reference's default) leaves them inexact, an exact LU does not.  The
    ice velocities are compared everywhere the ice is."""
    H = st["Hi_a"] > 0.0
    shape = st[f].shape[-2:]
    if shape == H.shape:
        m = H
    elif shape == (H.shape[0], H.shape[1] - 1):
        m = H[:, :-1] | H[:, 1:]
    else:
        m = H[:-1, :] | H[1:, :]
    return np.broadcast_to(m, st[f].shape)


def err(gpu, ref, f, kind, bc=False):
    a, b = gpu[f], ref[f]
    if bc:
        m = ice_mask(ref, f)
        a, b = np.where(m, a, 0.0), np.where(m, b, 0.0)
    if f in VEL_SCALED:
        uf, vf = vel_fields(kind)
        scale = max(np.abs(ref[uf]).max(), np.abs(ref[vf]).max())
        if f == "taub_cy":
            scale = max(np.abs(ref["taub_cx"]).max(), np.abs(ref["taub_cy"]).max())
        return relerr(a, b, scale)
    return relerr(a, b)


def vel_fields(kind):
    return ("u_SSA_cx", "v_SSA_cy") if kind == "ssa" else ("u_vav_cx", "v_vav_cy")


def cfg_copy(cfg):
    c2 = config.default_config()
    ctypes.memmove(ctypes.addressof(c2), ctypes.addressof(cfg), ctypes.sizeof(cfg))
    return c2


def gpu_solve(ctx, kind, g, cfg, st):
    return {"diva": ctx.solve_diva, "ssa": ctx.solve_ssa, "sia": ctx.solve_sia}[kind](g, cfg, st)


def tight_linear(st, rtol=1e-13, periodic=False):
    """KSP stop at rtol (relative to ||b|| under the reference Picard rule,
    to the initial linear residual under the residual rule); DIVA_SOR_nit
    caps the Krylov iterations of one linear solve.  With x-periodic rows the
    oracle solves with the reference SOR (its banded LU needs a band), so
    those cases give SOR a tight tolerance and room to meet it."""
    st.DIVA_PETSc_rtol = rtol
    st.DIVA_PETSc_abstol = 0.0
    st.DIVA_SOR_tol = 1e-11 if periodic else 1e-10
    st.DIVA_SOR_nit = 200000 if periodic else 3000


def is_periodic(cfg):
    return cfg.DIVA_boundary_BC_u_west == 2 or cfg.DIVA_boundary_BC_v_west == 2


def warm(kind, g, cfg, st, its=8):
    c2 = cfg_copy(cfg)
    config.set_config(c2, stop_rule=0, DIVA_visc_it_nit=its, DIVA_visc_it_norm_dUV_tol=0.0)
    O.solve(kind, g, c2, st, O.LIN_BANDED_LU)


@pytest.mark.parametrize("name", B.DIVA_CASES + ["limiter"] + B.SSA_CASES)
def test_one_iteration(ctx, name):
    kind, g, cfg, st = B.build(name)
    warm(kind, g, cfg, st)
    config.set_config(cfg, stop_rule=0, DIVA_visc_it_nit=1, DIVA_visc_it_norm_dUV_tol=0.0)
    tight_linear(st, rtol=1e-12 if is_periodic(cfg) else 1e-13, periodic=is_periodic(cfg))
    gpu, ref = st.copy(), st.copy()
    gpu_solve(ctx, kind, g, cfg, gpu)
    O.solve(kind, g, cfg, ref, O.LIN_BANDED_LU)
    fields = COEF + (("N_cx", "N_cy") if not cfg.include_SSADIVA_crossterms else ())
    if kind == "ssa":
        fields = tuple(f for f in fields if f != "F2_a")
    errs = {f: relerr(gpu[f], ref[f]) for f in fields}
    bad = {f: e for f, e in errs.items() if not e < 1e-10}
    assert not bad, (name, bad)
    for m in MASKS:
        np.testing.assert_array_equal(gpu[m], ref[m], err_msg=name + " " + m)
    for f in vel_fields(kind):
        e = err(gpu, ref, f, kind, bc=cfg.choice_ice_margin == 0)
        assert e < 1e-8, (name, f, e)


@pytest.mark.parametrize("name", FIXED_POINT)
def test_fixed_point(ctx, name):
    kind, g, cfg, st = B.build(name)
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-11, DIVA_visc_it_nit=3000)
    tight_linear(st, rtol=1e-2, periodic=is_periodic(cfg))
    gpu, ref = st.copy(), st.copy()
    gpu_solve(ctx, kind, g, cfg, gpu)
    O.solve(kind, g, cfg, ref, O.LIN_BANDED_LU)
    print(name, "picard gpu/oracle", gpu.stats["picard_iterations"], ref.stats["picard_iterations"],
          "residual", gpu.stats["nonlinear_residual"])
    assert gpu.stats["nonlinear_residual"] <= 1e-11
    names = vel_fields(kind) + ("u_3D_cx", "v_3D_cy", "w_3D_a", "u_surf_a", "uabs_vav_a", "beta_a")
    if kind == "diva":
        names += ("taub_cx", "taub_cy", "u_base_cx")
    for f in names:
        e = err(gpu, ref, f, kind, bc=cfg.choice_ice_margin == 0)
        assert e < 1e-6, (name, f, e)


@pytest.mark.parametrize("name", TRAJECTORY)
def test_fixed_number_of_picard_iterations(ctx, name):
    kind, g, cfg, st = B.build(name)
    K = 40 if name == "ssa_icestream" else 15
    config.set_config(cfg, stop_rule=0, DIVA_visc_it_nit=K, DIVA_visc_it_norm_dUV_tol=0.0)
    tight_linear(st)
    gpu, ref = st.copy(), st.copy()
    gpu_solve(ctx, kind, g, cfg, gpu)
    O.solve(kind, g, cfg, ref, O.LIN_BANDED_LU)
    assert gpu.stats["picard_iterations"] == ref.stats["picard_iterations"] == K
    u = vel_fields(kind)[0]
    if name == "limiter":
        # the limiter is active: |u| reaches DIVA_vel_max on the cx-grid
        assert np.abs(ref[u]).max() > 0.99 * cfg.DIVA_vel_max
    names = vel_fields(kind) + ("u_3D_cx", "w_3D_a", "beta_a")
    if kind == "diva":
        names += ("taub_cx", "taub_cy")
    for f in names:
        e = err(gpu, ref, f, kind)
        assert e < 1e-6, (name, f, e)


def test_ssa_icestream_matches_schoof_closed_form(ctx):
    """The GPU's SSA on the Schoof (2006) ice stream against the closed form
    (src/utilities_module.f90:836-877), independent of the oracle."""
    import test_oracle_pinning as P
    kind, g, cfg, st = B.build("ssa_icestream")
    ua = np.array([P._schoof(y, cfg.SSA_icestream_L, cfg.SSA_icestream_m, 3e-4, 2000.0, 1e-17) for y in g.y])
    st["u_SSA_cx"] = 0.5 * ua[:, None]
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-8, DIVA_visc_it_nit=3000)
    tight_linear(st, rtol=1e-2)
    ctx.solve_ssa(g, cfg, st)
    assert st.stats["nonlinear_residual"] <= 1e-8
    u = st["u_SSA_cx"][:, 2]
    assert np.abs(u - ua).max() <= 0.03 * ua.max()


@pytest.mark.parametrize("name", ["dome", "channel_small", "ssa_dome"])
def test_reference_picard_stop_rule(ctx, name):
    """IMAU_STOP_PICARD: the reference rule (src/ice_velocity_module.f90:1588-1596)
    with its default tolerance 1e-2 stops the GPU and the oracle after the
    same number of Picard iterations, at the same velocities."""
    from imau_vel import geometry
    if name == "dome":
        kind, (g, cfg, st) = "diva", geometry.dome()
    elif name == "channel_small":
        kind, (g, cfg, st) = "diva", geometry.channel(nx=61, ny=21, dx=8e3, seed=3)
    else:
        kind, g, cfg, st = B.build("ssa_dome")
    config.set_config(cfg, stop_rule=0, DIVA_visc_it_norm_dUV_tol=1e-2, DIVA_visc_it_nit=50)
    tight_linear(st)
    gpu, ref = st.copy(), st.copy()
    gpu_solve(ctx, kind, g, cfg, gpu)
    O.solve(kind, g, cfg, ref, O.LIN_BANDED_LU)
    print(name, "picard gpu/oracle", gpu.stats["picard_iterations"], ref.stats["picard_iterations"],
          gpu.stats["resid_UV"], ref.stats["resid_UV"])
    assert 1 < ref.stats["picard_iterations"] < 50
    assert gpu.stats["picard_iterations"] == ref.stats["picard_iterations"]
    assert abs(gpu.stats["resid_UV"] - ref.stats["resid_UV"]) <= 1e-6 * ref.stats["resid_UV"]
    for f in vel_fields(kind):
        assert err(gpu, ref, f, kind) < 1e-6, f


@pytest.mark.parametrize("name", B.SIA_CASES)
def test_sia(ctx, name):
    kind, g, cfg, st = B.build(name)
    gpu, ref = st.copy(), st.copy()
    ctx.solve_sia(g, cfg, gpu)
    O.solve("sia", g, cfg, ref)
    for f in ("u_3D_SIA_cx", "v_3D_SIA_cy", "u_3D_cx", "v_3D_cy", "u_vav_cx", "v_vav_cy", "u_surf_a", "w_3D_a"):
        e = relerr(gpu[f], ref[f])
        assert e < 1e-12, (name, f, e)


def test_hybrid_sia_ssa(ctx):
    """choice_ice_dynamics = 'SIA/SSA': solve_SIA then solve_SSA; the
    combined 3D and depth-averaged velocities
    (src/ice_velocity_module.f90:377-533) within 1e-6 of the oracle's."""
    kind, g, cfg, st = B.build("ssa_dome")
    config.set_config(cfg, choice_ice_dynamics="SIA/SSA", stop_rule=1, nonlinear_rtol=1e-11, DIVA_visc_it_nit=3000)
    tight_linear(st, rtol=1e-2)
    gpu, ref = st.copy(), st.copy()
    ctx.solve_sia(g, cfg, gpu)
    ctx.solve_ssa(g, cfg, gpu)
    O.solve("sia", g, cfg, ref)
    O.solve("ssa", g, cfg, ref, O.LIN_BANDED_LU)
    for f in ("u_SSA_cx", "v_SSA_cy", "u_3D_SIA_cx", "u_3D_cx", "v_3D_cy", "u_vav_cx", "v_vav_cy", "u_surf_a",
              "u_base_a", "w_3D_a"):
        e = relerr(gpu[f], ref[f])
        assert e < 1e-6, (f, e)


@pytest.mark.parametrize("law,param", [("Weertman", "Martin2011"), ("Coulomb", "MISMIP+")])
def test_reference_rejects_roughness_combinations(ctx, law, param):
    """calc_bed_roughness_Martin2011 / _MISMIPplus crash for the sliding laws
    they do not define (src/basal_conditions_and_sliding_module.f90:789-792,
    896-898); the library returns IMAU_ERR_CONFIG there."""
    from imau_vel import solver
    kind, g, cfg, st = B.sliding_case(law, "dry", "parameterised", param)
    with pytest.raises(solver.ImauError) as e:
        ctx.solve_diva(g, cfg, st)
    assert e.value.code == -1


@pytest.mark.parametrize("axis", ["x", "y"])
def test_periodic_slab_against_oracle_and_its_period(ctx, axis):
    """The SSA slab with a thickness wave and periodic rows along it (the
    oracle's pinned periodic case): the GPU solution repeats with the period
    of the boundary rows and meets the oracle's fixed point within 1e-6."""
    import test_oracle_pinning as P
    Pn, ref = P._periodic_slab(axis, 0)
    g = ref.grid
    along = ("west", "east") if axis == "x" else ("south", "north")
    across = ("south", "north") if axis == "x" else ("west", "east")
    bcs = {"DIVA_boundary_BC_%s_%s" % (c, s): "periodic" for c in ("u", "v") for s in along}
    bcs.update({"DIVA_boundary_BC_%s_%s" % (c, s): "infinite" for c in ("u", "v") for s in across})
    from imau_vel import geometry
    cfg = geometry.synthetic_config(choice_ice_dynamics="SSA", stop_rule=1, nonlinear_rtol=1e-12,
                                    DIVA_visc_it_nit=3000, **bcs)
    st = ref.copy()
    for f in ("u_SSA_cx", "v_SSA_cy", "taub_cx", "taub_cy", "visc_eff_3D_a"):
        st[f] = 0.0
    tight_linear(st, rtol=1e-2)
    ctx.solve_ssa(g, cfg, st)
    assert st.stats["nonlinear_residual"] <= 1e-12
    for f in ("u_SSA_cx", "v_SSA_cy", "beta_a", "u_surf_a"):
        e = relerr(st[f], ref[f])
        assert e < 1e-6, (f, e)
    u = st["u_SSA_cx"] if axis == "x" else st["v_SSA_cy"].T
    n = u.shape[1]
    assert np.abs(u[1:-1, Pn:] - u[1:-1, :n - Pn]).max() <= 1e-7 * np.abs(u).max()


def _visc_errs(u_new, u_old, chk):
    """estimate_visc_iter_UV_errors (src/ice_velocity_module.f90:1602-1652)."""
    return np.where(np.abs(chk) > 1e-6, 2.0 * np.abs(u_new - u_old) / np.abs(u_new + u_old + 1e-5), 0.0)


@pytest.mark.parametrize("name", ["no_sliding", "schoof_mismip", "ssa_dome"])
def test_visc_iter_error_estimate(ctx, name):
    """DIVA_err_cx / cy after K fixed Picard iterations of the reference loop
    hold iterate K-1 against K-2 (:355; :238 for solve_SSA, whose tolerance
    test reads the state's u_vav_cx).  Against the GPU's own run of K-1
    iterations (which ends on those two iterates) to rounding; against the
    oracle cell by cell within the bound the formula propagates from the 1e-6
    velocity bar delta: |d err| <= 2 delta (2 + err) / |u_new + u_old + 1e-5|."""
    kind, g, cfg, st = B.build(name)
    K = 6
    tight_linear(st)
    u, v = vel_fields(kind)
    if kind == "ssa":     # the state's vertical average: zero on half of each grid
        st["u_vav_cx"] = 1.0
        st["u_vav_cx"][: g.ny // 2] = 0.0
        st["v_vav_cy"] = 1.0
        st["v_vav_cy"][:, : g.nx // 2] = 0.0
    runs = {}
    for k in (K - 1, K):
        config.set_config(cfg, stop_rule=0, DIVA_visc_it_nit=k, DIVA_visc_it_norm_dUV_tol=0.0)
        gpu, ref = st.copy(), st.copy()
        gpu_solve(ctx, kind, g, cfg, gpu)
        O.solve(kind, g, cfg, ref, O.LIN_BANDED_LU)
        assert gpu.stats["picard_iterations"] == ref.stats["picard_iterations"] == k
        runs[k] = gpu, ref
    (g1, r1), (gK, rK) = runs[K - 1], runs[K]
    delta = 1e-6 * max(np.abs(r1[u]).max(), np.abs(r1[v]).max())
    for e, f, p, chk in (("DIVA_err_cx", u, "u_cx_prev", "u_vav_cx"), ("DIVA_err_cy", v, "v_cy_prev", "v_vav_cy")):
        cg = g1[f] if kind == "diva" else st[chk]
        np.testing.assert_allclose(gK[e], _visc_errs(g1[f], g1[p], cg), rtol=1e-13, atol=0.0, err_msg=e)
        cr = r1[f] if kind == "diva" else st[chk]
        np.testing.assert_allclose(rK[e], _visc_errs(r1[f], r1[p], cr), rtol=1e-13, atol=0.0, err_msg=e)
        s = np.abs(r1[f] + r1[p] + 1e-5)
        bound = 2.0 * delta * (2.0 + rK[e]) / np.maximum(s, 1e-300)
        away = np.abs(np.abs(cr) - 1e-6) > (delta if kind == "diva" else 0.0)   # off the test's threshold
        bad = away & (np.abs(gK[e] - rK[e]) > bound)
        assert not bad.any(), (e, np.argwhere(bad)[:5], gK[e][bad][:5], rK[e][bad][:5])
        assert (rK[e][away] > 0).any() and (rK[e][away] < 1e5).all()
        if kind == "ssa":
            assert (rK[e][np.abs(cr) <= 1e-6] == 0).all() and (np.abs(r1[f][np.abs(cr) <= 1e-6]) > 1e-3).any()


def test_visc_iter_error_estimate_of_a_batch(ctx):
    """Members of one batch that leave the reference Picard loop at different
    iterations keep the estimate of their own last continuing iteration: each
    equals its single solve; under the residual rule the estimate is the
    final iterate against u_cx_prev."""
    from imau_vel import geometry, solver
    cases = [geometry.dome(), geometry.random_geometry(45, 37, seed=4), geometry.dome(nx=21, ny=17)]
    cfg = cases[0][1]
    # the third has no ice: its solve returns before the loop (:273-279) and
    # leaves DIVA_err as the caller passed it
    empty = cases[2][2]
    empty["Hi_a"] = 0.0
    geometry.update_general_ice_model_data(empty, config.zeta_of(cfg))
    empty["DIVA_err_cx"] = 7.0
    empty["DIVA_err_cy"] = 8.0
    config.set_config(cfg, stop_rule=0, DIVA_visc_it_norm_dUV_tol=1e-2, DIVA_visc_it_nit=50)
    singles = []
    for g, _, st in cases:
        s = st.copy()
        ctx.solve_diva(g, cfg, s)
        singles.append(s)
    assert singles[0].stats["picard_iterations"] != singles[1].stats["picard_iterations"]
    batch = [st.copy() for _, _, st in cases]
    prob = solver.Problem(ctx, [c[0] for c in cases], cfg, batch)
    try:
        prob.solve()
        prob.download()
    finally:
        prob.close()
    for s1, sb in zip(singles, batch):
        assert s1.stats["picard_iterations"] == sb.stats["picard_iterations"]
        for e in ("DIVA_err_cx", "DIVA_err_cy"):
            np.testing.assert_array_equal(s1[e], sb[e], err_msg=e)
            assert (s1[e] < 1e5).all()
    assert (batch[2]["DIVA_err_cx"] == 7.0).all() and (batch[2]["DIVA_err_cy"] == 8.0).all()
    ref = empty.copy()
    O.solve("diva", cases[2][0], cfg, ref)
    assert (ref["DIVA_err_cx"] == 7.0).all() and (ref["u_vav_cx"] == 0.0).all()
    config.set_config(cfg, stop_rule=1, nonlinear_rtol=1e-6, DIVA_visc_it_nit=300)
    g, _, st = cases[0]
    ctx.solve_diva(g, cfg, st)
    for e, f, p in (("DIVA_err_cx", "u_vav_cx", "u_cx_prev"), ("DIVA_err_cy", "v_vav_cy", "v_cy_prev")):
        np.testing.assert_allclose(st[e], _visc_errs(st[f], st[p], st[f]), rtol=1e-13, atol=0.0, err_msg=e)


def test_ismip_hom_e_needs_51_columns(ctx):
    """The Arolla file has 51 records, one per grid column; another grid is
    rejected (IMAU_ERR_CONFIG) as by the oracle."""
    from imau_vel import geometry
    g, cfg, st = geometry.random_geometry(41, 37, seed=11, cfg=geometry.synthetic_config(
        choice_sliding_law="idealised", choice_idealised_sliding_law="ISMIP_HOM_E"))
    with pytest.raises(RuntimeError, match="51 lines"):
        ctx.solve_diva(g, cfg, st)
