"""Small inputs that reach each branch of the velocity path.

Every case is a (kind, grid, cfg, state) tuple, where kind is "diva", "ssa"
or "sia".  They are small enough for the oracle's exact banded LU to solve
them in about a second.  The GPU parity tests and the CPU oracle tests both
use them.

Branches and the reference lines they reach:
  sliding laws         src/basal_conditions_and_sliding_module.f90:908-1432
  hydrology            :90-238 (dry, saturated, Martin2011)
  roughness            :243-294, :769-903 (Martin2011, SSA_icestream, MISMIP+)
  basal freezing       src/ice_velocity_module.f90:1051-1055, 1162-1166
  BC-mode margin rows  :1762-1865 (shelf-land zero rows), 2192-2341 (fronts)
  'sans' rows          :2468-2589
  periodic boundaries  :1932-2191
  solve_SSA / SIA      :115-248 / :36-114
"""
from __future__ import annotations

import numpy as np

from imau_vel import config, geometry
from imau_vel.state import Grid, IceState


def _channel_small(cfg=None, seed=3):
    return geometry.channel(nx=61, ny=21, dx=8e3, seed=seed, cfg=cfg)


def _channel_cfg(**kw):
    return geometry.synthetic_config(DIVA_boundary_BC_u_west="zero", DIVA_boundary_BC_v_west="zero",
                                     DIVA_boundary_BC_v_south="zero", DIVA_boundary_BC_v_north="zero", **kw)


def _refresh(st, cfg):
    geometry.update_general_ice_model_data(st, config.zeta_of(cfg))


def sliding_case(law: str, hydrology: str = "saturated", roughness: str = "uniform", param: str = "Martin2011"):
    """The 61 x 21 marine channel (floating ice at its seaward end) under the
    given sliding law, hydrology and roughness choice.  Saturated hydrology
    makes the effective pressure fall to zero at the grounding line, so the
    plastic laws see both their frictional and their weak regime."""
    cfg = _channel_cfg(choice_sliding_law=law, choice_basal_hydrology=hydrology, choice_basal_roughness=roughness,
                       choice_param_basal_roughness=param)
    g, cfg, st = _channel_small(cfg)
    geometry.fill_uniform_roughness(st, cfg)
    if law in ("Coulomb", "Coulomb_regularised", "Zoet-Iverson"):
        # a weak till (5 degrees) so the plastic laws let the ice slide
        st["phi_fric_a"] = 5.0
    return "diva", g, cfg, st


def idealised_case(which: str):
    """ISMIP-HOM C / D / F friction fields (:1288-1432) on a random cap."""
    cfg = geometry.synthetic_config(choice_sliding_law="idealised", choice_idealised_sliding_law=which,
                                    ISMIP_HOM_L=80e3)
    g, cfg, st = geometry.random_geometry(41, 37, seed=11, cfg=cfg)
    return "diva", g, cfg, st


def arolla_like_rows():
    """51 records in the ISMIP-HOM E file format (x, Hb, Hs, slides): a
    synthetic valley-glacier flowline 5 km long at 100 m spacing, thickness
    20-300 m, with a slip zone for 2200 <= x <= 2500 m.  The shape of the
    Haut Glacier d'Arolla input, not its data."""
    x = np.arange(51) * 100.0
    Hb = 2600.0 - 0.12 * x + 40.0 * np.sin(2.0 * np.pi * x / 2500.0)
    Hs = Hb + 20.0 + 280.0 * np.sin(np.pi * x / 5000.0)
    slides = ((x >= 2200.0) & (x <= 2500.0)).astype(np.int32)
    return x, Hb, Hs, slides


def ismip_e_case():
    """ISMIP-HOM E (:1352-1399): beta 0 in the slip columns, 1e30 (clamped to
    DIVA_beta_max) elsewhere, on the 51-column flowline."""
    g, cfg, st = geometry.ismip_hom_e(arolla_like_rows(), ny=5)
    return "diva", g, cfg, st


def freezing_case():
    """include_basal_freezing: the western half of the bed is frozen (basal
    temperature 5 K below the pressure-melting point), the eastern half at
    the melting point (src/ice_velocity_module.f90:1051-1055, 1162-1166)."""
    cfg = geometry.synthetic_config(include_basal_freezing=1, deltaT_basal_freezing=2.0)
    g, cfg, st = geometry.random_geometry(41, 37, seed=12, cfg=cfg)
    X = np.meshgrid(g.x, g.y)[0]
    tpmp = 273.15 - 8.7e-4 * st["Hi_a"]
    st["Ti_pmp_a"] = tpmp[None, :, :]
    base = np.where(X < 0.0, tpmp - 5.0, tpmp)
    ti = np.repeat(base[None, :, :], cfg.nz, axis=0)
    ti[:-1] -= 10.0
    st["Ti_a"] = ti
    return "diva", g, cfg, st


def bc_margin_case(kind: str = "diva"):
    """choice_ice_margin = 'BC' on the 61 x 21 channel: the shelf calves two
    cells before the eastern domain edge (calving-front stress rows), and
    the floating ice next to the ice-free side walls gives shelf-land zero
    rows.  Only one column of ice-free ocean lies between the front and the
    domain boundary: further ice-free ocean would give rows with no
    coefficient at all (N = 0 and no friction), which the reference's own
    assembly leaves singular as well."""
    cfg = _channel_cfg(choice_ice_margin="BC", choice_ice_dynamics="SSA" if kind == "ssa" else "DIVA")
    g, cfg, st = _channel_small(cfg)
    H = st["Hi_a"].copy()
    H[:, g.nx - 2:] = 0.0
    st["Hi_a"] = H
    _refresh(st, cfg)
    geometry.fill_uniform_roughness(st, cfg)
    return kind, g, cfg, st


def sans_case(which: str = "random"):
    """include_SSADIVA_crossterms = .FALSE.: the normalised 'sans' rows (:2468-2589)."""
    if which == "random":
        cfg = geometry.synthetic_config(include_SSADIVA_crossterms=0)
        g, cfg, st = geometry.random_geometry(25, 21, seed=13, cfg=cfg)
    else:
        g, cfg, st = _channel_small(_channel_cfg(include_SSADIVA_crossterms=0))
    return "diva", g, cfg, st


def periodic_case():
    """Periodic domain boundaries on every side (:1932-2191): an inclined
    slab with a thickness bump, periodic in both directions (the surface
    slope is carried by the bed, which is periodic in its deviation only
    through the driving stress, so the slab is a true periodic problem)."""
    cfg = geometry.synthetic_config(**{"DIVA_boundary_BC_%s_%s" % (c, s): "periodic"
                                       for c in ("u", "v") for s in ("west", "east", "south", "north")})
    nx, ny, dx = 24, 20, 5e3
    g = Grid(nx, ny, dx)
    st = IceState(g, cfg.nz)
    X, Y = np.meshgrid(g.x, g.y)
    Lx, Ly = nx * dx, ny * dx
    st["Hi_a"] = 1000.0 + 200.0 * np.sin(2 * np.pi * X / Lx) * np.cos(2 * np.pi * Y / Ly)
    st["Hb_a"] = -5e-3 * X
    st["SL_a"] = -1e4
    st["A_flow_3D_a"] = 1e-16
    geometry.fill_uniform_roughness(st, cfg)
    _refresh(st, cfg)
    return "diva", g, cfg, st


def limiter_case():
    """Ice on a nearly frictionless bed (beta_sq = 20) whose velocities
    exceed DIVA_vel_max = 5000 m/yr, so apply_velocity_limits (:1432-1498)
    caps them on every Picard iteration."""
    cfg = geometry.synthetic_config()
    g, cfg, st = geometry.random_geometry(33, 29, seed=14, cfg=cfg, beta_sq=20.0)
    return "diva", g, cfg, st


def ssa_dome_case():
    cfg = geometry.synthetic_config(choice_ice_dynamics="SSA")
    g, cfg, st = geometry.dome(cfg=cfg)
    return "ssa", g, cfg, st


def ssa_icestream_case():
    """The Schoof (2006) ice stream: Coulomb sliding with the SSA_icestream
    till yield stress (:813-840; src/utilities_module.f90:836-877)."""
    L, mexp, tant, H, A = 100e3, 0.5, 3e-4, 2000.0, 1e-17
    ny, nx, dx = 121, 5, 5e3
    cfg = config.default_config(choice_ice_dynamics="SSA", choice_sliding_law="Coulomb",
                                choice_basal_roughness="parameterised", choice_param_basal_roughness="SSA_icestream",
                                choice_basal_hydrology="dry", include_basal_freezing=0, SSA_icestream_L=L,
                                SSA_icestream_m=mexp)
    g = Grid(nx, ny, dx, xmin=0.0, ymin=-(ny - 1) / 2 * dx)
    st = IceState(g, cfg.nz)
    st["Hi_a"] = H
    st["Hb_a"] = -tant * np.meshgrid(g.x, g.y)[0]
    st["SL_a"] = -1e5
    st["A_flow_3D_a"] = A
    _refresh(st, cfg)
    return "ssa", g, cfg, st


def ssa_channel_case():
    """solve_SSA on the marine channel with Zoet-Iverson sliding and Martin
    (2011) hydrology and till: the combination most shipped configurations
    run (choice_ice_dynamics = 'SIA/SSA' with Zoet-Iverson)."""
    cfg = _channel_cfg(choice_ice_dynamics="SSA", choice_sliding_law="Zoet-Iverson",
                       choice_basal_hydrology="Martin2011", choice_basal_roughness="parameterised",
                       choice_param_basal_roughness="Martin2011", Martin2011till_phi_min=2.0,
                       Martin2011till_phi_max=8.0)
    g, cfg, st = _channel_small(cfg)
    geometry.fill_uniform_roughness(st, cfg)
    return "ssa", g, cfg, st


def sia_case(which: str = "dome"):
    if which == "dome":
        cfg = geometry.synthetic_config(choice_ice_dynamics="SIA")
        g, cfg, st = geometry.dome(cfg=cfg)
    else:
        cfg = geometry.synthetic_config(choice_ice_dynamics="SIA")
        g, cfg, st = geometry.random_geometry(41, 37, seed=15, cfg=cfg)
        rng = np.random.default_rng(15)
        A3 = 10.0 ** rng.uniform(-17, -16, size=(cfg.nz, g.ny, g.nx))
        st["A_flow_3D_a"] = A3
        _refresh(st, cfg)
    return "sia", g, cfg, st


SLIDING_CASES = {
    "no_sliding": dict(law="no_sliding"),
    "coulomb_saturated": dict(law="Coulomb", hydrology="saturated"),
    "coulomb_reg_martin": dict(law="Coulomb_regularised", hydrology="Martin2011", roughness="parameterised",
                               param="Martin2011"),
    "tsai_saturated": dict(law="Tsai2015", hydrology="saturated"),
    "schoof_mismip": dict(law="Schoof2005", hydrology="saturated", roughness="parameterised", param="MISMIP+"),
    "weertman_mismip": dict(law="Weertman", hydrology="dry", roughness="parameterised", param="MISMIP+"),
    "zoet_iverson_martin": dict(law="Zoet-Iverson", hydrology="Martin2011", roughness="parameterised",
                                param="Martin2011"),
    "zoet_iverson_dry": dict(law="Zoet-Iverson", hydrology="dry"),
}


def build(name: str):
    if name in SLIDING_CASES:
        return sliding_case(**SLIDING_CASES[name])
    if name == "ismip_e":
        return ismip_e_case()
    if name.startswith("ismip_"):
        return idealised_case("ISMIP_HOM_" + name[-1].upper())
    table = {
        "freezing": freezing_case, "bc_margin": bc_margin_case, "bc_margin_ssa": lambda: bc_margin_case("ssa"),
        "sans_random": lambda: sans_case("random"), "sans_channel": lambda: sans_case("channel"),
        "periodic": periodic_case, "limiter": limiter_case, "ssa_dome": ssa_dome_case,
        "ssa_icestream": ssa_icestream_case, "ssa_channel": ssa_channel_case,
        "sia_dome": lambda: sia_case("dome"), "sia_random": lambda: sia_case("random"),
    }
    return table[name]()


DIVA_CASES = sorted(SLIDING_CASES) + ["ismip_c", "ismip_d", "ismip_e", "ismip_f", "freezing", "bc_margin", "sans_random",
                                      "sans_channel", "periodic"]
SSA_CASES = ["ssa_dome", "ssa_icestream", "ssa_channel", "bc_margin_ssa"]
SIA_CASES = ["sia_dome", "sia_random"]
