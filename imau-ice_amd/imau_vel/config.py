"""Velocity-solver configuration with the reference defaults.

Defaults follow src/configuration_module.f90 (velocity block :271-302,
sliding :349-377, zeta :168-171, rheology :425-426, benchmarks :70-85).
Every CHARACTER choice of the reference is an enum here (see include/imau_vel.h).
"""
from __future__ import annotations

from . import abi
from .abi import ImauCfg

# src/configuration_module.f90:168-171
DEFAULT_ZETA = [0.00, 0.10, 0.20, 0.30, 0.40, 0.50, 0.60, 0.70, 0.80, 0.90, 0.925, 0.95, 0.975, 0.99, 1.00]

DYNAMICS = {"none": abi.IMAU_DYN_NONE, "SIA": abi.IMAU_DYN_SIA, "SSA": abi.IMAU_DYN_SSA,
            "SIA/SSA": abi.IMAU_DYN_SIA_SSA, "DIVA": abi.IMAU_DYN_DIVA}
MARGIN = {"BC": abi.IMAU_MARGIN_BC, "infinite_slab": abi.IMAU_MARGIN_INFINITE_SLAB}
SLIDING = {"no_sliding": abi.IMAU_SLID_NO_SLIDING, "idealised": abi.IMAU_SLID_IDEALISED,
           "Coulomb": abi.IMAU_SLID_COULOMB, "Coulomb_regularised": abi.IMAU_SLID_COULOMB_REGULARISED,
           "Weertman": abi.IMAU_SLID_WEERTMAN, "Tsai2015": abi.IMAU_SLID_TSAI2015,
           "Schoof2005": abi.IMAU_SLID_SCHOOF2005, "Zoet-Iverson": abi.IMAU_SLID_ZOET_IVERSON}
IDEALISED = {"ISMIP_HOM_C": abi.IMAU_IDEAL_ISMIP_HOM_C, "ISMIP_HOM_D": abi.IMAU_IDEAL_ISMIP_HOM_D,
             "ISMIP_HOM_E": abi.IMAU_IDEAL_ISMIP_HOM_E, "ISMIP_HOM_F": abi.IMAU_IDEAL_ISMIP_HOM_F}
HYDROLOGY = {"dry": abi.IMAU_HYDRO_DRY, "saturated": abi.IMAU_HYDRO_SATURATED,
             "Martin2011": abi.IMAU_HYDRO_MARTIN2011}
ROUGHNESS = {"uniform": abi.IMAU_ROUGH_UNIFORM, "parameterised": abi.IMAU_ROUGH_PARAMETERISED,
             "prescribed": abi.IMAU_ROUGH_PRESCRIBED}
PARAM_ROUGHNESS = {"Martin2011": abi.IMAU_PROUGH_MARTIN2011, "SSA_icestream": abi.IMAU_PROUGH_SSA_ICESTREAM,
                   "MISMIP+": abi.IMAU_PROUGH_MISMIPPLUS}
BOUNDARY = {"infinite": abi.IMAU_BC_INFINITE, "zero": abi.IMAU_BC_ZERO, "periodic": abi.IMAU_BC_PERIODIC}
SOLVER = {"SOR": abi.IMAU_SOLVER_SOR, "PETSc": abi.IMAU_SOLVER_PETSC}

# uniform roughness values, src/configuration_module.f90:367-372
UNIFORM_WEERTMAN_BETA_SQ = 1.0e4
UNIFORM_COULOMB_PHI_FRIC = 15.0
UNIFORM_TSAI2015_ALPHA_SQ = 0.5
UNIFORM_TSAI2015_BETA_SQ = 1.0e4
UNIFORM_SCHOOF2005_ALPHA_SQ = 0.5
UNIFORM_SCHOOF2005_BETA_SQ = 1.0e4


def default_config(**overrides) -> ImauCfg:
    """Reference defaults; keyword overrides accept enum ints or the reference strings."""
    c = ImauCfg()
    c.choice_ice_dynamics = DYNAMICS["DIVA"]
    c.choice_ice_margin = MARGIN["infinite_slab"]
    c.choice_sliding_law = SLIDING["Coulomb_regularised"]
    c.choice_idealised_sliding_law = 0
    c.choice_basal_hydrology = HYDROLOGY["Martin2011"]
    c.choice_basal_roughness = ROUGHNESS["parameterised"]
    c.choice_param_basal_roughness = PARAM_ROUGHNESS["Martin2011"]
    for side in ("u_west", "u_east", "u_south", "u_north", "v_west", "v_east", "v_south", "v_north"):
        setattr(c, "DIVA_boundary_BC_" + side, BOUNDARY["infinite"])
    c.DIVA_choice_matrix_solver = SOLVER["PETSc"]
    c.include_SSADIVA_crossterms = 1
    c.do_GL_subgrid_friction = 1
    c.include_basal_freezing = 1
    c.nz = len(DEFAULT_ZETA)
    for k, z in enumerate(DEFAULT_ZETA):
        c.zeta[k] = z
    c.n_flow = 3.0
    c.DIVA_visc_it_norm_dUV_tol = 1e-2
    c.DIVA_visc_it_nit = 50
    c.DIVA_visc_it_relax = 0.4
    c.DIVA_epsilon_sq_0 = 1e-15
    c.DIVA_visc_eff_min = 1e3
    c.DIVA_beta_max = 1e20
    c.DIVA_vel_max = 5000.0
    c.slid_delta_v = 1.0e-3
    c.slid_Weertman_m = 3.0
    c.slid_Coulomb_reg_q_plastic = 0.3
    c.slid_Coulomb_reg_u_threshold = 100.0
    c.slid_ZI_ut = 200.0
    c.slid_ZI_p = 5.0
    c.deltaT_basal_freezing = 2.0
    c.Martin2011_hydro_Hb_min = 0.0
    c.Martin2011_hydro_Hb_max = 1000.0
    c.Martin2011till_phi_Hb_min = -1000.0
    c.Martin2011till_phi_Hb_max = 0.0
    c.Martin2011till_phi_min = 5.0
    c.Martin2011till_phi_max = 20.0
    c.SSA_icestream_L = 100000.0
    c.SSA_icestream_m = 0.5
    c.ISMIP_HOM_L = 160000.0
    c.precision = abi.IMAU_PREC_FP64
    c.stop_rule = abi.IMAU_STOP_PICARD
    c.nonlinear_rtol = 1e-4
    c.mg_levels = 0
    c.check_interval = 0
    c.picard_relax_max = 0.0
    c.mg_smoother = 0
    c.ksp_type = 0
    c.anderson_m = 0
    c.mg_rebuild = 1
    c.mg_graph = 1
    c.warm_visc_iters = 0
    c.anderson_beta = 1.0
    c.anderson_restart = 1.5
    c.warm_column_iters = 0
    c.stall_window = 0
    c.stall_drop = 10.0
    c.stall_relax = 0.0
    set_config(c, **overrides)
    return c


_STRING_MAPS = {
    "choice_ice_dynamics": DYNAMICS, "choice_ice_margin": MARGIN, "choice_sliding_law": SLIDING,
    "choice_idealised_sliding_law": IDEALISED, "choice_basal_hydrology": HYDROLOGY,
    "choice_basal_roughness": ROUGHNESS, "choice_param_basal_roughness": PARAM_ROUGHNESS,
    "DIVA_choice_matrix_solver": SOLVER,
    "precision": {"fp64": abi.IMAU_PREC_FP64, "mixed": abi.IMAU_PREC_MIXED,
                  "mixed_visc": abi.IMAU_PREC_MIXED_VISC},
    "ksp_type": {"bicgstab": 0, "gcr": 1},
}


def set_config(c: ImauCfg, **overrides) -> ImauCfg:
    for k, v in overrides.items():
        if k == "zeta":
            c.nz = len(v)
            for i, z in enumerate(v):
                c.zeta[i] = z
            continue
        if isinstance(v, str):
            table = _STRING_MAPS.get(k, BOUNDARY if k.startswith("DIVA_boundary_BC_") else None)
            if table is None or v not in table:
                raise ValueError("unknown choice %r for %s" % (v, k))
            v = table[v]
        if not hasattr(c, k):
            raise AttributeError("imau_cfg has no field %r" % k)
        setattr(c, k, v)
    return c


def zeta_of(c: ImauCfg):
    return [c.zeta[k] for k in range(c.nz)]


INTEGRATION = {"none": abi.IMAU_DHI_NONE, "explicit": abi.IMAU_DHI_EXPLICIT, "semi-implicit": abi.IMAU_DHI_SEMIIMPLICIT}
THICKNESS_BC = {"none": abi.IMAU_HBC_NONE, "zero": abi.IMAU_HBC_ZERO, "infinite": abi.IMAU_HBC_INFINITE,
                "periodic": abi.IMAU_HBC_PERIODIC, "fixed": abi.IMAU_HBC_FIXED}


def thickness_config(**overrides):
    """Reference defaults of the ice-thickness update (src/configuration_module.f90:53,
    308-314, 325-328); keyword overrides accept enum ints or the reference strings."""
    c = abi.ImauThicknessCfg()
    c.choice_ice_integration_method = INTEGRATION["explicit"]
    c.choice_ice_dynamics = DYNAMICS["DIVA"]
    for side in ("west", "east", "south", "north"):
        setattr(c, "ice_thickness_%s_BC" % side, THICKNESS_BC["zero"])
    c.dHi_choice_matrix_solver = SOLVER["SOR"]
    c.dHi_SOR_nit = 3000
    c.dHi_SOR_tol = 2.5
    c.dHi_SOR_omega = 1.3
    c.dHi_PETSc_rtol = 0.001
    c.dHi_PETSc_abstol = 0.001
    c.dt_max = 10.0
    for k, v in overrides.items():
        if isinstance(v, str):
            table = {"choice_ice_integration_method": INTEGRATION, "choice_ice_dynamics": DYNAMICS,
                     "dHi_choice_matrix_solver": SOLVER}.get(k, THICKNESS_BC if k.startswith("ice_thickness_") else None)
            if table is None or v not in table:
                raise ValueError("unknown choice %r for %s" % (v, k))
            v = table[v]
        if not hasattr(c, k):
            raise AttributeError("imau_thickness_cfg has no field %r" % k)
        setattr(c, k, v)
    return c


RHEOLOGY = {"uniform": abi.IMAU_RHEO_UNIFORM, "Huybrechts1992": abi.IMAU_RHEO_HUYBRECHTS1992,
            "MISMIP_mod": abi.IMAU_RHEO_MISMIP_MOD}


def producer_config(**overrides):
    """Reference defaults of the input producers (src/configuration_module.f90:168-171,
    273-274, 425-426, 702); keyword overrides accept enum ints or the reference strings."""
    c = abi.ImauProducerCfg()
    c.nz = len(DEFAULT_ZETA)
    for k, z in enumerate(DEFAULT_ZETA):
        c.zeta[k] = z
    c.choice_ice_margin = MARGIN["infinite_slab"]
    c.do_ocean_floodfill = 1
    c.choice_ice_rheology = RHEOLOGY["Huybrechts1992"]
    c.uniform_flow_factor = 1e-16
    c.m_enh_sheet = 1.0
    c.m_enh_shelf = 1.0
    for k, v in overrides.items():
        if k == "zeta":
            c.nz = len(v)
            for i, z in enumerate(v):
                c.zeta[i] = z
            continue
        if isinstance(v, str):
            table = {"choice_ice_margin": MARGIN, "choice_ice_rheology": RHEOLOGY}.get(k)
            if table is None or v not in table:
                raise ValueError("unknown choice %r for %s" % (v, k))
            v = table[v]
        if not hasattr(c, k):
            raise AttributeError("imau_producer_cfg has no field %r" % k)
        setattr(c, k, v)
    return c


CALVING = {"none": abi.DEFINES["IMAU_CALVING_NONE"],
           "threshold_thickness": abi.DEFINES["IMAU_CALVING_THRESHOLD_THICKNESS"]}


def timestep_config(**overrides):
    """Reference defaults of the time loop (src/configuration_module.f90:50-55),
    of the predictor-corrector step (:317-320) and of update_ice_thickness
    (:335-337, :412-417); choice_calving_law may
    be given by its reference name."""
    c = abi.ImauTimestepCfg()
    c.start_time_of_run = 0.0
    c.end_time_of_run = 50000.0
    c.dt_max = 10.0
    c.dt_min = 0.01
    c.dt_startup_phase = 10.0
    c.choice_calving_law = CALVING["threshold_thickness"]
    c.calving_threshold_thickness = 200.0
    c.continental_shelf_min_height = -2000.0
    c.pc_epsilon, c.pc_k_I, c.pc_k_p, c.pc_eta_min = 0.5, 0.2, 0.2, 1e-8   # :317-320
    if isinstance(overrides.get("choice_calving_law"), str):
        overrides["choice_calving_law"] = CALVING[overrides["choice_calving_law"]]
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise AttributeError("imau_timestep_cfg has no field %r" % k)
        setattr(c, k, v)
    return c


def timestep_state(time, t_next_SIA=None, t_next_SSA=None):
    """region%time and the dynamics timers; a timer left None is due at `time`."""
    s = abi.ImauTimestepState()
    s.time = time
    s.t_next_SIA = time if t_next_SIA is None else t_next_SIA
    s.t_next_SSA = time if t_next_SSA is None else t_next_SSA
    return s
