"""Host-side mirror of the velocity fields of type_ice_model.

Arrays are numpy, Fortran order, with the reference's shapes
(src/ice_dynamics_module.f90:1052-1196): a-grid (ny,nx), cx (ny,nx-1),
cy (ny-1,nx), b (ny-1,nx-1), 3D (nz,ny,nx) etc.  `IceState.io()` builds the
imau_ice_io struct that points at them.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi
from .abi import ImauGrid, ImauIceIO

# name -> (grid, is3d, dtype)
_F64, _I32 = np.float64, np.int32
FIELDS = {
    # inputs
    "Hi_a": ("a", 0, _F64), "Hb_a": ("a", 0, _F64), "SL_a": ("a", 0, _F64), "Hs_a": ("a", 0, _F64),
    "TAF_a": ("a", 0, _F64), "dHs_dx_a": ("a", 0, _F64), "dHs_dy_a": ("a", 0, _F64),
    "dHi_dx_a": ("a", 0, _F64), "dHi_dy_a": ("a", 0, _F64), "dHb_dt_a": ("a", 0, _F64),
    "A_flow_vav_a": ("a", 0, _F64),
    "mask_land_a": ("a", 0, _I32), "mask_ocean_a": ("a", 0, _I32), "mask_ice_a": ("a", 0, _I32),
    "mask_sheet_a": ("a", 0, _I32), "mask_shelf_a": ("a", 0, _I32),
    "Hi_cx": ("cx", 0, _F64), "dHs_dx_cx": ("cx", 0, _F64), "dHs_dy_cx": ("cx", 0, _F64),
    "Hi_cy": ("cy", 0, _F64), "dHs_dx_cy": ("cy", 0, _F64), "dHs_dy_cy": ("cy", 0, _F64),
    "Hi_b": ("b", 0, _F64),
    "A_flow_3D_a": ("a", 1, _F64), "A_flow_3D_cx": ("cx", 1, _F64), "A_flow_3D_cy": ("cy", 1, _F64),
    "Ti_a": ("a", 1, _F64), "Ti_pmp_a": ("a", 1, _F64), "ISMIP_HOM_E_slides": ("col", 0, _I32),
    "beta_sq_a": ("a", 0, _F64), "alpha_sq_a": ("a", 0, _F64), "phi_fric_a": ("a", 0, _F64), "tauc_a": ("a", 0, _F64),
    # in/out state
    "u_vav_cx": ("cx", 0, _F64), "v_vav_cy": ("cy", 0, _F64), "u_SSA_cx": ("cx", 0, _F64), "v_SSA_cy": ("cy", 0, _F64),
    "taub_cx": ("cx", 0, _F64), "taub_cy": ("cy", 0, _F64), "visc_eff_3D_a": ("a", 1, _F64),
    # outputs
    "u_3D_cx": ("cx", 1, _F64), "v_3D_cy": ("cy", 1, _F64), "u_3D_SIA_cx": ("cx", 1, _F64), "v_3D_SIA_cy": ("cy", 1, _F64),
    "u_3D_a": ("a", 1, _F64), "v_3D_a": ("a", 1, _F64), "w_3D_a": ("a", 1, _F64), "F1_3D_a": ("a", 1, _F64),
    "du_dz_3D_cx": ("cx", 1, _F64), "dv_dz_3D_cy": ("cy", 1, _F64),
    "u_vav_a": ("a", 0, _F64), "v_vav_a": ("a", 0, _F64), "u_surf_a": ("a", 0, _F64), "v_surf_a": ("a", 0, _F64),
    "u_base_a": ("a", 0, _F64), "v_base_a": ("a", 0, _F64), "uabs_vav_a": ("a", 0, _F64),
    "uabs_surf_a": ("a", 0, _F64), "uabs_base_a": ("a", 0, _F64), "w_surf_a": ("a", 0, _F64),
    "w_base_a": ("a", 0, _F64), "R_shear": ("a", 0, _F64),
    "u_surf_cx": ("cx", 0, _F64), "v_surf_cy": ("cy", 0, _F64), "u_base_cx": ("cx", 0, _F64), "v_base_cy": ("cy", 0, _F64),
    "beta_a": ("a", 0, _F64), "beta_eff_a": ("a", 0, _F64), "beta_eff_cx": ("cx", 0, _F64), "beta_eff_cy": ("cy", 0, _F64),
    "taudx_cx": ("cx", 0, _F64), "taudy_cy": ("cy", 0, _F64),
    "N_a": ("a", 0, _F64), "N_b": ("b", 0, _F64), "N_cx": ("cx", 0, _F64), "N_cy": ("cy", 0, _F64),
    "visc_eff_int_a": ("a", 0, _F64), "F2_a": ("a", 0, _F64),
    "f_grnd_a": ("a", 0, _F64), "f_grnd_cx": ("cx", 0, _F64), "f_grnd_cy": ("cy", 0, _F64), "f_grnd_b": ("b", 0, _F64),
    "overburden_pressure_a": ("a", 0, _F64), "pore_water_pressure_a": ("a", 0, _F64), "Neff_a": ("a", 0, _F64),
    "DIVA_mask_combi": ("a", 0, _I32), "DIVA_isfront_inner": ("a", 0, _I32), "DIVA_isfront_outer": ("a", 0, _I32),
    "u_cx_prev": ("cx", 0, _F64), "v_cy_prev": ("cy", 0, _F64),
    "DIVA_err_cx": ("cx", 0, _F64), "DIVA_err_cy": ("cy", 0, _F64),
}

STATE_FIELDS = ("u_vav_cx", "v_vav_cy", "u_SSA_cx", "v_SSA_cy", "taub_cx", "taub_cy", "visc_eff_3D_a")


def grid_shape(kind: str, nx: int, ny: int):
    return {"a": (ny, nx), "cx": (ny, nx - 1), "cy": (ny - 1, nx), "b": (ny - 1, nx - 1), "col": (nx,)}[kind]


class Grid:
    """type_grid subset: a-grid coordinates x(nx), y(ny) and spacing dx."""

    def __init__(self, nx: int, ny: int, dx: float, xmin: float = 0.0, ymin: float = 0.0):
        if nx < 4 or ny < 4:
            raise ValueError("grid must be at least 4x4")
        self.nx, self.ny, self.dx = int(nx), int(ny), float(dx)
        self.x = np.ascontiguousarray(xmin + dx * np.arange(nx, dtype=np.float64))
        self.y = np.ascontiguousarray(ymin + dx * np.arange(ny, dtype=np.float64))

    def struct(self) -> ImauGrid:
        g = ImauGrid()
        g.nx, g.ny, g.dx = self.nx, self.ny, self.dx
        g.x = self.x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        g.y = self.y.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return g


class IceState:
    """All velocity-path fields of one grid, zero-initialised like the reference's shared windows."""

    def __init__(self, grid: Grid, nz: int):
        self.grid, self.nz = grid, int(nz)
        self.f = {}
        for name, (kind, is3d, dt) in FIELDS.items():
            shp = grid_shape(kind, grid.nx, grid.ny)
            if is3d:
                shp = (self.nz,) + shp
            self.f[name] = np.zeros(shp, dtype=dt, order="F")
        self.DIVA_SOR_nit = 10000
        self.DIVA_SOR_tol = 2.5
        self.DIVA_SOR_omega = 1.3
        self.DIVA_PETSc_rtol = 0.01
        self.DIVA_PETSc_abstol = 2.5
        self.stats = {}

    def __getitem__(self, k):
        return self.f[k]

    def __setitem__(self, k, v):
        a = self.f[k]
        a[...] = v

    def copy(self) -> "IceState":
        s = IceState(self.grid, self.nz)
        for k, v in self.f.items():
            s.f[k][...] = v
        for k in ("DIVA_SOR_nit", "DIVA_SOR_tol", "DIVA_SOR_omega", "DIVA_PETSc_rtol", "DIVA_PETSc_abstol"):
            setattr(s, k, getattr(self, k))
        return s

    def io(self, outputs=None) -> ImauIceIO:
        """imau_ice_io pointing at this state's arrays.  `outputs` (a set of
        field names) restricts which output pointers are passed; inputs and
        in/out state are always passed."""
        io = ImauIceIO()
        for name in abi.IO_POINTER_FIELDS:
            if name not in self.f:
                continue
            if outputs is not None and name not in outputs and not _is_input(name):
                continue
            arr = self.f[name]
            ct = ctypes.c_int32 if arr.dtype == np.int32 else ctypes.c_double
            setattr(io, name, arr.ctypes.data_as(ctypes.POINTER(ct)))
        io.DIVA_SOR_nit = self.DIVA_SOR_nit
        io.DIVA_SOR_tol = self.DIVA_SOR_tol
        io.DIVA_SOR_omega = self.DIVA_SOR_omega
        io.DIVA_PETSc_rtol = self.DIVA_PETSc_rtol
        io.DIVA_PETSc_abstol = self.DIVA_PETSc_abstol
        return io

    def take_stats(self, io: ImauIceIO):
        self.stats = dict(picard_iterations=io.stat_picard_iterations, linear_sweeps=io.stat_linear_sweeps,
                          resid_UV=io.stat_resid_UV, nonlinear_residual=io.stat_nonlinear_residual,
                          seconds=io.stat_seconds)
        return self.stats


_INPUTS = set(list(FIELDS)[: list(FIELDS).index("u_vav_cx")]) | set(STATE_FIELDS)


def _is_input(name: str) -> bool:
    return name in _INPUTS


THICKNESS_FIELDS = {
    "Hi_a": ("a", _F64), "Hb_a": ("a", _F64), "SL_a": ("a", _F64), "u_vav_cx": ("cx", _F64), "v_vav_cy": ("cy", _F64),
    "SMB_year": ("a", _F64), "BMB": ("a", _F64), "float_margin_frac_a": ("a", _F64),
    "mask_ice_a": ("a", _I32), "mask_shelf_a": ("a", _I32), "mask_ocean_a": ("a", _I32), "mask_cf_a": ("a", _I32),
    "dHi_dt_a": ("a", _F64), "Hi_tplusdt_a": ("a", _F64), "Qx_cx": ("cx", _F64), "Qy_cy": ("cy", _F64),
}


class ThicknessState:
    """The fields of one ice-thickness update (imau_thickness_io), Fortran order."""

    def __init__(self, grid: Grid):
        self.grid = grid
        self.f = {name: np.zeros(grid_shape(kind, grid.nx, grid.ny), dtype=dt, order="F")
                  for name, (kind, dt) in THICKNESS_FIELDS.items()}

    def __getitem__(self, k):
        return self.f[k]

    def __setitem__(self, k, v):
        self.f[k][...] = v

    def copy(self) -> "ThicknessState":
        s = ThicknessState(self.grid)
        for k, v in self.f.items():
            s.f[k][...] = v
        return s

    def io(self):
        io = abi.ImauThicknessIO()
        for name, arr in self.f.items():
            ct = ctypes.c_int32 if arr.dtype == np.int32 else ctypes.c_double
            setattr(io, name, arr.ctypes.data_as(ctypes.POINTER(ct)))
        return io


# (grid, 3D, dtype) of every imau_producer_io field
PRODUCER_FIELDS = {
    "Hi_a": ("a", 0, _F64), "Hb_a": ("a", 0, _F64), "SL_a": ("a", 0, _F64),
    "dHi_dt_a": ("a", 0, _F64), "dHb_dt_a": ("a", 0, _F64), "Ti_a": ("a", 1, _F64),
    "Hs_a": ("a", 0, _F64), "TAF_a": ("a", 0, _F64), "dHs_dt_a": ("a", 0, _F64),
    "mask_land_a": ("a", 0, _I32), "mask_ocean_a": ("a", 0, _I32), "mask_ice_a": ("a", 0, _I32),
    "mask_sheet_a": ("a", 0, _I32), "mask_shelf_a": ("a", 0, _I32), "mask_coast_a": ("a", 0, _I32),
    "mask_margin_a": ("a", 0, _I32), "mask_gl_a": ("a", 0, _I32), "mask_cf_a": ("a", 0, _I32),
    "mask_a": ("a", 0, _I32),
    "Hi_cx": ("cx", 0, _F64), "Hi_cy": ("cy", 0, _F64), "Hi_b": ("b", 0, _F64),
    "A_flow_3D_a": ("a", 1, _F64), "A_flow_vav_a": ("a", 0, _F64),
    "A_flow_3D_cx": ("cx", 1, _F64), "A_flow_3D_cy": ("cy", 1, _F64),
    "A_flow_vav_cx": ("cx", 0, _F64), "A_flow_vav_cy": ("cy", 0, _F64),
    "dHi_dx_a": ("a", 0, _F64), "dHi_dy_a": ("a", 0, _F64), "dHs_dx_a": ("a", 0, _F64), "dHs_dy_a": ("a", 0, _F64),
    "dHs_dx_cx": ("cx", 0, _F64), "dHs_dy_cx": ("cx", 0, _F64), "dHs_dx_cy": ("cy", 0, _F64),
    "dHs_dy_cy": ("cy", 0, _F64),
    "float_margin_frac_a": ("a", 0, _F64), "Hi_eff_cf_a": ("a", 0, _F64),
}


class ProducerState:
    """The fields of update_general_ice_model_data / calc_ice_rheology
    (imau_producer_io), Fortran order; 3D fields are (nz, ny, nx)."""

    def __init__(self, grid: Grid, nz: int):
        self.grid, self.nz = grid, nz
        self.f = {}
        for name, (kind, is3d, dt) in PRODUCER_FIELDS.items():
            shape = grid_shape(kind, grid.nx, grid.ny)
            self.f[name] = np.zeros(((nz,) + shape) if is3d else shape, dtype=dt, order="F")

    def __getitem__(self, k):
        return self.f[k]

    def __setitem__(self, k, v):
        self.f[k][...] = v

    def copy(self) -> "ProducerState":
        s = ProducerState(self.grid, self.nz)
        for k, v in self.f.items():
            s.f[k][...] = v
        return s

    def io(self, omit=()):
        """imau_producer_io over these arrays; fields in `omit` are passed as NULL."""
        io = abi.ImauProducerIO()
        for name, arr in self.f.items():
            if name in omit:
                continue
            ct = ctypes.c_int32 if arr.dtype == np.int32 else ctypes.c_double
            setattr(io, name, arr.ctypes.data_as(ctypes.POINTER(ct)))
        return io


class RefGeo:
    """The fixed inputs of update_ice_thickness (src/ice_dynamics_module.f90:435-445):
    mask_noice and the present-day / GIA-equilibrium reference geometries,
    each (ny, nx) or None."""

    NAMES = ("mask_noice_a", "Hi_PD", "Hb_PD", "Hi_GIAeq", "Hb_GIAeq")

    def __init__(self, grid: Grid, **fields):
        self.f = {}
        for name, arr in fields.items():
            if name not in self.NAMES:
                raise KeyError("imau_refgeo_io has no field %r" % name)
            if arr is None:
                continue
            dt = np.int32 if name == "mask_noice_a" else np.float64
            self.f[name] = np.asfortranarray(np.broadcast_to(arr, (grid.ny, grid.nx)), dtype=dt)

    def io(self):
        io = abi.ImauRefgeoIO()
        for name, arr in self.f.items():
            ct = ctypes.c_int32 if arr.dtype == np.int32 else ctypes.c_double
            setattr(io, name, arr.ctypes.data_as(ctypes.POINTER(ct)))
        return io


class PcState:
    """The predictor-corrector state run_ice_dynamics_pc carries between time
    steps (imau_pc_state), as the reference initialises it at the start of a
    run (src/IMAU_ICE_main_model.f90:478-541, src/ice_dynamics_module.f90:993-997):
    pc_zeta = 1, pc_eta = pc_eta_prev = pc_epsilon, the time steps dt_min, the
    next velocity update at start_time_of_run, both rate fields zero."""

    def __init__(self, grid: Grid, scfg, time=None):
        self.dHidt_Hn_un = np.zeros((grid.ny, grid.nx), order="F")
        self.dHi_dt_a = np.zeros((grid.ny, grid.nx), order="F")
        s = abi.ImauPcState()
        s.pc_zeta = 1.0
        s.pc_eta = s.pc_eta_prev = scfg.pc_epsilon
        s.dt_crit_ice = s.dt_crit_ice_prev = s.dt = scfg.dt_min
        s.t_next = scfg.start_time_of_run if time is None else time   # the run's first step updates
        self.s = s
        self._bind()

    def _bind(self):
        dp = ctypes.POINTER(ctypes.c_double)
        self.s.dHidt_Hn_un = self.dHidt_Hn_un.ctypes.data_as(dp)
        self.s.dHi_dt_a = self.dHi_dt_a.ctypes.data_as(dp)

    def copy(self) -> "PcState":
        c = PcState.__new__(PcState)
        c.dHidt_Hn_un, c.dHi_dt_a = self.dHidt_Hn_un.copy(order="F"), self.dHi_dt_a.copy(order="F")
        c.s = abi.ImauPcState()
        ctypes.memmove(ctypes.addressof(c.s), ctypes.addressof(self.s), ctypes.sizeof(c.s))
        c._bind()
        return c
