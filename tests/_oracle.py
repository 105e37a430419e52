"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- the checker."""
import ctypes
import os

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

LIN_SOR, LIN_BANDED_LU = 0, 1
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _lib = ctypes.CDLL(LIB)
        _lib.orc_last_error.restype = ctypes.c_char_p
        _lib.orc_vertical_integrate.restype = ctypes.c_double
        _lib.orc_vertical_average.restype = ctypes.c_double
        _lib.orc_calc_fraction_above_zero.restype = ctypes.c_double
        _lib.orc_calc_fraction_above_zero.argtypes = [ctypes.c_double] * 4
        _lib.orc_sliding_beta.restype = ctypes.c_double
        _lib.orc_sliding_beta.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 10
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def solve(kind, grid, cfg, st, lin=LIN_BANDED_LU):
    """Run the oracle's solve_DIVA / solve_SSA / solve_SIA on IceState st in place."""
    L = lib()
    g = grid.struct()
    io = st.io()
    if kind == "diva":
        rc = L.orc_solve_diva(ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(io), lin)
    elif kind == "ssa":
        rc = L.orc_solve_ssa(ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(io), lin)
    elif kind == "sia":
        rc = L.orc_solve_sia(ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(io))
    else:
        raise ValueError(kind)
    if rc != 0:
        raise RuntimeError("oracle %s failed (%d): %s" % (kind, rc, L.orc_last_error().decode()))
    st.take_stats(io)
    return st


def nonlinear_residual(grid, cfg, st):
    L = lib()
    g = grid.struct()
    io = st.io()
    r = ctypes.c_double()
    L.orc_nonlinear_residual(ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(io), ctypes.byref(r))
    return r.value


def vertical_integrate(zeta, f):
    z = np.ascontiguousarray(zeta, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().orc_vertical_integrate(len(z), _ptr(z), _ptr(f))


def vertical_average(zeta, f):
    z = np.ascontiguousarray(zeta, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().orc_vertical_average(len(z), _ptr(z), _ptr(f))


def integrate_bottom_to_zeta(zeta, f):
    z = np.ascontiguousarray(zeta, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros_like(z)
    lib().orc_vertical_integration_from_bottom_to_zeta(len(z), _ptr(z), _ptr(f), _ptr(out))
    return out


def fraction_above_zero(nw, ne, sw, se):
    return lib().orc_calc_fraction_above_zero(nw, ne, sw, se)


def sliding_beta(cfg, u, v, beta_sq=1e4, alpha_sq=0.5, phi=15.0, tauc=0.0, neff=0.0, x=0.0, y=0.0, a_vav=1e-16):
    return lib().orc_sliding_beta(ctypes.addressof(cfg), u, v, beta_sq, alpha_sq, phi, tauc, neff, x, y, a_vav)


def assemble_csr(grid, cfg, st, u, v):
    m = (grid.nx - 1) * grid.ny + grid.nx * (grid.ny - 1)
    ptr = np.zeros(m + 1, np.int32)
    idx = np.zeros(9 * m, np.int32)
    val = np.zeros(9 * m)
    b = np.zeros(m)
    x = np.zeros(m)
    g = grid.struct()
    io = st.io()
    u = np.asfortranarray(u, dtype=np.float64)
    v = np.asfortranarray(v, dtype=np.float64)
    nnz = lib().orc_assemble_csr(ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(io), _ptr(u), _ptr(v),
                                 _ptr(ptr), _ptr(idx), _ptr(val), _ptr(b), _ptr(x))
    return ptr, idx[:nnz], val[:nnz], b, x


def sor(ptr, idx, val, b, x0, nit, tol, omega):
    x = np.array(x0, dtype=np.float64)
    it = ctypes.c_int()
    rc = lib().orc_sor(len(b), _ptr(ptr), _ptr(idx), _ptr(val), _ptr(b), _ptr(x), nit,
                       ctypes.c_double(tol), ctypes.c_double(omega), ctypes.byref(it))
    return rc, x, it.value


def banded_lu(ptr, idx, val, b):
    x = np.zeros(len(b))
    rc = lib().orc_banded_lu_solve(len(b), _ptr(ptr), _ptr(idx), _ptr(val), _ptr(b), _ptr(x))
    return rc, x


def thickness_update(grid, tcfg, ts, dt):
    """The oracle's calc_dHi_dt on ThicknessState ts in place."""
    L = lib()
    g = grid.struct()
    io = ts.io()
    rc = L.orc_thickness_update(ctypes.byref(g), ctypes.byref(tcfg), ctypes.byref(io), ctypes.c_double(dt))
    if rc != 0:
        raise RuntimeError("oracle thickness update failed (%d): %s" % (rc, L.orc_last_error().decode()))
    return ts


def critical_timestep_adv(grid, u, v, dt_max):
    L = lib()
    L.orc_critical_timestep_adv.restype = ctypes.c_double
    g = grid.struct()
    u = np.asfortranarray(u, dtype=np.float64)
    v = np.asfortranarray(v, dtype=np.float64)
    return L.orc_critical_timestep_adv(ctypes.byref(g), _ptr(u), _ptr(v), ctypes.c_double(dt_max))


def producers_update(grid, pcfg, ps, omit=()):
    """The oracle's update_general_ice_model_data (full masks) on ProducerState ps in place."""
    L = lib()
    g = grid.struct()
    io = ps.io(omit)
    rc = L.orc_producers_update(ctypes.byref(g), ctypes.byref(pcfg), ctypes.byref(io))
    if rc != 0:
        raise RuntimeError("oracle producers failed (%d): %s" % (rc, L.orc_last_error().decode()))
    return ps


def calc_ice_rheology(grid, pcfg, ps, time=0.0, omit=()):
    L = lib()
    g = grid.struct()
    io = ps.io(omit)
    rc = L.orc_calc_ice_rheology(ctypes.byref(g), ctypes.byref(pcfg), ctypes.byref(io), ctypes.c_double(time))
    if rc != 0:
        raise RuntimeError("oracle rheology failed (%d): %s" % (rc, L.orc_last_error().decode()))
    return ps


def run_ice_dynamics_direct(grid, cfg, tcfg, pcfg, scfg, st, SMB, BMB, tstate, t_end, geo=None, ref=None, pc=None):
    """The oracle's 'direct' time loop on IceState st (velocity outputs written
    into st); returns the final thickness.  geo: a ProducerState for the final
    derived fields, or None; ref: a RefGeo, or None; pc: a PcState runs the
    predictor-corrector loop (updated in place)."""
    L = lib()
    g = grid.struct()
    io = st.io()
    SMB = np.asfortranarray(np.broadcast_to(SMB, (grid.ny, grid.nx)), dtype=np.float64)
    BMB = np.asfortranarray(np.broadcast_to(BMB, (grid.ny, grid.nx)), dtype=np.float64)
    Hi = np.zeros((grid.ny, grid.nx), order="F")
    gio = geo.io(omit=("Hi_a", "Hb_a", "SL_a", "dHi_dt_a", "dHb_dt_a", "Ti_a")) if geo is not None else None
    rio = ref.io() if ref is not None else None
    args = (ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(tcfg), ctypes.byref(pcfg), ctypes.byref(scfg),
            ctypes.byref(io), ctypes.byref(rio) if rio is not None else None, _ptr(SMB), _ptr(BMB), _ptr(Hi),
            ctypes.byref(gio) if gio is not None else None, ctypes.byref(tstate))
    if pc is None:
        rc = L.orc_run_ice_dynamics_direct(*args, ctypes.c_double(t_end))
    else:
        rc = L.orc_run_ice_dynamics_pc(*args, ctypes.byref(pc.s), ctypes.c_double(t_end))
    if rc != 0:
        raise RuntimeError("oracle time loop failed (%d): %s" % (rc, L.orc_last_error().decode()))
    st.take_stats(io)
    return Hi


def floating_margin_fraction(grid, Hi, ice, cf, shelf, fmf, heff=None):
    L = lib()
    g = grid.struct()
    i32 = lambda a: np.asfortranarray(a, dtype=np.int32)
    ice, cf, shelf = i32(ice), i32(cf), i32(shelf)
    p32 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    Hi = np.asfortranarray(Hi, dtype=np.float64)
    L.orc_floating_margin_fraction(ctypes.byref(g), _ptr(Hi), p32(ice), p32(cf), p32(shelf), _ptr(fmf),
                                   _ptr(heff) if heff is not None else None)
    return fmf


def fix_geometry(grid, scfg, Hi, sheet, shelf, Ht):
    L = lib()
    g = grid.struct()
    i32 = lambda a: np.asfortranarray(a, dtype=np.int32)
    sheet, shelf = i32(sheet), i32(shelf)
    p32 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    Hi = np.asfortranarray(Hi, dtype=np.float64)
    L.orc_fix_geometry(ctypes.byref(g), ctypes.byref(scfg), _ptr(Hi), p32(sheet), p32(shelf), _ptr(Ht))
    return Ht


def remove_unconnected_shelves(grid, Hi, shelf, sheet):
    L = lib()
    g = grid.struct()
    i32 = lambda a: np.asfortranarray(a, dtype=np.int32)
    shelf, sheet = i32(shelf), i32(sheet)
    p32 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    L.orc_remove_unconnected_shelves(ctypes.byref(g), _ptr(Hi), p32(shelf), p32(sheet))
    return Hi


def critical_timestep_sia(grid, zeta, dt_max, st):
    L = lib()
    L.orc_critical_timestep_sia.restype = ctypes.c_double
    g = grid.struct()
    z = np.ascontiguousarray(zeta, dtype=np.float64)
    p32 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    return L.orc_critical_timestep_sia(ctypes.byref(g), len(z), _ptr(z), ctypes.c_double(dt_max),
                                       _ptr(st["u_3D_SIA_cx"]), _ptr(st["v_3D_SIA_cy"]), _ptr(st["u_3D_cx"]),
                                       _ptr(st["v_3D_cy"]), _ptr(st["Hi_cx"]), _ptr(st["Hi_cy"]),
                                       _ptr(st["dHs_dx_cx"]), _ptr(st["dHs_dy_cy"]), p32(st["mask_sheet_a"]))
