"""Python mirror of ice_velocity_module's call surface over libimauvel.so.

`Context.solve_diva(grid, cfg, state)` is the counterpart of
solve_DIVA(grid, ice) (src/ice_velocity_module.f90:249-374), likewise
solve_ssa / solve_sia.  A non-zero status raises ImauError with the library's
message, as the reference's crash() stops the run
(src/configuration_module.f90:3566-3638).  There is no fallback: without the
HIP library or a visible GPU every call raises.
"""
from __future__ import annotations

import ctypes
import os

from . import abi
from .abi import ImauCfg, ImauGrid, ImauIceIO

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libimauvel.so")

EXPORTED = (
    "imau_vel_init", "imau_vel_solve_sia", "imau_vel_solve_ssa", "imau_vel_solve_diva",
    "imau_vel_nonlinear_residual", "imau_vel_finalize", "imau_vel_last_error", "imau_vel_default_config",
    "imau_vel_problem_create", "imau_vel_problem_solve", "imau_vel_problem_reset_state",
    "imau_vel_problem_download", "imau_vel_problem_destroy", "imau_vel_solve_diva_batch",
    "imau_vel_kernel_timing", "imau_vel_problem_stats", "imau_vel_problem_residual_noalloc",
    "imau_vel_problem_restarts", "imau_vel_problem_thickness", "imau_vel_problem_autotune", "imau_vel_problem_set_tile_rows",
)
# the ice-thickness update that consumes the velocities
THICKNESS_EXPORTED = ("imau_thickness_default_config", "imau_thickness_update", "imau_critical_timestep_adv")
# the input producers the solve reads
PRODUCER_EXPORTED = ("imau_producer_default_config", "imau_update_general_ice_model_data", "imau_calc_ice_rheology")
# the time loop that calls them
TIMELOOP_EXPORTED = ("imau_run_ice_dynamics_direct", "imau_run_ice_dynamics_pc")


class ImauError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("imau_vel error %d: %s" % (code, msg))
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libimauvel.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("libimauvel.so not built (%s); run __graft_entry__.build()" % path)
    L = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    L.imau_vel_init.argtypes = [P(ImauGrid), P(ImauCfg), P(vp)]
    L.imau_vel_finalize.argtypes = [vp]
    L.imau_vel_finalize.restype = None
    L.imau_vel_last_error.argtypes = [vp]
    L.imau_vel_last_error.restype = ctypes.c_char_p
    L.imau_vel_default_config.argtypes = [P(ImauCfg)]
    L.imau_vel_default_config.restype = None
    for name in ("imau_vel_solve_sia", "imau_vel_solve_ssa", "imau_vel_solve_diva"):
        getattr(L, name).argtypes = [vp, P(ImauGrid), P(ImauCfg), P(ImauIceIO)]
    L.imau_vel_nonlinear_residual.argtypes = [vp, P(ImauGrid), P(ImauCfg), P(ImauIceIO), P(ctypes.c_double)]
    L.imau_vel_problem_create.argtypes = [vp, ctypes.c_int32, P(ImauGrid), P(ImauCfg), P(ImauIceIO), P(vp)]
    L.imau_vel_problem_solve.argtypes = [vp, vp, P(ImauCfg)]
    L.imau_vel_problem_reset_state.argtypes = [vp, vp]
    L.imau_vel_problem_download.argtypes = [vp, vp, P(ImauIceIO)]
    L.imau_vel_problem_destroy.argtypes = [vp]
    L.imau_vel_problem_destroy.restype = None
    L.imau_vel_solve_diva_batch.argtypes = [vp, ctypes.c_int32, P(ImauGrid), P(ImauCfg), P(ImauIceIO)]
    L.imau_vel_kernel_timing.argtypes = [vp, ctypes.c_int32, P(ctypes.c_int64), P(ctypes.c_double), P(ctypes.c_int64)]
    L.imau_vel_problem_stats.argtypes = [vp, vp, P(ctypes.c_int32), P(ctypes.c_double)]
    L.imau_vel_problem_restarts.argtypes = [vp, vp, P(ctypes.c_int32)]
    L.imau_vel_problem_autotune.argtypes = [vp, vp, P(ImauCfg), P(ctypes.c_int32)]
    L.imau_vel_problem_set_tile_rows.argtypes = [vp, vp, ctypes.c_int32]
    L.imau_vel_problem_thickness.argtypes = [vp, vp, ctypes.c_int32, P(abi.ImauThicknessCfg), P(abi.ImauThicknessIO),
                                             ctypes.c_double]
    L.imau_thickness_default_config.argtypes = [P(abi.ImauThicknessCfg)]
    L.imau_thickness_default_config.restype = None
    L.imau_thickness_update.argtypes = [vp, P(ImauGrid), P(abi.ImauThicknessCfg), P(abi.ImauThicknessIO), ctypes.c_double]
    L.imau_critical_timestep_adv.argtypes = [vp, P(ImauGrid), P(ctypes.c_double), P(ctypes.c_double), ctypes.c_double,
                                             P(ctypes.c_double)]
    L.imau_vel_problem_residual_noalloc.argtypes = [vp, vp, P(ImauCfg), vp]
    L.imau_producer_default_config.argtypes = [P(abi.ImauProducerCfg)]
    L.imau_producer_default_config.restype = None
    L.imau_update_general_ice_model_data.argtypes = [vp, P(ImauGrid), P(abi.ImauProducerCfg), P(abi.ImauProducerIO)]
    L.imau_calc_ice_rheology.argtypes = [vp, P(ImauGrid), P(abi.ImauProducerCfg), P(abi.ImauProducerIO),
                                         ctypes.c_double]
    L.imau_run_ice_dynamics_direct.argtypes = [vp, P(ImauGrid), P(ImauCfg), P(abi.ImauThicknessCfg),
                                               P(abi.ImauProducerCfg), P(abi.ImauTimestepCfg), P(ImauIceIO), vp, vp,
                                               vp, vp, vp, P(abi.ImauTimestepState), ctypes.c_double]
    L.imau_run_ice_dynamics_pc.argtypes = [vp, P(ImauGrid), P(ImauCfg), P(abi.ImauThicknessCfg),
                                           P(abi.ImauProducerCfg), P(abi.ImauTimestepCfg), P(ImauIceIO), vp, vp,
                                           vp, vp, vp, P(abi.ImauTimestepState), P(abi.ImauPcState), ctypes.c_double]
    _lib = L
    return L


class Context:
    """One velocity-solver context (device stream and workspaces)."""

    def __init__(self):
        self.L = load_library()
        self.h = ctypes.c_void_p()
        rc = self.L.imau_vel_init(None, None, ctypes.byref(self.h))
        if rc != 0:
            raise ImauError(rc, self.L.imau_vel_last_error(None).decode())

    def close(self):
        if self.h:
            self.L.imau_vel_finalize(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise ImauError(rc, self.L.imau_vel_last_error(self.h).decode())

    def _solve(self, fn, grid, cfg, state, outputs):
        g = grid.struct()
        io = state.io(outputs)
        self._check(fn(self.h, ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(io)))
        state.take_stats(io)
        return state

    def solve_diva(self, grid, cfg, state, outputs=None):
        return self._solve(self.L.imau_vel_solve_diva, grid, cfg, state, outputs)

    def solve_ssa(self, grid, cfg, state, outputs=None):
        return self._solve(self.L.imau_vel_solve_ssa, grid, cfg, state, outputs)

    def solve_sia(self, grid, cfg, state, outputs=None):
        return self._solve(self.L.imau_vel_solve_sia, grid, cfg, state, outputs)

    def nonlinear_residual(self, grid, cfg, state):
        g = grid.struct()
        io = state.io(set())
        r = ctypes.c_double()
        self._check(self.L.imau_vel_nonlinear_residual(self.h, ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(io),
                                                       ctypes.byref(r)))
        return r.value

    def thickness_update(self, grid, tcfg, ts, dt):
        """calc_dHi_dt (src/ice_thickness_module.f90:25-58) over one step dt on
        ThicknessState ts: dHi_dt_a, Hi_tplusdt_a (and the explicit fluxes)."""
        g = grid.struct()
        io = ts.io()
        self._check(self.L.imau_thickness_update(self.h, ctypes.byref(g), ctypes.byref(tcfg), ctypes.byref(io),
                                                 ctypes.c_double(dt)))
        return ts

    def update_general_ice_model_data(self, grid, pcfg, ps, omit=()):
        """update_general_ice_model_data (src/general_ice_model_data_module.f90:30-84)
        on ProducerState ps; fields named in `omit` are passed as NULL."""
        g = grid.struct()
        io = ps.io(omit)
        self._check(self.L.imau_update_general_ice_model_data(self.h, ctypes.byref(g), ctypes.byref(pcfg),
                                                              ctypes.byref(io)))
        return ps

    def calc_ice_rheology(self, grid, pcfg, ps, time=0.0, omit=()):
        """calc_ice_rheology (src/thermodynamics_module.f90:641-753) on ProducerState ps."""
        g = grid.struct()
        io = ps.io(omit)
        self._check(self.L.imau_calc_ice_rheology(self.h, ctypes.byref(g), ctypes.byref(pcfg), ctypes.byref(io),
                                                  ctypes.c_double(time)))
        return ps

    def run_ice_dynamics_direct(self, grid, cfg, tcfg, pcfg, scfg, st, SMB, BMB, tstate, t_end, geo=None, ref=None,
                                pc=None):
        """The 'direct' time loop (run_ice_dynamics_direct with
        update_ice_thickness) from tstate.time to t_end on IceState st, whose
        velocity outputs receive the last solve; returns the final thickness.
        geo: a ProducerState for the final derived fields, or None; ref: a
        state.RefGeo with the removal options' inputs, or None.  pc: a
        state.PcState runs the predictor-corrector loop instead
        (run_ice_dynamics_pc) and is updated in place."""
        import numpy as np
        g = grid.struct()
        io = st.io()
        shape = (grid.ny, grid.nx)
        SMB = np.asfortranarray(np.broadcast_to(SMB, shape), dtype=np.float64)
        BMB = np.asfortranarray(np.broadcast_to(BMB, shape), dtype=np.float64)
        Hi = np.zeros(shape, order="F")
        gio = geo.io(omit=("Hi_a", "Hb_a", "SL_a", "dHi_dt_a", "dHb_dt_a", "Ti_a")) if geo is not None else None
        vp = ctypes.c_void_p
        rio = ref.io() if ref is not None else None
        args = (self.h, ctypes.byref(g), ctypes.byref(cfg), ctypes.byref(tcfg), ctypes.byref(pcfg), ctypes.byref(scfg),
                ctypes.byref(io), ctypes.cast(ctypes.pointer(rio), vp) if rio is not None else None,
                vp(SMB.ctypes.data), vp(BMB.ctypes.data), vp(Hi.ctypes.data),
                ctypes.cast(ctypes.pointer(gio), vp) if gio is not None else None, ctypes.byref(tstate))
        if pc is None:
            self._check(self.L.imau_run_ice_dynamics_direct(*args, ctypes.c_double(t_end)))
        else:
            self._check(self.L.imau_run_ice_dynamics_pc(*args, ctypes.byref(pc.s), ctypes.c_double(t_end)))
        st.take_stats(io)
        return Hi

    def critical_timestep_adv(self, grid, u, v, dt_max):
        """calc_critical_timestep_adv (src/ice_dynamics_module.f90:726-762)."""
        import numpy as np
        g = grid.struct()
        u = np.asfortranarray(u, dtype=np.float64)
        v = np.asfortranarray(v, dtype=np.float64)
        out = ctypes.c_double()
        dp = ctypes.POINTER(ctypes.c_double)
        self._check(self.L.imau_critical_timestep_adv(self.h, ctypes.byref(g), u.ctypes.data_as(dp),
                                                      v.ctypes.data_as(dp), ctypes.c_double(dt_max),
                                                      ctypes.byref(out)))
        return out.value

    def kernel_timing(self, what):
        """(launches, summed ms, summed cells of the processed members) timed since
        the last call; then times `what` next: False/"off", True/"smoother" (one
        fine-level smoother sweep) or "visc" (the 3D viscosity kernel)."""
        code = {False: abi.IMAU_TIME_OFF, True: abi.IMAU_TIME_SMOOTHER, "off": abi.IMAU_TIME_OFF,
                "smoother": abi.IMAU_TIME_SMOOTHER, "visc": abi.IMAU_TIME_VISC}[what]
        n, ms, cells = ctypes.c_int64(), ctypes.c_double(), ctypes.c_int64()
        self._check(self.L.imau_vel_kernel_timing(self.h, code, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(cells)))
        return n.value, ms.value, cells.value

    def problem(self, grids, cfg, states):
        return Problem(self, grids, cfg, states)


class Problem:
    """Device-resident grid or ensemble (inputs and state stay in HBM)."""

    def __init__(self, ctx: Context, grids, cfg, states):
        self.ctx, self.cfg = ctx, cfg
        self.grids, self.states = list(grids), list(states)
        n = len(self.grids)
        self._g = (ImauGrid * n)(*[g.struct() for g in self.grids])
        self._io = (ImauIceIO * n)(*[s.io() for s in self.states])
        self.h = ctypes.c_void_p()
        ctx._check(ctx.L.imau_vel_problem_create(ctx.h, n, self._g, ctypes.byref(cfg), self._io, ctypes.byref(self.h)))

    def solve(self, cfg=None):
        self.ctx._check(self.ctx.L.imau_vel_problem_solve(self.ctx.h, self.h, ctypes.byref(cfg or self.cfg)))

    def reset_state(self):
        self.ctx._check(self.ctx.L.imau_vel_problem_reset_state(self.ctx.h, self.h))

    def download(self, outputs=None):
        n = len(self.states)
        io = (ImauIceIO * n)(*[s.io(outputs) for s in self.states])
        self.ctx._check(self.ctx.L.imau_vel_problem_download(self.ctx.h, self.h, io))
        for s, one in zip(self.states, io):
            s.take_stats(one)
        return self.states

    def stats(self):
        """(Picard iterations, last FP64 nonlinear residual) per member of the
        last solve, without downloading any field."""
        n = len(self.grids)
        its = (ctypes.c_int32 * n)()
        res = (ctypes.c_double * n)()
        self.ctx._check(self.ctx.L.imau_vel_problem_stats(self.ctx.h, self.h, its, res))
        return list(its), list(res)

    def thickness(self, g, tcfg, ts, dt):
        """The thickness update of member g with the velocities of the last
        solve, read on the device; ts supplies the other inputs."""
        io = ts.io()
        self.ctx._check(self.ctx.L.imau_vel_problem_thickness(self.ctx.h, self.h, g, ctypes.byref(tcfg),
                                                              ctypes.byref(io), ctypes.c_double(dt)))
        return ts

    def autotune(self, cfg=None):
        """Pick the tile height of the 3D viscosity kernel by timing the
        candidates on this problem; returns the chosen height."""
        rows = ctypes.c_int32()
        self.ctx._check(self.ctx.L.imau_vel_problem_autotune(self.ctx.h, self.h, ctypes.byref(cfg or self.cfg),
                                                             ctypes.byref(rows)))
        return rows.value

    def set_tile_rows(self, rows):
        """Fix the tile height of the 3D viscosity kernel (0: tiled kernel;
        4, 8, 16, 32: row-marching kernel)."""
        self.ctx._check(self.ctx.L.imau_vel_problem_set_tile_rows(self.ctx.h, self.h, rows))

    def restarts(self):
        """Stall restarts per member of the last solve (imau_cfg.stall_window)."""
        n = len(self.grids)
        out = (ctypes.c_int32 * n)()
        self.ctx._check(self.ctx.L.imau_vel_problem_restarts(self.ctx.h, self.h, out))
        return list(out)

    def residual_noalloc(self, out_ptr, cfg=None):
        """FP64 relative nonlinear residual of every member at the current state,
        written asynchronously to the device array at out_ptr (ngrids doubles)
        by the entry point that allocates nothing."""
        self.ctx._check(self.ctx.L.imau_vel_problem_residual_noalloc(self.ctx.h, self.h, ctypes.byref(cfg or self.cfg),
                                                                     ctypes.c_void_p(out_ptr)))

    def close(self):
        if self.h:
            self.ctx.L.imau_vel_problem_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
