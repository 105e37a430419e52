"""The reference's NetCDF files on the host side of the velocity path.

IMAU-ICE writes its restart and help-field files with nf90_create(...,
IOR(nf90_clobber, nf90_share)) (src/netcdf_module.f90:1072, :1217), i.e. in
the NetCDF-3 classic format, which scipy.io.netcdf_file reads and writes.
A field the model holds as d(j,i) (2D) or d(k,j,i) (3D) is written with
Fortran dimensions [x, y, (zeta,) time] (write_data_to_file_dp_2D/_3D,
:6454-6520), so in the file's C order a time frame is a (ny, nx) or
(nz, ny, nx) array with the same [j, i] / [k, j, i] indexing as the model.

read_restart follows inquire_restart_file_geometry / _temperature
(:2596-2702) and read_restart_file_geometry / _temperature (:2807-2954):
every required dimension and variable must be present with the reference's
dimensions, the requested time must lie inside the file's time range, the
closest frame is read (first one on a tie), zeta must match the
configuration's within 1e-4.  The reference then maps the file's grid onto
the model grid (map_square_to_square_cons_2nd_order_2D,
src/reference_fields_module.f90:281); here the file's grid must be the model
grid, and a different grid is an error.

write_restart and write_help_fields produce the reference's layouts
(create_restart_file :1040-1130, create_help_field :1307-1571 for the
velocity fields) so a run on this path leaves files the reference's readers
and plotting scripts take.
"""
from __future__ import annotations

import warnings

import numpy as np
from scipy.io import netcdf_file

from .state import Grid, IceState

GEOMETRY = ("Hi", "Hb", "Hs", "SL", "dHb")      # read_restart_file_geometry
TEMPERATURE = ("Ti",)                           # read_restart_file_temperature
_LONG = {  # create_restart_file, src/netcdf_module.f90:1080-1095
    "Hi": ("Ice thickness", "m"), "Hb": ("Bedrock elevation", "m"), "Hs": ("Surface elevation", "m"),
    "Ti": ("Ice temperature", "K"), "SL": ("Sea surface change", "m"), "dHb": ("Bedrock deformation", "m"),
}
# velocity help fields (create_help_field :1429-1454): file name -> (model field, 3D, long name)
HELP_VELOCITY = {
    "u_3D": ("u_3D_a", True, "3D ice x-velocity"),
    "v_3D": ("v_3D_a", True, "3D ice y-velocity"),
    "w_3D": ("w_3D_a", True, "3D ice z-velocity"),
    "u_vav": ("u_vav_a", False, "Vertically averaged ice x-velocity"),
    "v_vav": ("v_vav_a", False, "Vertically averaged ice x-velocity"),
    "uabs_vav": ("uabs_vav_a", False, "Vertically averaged ice velocity"),
    "u_surf": ("u_surf_a", False, "Surface ice x-velocity"),
    "v_surf": ("v_surf_a", False, "Surface ice y-velocity"),
    "uabs_surf": ("uabs_surf_a", False, "Surface ice velocity"),
    "u_base": ("u_base_a", False, "Basal ice x-velocity"),
    "v_base": ("v_base_a", False, "Basal ice y-velocity"),
    "uabs_base": ("uabs_base_a", False, "Basal ice velocity"),
    "R_shear": ("R_shear", False, "Shearing ratio; 1 = full shearing, 0 = full sliding"),
}


class RestartError(ValueError):
    """A restart file the reference's reader would reject (it crashes)."""


class Restart:
    """One time frame of a restart file: x (nx), y (ny), zeta (nz or None),
    the file's times, the frame read (ti, time) and the fields, each (ny, nx)
    or (nz, ny, nx) in Fortran order like the model's arrays."""

    def __init__(self):
        self.x = self.y = self.zeta = self.times = None
        self.ti, self.time = -1, float("nan")
        self.fields = {}

    def __getitem__(self, k):
        return self.fields[k]

    def grid(self) -> Grid:
        """The model grid of the file (square cells, uniform spacing, as type_grid)."""
        nx, ny = len(self.x), len(self.y)
        dx = float(self.x[1] - self.x[0])
        g = Grid(nx, ny, dx, xmin=float(self.x[0]), ymin=float(self.y[0]))
        if not (np.allclose(g.x, self.x, rtol=0, atol=1e-6 * abs(dx)) and
                np.allclose(g.y, self.y, rtol=0, atol=1e-6 * abs(dx))):
            raise RestartError("restart grid is not uniform with square cells")
        return g


def _inquire(f, dims, var, want):
    """inquire_double_var: `var` exists as a double with the dimensions `want`
    (the reference's Fortran order), i.e. reversed in the file."""
    if var not in f.variables:
        raise RestartError("variable '%s' not found in restart file" % var)
    v = f.variables[var]
    if tuple(v.dimensions) != tuple(reversed(want)):
        raise RestartError("variable '%s' has dimensions %s, expected %s" % (var, tuple(v.dimensions),
                                                                         tuple(reversed(want))))
    if v.typecode() != "d":
        raise RestartError("variable '%s' is not double precision" % var)
    return v


def read_restart(path, time_to_restart_from, zeta=None, start_time_of_run=None, temperature=None) -> Restart:
    """Read the frame of `path` closest to time_to_restart_from.  Geometry
    fields always; Ti when temperature is True, or when it is None and the
    file has it.  zeta (the configuration's) is required to read Ti."""
    r = Restart()
    with netcdf_file(path, "r", mmap=False) as f:
        for d in ("x", "y", "time"):
            if d not in f.dimensions:
                raise RestartError("dimension '%s' not found in restart file" % d)
        _inquire(f, None, "x", ("x",))
        _inquire(f, None, "y", ("y",))
        _inquire(f, None, "time", ("time",))
        r.x = np.array(f.variables["x"][:], dtype=np.float64)
        r.y = np.array(f.variables["y"][:], dtype=np.float64)
        r.times = np.array(f.variables["time"][:], dtype=np.float64)
        t0 = float(time_to_restart_from)
        if r.times.size == 0 or t0 < r.times.min() or t0 > r.times.max():
            raise RestartError("time_to_restart_from = %g outside range of restart file!" % t0)
        # closest frame; the first one on a tie (strict <, :2845-2853)
        ti, dmin = -1, 1e8
        for k, t in enumerate(r.times):
            d = abs(t - t0)
            if d < dmin:
                ti, dmin = k, d
        r.ti, r.time = ti, float(r.times[ti])
        if dmin > 0.0:
            warnings.warn("no exact match for time_to_restart_from = %g yr in restart file! "
                          "Reading closest match = %g yr instead." % (t0, r.time))
        if start_time_of_run is not None and t0 != start_time_of_run:
            warnings.warn("starting run at t = %g yr with restart data at t = %g yr!" % (start_time_of_run, t0))
        for name in GEOMETRY:
            v = _inquire(f, None, name, ("x", "y", "time"))
            r.fields[name] = np.asfortranarray(np.array(v[ti], dtype=np.float64))
        if temperature is None:
            temperature = "Ti" in f.variables
        if temperature:
            if "zeta" not in f.dimensions:
                raise RestartError("dimension 'zeta' not found in restart file")
            r.zeta = np.array(_inquire(f, None, "zeta", ("zeta",))[:], dtype=np.float64)
            if zeta is None:
                raise RestartError("reading Ti needs the configuration's zeta")
            zeta = np.asarray(zeta, dtype=np.float64)
            if len(r.zeta) != len(zeta) or np.any(np.abs(zeta - r.zeta) > 1e-4):
                raise RestartError("vertical coordinate zeta in restart file doesnt match zeta in config!")
            v = _inquire(f, None, "Ti", ("x", "y", "zeta", "time"))
            r.fields["Ti"] = np.asfortranarray(np.array(v[ti], dtype=np.float64))
        for name, a in r.fields.items():
            if not np.all(np.isfinite(a)):
                raise RestartError("NaN or Inf in restart field '%s'" % name)
    return r


def restart_state(r: Restart, nz: int, grid: Grid | None = None) -> tuple[Grid, IceState]:
    """An IceState with the restart frame's geometry (Hi_a, Hb_a, SL_a, and
    Ti_a when read); grid, when given, must be the file's grid."""
    g = r.grid()
    if grid is not None:
        if grid.nx != g.nx or grid.ny != g.ny or not (np.allclose(grid.x, r.x) and np.allclose(grid.y, r.y)):
            raise RestartError("restart grid differs from the model grid (no remapping on this path)")
        g = grid
    st = IceState(g, nz)
    st["Hi_a"] = r["Hi"]
    st["Hb_a"] = r["Hb"]
    st["SL_a"] = r["SL"]
    if "Ti" in r.fields:
        if r["Ti"].shape[0] != nz:
            raise RestartError("restart Ti has %d levels, the model %d" % (r["Ti"].shape[0], nz))
        st["Ti_a"] = r["Ti"]
    return g, st


def _coords(f, x, y, zeta, months=False):
    # the reference defines time last; scipy's writer needs the record
    # dimension first (dimension ids differ, names and shapes do not)
    f.createDimension("time", None)
    f.createDimension("x", len(x))
    f.createDimension("y", len(y))
    if zeta is not None:
        f.createDimension("zeta", len(zeta))
    if months:
        f.createDimension("month", 12)
    for name, vals, ln, un in (("x", x, "X-coordinate", "m"), ("y", y, "Y-coordinate", "m")):
        v = f.createVariable(name, "d", (name,))
        v[:] = vals
        v.long_name, v.units = ln, un
    if zeta is not None:
        v = f.createVariable("zeta", "d", ("zeta",))
        v[:] = zeta
        v.long_name, v.units = "Vertical scaled coordinate", "unitless"
    if months:
        v = f.createVariable("month", "d", ("month",))
        v[:] = np.arange(1, 13, dtype=np.float64)
        v.long_name, v.units = "Month", "1-12"
    t = f.createVariable("time", "d", ("time",))
    t.long_name, t.units = "Time", "years"
    return t


def write_restart(path, grid: Grid, zeta, frames):
    """A restart file in the reference's layout (create_restart_file,
    src/netcdf_module.f90:1040-1130, the geometry and temperature variables):
    frames is a list of (time, {"Hi", "Hb", "Hs", "SL", "dHb": (ny, nx), "Ti": (nz, ny, nx)})."""
    with netcdf_file(path, "w", version=1) as f:
        t = _coords(f, grid.x, grid.y, zeta, months=True)
        vars_ = {}
        for name in GEOMETRY + TEMPERATURE:
            dims = ("time", "zeta", "y", "x") if name == "Ti" else ("time", "y", "x")
            v = f.createVariable(name, "d", dims)
            v.long_name, v.units = _LONG[name]
            vars_[name] = v
        for k, (time, flds) in enumerate(frames):
            t[k] = time
            for name, v in vars_.items():
                v[k] = np.asarray(flds[name], dtype=np.float64)


def write_help_fields(path, grid: Grid, zeta, frames, names=tuple(HELP_VELOCITY)):
    """Velocity help fields (create_help_field / write_help_field,
    src/netcdf_module.f90:339-364, :1429-1454) of IceStates: frames is a list
    of (time, IceState after a solve)."""
    with netcdf_file(path, "w", version=1) as f:
        t = _coords(f, grid.x, grid.y, zeta)
        vars_ = {}
        for name in names:
            field, is3d, ln = HELP_VELOCITY[name]
            v = f.createVariable(name, "d", ("time", "zeta", "y", "x") if is3d else ("time", "y", "x"))
            v.long_name = ln
            if name != "R_shear":
                v.units = "m/yr"
            vars_[name] = (v, field)
        for k, (time, st) in enumerate(frames):
            t[k] = time
            for v, field in vars_.values():
                v[k] = st[field]


def read_help_field(path, name, time):
    """One frame (the closest in time) of a help field written above or by the reference."""
    with netcdf_file(path, "r", mmap=False) as f:
        times = np.array(f.variables["time"][:], dtype=np.float64)
        ti = int(np.argmin(np.abs(times - time)))
        return np.asfortranarray(np.array(f.variables[name][ti], dtype=np.float64))
