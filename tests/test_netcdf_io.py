"""The reference's NetCDF-3 restart and help-field files (src/netcdf_module.f90):
layout, frame choice and rejections of read_restart_file_geometry /
_temperature (:2596-2954), and a solve from a restart frame."""
import warnings

import numpy as np
import pytest
from scipy.io import netcdf_file

import _oracle as O
from imau_vel import config, geometry, netcdf_io
from imau_vel.state import Grid


def frame(g, nz, seed):
    rng = np.random.default_rng(seed)
    f = {k: rng.standard_normal((g.ny, g.nx)) * 100 for k in netcdf_io.GEOMETRY}
    f["Ti"] = 250 + rng.standard_normal((nz, g.ny, g.nx))
    return f


@pytest.fixture
def grid():
    return Grid(9, 6, 4e3, xmin=-16e3, ymin=100e3)     # nx != ny: a transposition cannot pass


def test_round_trip_is_exact(tmp_path, grid):
    zeta = np.array(config.DEFAULT_ZETA)
    frames = [(t, frame(grid, len(zeta), s)) for s, t in enumerate((-100.0, 0.0, 250.0))]
    p = tmp_path / "restart.nc"
    netcdf_io.write_restart(p, grid, zeta, frames)
    for t, f in frames:
        r = netcdf_io.read_restart(p, t, zeta=zeta)
        assert r.time == t and r.grid().nx == grid.nx and r.grid().dx == grid.dx
        for k, a in f.items():
            np.testing.assert_array_equal(r[k], a, err_msg=k)
            assert r[k].flags.f_contiguous


def test_file_layout_is_the_references(tmp_path, grid):
    """write_data_to_file_dp_2D/_3D write d(j,i) / d(k,j,i) with Fortran
    dimensions [x, y, (zeta,) time]: the C-order variable is [time][y][x]
    (or [time][zeta][y][x]); a field that varies differently along x and y
    shows the orientation."""
    zeta = np.array(config.DEFAULT_ZETA)
    f = frame(grid, len(zeta), 0)
    f["Hi"] = grid.x[None, :] + 1000.0 * grid.y[:, None]
    f["Ti"] = np.arange(len(zeta))[:, None, None] * 10.0 + f["Hi"][None]
    p = tmp_path / "restart.nc"
    netcdf_io.write_restart(p, grid, zeta, [(0.0, f)])
    with netcdf_file(p, "r", mmap=False) as nc:
        assert nc.version_byte == 1                       # NetCDF-3 classic, as nf90_create without NF90_NETCDF4
        assert nc.variables["Hi"].dimensions == ("time", "y", "x")
        assert nc.variables["Ti"].dimensions == ("time", "zeta", "y", "x")
        assert nc.dimensions["time"] is None              # unlimited, as nf90_unlimited
        assert nc.dimensions["month"] == 12
        Hi = np.array(nc.variables["Hi"][0])
        assert Hi[2, 5] == grid.x[5] + 1000.0 * grid.y[2]
        assert nc.variables["Hi"].long_name == b"Ice thickness" and nc.variables["Hi"].units == b"m"
    r = netcdf_io.read_restart(p, 0.0, zeta=zeta)
    assert r["Ti"][3, 2, 5] == 30.0 + grid.x[5] + 1000.0 * grid.y[2]


def test_closest_frame_and_tie(tmp_path, grid):
    """The first frame with the smallest |t - t0| (strict <), with the warning
    of an inexact match."""
    zeta = np.array(config.DEFAULT_ZETA)
    frames = [(t, frame(grid, len(zeta), s)) for s, t in enumerate((0.0, 10.0, 20.0))]
    p = tmp_path / "restart.nc"
    netcdf_io.write_restart(p, grid, zeta, frames)
    for t0, want in ((14.0, 10.0), (15.0, 10.0), (16.0, 20.0), (20.0, 20.0)):
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            r = netcdf_io.read_restart(p, t0, temperature=False)
        assert r.time == want
        np.testing.assert_array_equal(r["Hb"], frames[int(want / 10)][1]["Hb"])
        assert any("no exact match" in str(x.message) for x in w) == (t0 != want)
    with pytest.warns(UserWarning, match="starting run at"):
        netcdf_io.read_restart(p, 10.0, temperature=False, start_time_of_run=0.0)


def test_rejections(tmp_path, grid):
    zeta = np.array(config.DEFAULT_ZETA)
    p = tmp_path / "restart.nc"
    netcdf_io.write_restart(p, grid, zeta, [(0.0, frame(grid, len(zeta), 0)), (5.0, frame(grid, len(zeta), 1))])
    with pytest.raises(netcdf_io.RestartError, match="outside range"):
        netcdf_io.read_restart(p, 5.5, zeta=zeta)
    with pytest.raises(netcdf_io.RestartError, match="outside range"):
        netcdf_io.read_restart(p, -0.1, zeta=zeta)
    with pytest.raises(netcdf_io.RestartError, match="zeta"):
        netcdf_io.read_restart(p, 0.0, zeta=zeta[:-1])
    z2 = zeta.copy()
    z2[3] += 2e-4
    with pytest.raises(netcdf_io.RestartError, match="zeta"):
        netcdf_io.read_restart(p, 0.0, zeta=z2)
    z2[3] = zeta[3] + 5e-5                                   # within the reference's 1e-4
    assert netcdf_io.read_restart(p, 0.0, zeta=z2).ti == 0
    # a file without SL, one with Hi on transposed dimensions, one with a NaN
    q = tmp_path / "bad.nc"
    for bad in ("missing", "transposed", "nan"):
        with netcdf_file(q, "w", version=1) as f:
            f.createDimension("time", None)
            f.createDimension("x", grid.nx)
            f.createDimension("y", grid.ny)
            for n, d in (("x", grid.x), ("y", grid.y)):
                f.createVariable(n, "d", (n,))[:] = d
            f.createVariable("time", "d", ("time",))[0] = 0.0
            for n in netcdf_io.GEOMETRY:
                if bad == "missing" and n == "SL":
                    continue
                dims = ("time", "x", "y") if (bad == "transposed" and n == "Hi") else ("time", "y", "x")
                v = f.createVariable(n, "d", dims)
                a = np.zeros((1,) + tuple(f.dimensions[d] for d in dims[1:]))
                if bad == "nan" and n == "Hb":
                    a[0, 1, 1] = np.nan
                v[:] = a
        with pytest.raises(netcdf_io.RestartError, match={"missing": "'SL' not found", "transposed": "'Hi' has",
                                                            "nan": "NaN"}[bad]):
            netcdf_io.read_restart(q, 0.0, temperature=False)


def test_solve_from_a_restart_frame(tmp_path):
    """A dome written to a restart file and read back is the same geometry:
    its solve (the oracle's DIVA) gives the same velocities, and its velocity
    help fields read back as written."""
    g, cfg, st = geometry.dome(nx=17, ny=15, dx=50e3)
    zeta = np.array(cfg.zeta[:cfg.nz])
    st["Ti_a"] = 260.0 + np.arange(cfg.nz)[:, None, None] * 0.5
    f = {"Hi": st["Hi_a"], "Hb": st["Hb_a"], "Hs": st["Hb_a"] + st["Hi_a"], "SL": st["SL_a"],
         "dHb": np.zeros((g.ny, g.nx)), "Ti": st["Ti_a"]}
    p = tmp_path / "restart.nc"
    netcdf_io.write_restart(p, g, zeta, [(0.0, f)])
    g2, st2 = netcdf_io.restart_state(netcdf_io.read_restart(p, 0.0, zeta=zeta), cfg.nz, grid=g)
    full = st.copy()
    for k in ("Hi_a", "Hb_a", "SL_a", "Ti_a"):
        np.testing.assert_array_equal(st2[k], st[k], err_msg=k)
        full[k] = st2[k]
    O.solve("diva", g, cfg, st)
    O.solve("diva", g, cfg, full)
    for k in ("u_vav_cx", "v_vav_cy", "u_surf_a", "w_3D_a"):
        np.testing.assert_array_equal(full[k], st[k], err_msg=k)
    h = tmp_path / "help_fields.nc"
    netcdf_io.write_help_fields(h, g, zeta, [(0.0, full), (10.0, st)], names=("u_surf", "uabs_vav", "w_3D", "R_shear"))
    np.testing.assert_array_equal(netcdf_io.read_help_field(h, "u_surf", 0.0), full["u_surf_a"])
    np.testing.assert_array_equal(netcdf_io.read_help_field(h, "w_3D", 9.0), st["w_3D_a"])
    with netcdf_file(h, "r", mmap=False) as nc:
        assert nc.variables["uabs_vav"].dimensions == ("time", "y", "x")
        assert nc.variables["w_3D"].dimensions == ("time", "zeta", "y", "x")
        assert nc.variables["u_surf"].units == b"m/yr"
