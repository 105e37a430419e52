"""Synthetic ice-sheet geometries and the host-side input producers.

`update_general_ice_model_data` restates the producer the reference runs
before every velocity solve (src/general_ice_model_data_module.f90:30-84,
masks :87-346 without ocean flood-fill; flotation helpers
src/utilities_module.f90:126-158; A vertical average
src/thermodynamics_module.f90:737-742).  The generators build the workloads
the velocity path is measured on:

  dome      29x29 at 50 km, flat bed, H = 3000 (1-(r/750 km)^(4/3))^(3/8)
  channel   401x41 at 2 km marine channel with Gaussian side walls
  random    n_x x n_y random geometries (Matern bed, dome + fractal noise)
  ensemble  uneven-size members drawn from `random`
  ismip_hom_e  the ISMIP-HOM E flowline from a file in the Arolla format
"""
from __future__ import annotations

import numpy as np

from .config import (UNIFORM_COULOMB_PHI_FRIC, UNIFORM_SCHOOF2005_ALPHA_SQ, UNIFORM_SCHOOF2005_BETA_SQ,
                     UNIFORM_TSAI2015_ALPHA_SQ, UNIFORM_TSAI2015_BETA_SQ, UNIFORM_WEERTMAN_BETA_SQ,
                     default_config, zeta_of)
from .state import Grid, IceState

ICE_DENSITY = 910.0        # src/parameters_module.f90:20
SEAWATER_DENSITY = 1028.0  # src/parameters_module.f90:21
TINY = np.finfo(np.float64).tiny


# --------------------------------------------------------------------------
# staggered-grid operators used by the producer (derivatives_and_grids_module)
# --------------------------------------------------------------------------
def _ddx_a_to_a(d, dx):
    o = np.empty_like(d)
    o[:, 1:-1] = (d[:, 2:] - d[:, :-2]) / (2 * dx)
    o[:, 0] = (d[:, 1] - d[:, 0]) / dx
    o[:, -1] = (d[:, -1] - d[:, -2]) / dx
    return o


def _ddy_a_to_a(d, dx):
    o = np.empty_like(d)
    o[1:-1, :] = (d[2:, :] - d[:-2, :]) / (2 * dx)
    o[0, :] = (d[1, :] - d[0, :]) / dx
    o[-1, :] = (d[-1, :] - d[-2, :]) / dx
    return o


def _ddx_a_to_cx(d, dx):
    return (d[:, 1:] - d[:, :-1]) / dx


def _ddy_a_to_cy(d, dx):
    return (d[1:, :] - d[:-1, :]) / dx


def _ddy_a_to_cx(d, dx):
    ny = d.shape[0]
    o = np.empty((ny, d.shape[1] - 1))
    o[1:-1, :] = (d[2:, :-1] + d[2:, 1:] - d[:-2, :-1] - d[:-2, 1:]) / (4.0 * dx)
    o[0, :] = (d[1, :-1] + d[1, 1:] - d[0, :-1] - d[0, 1:]) / (2.0 * dx)
    o[-1, :] = (d[-1, :-1] + d[-1, 1:] - d[-2, :-1] - d[-2, 1:]) / (2.0 * dx)
    return o


def _ddx_a_to_cy(d, dx):
    nx = d.shape[1]
    o = np.empty((d.shape[0] - 1, nx))
    o[:, 1:-1] = (d[:-1, 2:] + d[1:, 2:] - d[:-1, :-2] - d[1:, :-2]) / (4.0 * dx)
    o[:, 0] = (d[:-1, 1] + d[1:, 1] - d[:-1, 0] - d[1:, 0]) / (2.0 * dx)
    o[:, -1] = (d[:-1, -1] + d[1:, -1] - d[:-1, -2] - d[1:, -2]) / (2.0 * dx)
    return o


def _vertical_average(f3, zeta):
    z = np.asarray(zeta)
    s = np.zeros(f3.shape[1:])
    for k in range(len(z) - 1):
        s = s + 0.5 * (f3[k + 1] + f3[k]) * (z[k + 1] - z[k])
    return s


def update_general_ice_model_data(st: IceState, zeta) -> None:
    """Fill Hs, TAF, masks, staggered H and A, and the surface slopes from
    Hi_a, Hb_a, SL_a and A_flow_3D_a (src/general_ice_model_data_module.f90:30-84)."""
    dx = st.grid.dx
    Hi, Hb, SL = st["Hi_a"], st["Hb_a"], st["SL_a"]
    st["Hs_a"] = Hi + np.maximum(SL - ICE_DENSITY / SEAWATER_DENSITY * Hi, Hb)
    st["TAF_a"] = Hi - np.maximum(0.0, (SL - Hb) * (SEAWATER_DENSITY / ICE_DENSITY))
    floating = Hi < (SL - Hb) * SEAWATER_DENSITY / ICE_DENSITY
    st["mask_land_a"] = np.where(floating, 0, 1)
    st["mask_ocean_a"] = np.where(floating, 1, 0)
    ice = Hi > TINY
    st["mask_ice_a"] = ice.astype(np.int32)
    st["mask_sheet_a"] = (ice & ~floating).astype(np.int32)
    st["mask_shelf_a"] = (ice & floating).astype(np.int32)
    st["Hi_cx"] = (Hi[:, :-1] + Hi[:, 1:]) / 2.0
    st["Hi_cy"] = (Hi[:-1, :] + Hi[1:, :]) / 2.0
    st["Hi_b"] = (Hi[:-1, :-1] + Hi[:-1, 1:] + Hi[1:, :-1] + Hi[1:, 1:]) / 4.0
    A3 = st["A_flow_3D_a"]
    st["A_flow_3D_cx"] = (A3[:, :, :-1] + A3[:, :, 1:]) / 2.0
    st["A_flow_3D_cy"] = (A3[:, :-1, :] + A3[:, 1:, :]) / 2.0
    st["A_flow_vav_a"] = _vertical_average(A3, zeta)
    Hs = st["Hs_a"]
    st["dHi_dx_a"] = _ddx_a_to_a(Hi, dx)
    st["dHi_dy_a"] = _ddy_a_to_a(Hi, dx)
    st["dHs_dx_a"] = _ddx_a_to_a(Hs, dx)
    st["dHs_dy_a"] = _ddy_a_to_a(Hs, dx)
    st["dHs_dx_cx"] = _ddx_a_to_cx(Hs, dx)
    st["dHs_dy_cx"] = _ddy_a_to_cx(Hs, dx)
    st["dHs_dx_cy"] = _ddx_a_to_cy(Hs, dx)
    st["dHs_dy_cy"] = _ddy_a_to_cy(Hs, dx)


def fill_uniform_roughness(st: IceState, cfg) -> None:
    """initialise_bed_roughness 'uniform' (src/basal_conditions_and_sliding_module.f90:336-356)."""
    st["beta_sq_a"] = UNIFORM_WEERTMAN_BETA_SQ
    st["phi_fric_a"] = UNIFORM_COULOMB_PHI_FRIC
    if cfg.choice_sliding_law == 5:  # Tsai2015
        st["alpha_sq_a"] = UNIFORM_TSAI2015_ALPHA_SQ
        st["beta_sq_a"] = UNIFORM_TSAI2015_BETA_SQ
    else:
        st["alpha_sq_a"] = UNIFORM_SCHOOF2005_ALPHA_SQ
        if cfg.choice_sliding_law == 6:
            st["beta_sq_a"] = UNIFORM_SCHOOF2005_BETA_SQ


def _finish(st: IceState, cfg, A) -> IceState:
    A = np.asarray(A, dtype=np.float64)
    if A.ndim == 0:
        st["A_flow_3D_a"] = float(A)
    else:
        st["A_flow_3D_a"] = np.broadcast_to(A, st["A_flow_3D_a"].shape[1:])[None, :, :]
    fill_uniform_roughness(st, cfg)
    update_general_ice_model_data(st, zeta_of(cfg))
    return st


def synthetic_config(**overrides):
    """Shared settings of the synthetic workloads: DIVA, uniform rheology,
    Weertman sliding with uniform beta_sq = 1e4, no basal freezing."""
    kw = dict(choice_ice_dynamics="DIVA", choice_sliding_law="Weertman", choice_basal_roughness="uniform",
              choice_basal_hydrology="dry", include_basal_freezing=0)
    kw.update(overrides)
    return default_config(**kw)


def dome(nx: int = 29, ny: int = 29, dx: float = 50e3, A: float = 1e-16, cfg=None):
    """Radially symmetric dome on a flat bed at 0 m (centre cell at the origin)."""
    cfg = cfg or synthetic_config()
    cx, cy = (nx - 1) // 2, (ny - 1) // 2
    g = Grid(nx, ny, dx, xmin=-cx * dx, ymin=-cy * dx)
    st = IceState(g, cfg.nz)
    X, Y = np.meshgrid(g.x, g.y)          # (ny,nx)
    r = np.sqrt(X ** 2 + Y ** 2)
    st["Hi_a"] = 3000.0 * np.maximum(0.0, 1.0 - (r / 750e3) ** (4.0 / 3.0)) ** (3.0 / 8.0)
    st["Hb_a"] = 0.0
    st["SL_a"] = 0.0
    return g, cfg, _finish(st, cfg, A)


def channel(nx: int = 401, ny: int = 41, dx: float = 2e3, seed: int = 0, A: float = 2e-17, cfg=None):
    """Marine channel 800x80 km: bed -150 - 0.5 x[km] plus Gaussian side walls
    500 m high with e-folding width 10 km (exp(-(d/10 km)^2)) centred on the
    walls at y = +-40 km; thickness a seeded smooth profile P(x) from 1500 m
    to 300 m along flow, thinned by the wall height, H = max(0, P(x) - wall(y)),
    so that the grounded surface is level across flow; MISMIP+ style 'zero'
    boundaries for u_west, v_west, v_south, v_north."""
    if cfg is None:
        cfg = synthetic_config(DIVA_boundary_BC_u_west="zero", DIVA_boundary_BC_v_west="zero",
                               DIVA_boundary_BC_v_south="zero", DIVA_boundary_BC_v_north="zero")
    g = Grid(nx, ny, dx, xmin=0.0, ymin=-(ny - 1) // 2 * dx)
    st = IceState(g, cfg.nz)
    X, Y = np.meshgrid(g.x, g.y)
    xkm = X / 1e3
    wall = 500.0 * (np.exp(-((Y - 40e3) / 10e3) ** 2) + np.exp(-((Y + 40e3) / 10e3) ** 2))
    st["Hb_a"] = -150.0 - 0.5 * xkm + wall
    rng = np.random.default_rng(seed)
    xi = (g.x - g.x[0]) / (g.x[-1] - g.x[0])
    amp = rng.uniform(-60.0, 60.0, size=3)
    prof = 1500.0 + (300.0 - 1500.0) * xi
    for k, a in enumerate(amp, start=1):
        prof = prof + a * np.sin(k * np.pi * xi)
    st["Hi_a"] = np.maximum(0.0, prof[None, :] - wall)
    st["SL_a"] = 0.0
    return g, cfg, _finish(st, cfg, A)


class ArollaFileError(ValueError):
    """A file the reference's ISMIP-HOM E readers reject (they crash)."""


def read_ismip_hom_e(path):
    """The ISMIP-HOM E input file (C%ISMIP_HOM_E_Arolla_filename): 51 records
    of `x, Hb, Hs, slides`, read list-directed as in
    initialise_reference_geometry_idealised_ISMIP_HOM_E
    (src/reference_fields_module.f90:751-762) and
    calc_sliding_law_idealised_ISMIP_HOM_E
    (src/basal_conditions_and_sliding_module.f90:1373-1389); records past
    the 51st are not read.  Returns (x, Hb, Hs, slides) arrays of 51."""
    rows = []
    with open(path) as f:
        for line in f:
            tok = line.replace(",", " ").split()
            if not tok:
                continue
            if len(rows) == 51:
                break
            if len(tok) < 4:
                raise ArollaFileError('length of text file "%s" should be 51 lines!' % path)
            rows.append((float(tok[0]), float(tok[1]), float(tok[2]), int(float(tok[3]))))
    if len(rows) != 51:
        raise ArollaFileError('length of text file "%s" should be 51 lines!' % path)
    a = np.array(rows, dtype=np.float64)
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy(), a[:, 3].astype(np.int32)


def ismip_hom_e(source, ny: int = 5, A: float = 1e-16, cfg=None):
    """ISMIP-HOM experiment E: the flowline of `source` (a path to a file in
    the Arolla format, or its (x, Hb, Hs, slides) arrays) on nx = 51 columns
    (the reference forces nsx = 25, src/IMAU_ICE_main_model.f90:704) and ny
    identical rows: Hb and Hi = Hs - Hb per column
    (src/reference_fields_module.f90:753-758), DIVA with the ISMIP_HOM_E
    sliding law, whose slip columns are the file's fourth column."""
    x, Hb, Hs, slides = read_ismip_hom_e(source) if isinstance(source, (str, bytes)) or hasattr(source, "__fspath__") \
        else (np.asarray(a) for a in source)
    if len(x) != 51:
        raise ArollaFileError("an ISMIP-HOM E flowline has 51 columns")
    cfg = cfg or synthetic_config(choice_sliding_law="idealised", choice_idealised_sliding_law="ISMIP_HOM_E")
    dx = float(x[1] - x[0])
    g = Grid(51, ny, dx, xmin=float(x[0]), ymin=-(ny - 1) // 2 * dx)
    st = IceState(g, cfg.nz)
    st["Hb_a"] = np.broadcast_to(Hb[None, :], (ny, 51))
    st["Hi_a"] = np.broadcast_to((Hs - Hb)[None, :], (ny, 51))
    st["SL_a"] = 0.0
    st["ISMIP_HOM_E_slides"] = slides
    return g, cfg, _finish(st, cfg, A)


def _matern_field(ny, nx, dx, sigma, ell, rng, nu=1.5):
    ky = np.fft.fftfreq(ny, d=dx) * 2 * np.pi
    kx = np.fft.fftfreq(nx, d=dx) * 2 * np.pi
    K2 = ky[:, None] ** 2 + kx[None, :] ** 2
    spec = (1.0 + K2 * ell ** 2) ** (-(nu + 1.0) / 2.0)
    w = rng.standard_normal((ny, nx))
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * spec))
    f = f - f.mean()
    s = f.std()
    return sigma * f / s if s > 0 else f


def _fractal_field(ny, nx, rng, slope=2.0):
    ky = np.fft.fftfreq(ny)
    kx = np.fft.fftfreq(nx)
    K = np.sqrt(ky[:, None] ** 2 + kx[None, :] ** 2)
    K[0, 0] = 1.0
    amp = K ** (-slope / 2.0 - 1.0)
    amp[0, 0] = 0.0
    w = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
    f = np.real(np.fft.ifft2(w * amp))
    s = f.std()
    return f / s if s > 0 else f


def random_geometry(nx: int, ny: int, dx: float = 4e3, seed: int = 0, A=None, beta_sq=None,
                    beta_sigma: float = 0.3, beta_scale: float = 1.0, u_slide_max: float = 1500.0,
                    tau0: float = 5e4, cfg=None):
    """Seeded random land-terminating ice cap on a rough bed.

    * bed: Matern field (nu 3/2, sigma 400 m, correlation 200 km) around +100 m;
    * surface: a parabolic cap s = 100 m + H0 (1 - r^2) over an ellipse with
      semi-axes 0.42 x the domain, H0 = sqrt(2 tau0 R / (rho_i g)) (R the shorter
      semi-axis) so the driving stress stays near tau0 whatever the domain
      size, plus seeded fractal noise of 1% of H0; H = max(0, s - bed);
    * no floating ice and no ice-free ocean: cells whose ice would float are
      made ice-free and ice-free cells get a bed of at least 10 m, so every
      margin is a land margin (ice-free ocean cells next to a marine cliff carry
      no friction and run into the 5000 m/yr velocity limiter, which leaves a
      nonlinear residual that cannot vanish);
    * A log-uniform in [1e-17, 1e-16] unless given;
    * beta_sq: beta_scale x 1e4 x exp(beta_sigma x Matern(100 km)), floored at
      tau_d / u_slide_max^(1/3) where tau_d = rho_i g H |grad s| on the a-grid,
      so that Weertman sliding under the local driving stress stays below
      u_slide_max (a calibrated friction field, as a basal inversion would give).
      A scalar beta_sq replaces the field (no floor)."""
    cfg = cfg or synthetic_config()
    rng = np.random.default_rng(seed)
    g = Grid(nx, ny, dx, xmin=-(nx - 1) / 2 * dx, ymin=-(ny - 1) / 2 * dx)
    st = IceState(g, cfg.nz)
    X, Y = np.meshgrid(g.x, g.y)
    bed = _matern_field(ny, nx, dx, 400.0, 200e3, rng) + 100.0
    Rx, Ry = 0.42 * (nx - 1) * dx, 0.42 * (ny - 1) * dx
    r = np.sqrt((X / Rx) ** 2 + (Y / Ry) ** 2)
    H0 = np.sqrt(2.0 * tau0 * min(Rx, Ry) / (910.0 * 9.81))
    prof = H0 * np.maximum(0.0, 1.0 - r * r)
    surf = 100.0 + prof + 0.01 * H0 * _fractal_field(ny, nx, rng) * np.minimum(1.0, prof / 300.0)
    H = np.where(prof > 0.0, np.maximum(0.0, surf - bed), 0.0)
    H = np.where(H * 910.0 < -bed * 1028.0 * 1.05, 0.0, H)
    st["Hb_a"] = np.where(H > 0.0, bed, np.maximum(bed, 10.0))
    st["Hi_a"] = H
    st["SL_a"] = 0.0
    if A is None:
        A = 10.0 ** rng.uniform(-17.0, -16.0)
    _finish(st, cfg, A)
    field = beta_scale * 1.0e4 * np.exp(beta_sigma * _matern_field(ny, nx, dx, 1.0, 100e3, rng))
    if beta_sq is None:
        taud = 910.0 * 9.81 * st["Hi_a"] * np.hypot(st["dHs_dx_a"], st["dHs_dy_a"])
        st["beta_sq_a"] = np.maximum(field, taud / u_slide_max ** (1.0 / 3.0))
    else:
        st["beta_sq_a"] = beta_sq
    st.meta = dict(seed=seed, A=float(np.asarray(A).mean()), beta_scale=beta_scale)
    return g, cfg, st


def ensemble(n_members: int = 64, nmin: int = 96, nmax: int = 512, seed: int = 0, dx: float = 4e3, cfg=None):
    """Uneven-size ensemble: member sizes n ~ U[nmin, nmax] rounded to odd,
    A and beta_sq scale log-uniform over one decade."""
    cfg = cfg or synthetic_config()
    rng = np.random.default_rng(seed)
    members = []
    for m in range(n_members):
        n = int(rng.integers(nmin, nmax + 1)) | 1
        A = 10.0 ** rng.uniform(-17.0, -16.0)
        bscale = 10.0 ** rng.uniform(0.0, 1.0)
        g, c, st = random_geometry(n, n, dx=dx, seed=seed * 1000 + m + 1, A=A, beta_scale=bscale, cfg=cfg)
        members.append((g, st))
    return cfg, members
