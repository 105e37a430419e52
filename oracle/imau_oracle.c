/*
 * imau_oracle.c -- CPU FP64 restatement of IMAU-ICE's ice-velocity path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke test and
 * bench.py's cpu_baseline leg use it; the product never links it.
 *
 * Each routine cites the reference routine it restates (paths relative to the
 * IMAU-ICE source tree).  Arrays use the reference layout (Fortran
 * column-major, see include/imau_vel.h).  Deliberate, documented choices:
 *   - MAX/MIN that may see a NaN use fmax/fmin (NaN dropped), which is how the
 *     reference's cold start (0/0 in du/dz, src/ice_velocity_module.f90:692)
 *     resolves to visc_eff_min at :802.
 *   - the PETSc KSP solve is replaced by an exact banded LU (partial pivoting).
 */
#include "imau_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* src/parameters_module.f90:11-24 */
#define ICE_DENSITY      910.0
#define SEAWATER_DENSITY 1028.0
#define GRAV             9.81
#define PI_D             3.141592653589793

static char g_err[512];
const char *orc_last_error(void) { return g_err; }
static int fail(int code, const char *msg) { snprintf(g_err, sizeof g_err, "%s", msg); return code; }

/* ------------------------------------------------------------------------ */
/* vertical integrals, src/utilities_module.f90:21-123                      */
/* ------------------------------------------------------------------------ */
void orc_vertical_integration_from_bottom_to_zeta(int nz, const double *z, const double *f, double *out)
{
  /* src/utilities_module.f90:21-46 */
  out[nz - 1] = 0.0;
  for (int k = nz - 2; k >= 0; --k)
    out[k] = out[k + 1] - 0.5 * (f[k + 1] + f[k]) * (z[k + 1] - z[k]);
}
double orc_vertical_integrate(int nz, const double *z, const double *f)
{
  /* src/utilities_module.f90:73-96 */
  double s = 0.0;
  for (int k = 1; k < nz; ++k) s += 0.5 * (f[k] + f[k - 1]) * (z[k] - z[k - 1]);
  return s;
}
double orc_vertical_average(int nz, const double *z, const double *f)
{
  /* src/utilities_module.f90:97-123 */
  double s = 0.0;
  for (int k = 0; k < nz - 1; ++k) s += 0.5 * (f[k + 1] + f[k]) * (z[k + 1] - z[k]);
  return s;
}

/* ------------------------------------------------------------------------ */
/* grid helpers                                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
  int nx, ny, nz;
  double dx;
  const double *x, *y;
  const imau_cfg *C;
} G;

#define A2(j, i)   ((size_t)(i) * g->ny + (j))           /* a-grid and cx-grid */
#define CY(j, i)   ((size_t)(i) * (g->ny - 1) + (j))     /* cy-grid and b-grid */
#define A3I(k, j, i) (((size_t)(i) * g->ny + (j)) * g->nz + (k))
#define CY3(k, j, i) (((size_t)(i) * (g->ny - 1) + (j)) * g->nz + (k))

static double *dalloc(size_t n) { double *p = (double *)calloc(n ? n : 1, sizeof(double)); return p; }
static int32_t *ialloc(size_t n) { int32_t *p = (int32_t *)calloc(n ? n : 1, sizeof(int32_t)); return p; }

/* src/derivatives_and_grids_module.f90:24-58 */
static void ddx_a_to_a(const G *g, const double *d, double *o)
{
  for (int i = 1; i < g->nx - 1; ++i)
    for (int j = 0; j < g->ny; ++j) o[A2(j, i)] = (d[A2(j, i + 1)] - d[A2(j, i - 1)]) / (2 * g->dx);
  for (int j = 0; j < g->ny; ++j) {
    o[A2(j, 0)] = (d[A2(j, 1)] - d[A2(j, 0)]) / g->dx;
    o[A2(j, g->nx - 1)] = (d[A2(j, g->nx - 1)] - d[A2(j, g->nx - 2)]) / g->dx;
  }
}
/* src/derivatives_and_grids_module.f90:59-93 */
static void ddy_a_to_a(const G *g, const double *d, double *o)
{
  for (int i = 0; i < g->nx; ++i) {
    for (int j = 1; j < g->ny - 1; ++j) o[A2(j, i)] = (d[A2(j + 1, i)] - d[A2(j - 1, i)]) / (2 * g->dx);
    o[A2(0, i)] = (d[A2(1, i)] - d[A2(0, i)]) / g->dx;
    o[A2(g->ny - 1, i)] = (d[A2(g->ny - 1, i)] - d[A2(g->ny - 2, i)]) / g->dx;
  }
}
/* src/derivatives_and_grids_module.f90:478-506 */
static void ddx_a_to_cx(const G *g, const double *d, double *o)
{
  for (int i = 0; i < g->nx - 1; ++i)
    for (int j = 0; j < g->ny; ++j) o[A2(j, i)] = (d[A2(j, i + 1)] - d[A2(j, i)]) / g->dx;
}
/* src/derivatives_and_grids_module.f90:507-535 */
static void ddy_a_to_cy(const G *g, const double *d, double *o)
{
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny - 1; ++j) o[CY(j, i)] = (d[A2(j + 1, i)] - d[A2(j, i)]) / g->dx;
}
/* src/derivatives_and_grids_module.f90:536-572 */
static void ddx_a_to_cy(const G *g, const double *d, double *o)
{
  for (int i = 1; i < g->nx - 1; ++i)
    for (int j = 0; j < g->ny - 1; ++j)
      o[CY(j, i)] = (d[A2(j, i + 1)] + d[A2(j + 1, i + 1)] - d[A2(j, i - 1)] - d[A2(j + 1, i - 1)]) / (4.0 * g->dx);
  int n = g->nx - 1;
  for (int j = 0; j < g->ny - 1; ++j) {
    o[CY(j, 0)] = (d[A2(j, 1)] + d[A2(j + 1, 1)] - d[A2(j, 0)] - d[A2(j + 1, 0)]) / (2.0 * g->dx);
    o[CY(j, n)] = (d[A2(j, n)] + d[A2(j + 1, n)] - d[A2(j, n - 1)] - d[A2(j + 1, n - 1)]) / (2.0 * g->dx);
  }
}
/* src/derivatives_and_grids_module.f90:573-609 */
static void ddy_a_to_cx(const G *g, const double *d, double *o)
{
  int n = g->ny - 1;
  for (int i = 0; i < g->nx - 1; ++i) {
    for (int j = 1; j < g->ny - 1; ++j)
      o[A2(j, i)] = (d[A2(j + 1, i)] + d[A2(j + 1, i + 1)] - d[A2(j - 1, i)] - d[A2(j - 1, i + 1)]) / (4.0 * g->dx);
    o[A2(0, i)] = (d[A2(1, i)] + d[A2(1, i + 1)] - d[A2(0, i)] - d[A2(0, i + 1)]) / (2.0 * g->dx);
    o[A2(n, i)] = (d[A2(n, i)] + d[A2(n, i + 1)] - d[A2(n - 1, i)] - d[A2(n - 1, i + 1)]) / (2.0 * g->dx);
  }
}
/* src/derivatives_and_grids_module.f90:755-787 */
static void ddx_cx_to_a(const G *g, const double *d, double *o)
{
  for (int i = 1; i < g->nx - 1; ++i)
    for (int j = 0; j < g->ny; ++j) o[A2(j, i)] = (d[A2(j, i)] - d[A2(j, i - 1)]) / g->dx;
  for (int j = 0; j < g->ny; ++j) {
    o[A2(j, 0)] = o[A2(j, 1)];
    o[A2(j, g->nx - 1)] = o[A2(j, g->nx - 2)];
  }
}
/* src/derivatives_and_grids_module.f90:788-820 */
static void ddy_cy_to_a(const G *g, const double *d, double *o)
{
  for (int i = 0; i < g->nx; ++i) {
    for (int j = 1; j < g->ny - 1; ++j) o[A2(j, i)] = (d[CY(j, i)] - d[CY(j - 1, i)]) / g->dx;
    o[A2(0, i)] = o[A2(1, i)];
    o[A2(g->ny - 1, i)] = o[A2(g->ny - 2, i)];
  }
}
/* src/derivatives_and_grids_module.f90:821-879 */
static void ddy_cx_to_a(const G *g, const double *d, double *o)
{
  const int nx = g->nx, ny = g->ny;
  const double dx = g->dx;
  for (int i = 1; i < nx - 1; ++i)
    for (int j = 1; j < ny - 1; ++j)
      o[A2(j, i)] = (d[A2(j + 1, i - 1)] + d[A2(j + 1, i)] - d[A2(j - 1, i - 1)] - d[A2(j - 1, i)]) / (4.0 * dx);
  for (int i = 1; i < nx - 1; ++i) {
    int j = 0;
    o[A2(j, i)] = (d[A2(j + 1, i - 1)] + d[A2(j + 1, i)] - d[A2(j, i - 1)] - d[A2(j, i)]) / (4.0 * dx);
    j = ny - 1;
    o[A2(j, i)] = (d[A2(j, i - 1)] + d[A2(j, i)] - d[A2(j - 1, i - 1)] - d[A2(j - 1, i)]) / (4.0 * dx);
  }
  for (int j = 1; j < ny - 1; ++j) {
    o[A2(j, 0)] = (d[A2(j + 1, 0)] - d[A2(j - 1, 0)]) / (2.0 * dx);
    o[A2(j, nx - 1)] = (d[A2(j + 1, nx - 2)] - d[A2(j - 1, nx - 2)]) / (2.0 * dx);
  }
  o[A2(0, 0)] = (d[A2(1, 0)] - d[A2(0, 0)]) / dx;
  o[A2(0, nx - 1)] = (d[A2(1, nx - 2)] - d[A2(0, nx - 2)]) / dx;
  o[A2(ny - 1, 0)] = (d[A2(ny - 1, 0)] - d[A2(ny - 2, 0)]) / dx;
  o[A2(ny - 1, nx - 1)] = (d[A2(ny - 1, nx - 2)] - d[A2(ny - 2, nx - 2)]) / dx;
}
/* src/derivatives_and_grids_module.f90:880-938 */
static void ddx_cy_to_a(const G *g, const double *d, double *o)
{
  const int nx = g->nx, ny = g->ny;
  const double dx = g->dx;
  for (int i = 1; i < nx - 1; ++i)
    for (int j = 1; j < ny - 1; ++j)
      o[A2(j, i)] = (d[CY(j - 1, i + 1)] + d[CY(j, i + 1)] - d[CY(j - 1, i - 1)] - d[CY(j, i - 1)]) / (4.0 * dx);
  for (int j = 1; j < ny - 1; ++j) {
    int i = 0;
    o[A2(j, i)] = (d[CY(j - 1, i + 1)] + d[CY(j, i + 1)] - d[CY(j - 1, i)] - d[CY(j, i)]) / (4.0 * dx);
    i = nx - 1;
    o[A2(j, i)] = (d[CY(j - 1, i)] + d[CY(j, i)] - d[CY(j - 1, i - 1)] - d[CY(j, i - 1)]) / (4.0 * dx);
  }
  for (int i = 1; i < nx - 1; ++i) {
    o[A2(0, i)] = (d[CY(0, i + 1)] - d[CY(0, i - 1)]) / (2.0 * dx);
    o[A2(ny - 1, i)] = (d[CY(ny - 2, i + 1)] - d[CY(ny - 2, i - 1)]) / (2.0 * dx);
  }
  o[A2(0, 0)] = (d[CY(0, 1)] - d[CY(0, 0)]) / dx;
  o[A2(0, nx - 1)] = (d[CY(0, nx - 1)] - d[CY(0, nx - 2)]) / dx;
  o[A2(ny - 1, 0)] = (d[CY(ny - 2, 1)] - d[CY(ny - 2, 0)]) / dx;
  o[A2(ny - 1, nx - 1)] = (d[CY(ny - 2, nx - 1)] - d[CY(ny - 2, nx - 2)]) / dx;
}
/* src/derivatives_and_grids_module.f90:1454-1482 */
static void map_a_to_cx(const G *g, const double *d, double *o)
{
  for (int i = 0; i < g->nx - 1; ++i)
    for (int j = 0; j < g->ny; ++j) o[A2(j, i)] = (d[A2(j, i)] + d[A2(j, i + 1)]) / 2.0;
}
/* src/derivatives_and_grids_module.f90:1483-1511 */
static void map_a_to_cy(const G *g, const double *d, double *o)
{
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny - 1; ++j) o[CY(j, i)] = (d[A2(j, i)] + d[A2(j + 1, i)]) / 2.0;
}
/* src/derivatives_and_grids_module.f90:1798-1830 */
static void map_a_to_b(const G *g, const double *d, double *o)
{
  for (int i = 0; i < g->nx - 1; ++i)
    for (int j = 0; j < g->ny - 1; ++j)
      o[CY(j, i)] = (d[A2(j, i)] + d[A2(j, i + 1)] + d[A2(j + 1, i)] + d[A2(j + 1, i + 1)]) / 4.0;
}
/* src/derivatives_and_grids_module.f90:1579-1611 */
static void map_cx_to_a(const G *g, const double *d, double *o)
{
  for (int i = 1; i < g->nx - 1; ++i)
    for (int j = 0; j < g->ny; ++j) o[A2(j, i)] = (d[A2(j, i - 1)] + d[A2(j, i)]) / 2.0;
  for (int j = 0; j < g->ny; ++j) {
    o[A2(j, 0)] = d[A2(j, 0)];
    o[A2(j, g->nx - 1)] = d[A2(j, g->nx - 2)];
  }
}
/* src/derivatives_and_grids_module.f90:1612-1644 */
static void map_cy_to_a(const G *g, const double *d, double *o)
{
  for (int i = 0; i < g->nx; ++i) {
    for (int j = 1; j < g->ny - 1; ++j) o[A2(j, i)] = (d[CY(j - 1, i)] + d[CY(j, i)]) / 2.0;
    o[A2(0, i)] = d[CY(0, i)];
    o[A2(g->ny - 1, i)] = d[CY(g->ny - 2, i)];
  }
}
/* src/derivatives_and_grids_module.f90:1718-1756 */
static void map_cx_to_cy(const G *g, const double *d, double *o)
{
  for (int i = 1; i < g->nx - 1; ++i)
    for (int j = 0; j < g->ny - 1; ++j)
      o[CY(j, i)] = (d[A2(j, i - 1)] + d[A2(j, i)] + d[A2(j + 1, i - 1)] + d[A2(j + 1, i)]) / 4.0;
  for (int j = 0; j < g->ny - 1; ++j) {
    o[CY(j, 0)] = (d[A2(j, 0)] + d[A2(j + 1, 0)]) / 2.0;
    o[CY(j, g->nx - 1)] = (d[A2(j, g->nx - 2)] + d[A2(j + 1, g->nx - 2)]) / 2.0;
  }
}
/* src/derivatives_and_grids_module.f90:1757-1797 */
static void map_cy_to_cx(const G *g, const double *d, double *o)
{
  for (int i = 0; i < g->nx - 1; ++i) {
    for (int j = 1; j < g->ny - 1; ++j)
      o[A2(j, i)] = (d[CY(j - 1, i)] + d[CY(j - 1, i + 1)] + d[CY(j, i)] + d[CY(j, i + 1)]) / 4.0;
    o[A2(0, i)] = (d[CY(0, i)] + d[CY(0, i + 1)]) / 2.0;
    o[A2(g->ny - 1, i)] = (d[CY(g->ny - 2, i)] + d[CY(g->ny - 2, i + 1)]) / 2.0;
  }
}
/* 3D maps, src/derivatives_and_grids_module.f90:1513-1578 (a->cx, a->cy) */
static void map_a_to_cx_3d(const G *g, const double *d, double *o)
{
  for (int i = 0; i < g->nx - 1; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) o[A3I(k, j, i)] = (d[A3I(k, j, i)] + d[A3I(k, j, i + 1)]) / 2.0;
}
static void map_a_to_cy_3d(const G *g, const double *d, double *o)
{
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny - 1; ++j)
      for (int k = 0; k < g->nz; ++k) o[CY3(k, j, i)] = (d[A3I(k, j, i)] + d[A3I(k, j + 1, i)]) / 2.0;
}

/* ------------------------------------------------------------------------ */
/* producers, src/general_ice_model_data_module.f90:30-346 and utilities    */
/* ------------------------------------------------------------------------ */
static int is_floating(double Hi, double Hb, double SL)
{ /* src/utilities_module.f90:126-136 */
  return Hi < (SL - Hb) * SEAWATER_DENSITY / ICE_DENSITY;
}
static double surface_elevation(double Hi, double Hb, double SL)
{ /* src/utilities_module.f90:138-147 */
  return Hi + fmax(SL - ICE_DENSITY / SEAWATER_DENSITY * Hi, Hb);
}
static double thickness_above_floatation(double Hi, double Hb, double SL)
{ /* src/utilities_module.f90:149-158 */
  return Hi - fmax(0.0, (SL - Hb) * (SEAWATER_DENSITY / ICE_DENSITY));
}

void orc_update_general_ice_model_data(const imau_grid *gr, const imau_cfg *c,
                                       const double *Hi, const double *Hb, const double *SL,
                                       const double *A3,
                                       double *Hs, double *TAF, int32_t *mask_land, int32_t *mask_ocean,
                                       int32_t *mask_ice, int32_t *mask_sheet, int32_t *mask_shelf,
                                       double *Hi_cx, double *Hi_cy, double *Hi_b,
                                       double *A3_cx, double *A3_cy, double *A_vav,
                                       double *dHi_dx_a, double *dHi_dy_a, double *dHs_dx_a, double *dHs_dy_a,
                                       double *dHs_dx_cx, double *dHs_dy_cx, double *dHs_dx_cy, double *dHs_dy_cy)
{
  /* src/general_ice_model_data_module.f90:30-84 (no ocean flood-fill: do_ocean_floodfill=.FALSE.) */
  G gg = {gr->nx, gr->ny, c->nz, gr->dx, gr->x, gr->y, c};
  const G *g = &gg;
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny; ++j) {
      size_t p = A2(j, i);
      Hs[p] = surface_elevation(Hi[p], Hb[p], SL[p]);
      TAF[p] = thickness_above_floatation(Hi[p], Hb[p], SL[p]);
      /* src/general_ice_model_data_module.f90:109-145 */
      int flt = is_floating(Hi[p], Hb[p], SL[p]);
      mask_land[p] = flt ? 0 : 1;
      mask_ocean[p] = flt ? 1 : 0;
      /* src/general_ice_model_data_module.f90:187 : Hi > TINY(Hi) */
      mask_ice[p] = (Hi[p] > 2.2250738585072014e-308) ? 1 : 0;
      mask_sheet[p] = (mask_ice[p] == 1 && mask_land[p] == 1) ? 1 : 0;
      mask_shelf[p] = (mask_ice[p] == 1 && mask_ocean[p] == 1) ? 1 : 0;
    }
  map_a_to_cx(g, Hi, Hi_cx);
  map_a_to_cy(g, Hi, Hi_cy);
  map_a_to_b(g, Hi, Hi_b);
  if (A3 && A3_cx) map_a_to_cx_3d(g, A3, A3_cx);
  if (A3 && A3_cy) map_a_to_cy_3d(g, A3, A3_cy);
  if (A3 && A_vav) /* calc_ice_rheology vertical average, src/thermodynamics_module.f90:737-742 */
    for (size_t p = 0; p < (size_t)g->nx * g->ny; ++p) A_vav[p] = orc_vertical_average(g->nz, c->zeta, A3 + p * g->nz);
  ddx_a_to_a(g, Hi, dHi_dx_a);
  ddy_a_to_a(g, Hi, dHi_dy_a);
  ddx_a_to_a(g, Hs, dHs_dx_a);
  ddy_a_to_a(g, Hs, dHs_dy_a);
  ddx_a_to_cx(g, Hs, dHs_dx_cx);
  ddy_a_to_cx(g, Hs, dHs_dy_cx);
  ddx_a_to_cy(g, Hs, dHs_dx_cy);
  ddy_a_to_cy(g, Hs, dHs_dy_cy);
}

/* ------------------------------------------------------------------------ */
/* input producers with the full masks: update_general_ice_model_data        */
/* (src/general_ice_model_data_module.f90:30-84), determine_masks (:87-346), */
/* ocean_floodfill (:868-986); calc_ice_rheology                             */
/* (src/thermodynamics_module.f90:641-753)                                   */
/* ------------------------------------------------------------------------ */
#define SETI(ptr, p, v) do { if (ptr) (ptr)[p] = (v); } while (0)

/* ocean_floodfill, :868-986: a depth-first fill from every edge cell through
 * the 8 neighbours of floating cells, with the reference's stack order */
static void ocean_floodfill(const G *g, const double *Hi, const double *Hb, const double *SL, int32_t *land,
                            int32_t *ocean)
{
  const int nx = g->nx, ny = g->ny;
  int32_t *map = ialloc((size_t)nx * ny);
  int32_t *stack = ialloc(2 * (size_t)nx * ny + 4 * (size_t)(nx + ny));
  long n = 0;
  for (size_t p = 0; p < (size_t)nx * ny; ++p) { land[p] = 1; ocean[p] = 0; }
  for (int i = 0; i < nx; ++i) {        /* :905-914 */
    int j = 0;
    stack[2 * n] = i; stack[2 * n + 1] = j; ++n; map[A2(j, i)] = 1;
    j = ny - 1;
    stack[2 * n] = i; stack[2 * n + 1] = j; ++n; map[A2(j, i)] = 1;
  }
  for (int j = 0; j < ny; ++j) {        /* :915-924 */
    int i = 0;
    stack[2 * n] = i; stack[2 * n + 1] = j; ++n; map[A2(j, i)] = 1;
    i = nx - 1;
    stack[2 * n] = i; stack[2 * n + 1] = j; ++n; map[A2(j, i)] = 1;
  }
  while (n > 0) {                       /* :927-968 */
    --n;
    const int i = stack[2 * n], j = stack[2 * n + 1];
    if (!is_floating(Hi[A2(j, i)], Hb[A2(j, i)], SL[A2(j, i)])) continue;
    map[A2(j, i)] = 2;
    ocean[A2(j, i)] = 1;
    land[A2(j, i)] = 0;
    for (int ii = i - 1; ii <= i + 1; ++ii)
      for (int jj = j - 1; jj <= j + 1; ++jj) {
        if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || (ii == i && jj == j)) continue;
        if (map[A2(jj, ii)] == 0) {
          map[A2(jj, ii)] = 1;
          stack[2 * n] = ii; stack[2 * n + 1] = jj; ++n;
        }
      }
  }
  free(map);
  free(stack);
}

int orc_producers_update(const imau_grid *gr, const imau_producer_cfg *c, imau_producer_io *io)
{
  G gg = {gr->nx, gr->ny, c->nz, gr->dx, gr->x, gr->y, NULL};
  const G *g = &gg;
  const int nx = g->nx, ny = g->ny;
  const size_t na = (size_t)nx * ny;
  const double *Hi = io->Hi_a, *Hb = io->Hb_a, *SL = io->SL_a;
  double *Hs = io->Hs_a ? io->Hs_a : dalloc(na);
  int32_t *land = ialloc(na), *ocean = ialloc(na), *ice = ialloc(na), *sheet = ialloc(na), *shelf = ialloc(na);
  /* :45-56 */
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const size_t p = A2(j, i);
      Hs[p] = surface_elevation(Hi[p], Hb[p], SL[p]);
      if (io->TAF_a) io->TAF_a[p] = thickness_above_floatation(Hi[p], Hb[p], SL[p]);
    }
  /* :47-52, before the masks are updated: the land mask as it is on entry
   * (the previous call's) */
  if (io->dHi_dt_a && io->dHb_dt_a && io->dHs_dt_a && io->mask_land_a)
    for (size_t p = 0; p < na; ++p)
      io->dHs_dt_a[p] = io->mask_land_a[p] == 1 ? io->dHb_dt_a[p] + io->dHi_dt_a[p]
                                                : io->dHi_dt_a[p] * (1.0 - 910.0 / 1028.0);
  /* determine_masks_landocean, :109-151 */
  if (c->do_ocean_floodfill) ocean_floodfill(g, Hi, Hb, SL, land, ocean);
  else
    for (size_t p = 0; p < na; ++p) {
      const int flt = is_floating(Hi[p], Hb[p], SL[p]);
      land[p] = flt ? 0 : 1;
      ocean[p] = flt ? 1 : 0;
    }
  /* determine_masks_ice, :152-201 */
  for (size_t p = 0; p < na; ++p) {
    ice[p] = Hi[p] > 2.2250738585072014e-308 ? 1 : 0;
    sheet[p] = (ice[p] == 1 && land[p] == 1) ? 1 : 0;
    shelf[p] = (ice[p] == 1 && ocean[p] == 1) ? 1 : 0;
  }
  /* determine_masks_transitions, :202-281 (interior cells only) */
  for (size_t p = 0; p < na; ++p) {
    SETI(io->mask_coast_a, p, 0); SETI(io->mask_margin_a, p, 0);
    SETI(io->mask_gl_a, p, 0); SETI(io->mask_cf_a, p, 0);
  }
  int32_t *coast = ialloc(na), *margin = ialloc(na), *gl = ialloc(na), *cf = ialloc(na);
  for (int i = 1; i < nx - 1; ++i)
    for (int j = 1; j < ny - 1; ++j) {
      const size_t p = A2(j, i), s = A2(j - 1, i), n = A2(j + 1, i), w = A2(j, i - 1), e = A2(j, i + 1);
      if (land[p] == 1 && ice[p] == 0 && (ocean[s] == 1 || ocean[n] == 1 || ocean[w] == 1 || ocean[e] == 1))
        coast[p] = 1;
      if (ice[p] == 1 && (ice[s] == 0 || ice[n] == 0 || ice[w] == 0 || ice[e] == 0)) margin[p] = 1;
      if (sheet[p] == 1 && (shelf[s] == 1 || shelf[n] == 1 || shelf[w] == 1 || shelf[e] == 1)) gl[p] = 1;
      if (ice[p] == 1 && ((ocean[s] == 1 && ice[s] == 0) || (ocean[n] == 1 && ice[n] == 0) ||
                          (ocean[w] == 1 && ice[w] == 0) || (ocean[e] == 1 && ice[e] == 0)))
        cf[p] = 1;
    }
  /* determine_masks_total, :282-346 */
  for (size_t p = 0; p < na; ++p) {
    int32_t t = land[p] == 1 ? IMAU_TYPE_LAND : IMAU_TYPE_OCEAN;
    if (coast[p] == 1) t = IMAU_TYPE_COAST;
    if (sheet[p] == 1) t = IMAU_TYPE_SHEET;
    else if (shelf[p] == 1) t = IMAU_TYPE_SHELF;
    if (margin[p] == 1) t = IMAU_TYPE_MARGIN;
    if (gl[p] == 1) t = IMAU_TYPE_GROUNDINGLINE;
    if (cf[p] == 1) t = IMAU_TYPE_CALVINGFRONT;
    SETI(io->mask_a, p, t);
    SETI(io->mask_land_a, p, land[p]); SETI(io->mask_ocean_a, p, ocean[p]);
    SETI(io->mask_ice_a, p, ice[p]); SETI(io->mask_sheet_a, p, sheet[p]); SETI(io->mask_shelf_a, p, shelf[p]);
    SETI(io->mask_coast_a, p, coast[p]); SETI(io->mask_margin_a, p, margin[p]);
    SETI(io->mask_gl_a, p, gl[p]); SETI(io->mask_cf_a, p, cf[p]);
  }
  /* :62-80 */
  if (io->Hi_cx) map_a_to_cx(g, Hi, io->Hi_cx);
  if (io->Hi_cy) map_a_to_cy(g, Hi, io->Hi_cy);
  if (io->Hi_b) map_a_to_b(g, Hi, io->Hi_b);
  if (io->A_flow_3D_a && io->A_flow_3D_cx) map_a_to_cx_3d(g, io->A_flow_3D_a, io->A_flow_3D_cx);
  if (io->A_flow_3D_a && io->A_flow_3D_cy) map_a_to_cy_3d(g, io->A_flow_3D_a, io->A_flow_3D_cy);
  if (io->A_flow_vav_a && io->A_flow_vav_cx) map_a_to_cx(g, io->A_flow_vav_a, io->A_flow_vav_cx);
  if (io->A_flow_vav_a && io->A_flow_vav_cy) map_a_to_cy(g, io->A_flow_vav_a, io->A_flow_vav_cy);
  if (io->dHi_dx_a) ddx_a_to_a(g, Hi, io->dHi_dx_a);
  if (io->dHi_dy_a) ddy_a_to_a(g, Hi, io->dHi_dy_a);
  if (io->dHs_dx_a) ddx_a_to_a(g, Hs, io->dHs_dx_a);
  if (io->dHs_dy_a) ddy_a_to_a(g, Hs, io->dHs_dy_a);
  if (io->dHs_dx_cx) ddx_a_to_cx(g, Hs, io->dHs_dx_cx);
  if (io->dHs_dy_cx) ddy_a_to_cx(g, Hs, io->dHs_dy_cx);
  if (io->dHs_dx_cy) ddx_a_to_cy(g, Hs, io->dHs_dx_cy);
  if (io->dHs_dy_cy) ddy_a_to_cy(g, Hs, io->dHs_dy_cy);
  if (!io->Hs_a) free(Hs);
  free(land); free(ocean); free(ice); free(sheet); free(shelf);
  free(coast); free(margin); free(gl); free(cf);
  return 0;
}

int orc_calc_ice_rheology(const imau_grid *gr, const imau_producer_cfg *c, imau_producer_io *io, double time)
{
  G gg = {gr->nx, gr->ny, c->nz, gr->dx, gr->x, gr->y, NULL};
  const G *g = &gg;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const size_t na = (size_t)nx * ny;
  /* :654-658 */
  const double A_low = 1.14e-05, A_high = 5.47e+10, Q_low = 6.0e+04, Q_high = 13.9e+04, R_gas = 8.314;
  double *A3 = io->A_flow_3D_a, *Av = io->A_flow_vav_a;
  if (!A3 || !Av) { snprintf(g_err, sizeof g_err, "calc_ice_rheology needs A_flow_3D_a and A_flow_vav_a"); return -1; }
  if (c->choice_ice_rheology == IMAU_RHEO_UNIFORM) {          /* :664-669 */
    for (size_t q = 0; q < na * nz; ++q) A3[q] = c->uniform_flow_factor;
  } else if (c->choice_ice_rheology == IMAU_RHEO_HUYBRECHTS1992) {  /* :670-699 */
    if (!io->Ti_a || !io->mask_ice_a) { snprintf(g_err, sizeof g_err, "Huybrechts1992 needs Ti_a and mask_ice_a"); return -1; }
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const size_t q = A3I(k, j, i);
          if (io->mask_ice_a[A2(j, i)] == 1) {
            const double T = io->Ti_a[q];
            A3[q] = T < 263.15 ? A_low * exp(-Q_low / (R_gas * T)) : A_high * exp(-Q_high / (R_gas * T));
          } else {
            A3[q] = c->choice_ice_margin == IMAU_MARGIN_BC ? 0.0 : A_low * exp(-Q_low / (R_gas * 263.15));
          }
        }
  } else if (c->choice_ice_rheology == IMAU_RHEO_MISMIP_MOD) {     /* :701-713 */
    double A = 1.0e-16;
    if (time < 15000.0) A = 1.0e-16;
    else if (time < 30000.0) A = 1.0e-17;
    else if (time < 45000.0) A = 1.0e-16;
    for (size_t q = 0; q < na * nz; ++q) A3[q] = A;
  } else {
    snprintf(g_err, sizeof g_err, "unknown choice_ice_rheology %d", c->choice_ice_rheology);
    return -1;
  }
  /* :719-729 enhancement factors */
  if (io->mask_sheet_a && io->mask_shelf_a)
    for (size_t p = 0; p < na; ++p) {
      if (io->mask_sheet_a[p] != 1 && io->mask_shelf_a[p] != 1) continue;
      const double f = io->mask_sheet_a[p] == 1 ? c->m_enh_sheet : c->m_enh_shelf;
      for (int k = 0; k < nz; ++k) A3[p * nz + k] *= f;
    }
  /* :731-737 */
  for (size_t p = 0; p < na; ++p) Av[p] = orc_vertical_average(nz, c->zeta, A3 + p * nz);
  /* :739-743 */
  if (io->A_flow_3D_cx) map_a_to_cx_3d(g, A3, io->A_flow_3D_cx);
  if (io->A_flow_3D_cy) map_a_to_cy_3d(g, A3, io->A_flow_3D_cy);
  if (io->A_flow_vav_cx) map_a_to_cx(g, Av, io->A_flow_vav_cx);
  if (io->A_flow_vav_cy) map_a_to_cy(g, Av, io->A_flow_vav_cy);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* grounded fractions, src/general_ice_model_data_module.f90:349-781         */
/* ------------------------------------------------------------------------ */
static void rotate_quad(double *NW, double *NE, double *SW, double *SE)
{ /* src/general_ice_model_data_module.f90:764-781 (anticlockwise by 90 deg) */
  double fNW = *NW, fNE = *NE, fSE = *SE, fSW = *SW;
  *NW = fNE; *NE = fSE; *SE = fSW; *SW = fNW;
}
static int rotate_quad_until_match(double *NW, double *NE, double *SW, double *SE)
{ /* src/general_ice_model_data_module.f90:712-763 */
  for (int nit = 1; nit <= 5; ++nit) {
    rotate_quad(NW, NE, SW, SE);
    if (*SW > 0 && *SE < 0 && *NE < 0 && *NW < 0) return 1;
    if (*SW < 0 && *SE > 0 && *NE > 0 && *NW > 0) return 2;
    if (*SW > 0 && *SE > 0 && *NE < 0 && *NW < 0) return 3;
    if (*SW > 0 && *SE < 0 && *NE > 0 && *NW < 0) return 4;
  }
  return 0;
}
static double perturb(double f, double s)
{ /* src/general_ice_model_data_module.f90:539-574 */
  const double ftol = 1e-4;
  if (f == 0.0) return ftol * s;
  if (f > 0.0) return fmax(ftol * s, f);
  if (f < 0.0) return fmin(-ftol * s, f);
  return f;
}
static double corner_area(double aa, double bb, double cc, double dd)
{
  return ((bb * cc - aa * dd) * log(fabs(1.0 - (aa * dd) / (bb * cc))) + aa * dd) / (dd * dd);
}
static double strip_prim(double aa, double bb, double cc, double dd, double x)
{
  return ((bb * cc - aa * dd) * log(fabs(cc + dd * x)) - bb * dd * x) / (dd * dd);
}
double orc_calc_fraction_above_zero(double f_NW, double f_NE, double f_SW, double f_SE)
{
  /* src/general_ice_model_data_module.f90:521-711 */
  double NW = perturb(f_NW, 1.1), NE = perturb(f_NE, 1.21), SW = perturb(f_SW, 1.13), SE = perturb(f_SE, 0.97);
  double phi;
  if (NW >= 0 && NE >= 0 && SW >= 0 && SE >= 0) phi = 1.0;
  else if (NW <= 0 && NE <= 0 && SW <= 0 && SE <= 0) phi = 0.0;
  else {
    int scen = rotate_quad_until_match(&NW, &NE, &SW, &SE);
    double aa, bb, cc, dd;
    if (scen == 1) {
      aa = SW; bb = SE - SW; cc = NW - SW; dd = NE + SW - NW - SE;
      phi = corner_area(aa, bb, cc, dd);
      if (fabs(dd) < 1e-4) {
        SW = (SW > 0) ? SW + 0.1 : SW - 0.1;
        aa = SW; bb = SE - SW; cc = NW - SW; dd = NE + SW - NW - SE;
        phi = corner_area(aa, bb, cc, dd);
      }
    } else if (scen == 2) {
      aa = -SW; bb = -(SE - SW); cc = -(NW - SW); dd = -(NE + SW - NW - SE);
      phi = 1.0 - corner_area(aa, bb, cc, dd);
      if (fabs(dd) < 1e-4) {
        SW = (SW > 0) ? SW + 0.1 : SW - 0.1;
        aa = -SW; bb = -(SE - SW); cc = -(NW - SW); dd = -(NE + SW - NW - SE);
        phi = 1.0 - corner_area(aa, bb, cc, dd);
      }
    } else if (scen == 3) {
      if (fabs(1.0 - NW / NE) < 1e-6 && fabs(1.0 - SW / SE) < 1e-6) {
        phi = SW / (SW - NW);
      } else {
        aa = SW; bb = SE - SW; cc = NW - SW; dd = NE + SW - NW - SE;
        phi = strip_prim(aa, bb, cc, dd, 1.0) - strip_prim(aa, bb, cc, dd, 0.0);
        if (fabs(dd) < 1e-4) {
          SW = (SW > 0) ? SW + 0.1 : SW - 0.1;
          aa = SW; bb = SE - SW; cc = NW - SW; dd = NE + SW - NW - SE;
          phi = strip_prim(aa, bb, cc, dd, 1.0) - strip_prim(aa, bb, cc, dd, 0.0);
        }
      }
    } else if (scen == 4) {
      aa = SW; bb = SE - SW; cc = NW - SW; dd = NE + SW - NW - SE;
      phi = corner_area(aa, bb, cc, dd);
      rotate_quad(&NW, &NE, &SW, &SE);
      rotate_quad(&NW, &NE, &SW, &SE);
      aa = SW; bb = SE - SW; cc = NW - SW; dd = NE + SW - NW - SE;
      phi += corner_area(aa, bb, cc, dd);
    } else {
      phi = NAN; /* the reference crashes: 'couldnt find matching scenario' */
    }
  }
  return fmax(0.0, fmin(1.0, phi));
}

static void determine_grounded_fractions(const G *g, const double *TAF, double *fa, double *fcx, double *fcy, double *fb)
{
  /* src/general_ice_model_data_module.f90:349-520 */
  const int nx = g->nx, ny = g->ny;
  const int ex = ny + 2;
  double *T = dalloc((size_t)(nx + 2) * (ny + 2));
#define TE(jj, ii) T[(size_t)(ii) * ex + (jj)]
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) TE(j + 1, i + 1) = TAF[A2(j, i)];
  for (int j = 0; j < ny; ++j) { TE(j + 1, 0) = TAF[A2(j, 0)]; TE(j + 1, nx + 1) = TAF[A2(j, nx - 1)]; }
  for (int i = 0; i < nx; ++i) { TE(0, i + 1) = TAF[A2(0, i)]; TE(ny + 1, i + 1) = TAF[A2(ny - 1, i)]; }
  TE(0, 0) = TAF[A2(0, 0)];
  TE(0, nx + 1) = TAF[A2(0, nx - 1)];
  TE(ny + 1, 0) = TAF[A2(ny - 1, 0)];
  TE(ny + 1, nx + 1) = TAF[A2(ny - 1, nx - 1)];
  double *qNW = dalloc((size_t)nx * ny), *qNE = dalloc((size_t)nx * ny), *qSW = dalloc((size_t)nx * ny), *qSE = dalloc((size_t)nx * ny);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      int ii = i + 1, jj = j + 1;
      double f_NW = 0.25 * (TE(jj + 1, ii - 1) + TE(jj + 1, ii) + TE(jj, ii - 1) + TE(jj, ii));
      double f_N = 0.5 * (TE(jj + 1, ii) + TE(jj, ii));
      double f_NE = 0.25 * (TE(jj + 1, ii) + TE(jj + 1, ii + 1) + TE(jj, ii) + TE(jj, ii + 1));
      double f_W = 0.5 * (TE(jj, ii - 1) + TE(jj, ii));
      double f_m = TE(jj, ii);
      double f_E = 0.5 * (TE(jj, ii) + TE(jj, ii + 1));
      double f_SW = 0.25 * (TE(jj, ii - 1) + TE(jj, ii) + TE(jj - 1, ii - 1) + TE(jj - 1, ii));
      double f_S = 0.5 * (TE(jj, ii) + TE(jj - 1, ii));
      double f_SE = 0.25 * (TE(jj, ii) + TE(jj, ii + 1) + TE(jj - 1, ii) + TE(jj - 1, ii + 1));
      size_t p = A2(j, i);
      qNW[p] = orc_calc_fraction_above_zero(f_NW, f_N, f_W, f_m);
      qNE[p] = orc_calc_fraction_above_zero(f_N, f_NE, f_m, f_E);
      qSW[p] = orc_calc_fraction_above_zero(f_W, f_m, f_SW, f_S);
      qSE[p] = orc_calc_fraction_above_zero(f_m, f_E, f_S, f_SE);
    }
#undef TE
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      size_t p = A2(j, i);
      fa[p] = 0.25 * (qNW[p] + qNE[p] + qSW[p] + qSE[p]);
      if (i < nx - 1) fcx[p] = 0.25 * (qNE[p] + qSE[p] + qNW[A2(j, i + 1)] + qSW[A2(j, i + 1)]);
      if (j < ny - 1) fcy[CY(j, i)] = 0.25 * (qNE[p] + qNW[p] + qSE[A2(j + 1, i)] + qSW[A2(j + 1, i)]);
      if (i < nx - 1 && j < ny - 1)
        fb[CY(j, i)] = 0.25 * (qNE[p] + qNW[A2(j, i + 1)] + qSE[A2(j + 1, i)] + qSW[A2(j + 1, i + 1)]);
    }
  free(T); free(qNW); free(qNE); free(qSW); free(qSE);
}

/* ------------------------------------------------------------------------ */
/* sliding laws, src/basal_conditions_and_sliding_module.f90:908-1432        */
/* ------------------------------------------------------------------------ */
double orc_sliding_beta(const imau_cfg *C, double u_a, double v_a, double beta_sq, double alpha_sq,
                        double phi_fric, double tauc, double Neff, double x, double y, double A_vav)
{
  double uabs = sqrt(C->slid_delta_v * C->slid_delta_v + u_a * u_a + v_a * v_a);
  switch (C->choice_sliding_law) {
  case IMAU_SLID_NO_SLIDING: return 0.0;                                               /* :926-928 */
  case IMAU_SLID_WEERTMAN:                                                             /* :984-987 */
    return beta_sq * pow(uabs, 1.0 / C->slid_Weertman_m - 1.0);
  case IMAU_SLID_COULOMB:                                                              /* :1039-1041 */
    return tauc / uabs;
  case IMAU_SLID_COULOMB_REGULARISED:                                                  /* :1078-1083 */
    return tauc * pow(uabs, C->slid_Coulomb_reg_q_plastic - 1.0) /
           pow(C->slid_Coulomb_reg_u_threshold, C->slid_Coulomb_reg_q_plastic);
  case IMAU_SLID_TSAI2015:                                                             /* :1128-1131 */
    return fmin(alpha_sq * Neff, beta_sq * pow(uabs, 1.0 / C->slid_Weertman_m)) * pow(uabs, -1.0);
  case IMAU_SLID_SCHOOF2005: {                                                         /* :1175-1183 */
    double m = C->slid_Weertman_m;
    return ((beta_sq * pow(uabs, 1.0 / m) * alpha_sq * Neff) /
            pow(pow(beta_sq, m) * uabs + pow(alpha_sq * Neff, m), 1.0 / m)) * pow(uabs, -1.0);
  }
  case IMAU_SLID_ZOET_IVERSON:                                                         /* :1220-1223 */
    return Neff * tan((PI_D / 180.0) * phi_fric) * pow(uabs, 1.0 / C->slid_ZI_p - 1.0) *
           pow(uabs + C->slid_ZI_ut, -1.0 / C->slid_ZI_p);
  case IMAU_SLID_IDEALISED:                                                            /* :1288-1432 */
    if (C->choice_idealised_sliding_law == IMAU_IDEAL_ISMIP_HOM_C)
      return 1000.0 + 1000.0 * sin(2.0 * PI_D * x / C->ISMIP_HOM_L) * sin(2.0 * PI_D * y / C->ISMIP_HOM_L);
    if (C->choice_idealised_sliding_law == IMAU_IDEAL_ISMIP_HOM_D)
      return 1000.0 + 1000.0 * sin(2.0 * PI_D * x / C->ISMIP_HOM_L);
    if (C->choice_idealised_sliding_law == IMAU_IDEAL_ISMIP_HOM_F)
      return pow(A_vav * 1000.0, -1.0);
    return NAN;
  default: return NAN;
  }
}

void orc_ssa_schoof2006(double n_flow, double L, double m, double tantheta, double h0,
                        double A_flow, double y, double *U, double *tauc)
{
  /* src/utilities_module.f90:836-877 */
  double B = pow(A_flow, -1.0 / n_flow);
  double f = ICE_DENSITY * GRAV * h0 * tantheta;
  double W = L * pow(m + 1.0, 1.0 / m);
  *tauc = f * pow(fabs(y / L), m);
  double ua = -2.0 * f * f * f * L * L * L * L / (B * B * B * h0 * h0 * h0);
  double ub = (1.0 / 4.0) * (pow(y / L, 4.0) - pow(m + 1.0, 4.0 / m));
  double uc = (-3.0 / ((m + 1.0) * (m + 4.0))) * (pow(fabs(y / L), m + 4.0) - pow(m + 1.0, 1.0 + 4.0 / m));
  double ud = (3.0 / (pow(m + 1.0, 2) * (2.0 * m + 4.0))) * (pow(fabs(y / L), 2 * m + 4.0) - pow(m + 1.0, 2.0 + 4.0 / m));
  double ue = (-1.0 / (pow(m + 1.0, 3) * (3.0 * m + 4.0))) * (pow(fabs(y / L), 3 * m + 4.0) - pow(m + 1.0, 3.0 + 4.0 / m));
  *U = ua * (ub + uc + ud + ue);
  if (fabs(y) > W) *U = 0.0;
}

/* ------------------------------------------------------------------------ */
/* solver state                                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
  G g;
  size_t na, ncx, ncy, nb;
  /* inputs */
  const imau_ice_io *io;
  /* working fields */
  double *u, *v;                   /* the velocity being solved (vav or SSA) */
  double *taub_cx, *taub_cy;
  double *visc3;                   /* visc_eff_3D_a */
  double *du_dz, *dv_dz;           /* (nz,ny,nx-1), (nz,ny-1,nx) */
  double *N_a, *N_b, *N_cx, *N_cy, *visc_int_a, *visc_int_b, *F2_a;
  double *beta_a, *beta_eff_a, *beta_eff_cx, *beta_eff_cy;
  double *taudx, *taudy;
  double *fg_a, *fg_cx, *fg_cy, *fg_b;
  double *overb, *pore, *Neff;
  double *beta_sq, *alpha_sq, *phi_fric, *tauc;
  double *u_prev, *v_prev;
  double *u_base_cx, *v_base_cy;
  int32_t *combi, *fin, *fout;
  int32_t *ij2n_u, *ij2n_v, *n2ij;
  int m;
  int32_t *ptr, *idx;
  double *val, *b, *x;
} S;

static void st_free(S *s)
{
  free(s->u); free(s->v); free(s->taub_cx); free(s->taub_cy); free(s->visc3); free(s->du_dz); free(s->dv_dz);
  free(s->N_a); free(s->N_b); free(s->N_cx); free(s->N_cy); free(s->visc_int_a); free(s->visc_int_b); free(s->F2_a);
  free(s->beta_a); free(s->beta_eff_a); free(s->beta_eff_cx); free(s->beta_eff_cy); free(s->taudx); free(s->taudy);
  free(s->fg_a); free(s->fg_cx); free(s->fg_cy); free(s->fg_b); free(s->overb); free(s->pore); free(s->Neff);
  free(s->beta_sq); free(s->alpha_sq); free(s->phi_fric); free(s->tauc); free(s->u_prev); free(s->v_prev);
  free(s->u_base_cx); free(s->v_base_cy); free(s->combi); free(s->fin); free(s->fout);
  free(s->ij2n_u); free(s->ij2n_v); free(s->n2ij); free(s->ptr); free(s->idx); free(s->val); free(s->b); free(s->x);
}

static void cp_in(double *dst, const double *src, size_t n) { if (src) memcpy(dst, src, n * sizeof(double)); }
static void cp_out(double *dst, const double *src, size_t n) { if (dst) memcpy(dst, src, n * sizeof(double)); }

static int st_init(S *s, const imau_grid *gr, const imau_cfg *c, const imau_ice_io *io, int ssa)
{
  memset(s, 0, sizeof *s);
  s->g.nx = gr->nx; s->g.ny = gr->ny; s->g.nz = c->nz; s->g.dx = gr->dx; s->g.x = gr->x; s->g.y = gr->y; s->g.C = c;
  s->io = io;
  const G *g = &s->g;
  size_t nx = g->nx, ny = g->ny, nz = g->nz;
  s->na = nx * ny; s->ncx = ny * (nx - 1); s->ncy = (ny - 1) * nx; s->nb = (ny - 1) * (nx - 1);
  s->u = dalloc(s->ncx); s->v = dalloc(s->ncy);
  cp_in(s->u, ssa ? io->u_SSA_cx : io->u_vav_cx, s->ncx);
  cp_in(s->v, ssa ? io->v_SSA_cy : io->v_vav_cy, s->ncy);
  s->taub_cx = dalloc(s->ncx); s->taub_cy = dalloc(s->ncy);
  cp_in(s->taub_cx, io->taub_cx, s->ncx); cp_in(s->taub_cy, io->taub_cy, s->ncy);
  s->visc3 = dalloc(nz * s->na); cp_in(s->visc3, io->visc_eff_3D_a, nz * s->na);
  s->du_dz = dalloc(nz * s->ncx); s->dv_dz = dalloc(nz * s->ncy);
  s->N_a = dalloc(s->na); s->N_b = dalloc(s->nb); s->N_cx = dalloc(s->ncx); s->N_cy = dalloc(s->ncy);
  s->visc_int_a = dalloc(s->na); s->visc_int_b = dalloc(s->nb); s->F2_a = dalloc(s->na);
  cp_in(s->N_a, io->N_a, s->na); cp_in(s->visc_int_a, io->visc_eff_int_a, s->na);
  s->beta_a = dalloc(s->na); s->beta_eff_a = dalloc(s->na); s->beta_eff_cx = dalloc(s->ncx); s->beta_eff_cy = dalloc(s->ncy);
  s->taudx = dalloc(s->ncx); s->taudy = dalloc(s->ncy);
  s->fg_a = dalloc(s->na); s->fg_cx = dalloc(s->ncx); s->fg_cy = dalloc(s->ncy); s->fg_b = dalloc(s->nb);
  s->overb = dalloc(s->na); s->pore = dalloc(s->na); s->Neff = dalloc(s->na);
  s->beta_sq = dalloc(s->na); s->alpha_sq = dalloc(s->na); s->phi_fric = dalloc(s->na); s->tauc = dalloc(s->na);
  cp_in(s->beta_sq, io->beta_sq_a, s->na); cp_in(s->alpha_sq, io->alpha_sq_a, s->na);
  cp_in(s->phi_fric, io->phi_fric_a, s->na); cp_in(s->tauc, io->tauc_a, s->na);
  s->u_prev = dalloc(s->ncx); s->v_prev = dalloc(s->ncy);
  s->u_base_cx = dalloc(s->ncx); s->v_base_cy = dalloc(s->ncy);
  s->combi = ialloc(s->na); s->fin = ialloc(s->na); s->fout = ialloc(s->na);
  if (io->DIVA_isfront_inner) memcpy(s->fin, io->DIVA_isfront_inner, s->na * sizeof(int32_t));
  /* src/ice_velocity_module.f90:2814-2860 */
  s->m = (int)((nx - 1) * ny + nx * (ny - 1));
  s->ij2n_u = ialloc(s->ncx); s->ij2n_v = ialloc(s->ncy); s->n2ij = ialloc((size_t)s->m * 4);
  int n = 0;
  for (size_t i = 0; i < nx; ++i)
    for (size_t j = 0; j < ny; ++j) {
      if (i < nx - 1) { s->ij2n_u[A2(j, i)] = n; s->n2ij[(size_t)n * 4 + 0] = (int)i; s->n2ij[(size_t)n * 4 + 1] = (int)j; s->n2ij[(size_t)n * 4 + 2] = -1; n++; }
      if (j < ny - 1) { s->ij2n_v[CY(j, i)] = n; s->n2ij[(size_t)n * 4 + 0] = -1; s->n2ij[(size_t)n * 4 + 2] = (int)i; s->n2ij[(size_t)n * 4 + 3] = (int)j; n++; }
    }
  s->ptr = ialloc((size_t)s->m + 1); s->idx = ialloc((size_t)s->m * 9); s->val = dalloc((size_t)s->m * 9);
  s->b = dalloc(s->m); s->x = dalloc(s->m);
  return 0;
}

/* calc_staggered_margin, src/ice_velocity_module.f90:1300-1319 */
static double stag_margin(double v0, double v1, double H0, double H1)
{
  if (H0 > 0.0 && H1 == 0.0) return v0;
  if (H0 == 0.0 && H1 > 0.0) return v1;
  return 0.5 * (v0 + v1);
}

/* src/ice_velocity_module.f90:664-721 */
static void calc_vertical_shear_strain_rates(S *s)
{
  const G *g = &s->g;
  const double *Hi = s->io->Hi_a;
  const int32_t *mi = s->io->mask_ice_a;
  const double *z = g->C->zeta;
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny; ++j) {
      if (i < g->nx - 1) {
        if (mi[A2(j, i)] == 0 && mi[A2(j, i + 1)] == 0) {
          for (int k = 0; k < g->nz; ++k) s->du_dz[A3I(k, j, i)] = 0.0;
        } else {
          for (int k = 0; k < g->nz; ++k) {
            double ve = stag_margin(s->visc3[A3I(k, j, i)], s->visc3[A3I(k, j, i + 1)], Hi[A2(j, i)], Hi[A2(j, i + 1)]);
            s->du_dz[A3I(k, j, i)] = (s->taub_cx[A2(j, i)] / ve) * z[k];
          }
        }
      }
      if (j < g->ny - 1) {
        if (mi[A2(j, i)] == 0 && mi[A2(j + 1, i)] == 0) {
          for (int k = 0; k < g->nz; ++k) s->dv_dz[CY3(k, j, i)] = 0.0;
        } else {
          for (int k = 0; k < g->nz; ++k) {
            double ve = stag_margin(s->visc3[A3I(k, j, i)], s->visc3[A3I(k, j + 1, i)], Hi[A2(j, i)], Hi[A2(j + 1, i)]);
            s->dv_dz[CY3(k, j, i)] = (s->taub_cy[CY(j, i)] / ve) * z[k];
          }
        }
      }
    }
}

/* src/ice_velocity_module.f90:722-899 */
static void calc_effective_viscosity(S *s, const double *u_cx, const double *v_cy)
{
  const G *g = &s->g;
  const imau_cfg *C = g->C;
  const double *Hi = s->io->Hi_a, *Hib = s->io->Hi_b, *A3 = s->io->A_flow_3D_a;
  const int32_t *mi = s->io->mask_ice_a;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  double *dudx = dalloc(s->na), *dudy = dalloc(s->na), *dvdx = dalloc(s->na), *dvdy = dalloc(s->na);
  ddx_cx_to_a(g, u_cx, dudx);
  ddy_cx_to_a(g, u_cx, dudy);
  ddx_cy_to_a(g, v_cy, dvdx);
  ddy_cy_to_a(g, v_cy, dvdy);
  const double n = C->n_flow;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      size_t p = A2(j, i);
      if (C->choice_ice_margin == IMAU_MARGIN_BC && mi[p] == 0) {
        for (int k = 0; k < nz; ++k) s->visc3[A3I(k, j, i)] = 0.0;
        /* visc_eff_int_a and N_a keep their previous values (CYCLE before :807-814) */
        continue;
      }
      for (int k = 0; k < nz; ++k) {
        double du_dz_a, dv_dz_a;
        if (i == 0) du_dz_a = s->du_dz[A3I(k, j, 0)];
        else if (i == nx - 1) du_dz_a = s->du_dz[A3I(k, j, nx - 2)];
        else du_dz_a = stag_margin(s->du_dz[A3I(k, j, i - 1)], s->du_dz[A3I(k, j, i)], Hi[A2(j, i - 1)], Hi[A2(j, i)]);
        if (j == 0) dv_dz_a = s->dv_dz[CY3(k, 0, i)];
        else if (j == ny - 1) dv_dz_a = s->dv_dz[CY3(k, ny - 2, i)];
        else dv_dz_a = stag_margin(s->dv_dz[CY3(k, j - 1, i)], s->dv_dz[CY3(k, j, i)], Hi[A2(j - 1, i)], Hi[A2(j, i)]);
        double eps_sq = dudx[p] * dudx[p] + dvdy[p] * dvdy[p] + dudx[p] * dvdy[p] +
                        0.25 * (dudy[p] + dvdx[p]) * (dudy[p] + dvdx[p]) +
                        0.25 * (du_dz_a * du_dz_a + dv_dz_a * dv_dz_a) + C->DIVA_epsilon_sq_0;
        double ve = 0.5 * pow(A3[A3I(k, j, i)], -1.0 / n) * pow(eps_sq, (1.0 - n) / (2.0 * n));
        s->visc3[A3I(k, j, i)] = fmax(ve, C->DIVA_visc_eff_min); /* NaN -> visc_eff_min */
      }
      s->visc_int_a[p] = orc_vertical_integrate(nz, C->zeta, s->visc3 + (size_t)p * nz);
      if (C->choice_ice_margin == IMAU_MARGIN_BC) s->N_a[p] = s->visc_int_a[p] * Hi[p];
      else s->N_a[p] = s->visc_int_a[p] * fmax(0.1, Hi[p]);
    }
  double *col = dalloc(nz);
  for (int i = 0; i < nx - 1; ++i)
    for (int j = 0; j < ny - 1; ++j) {
      if (C->choice_ice_margin == IMAU_MARGIN_BC) {
        double w = 0.0;
        for (int k = 0; k < nz; ++k) col[k] = 0.0;
        const int jj[4] = {j, j, j + 1, j + 1}, ii[4] = {i, i + 1, i, i + 1};
        for (int q = 0; q < 4; ++q)
          if (mi[A2(jj[q], ii[q])] == 1) {
            for (int k = 0; k < nz; ++k) col[k] += s->visc3[A3I(k, jj[q], ii[q])];
            w += 1.0;
          }
        if (w > 0.0) for (int k = 0; k < nz; ++k) col[k] /= w;
      } else {
        for (int k = 0; k < nz; ++k)
          col[k] = (s->visc3[A3I(k, j, i)] + s->visc3[A3I(k, j, i + 1)] + s->visc3[A3I(k, j + 1, i)] + s->visc3[A3I(k, j + 1, i + 1)]) / 4.0;
      }
      s->visc_int_b[CY(j, i)] = orc_vertical_integrate(nz, C->zeta, col);
      if (C->choice_ice_margin == IMAU_MARGIN_BC) s->N_b[CY(j, i)] = s->visc_int_b[CY(j, i)] * Hib[CY(j, i)];
      else s->N_b[CY(j, i)] = s->visc_int_b[CY(j, i)] * fmax(0.1, Hib[CY(j, i)]);
    }
  free(col);
  if (!C->include_SSADIVA_crossterms) {
    map_a_to_cx(g, s->N_a, s->N_cx);
    map_a_to_cy(g, s->N_a, s->N_cy);
  }
  free(dudx); free(dudy); free(dvdx); free(dvdy);
}

/* src/ice_velocity_module.f90:900-959 */
static void calc_sliding_term_beta(S *s, const double *u_cx, const double *v_cy)
{
  const G *g = &s->g;
  const imau_cfg *C = g->C;
  double *ua = dalloc(s->na), *va = dalloc(s->na);
  map_cx_to_a(g, u_cx, ua);
  map_cy_to_a(g, v_cy, va);
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny; ++j) {
      size_t p = A2(j, i);
      double tauc = s->tauc[p];
      if (C->choice_sliding_law == IMAU_SLID_COULOMB_REGULARISED ||
          (C->choice_sliding_law == IMAU_SLID_COULOMB &&
           !(C->choice_basal_roughness == IMAU_ROUGH_PARAMETERISED && C->choice_param_basal_roughness == IMAU_PROUGH_SSA_ICESTREAM))) {
        /* src/basal_conditions_and_sliding_module.f90:1031-1037, 1079 */
        tauc = tan((PI_D / 180.0) * s->phi_fric[p]) * s->Neff[p];
        s->tauc[p] = tauc;
      }
      double b;
      if (C->choice_sliding_law == IMAU_SLID_IDEALISED && C->choice_idealised_sliding_law == IMAU_IDEAL_ISMIP_HOM_E)
        b = s->io->ISMIP_HOM_E_slides[i] == 1 ? 0.0 : 1.0e30;  /* line i of the Arolla file, :1372-1384 */
      else
        b = orc_sliding_beta(C, ua[p], va[p], s->beta_sq[p], s->alpha_sq[p], s->phi_fric[p], tauc, s->Neff[p],
                             g->x[i], g->y[j], s->io->A_flow_vav_a ? s->io->A_flow_vav_a[p] : 0.0);
      b = fmin(b, C->DIVA_beta_max);
      if (!C->do_GL_subgrid_friction && s->io->mask_ocean_a[p] == 1) b = 0.0;
      s->beta_a[p] = b;
    }
  free(ua); free(va);
}

/* src/ice_velocity_module.f90:960-1013 */
static void calc_F_integral(S *s)
{
  const G *g = &s->g;
  const imau_cfg *C = g->C;
  const int nz = g->nz;
  double f[IMAU_NZ_MAX];
  for (int k = 0; k < nz; ++k) f[k] = (1.0 / C->DIVA_visc_eff_min) * pow(C->zeta[k], 2.0);
  double Fmin = orc_vertical_integrate(nz, C->zeta, f);
  for (size_t p = 0; p < s->na; ++p) {
    if (s->io->mask_ice_a[p] == 1) {
      for (int k = 0; k < nz; ++k) f[k] = (s->io->Hi_a[p] / s->visc3[p * nz + k]) * pow(C->zeta[k], 2.0);
      s->F2_a[p] = fmax(Fmin, orc_vertical_integrate(nz, C->zeta, f));
    } else {
      s->F2_a[p] = (C->choice_ice_margin == IMAU_MARGIN_BC) ? 0.0 : Fmin;
    }
  }
}

static int frozen_bed(const S *s, size_t p)
{
  const imau_cfg *C = s->g.C;
  if (!C->include_basal_freezing) return 0;
  size_t kb = p * s->g.nz + (s->g.nz - 1);
  return (s->io->Ti_pmp_a[kb] - s->io->Ti_a[kb]) > C->deltaT_basal_freezing;
}

/* src/ice_velocity_module.f90:1014-1086 */
static void calc_beta_eff(S *s)
{
  const G *g = &s->g;
  const imau_cfg *C = g->C;
  for (size_t p = 0; p < s->na; ++p) {
    if (C->choice_sliding_law == IMAU_SLID_NO_SLIDING) s->beta_eff_a[p] = 1.0 / s->F2_a[p];
    else {
      s->beta_eff_a[p] = s->beta_a[p] / (1.0 + s->beta_a[p] * s->F2_a[p]);
      if (frozen_bed(s, p)) s->beta_eff_a[p] = 1.0 / s->F2_a[p];
    }
  }
  map_a_to_cx(g, s->beta_eff_a, s->beta_eff_cx);
  map_a_to_cy(g, s->beta_eff_a, s->beta_eff_cy);
  if (C->do_GL_subgrid_friction) {
    for (size_t p = 0; p < s->ncx; ++p) s->beta_eff_cx[p] *= s->fg_cx[p] * s->fg_cx[p];
    for (size_t p = 0; p < s->ncy; ++p) s->beta_eff_cy[p] *= s->fg_cy[p] * s->fg_cy[p];
  }
}

/* src/ice_velocity_module.f90:1087-1120 */
static void calc_basal_stress(S *s)
{
  for (size_t p = 0; p < s->ncx; ++p) s->taub_cx[p] = s->beta_eff_cx[p] * s->u[p];
  for (size_t p = 0; p < s->ncy; ++p) s->taub_cy[p] = s->beta_eff_cy[p] * s->v[p];
}

/* src/ice_velocity_module.f90:1121-1199 */
static void calc_basal_velocities(S *s)
{
  const G *g = &s->g;
  const imau_cfg *C = g->C;
  const double *Hi = s->io->Hi_a;
  if (C->choice_sliding_law == IMAU_SLID_NO_SLIDING) {
    memset(s->u_base_cx, 0, s->ncx * sizeof(double));
    memset(s->v_base_cy, 0, s->ncy * sizeof(double));
    return;
  }
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny; ++j) {
      if (i < g->nx - 1) {
        double F2s = stag_margin(s->F2_a[A2(j, i)], s->F2_a[A2(j, i + 1)], Hi[A2(j, i)], Hi[A2(j, i + 1)]);
        s->u_base_cx[A2(j, i)] = s->u[A2(j, i)] - s->taub_cx[A2(j, i)] * F2s;
        if (frozen_bed(s, A2(j, i))) s->u_base_cx[A2(j, i)] = 0.0;
      }
      if (j < g->ny - 1) {
        double F2s = stag_margin(s->F2_a[A2(j, i)], s->F2_a[A2(j + 1, i)], Hi[A2(j, i)], Hi[A2(j + 1, i)]);
        s->v_base_cy[CY(j, i)] = s->v[CY(j, i)] - s->taub_cy[CY(j, i)] * F2s;
        if (frozen_bed(s, A2(j, i))) s->v_base_cy[CY(j, i)] = 0.0;
      }
    }
}

/* src/basal_conditions_and_sliding_module.f90:31-55 (+ 90-136, 243-294, 769-903) */
static void calc_basal_conditions(S *s)
{
  const G *g = &s->g;
  const imau_cfg *C = g->C;
  const imau_ice_io *io = s->io;
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny; ++j) {
      size_t p = A2(j, i);
      if (C->choice_basal_hydrology == IMAU_HYDRO_DRY) s->pore[p] = 0.0;
      else if (C->choice_basal_hydrology == IMAU_HYDRO_SATURATED) s->pore[p] = -SEAWATER_DENSITY * GRAV * (io->SL_a[p] - io->Hb_a[p]);
      else {
        double lp = fmin(1.0, fmax(0.0, 1.0 - (io->Hb_a[p] - io->SL_a[p] - C->Martin2011_hydro_Hb_min) /
                                                 (C->Martin2011_hydro_Hb_max - C->Martin2011_hydro_Hb_min)));
        s->pore[p] = 0.96 * ICE_DENSITY * GRAV * io->Hi_a[p] * lp;
      }
      s->overb[p] = ICE_DENSITY * GRAV * io->Hi_a[p];
      s->Neff[p] = fmax(0.0, s->overb[p] - s->pore[p]);
    }
  if (C->choice_sliding_law == IMAU_SLID_NO_SLIDING || C->choice_sliding_law == IMAU_SLID_IDEALISED) return;
  if (C->choice_basal_roughness != IMAU_ROUGH_PARAMETERISED) return;
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny; ++j) {
      size_t p = A2(j, i);
      if (C->choice_param_basal_roughness == IMAU_PROUGH_MARTIN2011) {
        double w = fmin(1.0, fmax(0.0, (io->Hb_a[p] - C->Martin2011till_phi_Hb_min) / (C->Martin2011till_phi_Hb_max - C->Martin2011till_phi_Hb_min)));
        s->phi_fric[p] = (1.0 - w) * C->Martin2011till_phi_min + w * C->Martin2011till_phi_max;
      } else if (C->choice_param_basal_roughness == IMAU_PROUGH_SSA_ICESTREAM) {
        double U;
        orc_ssa_schoof2006(C->n_flow, C->SSA_icestream_L, C->SSA_icestream_m, fabs(io->dHs_dx_a[p]), io->Hi_a[p],
                           io->A_flow_vav_a[p], g->y[j], &U, &s->tauc[p]);
      } else {
        /* src/basal_conditions_and_sliding_module.f90:846-899: Weertman sets beta_sq only */
        s->beta_sq[p] = 1.0e4;
        if (C->choice_sliding_law != IMAU_SLID_WEERTMAN) s->alpha_sq[p] = 0.5;
      }
    }
}

/* calc_basal_conditions alone, for unit tests: the pressures and Neff into
 * io's outputs, the friction angle and till yield stress into phi / tauc */
int orc_calc_basal_conditions(const imau_grid *gr, const imau_cfg *c, imau_ice_io *io, double *phi, double *tauc)
{
  S s;
  st_init(&s, gr, c, io, 0);
  calc_basal_conditions(&s);
  cp_out(io->overburden_pressure_a, s.overb, s.na);
  cp_out(io->pore_water_pressure_a, s.pore, s.na);
  cp_out(io->Neff_a, s.Neff, s.na);
  cp_out(phi, s.phi_fric, s.na);
  cp_out(tauc, s.tauc, s.na);
  st_free(&s);
  return 0;
}

/* src/ice_velocity_module.f90:1322-1431 */
static void set_sico_masks(S *s)
{
  const G *g = &s->g;
  const double *Hi = s->io->Hi_a, *TAF = s->io->TAF_a;
  const int nx = g->nx, ny = g->ny;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      size_t p = A2(j, i);
      int flt = TAF[p] < 0.0;
      if (Hi[p] == 0.0) s->combi[p] = flt ? 2 : 1;
      else s->combi[p] = flt ? 3 : 0;
    }
  /* the reference clears isfront_inner only on columns 1:MIN(ny-1,nx) (ny/nx quirk Q3);
   * every point is re-evaluated below, so stale values can only survive where
   * the conditions are false -- they are set, never cleared, below. */
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      if (i <= ny - 2) s->fin[A2(j, i)] = 0;
      s->fout[A2(j, i)] = 0;
    }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      int i1 = i > 0 ? i - 1 : 0, i2 = i < nx - 1 ? i + 1 : nx - 1;
      int j1 = j > 0 ? j - 1 : 0, j2 = j < ny - 1 ? j + 1 : ny - 1;
      size_t p = A2(j, i);
      if (Hi[p] > 0.0 && (Hi[A2(j, i1)] == 0.0 || Hi[A2(j, i2)] == 0.0 || Hi[A2(j1, i)] == 0.0 || Hi[A2(j2, i)] == 0.0))
        s->fin[p] = 1;
      if (Hi[p] == 0.0 && (Hi[A2(j, i1)] > 0.0 || Hi[A2(j, i2)] > 0.0 || Hi[A2(j1, i)] > 0.0 || Hi[A2(j2, i)] > 0.0))
        s->fout[p] = 1;
      if (s->fin[p] == 1 && s->combi[p] == 0) s->fin[p] = 0;
    }
}

/* ------------------------------------------------------------------------ */
/* CSR assembly, src/ice_velocity_module.f90:1656-2589                       */
/* ------------------------------------------------------------------------ */
typedef struct {
  const G *g;
  const int32_t *ij2n_u, *ij2n_v, *n2ij;
  const double *N_a, *N_b, *N_cx, *N_cy, *bcx, *bcy, *taudx, *taudy;
  const int32_t *combi, *fin, *fout;
  const double *Hi, *Hb, *SL;
  const double *u, *v;
  int32_t *ptr, *idx;
  double *val, *b, *x;
  int k;
} ASM;

static void put(ASM *a, int col, double v) { a->idx[a->k] = col; a->val[a->k] = v; a->k++; }

static int corner_bc(int b1, int b2)
{ /* src/ice_velocity_module.f90:1954-1989 */
  if (b1 == IMAU_BC_ZERO || b2 == IMAU_BC_ZERO) return IMAU_BC_ZERO;
  if (b1 == IMAU_BC_INFINITE || b2 == IMAU_BC_INFINITE) return IMAU_BC_INFINITE;
  return IMAU_BC_PERIODIC;
}

/* src/ice_velocity_module.f90:1932-2061 (u) and 2062-2191 (v) */
static void eq_boundary(ASM *a, int n, int isu)
{
  const G *g = a->g;
  const imau_cfg *C = g->C;
  int i, j, imax, jmax;
  int bw, be, bs, bn;
  if (isu) {
    i = a->n2ij[(size_t)n * 4 + 0]; j = a->n2ij[(size_t)n * 4 + 1];
    imax = g->nx - 2; jmax = g->ny - 1;
    bw = C->DIVA_boundary_BC_u_west; be = C->DIVA_boundary_BC_u_east; bs = C->DIVA_boundary_BC_u_south; bn = C->DIVA_boundary_BC_u_north;
  } else {
    i = a->n2ij[(size_t)n * 4 + 2]; j = a->n2ij[(size_t)n * 4 + 3];
    imax = g->nx - 1; jmax = g->ny - 2;
    bw = C->DIVA_boundary_BC_v_west; be = C->DIVA_boundary_BC_v_east; bs = C->DIVA_boundary_BC_v_south; bn = C->DIVA_boundary_BC_v_north;
  }
  int BC;
  if (i == 0 && j == 0) BC = corner_bc(bs, bw);
  else if (i == imax && j == 0) BC = corner_bc(bs, be);
  else if (i == 0 && j == jmax) BC = corner_bc(bn, bw);
  else if (i == imax && j == jmax) BC = corner_bc(bn, be);
  else if (i == 0) BC = bw;
  else if (i == imax) BC = be;
  else if (j == 0) BC = bs;
  else BC = bn;
  int i_in, i_opp, j_in, j_opp;
  if (i == 0) { i_in = 1; i_opp = imax - 1; }
  else if (i == imax) { i_in = imax - 1; i_opp = 1; }
  else { i_in = i; i_opp = i; }
  if (j == 0) { j_in = 1; j_opp = jmax - 1; }
  else if (j == jmax) { j_in = jmax - 1; j_opp = 1; }
  else { j_in = j; j_opp = j; }
  int nu, nin, nopp;
  double cur;
  if (isu) { nu = a->ij2n_u[A2(j, i)]; nin = a->ij2n_u[A2(j_in, i_in)]; nopp = a->ij2n_u[A2(j_opp, i_opp)]; cur = a->u[A2(j, i)]; }
  else { nu = a->ij2n_v[CY(j, i)]; nin = a->ij2n_v[CY(j_in, i_in)]; nopp = a->ij2n_v[CY(j_opp, i_opp)]; cur = a->v[CY(j, i)]; }
  if (BC == IMAU_BC_INFINITE) { put(a, nu, 1.0); put(a, nin, -1.0); a->b[n] = 0.0; a->x[n] = cur; }
  else if (BC == IMAU_BC_ZERO) { put(a, nu, 1.0); a->b[n] = 0.0; a->x[n] = 0.0; }
  else { put(a, nu, 1.0); put(a, nopp, -1.0); a->b[n] = 0.0; a->x[n] = cur; }
}

static void zero_row(ASM *a, int n) { put(a, n, 1.0); a->b[n] = 0.0; a->x[n] = 0.0; }

static double front_rhs(const ASM *a, size_t p)
{ /* src/ice_velocity_module.f90:2250-2261 */
  double H_ice = a->Hi[p], H_ocn;
  if (a->SL[p] > a->Hb[p]) H_ocn = fmin(ICE_DENSITY / SEAWATER_DENSITY * a->Hi[p], a->SL[p] - a->Hb[p]);
  else H_ocn = 0.0;
  return 0.5 * ICE_DENSITY * GRAV * H_ice * H_ice - 0.5 * SEAWATER_DENSITY * GRAV * H_ocn * H_ocn;
}

/* src/ice_velocity_module.f90:2192-2266 */
static void eq1_margin(ASM *a, int n, int i, int j)
{
  const G *g = a->g;
  double inv_dx = 1.0 / g->dx;
  int i1 = (a->fin[A2(j, i)] == 1) ? i : i + 1;
  if (a->fout[A2(j, i1 - 1)] == 1 && a->fout[A2(j, i1 + 1)] == 1) { zero_row(a, n); return; }
  double Na = a->N_a[A2(j, i1)];
  put(a, a->ij2n_u[A2(j, i1 - 1)], -4.0 * inv_dx * Na);
  put(a, a->ij2n_u[A2(j, i1)], 4.0 * inv_dx * Na);
  put(a, a->ij2n_v[CY(j - 1, i1)], -2.0 * inv_dx * Na);
  put(a, a->ij2n_v[CY(j, i1)], 2.0 * inv_dx * Na);
  a->b[n] = front_rhs(a, A2(j, i1));
  a->x[n] = a->u[A2(j, i)];
}
/* src/ice_velocity_module.f90:2267-2341 */
static void eq2_margin(ASM *a, int n, int i, int j)
{
  const G *g = a->g;
  double inv_dx = 1.0 / g->dx;
  int j1 = (a->fin[A2(j, i)] == 1) ? j : j + 1;
  if (a->fout[A2(j1 - 1, i)] == 1 && a->fout[A2(j1 + 1, i)] == 1) { zero_row(a, n); return; }
  double Na = a->N_a[A2(j1, i)];
  put(a, a->ij2n_v[CY(j1 - 1, i)], -4.0 * inv_dx * Na);
  put(a, a->ij2n_v[CY(j1, i)], 4.0 * inv_dx * Na);
  put(a, a->ij2n_u[A2(j1, i - 1)], -2.0 * inv_dx * Na);
  put(a, a->ij2n_u[A2(j1, i)], 2.0 * inv_dx * Na);
  a->b[n] = front_rhs(a, A2(j1, i));
  a->x[n] = a->v[CY(j, i)];
}

/* src/ice_velocity_module.f90:2342-2404 and 2468-2528 */
static void eq1_free(ASM *a, int n, int i, int j, int cross)
{
  const G *g = a->g;
  const double dx2 = g->dx * g->dx;
  const double *Na = a->N_a, *Nb = a->N_b;
  if (cross) {
    put(a, a->ij2n_u[A2(j, i + 1)], 4.0 * Na[A2(j, i + 1)]);
    put(a, a->ij2n_u[A2(j, i - 1)], 4.0 * Na[A2(j, i)]);
    put(a, a->ij2n_u[A2(j + 1, i)], Nb[CY(j, i)]);
    put(a, a->ij2n_u[A2(j - 1, i)], Nb[CY(j - 1, i)]);
    put(a, a->ij2n_u[A2(j, i)], -4.0 * (Na[A2(j, i + 1)] + Na[A2(j, i)]) + -1.0 * (Nb[CY(j, i)] + Nb[CY(j - 1, i)]) - a->bcx[A2(j, i)] * dx2);
    put(a, a->ij2n_v[CY(j, i)], -2.0 * Na[A2(j, i)] - Nb[CY(j, i)]);
    put(a, a->ij2n_v[CY(j, i + 1)], 2.0 * Na[A2(j, i + 1)] + Nb[CY(j, i)]);
    put(a, a->ij2n_v[CY(j - 1, i)], 2.0 * Na[A2(j, i)] + Nb[CY(j - 1, i)]);
    put(a, a->ij2n_v[CY(j - 1, i + 1)], -2.0 * Na[A2(j, i + 1)] - Nb[CY(j - 1, i)]);
    a->b[n] = -a->taudx[A2(j, i)] * dx2;
  } else {
    double Ncx = a->N_cx[A2(j, i)];
    put(a, a->ij2n_u[A2(j, i + 1)], 4.0);
    put(a, a->ij2n_u[A2(j, i - 1)], 4.0);
    put(a, a->ij2n_u[A2(j + 1, i)], 1.0);
    put(a, a->ij2n_u[A2(j - 1, i)], 1.0);
    put(a, a->ij2n_u[A2(j, i)], -10.0 - dx2 * a->bcx[A2(j, i)] / Ncx);
    put(a, a->ij2n_v[CY(j, i)], -3.0);
    put(a, a->ij2n_v[CY(j, i + 1)], 3.0);
    put(a, a->ij2n_v[CY(j - 1, i)], 3.0);
    put(a, a->ij2n_v[CY(j - 1, i + 1)], -3.0);
    a->b[n] = -a->taudx[A2(j, i)] * dx2 / Ncx;
  }
  a->x[n] = a->u[A2(j, i)];
}
/* src/ice_velocity_module.f90:2405-2467 and 2529-2589 */
static void eq2_free(ASM *a, int n, int i, int j, int cross)
{
  const G *g = a->g;
  const double dx2 = g->dx * g->dx;
  const double *Na = a->N_a, *Nb = a->N_b;
  if (cross) {
    put(a, a->ij2n_v[CY(j + 1, i)], 4.0 * Na[A2(j + 1, i)]);
    put(a, a->ij2n_v[CY(j - 1, i)], 4.0 * Na[A2(j, i)]);
    put(a, a->ij2n_v[CY(j, i + 1)], Nb[CY(j, i)]);
    put(a, a->ij2n_v[CY(j, i - 1)], Nb[CY(j, i - 1)]);
    put(a, a->ij2n_v[CY(j, i)], -4.0 * (Na[A2(j + 1, i)] + Na[A2(j, i)]) + -1.0 * (Nb[CY(j, i)] + Nb[CY(j, i - 1)]) - a->bcy[CY(j, i)] * dx2);
    put(a, a->ij2n_u[A2(j + 1, i - 1)], -2.0 * Na[A2(j + 1, i)] - Nb[CY(j, i - 1)]);
    put(a, a->ij2n_u[A2(j + 1, i)], 2.0 * Na[A2(j + 1, i)] + Nb[CY(j, i)]);
    put(a, a->ij2n_u[A2(j, i - 1)], 2.0 * Na[A2(j, i)] + Nb[CY(j, i - 1)]);
    put(a, a->ij2n_u[A2(j, i)], -2.0 * Na[A2(j, i)] - Nb[CY(j, i)]);
    a->b[n] = -a->taudy[CY(j, i)] * dx2;
  } else {
    double Ncy = a->N_cy[CY(j, i)];
    put(a, a->ij2n_v[CY(j + 1, i)], 4.0);
    put(a, a->ij2n_v[CY(j - 1, i)], 4.0);
    put(a, a->ij2n_v[CY(j, i + 1)], 1.0);
    put(a, a->ij2n_v[CY(j, i - 1)], 1.0);
    put(a, a->ij2n_v[CY(j, i)], -10.0 - dx2 * a->bcy[CY(j, i)] / Ncy);
    put(a, a->ij2n_u[A2(j + 1, i - 1)], -3.0);
    put(a, a->ij2n_u[A2(j + 1, i)], 3.0);
    put(a, a->ij2n_u[A2(j, i - 1)], 3.0);
    put(a, a->ij2n_u[A2(j, i)], -3.0);
    a->b[n] = -a->taudy[CY(j, i)] * dx2 / Ncy;
  }
  a->x[n] = a->v[CY(j, i)];
}

static int assemble(ASM *a)
{
  /* src/ice_velocity_module.f90:1683-1727 */
  const G *g = a->g;
  const imau_cfg *C = g->C;
  const int m = (g->nx - 1) * g->ny + g->nx * (g->ny - 1);
  a->k = 0;
  a->ptr[0] = 0;
  for (int n = 0; n < m; ++n) {
    int isu = a->n2ij[(size_t)n * 4 + 0] >= 0;
    if (isu) {
      int i = a->n2ij[(size_t)n * 4 + 0], j = a->n2ij[(size_t)n * 4 + 1];
      if (i == 0 || i == g->nx - 2 || j == 0 || j == g->ny - 1) eq_boundary(a, n, 1);
      else if (C->choice_ice_margin == IMAU_MARGIN_BC &&
               ((a->combi[A2(j, i)] == 3 && a->combi[A2(j, i + 1)] == 1) || (a->combi[A2(j, i)] == 1 && a->combi[A2(j, i + 1)] == 3)))
        zero_row(a, n);                                                             /* :1787-1796 */
      else if (C->choice_ice_margin == IMAU_MARGIN_BC &&
               ((a->fin[A2(j, i)] == 1 && a->fout[A2(j, i + 1)] == 1) || (a->fout[A2(j, i)] == 1 && a->fin[A2(j, i + 1)] == 1)))
        eq1_margin(a, n, i, j);                                                     /* :1799-1804 */
      else eq1_free(a, n, i, j, C->include_SSADIVA_crossterms);
    } else {
      int i = a->n2ij[(size_t)n * 4 + 2], j = a->n2ij[(size_t)n * 4 + 3];
      if (i == 0 || i == g->nx - 1 || j == 0 || j == g->ny - 2) eq_boundary(a, n, 0);
      else if (C->choice_ice_margin == IMAU_MARGIN_BC &&
               ((a->combi[A2(j, i)] == 3 && a->combi[A2(j + 1, i)] == 1) || (a->combi[A2(j, i)] == 1 && a->combi[A2(j + 1, i)] == 3)))
        zero_row(a, n);
      else if (C->choice_ice_margin == IMAU_MARGIN_BC &&
               ((a->fin[A2(j, i)] == 1 && a->fout[A2(j + 1, i)] == 1) || (a->fout[A2(j, i)] == 1 && a->fin[A2(j + 1, i)] == 1)))
        eq2_margin(a, n, i, j);
      else eq2_free(a, n, i, j, C->include_SSADIVA_crossterms);
    }
    a->ptr[n + 1] = a->k;
  }
  return a->k;
}

static void asm_setup(ASM *a, S *s, const double *u, const double *v)
{
  memset(a, 0, sizeof *a);
  a->g = &s->g;
  a->ij2n_u = s->ij2n_u; a->ij2n_v = s->ij2n_v; a->n2ij = s->n2ij;
  a->N_a = s->N_a; a->N_b = s->N_b; a->N_cx = s->N_cx; a->N_cy = s->N_cy;
  a->bcx = s->beta_eff_cx; a->bcy = s->beta_eff_cy; a->taudx = s->taudx; a->taudy = s->taudy;
  a->combi = s->combi; a->fin = s->fin; a->fout = s->fout;
  a->Hi = s->io->Hi_a; a->Hb = s->io->Hb_a; a->SL = s->io->SL_a;
  a->u = u; a->v = v;
  a->ptr = s->ptr; a->idx = s->idx; a->val = s->val; a->b = s->b; a->x = s->x;
}

/* src/utilities_module.f90:644-718 (single rank: i1:i2 = 1:m) */
int orc_sor(int m, const int32_t *ptr, const int32_t *idx, const double *val, const double *b,
            double *x, int nit, double tol, double omega, int *iterations)
{
  double omega_dyn = omega, res_max = tol * 2.0;
  int it = 0, total = 0;
  while (res_max > tol && it < nit) {
    it++; total++;
    res_max = 0.0;
    for (int i = 0; i < m; ++i) {
      double lhs = 0.0, cij = 0.0;
      for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
        int j = idx[k];
        lhs += val[k] * x[j];
        if (j == i) cij = val[k];
      }
      double res = (lhs - b[i]) / cij;
      res_max = fmax(res_max, fabs(res));
      x[i] -= omega_dyn * res;
    }
    if (it > 100 && res_max > 1e3) {
      omega_dyn -= 0.1;
      it = 0;
      for (int i = 0; i < m; ++i) x[i] = 0.0;
      if (omega_dyn <= 0.1) { if (iterations) *iterations = total; return fail(IMAU_ERR_DIVERGED, "SOR divergence even with extremely low relaxation parameter"); }
    }
  }
  if (iterations) *iterations = total;
  return 0;
}

/* Exact solve standing in for PETSc: banded Gaussian elimination with
 * partial pivoting (row r stores columns r-kl .. r+kl+ku). */
int orc_banded_lu_solve(int m, const int32_t *ptr, const int32_t *idx, const double *val,
                        const double *b, double *x)
{
  int kl = 0, ku = 0;
  for (int r = 0; r < m; ++r)
    for (int k = ptr[r]; k < ptr[r + 1]; ++k) {
      int d = idx[k] - r;
      if (d < 0 && -d > kl) kl = -d;
      if (d > 0 && d > ku) ku = d;
    }
  const int W = 2 * kl + ku + 1;
  double *A = (double *)calloc((size_t)m * W, sizeof(double));
  double *rhs = (double *)malloc((size_t)m * sizeof(double));
  if (!A || !rhs) { free(A); free(rhs); return fail(IMAU_ERR_ARG, "banded LU: out of memory"); }
#define AB(r, c) A[(size_t)(r) * W + ((c) - (r) + kl)]
  /* rows are equilibrated first: the domain-boundary rows (u - u_in = 0) have
   * unit entries while free rows carry N ~ 1e9..1e13, and unscaled
   * elimination would resolve the boundary rows only to ~1e-4 */
  for (int r = 0; r < m; ++r) {
    double smax = 0.0;
    for (int k = ptr[r]; k < ptr[r + 1]; ++k) smax = fmax(smax, fabs(val[k]));
    double sc = smax > 0.0 ? 1.0 / smax : 1.0;
    rhs[r] = b[r] * sc;
    for (int k = ptr[r]; k < ptr[r + 1]; ++k) AB(r, idx[k]) += val[k] * sc;
  }
  for (int k = 0; k < m; ++k) {
    int rmax = k + kl < m - 1 ? k + kl : m - 1;
    int cmax = k + kl + ku < m - 1 ? k + kl + ku : m - 1;
    int p = k;
    double best = fabs(AB(k, k));
    for (int r = k + 1; r <= rmax; ++r) if (fabs(AB(r, k)) > best) { best = fabs(AB(r, k)); p = r; }
    if (best == 0.0) { free(A); free(rhs); return fail(IMAU_ERR_NAN, "banded LU: singular matrix"); }
    if (p != k) {
      for (int c = k; c <= cmax; ++c) { double t = AB(k, c); AB(k, c) = AB(p, c); AB(p, c) = t; }
      double t = rhs[k]; rhs[k] = rhs[p]; rhs[p] = t;
    }
    double piv = AB(k, k);
    for (int r = k + 1; r <= rmax; ++r) {
      double f = AB(r, k) / piv;
      if (f == 0.0) continue;
      for (int c = k; c <= cmax; ++c) AB(r, c) -= f * AB(k, c);
      rhs[r] -= f * rhs[k];
    }
  }
  for (int k = m - 1; k >= 0; --k) {
    int cmax = k + kl + ku < m - 1 ? k + kl + ku : m - 1;
    double s = rhs[k];
    for (int c = k + 1; c <= cmax; ++c) s -= AB(k, c) * x[c];
    x[k] = s / AB(k, k);
  }
#undef AB
  free(A); free(rhs);
  return 0;
}

static int has_x_periodic(const imau_cfg *C)
{
  return C->DIVA_boundary_BC_u_west == IMAU_BC_PERIODIC || C->DIVA_boundary_BC_u_east == IMAU_BC_PERIODIC ||
         C->DIVA_boundary_BC_v_west == IMAU_BC_PERIODIC || C->DIVA_boundary_BC_v_east == IMAU_BC_PERIODIC;
}

/* src/ice_velocity_module.f90:1656-1761 */
static int solve_linearised(S *s, double *u, double *v, int lin)
{
  const G *g = &s->g;
  set_sico_masks(s);
  ASM a;
  asm_setup(&a, s, u, v);
  assemble(&a);
  int rc, its = 0;
  if (lin == ORC_LIN_BANDED_LU && !has_x_periodic(g->C))
    rc = orc_banded_lu_solve(s->m, s->ptr, s->idx, s->val, s->b, s->x);
  else
    rc = orc_sor(s->m, s->ptr, s->idx, s->val, s->b, s->x, s->io->DIVA_SOR_nit, s->io->DIVA_SOR_tol, s->io->DIVA_SOR_omega, &its);
  if (rc) return rc;
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny; ++j) {
      if (i < g->nx - 1) u[A2(j, i)] = s->x[s->ij2n_u[A2(j, i)]];
      if (j < g->ny - 1) v[CY(j, i)] = s->x[s->ij2n_v[CY(j, i)]];
    }
  return 0;
}

/* src/ice_velocity_module.f90:1432-1498 */
static void apply_velocity_limits(S *s, double *u, double *v)
{
  const G *g = &s->g;
  double vmax = g->C->DIVA_vel_max;
  double *u_cy = dalloc(s->ncy), *v_cx = dalloc(s->ncx), *ua_cx = dalloc(s->ncx), *ua_cy = dalloc(s->ncy);
  map_cx_to_cy(g, u, u_cy);
  map_cy_to_cx(g, v, v_cx);
  for (size_t p = 0; p < s->ncx; ++p) ua_cx[p] = sqrt(u[p] * u[p] + v_cx[p] * v_cx[p]);
  for (size_t p = 0; p < s->ncy; ++p) ua_cy[p] = sqrt(u_cy[p] * u_cy[p] + v[p] * v[p]);
  for (size_t p = 0; p < s->ncx; ++p) if (ua_cx[p] > vmax) u[p] = u[p] * vmax / ua_cx[p];
  for (size_t p = 0; p < s->ncy; ++p) if (ua_cy[p] > vmax) v[p] = v[p] * vmax / ua_cy[p];
  free(u_cy); free(v_cx); free(ua_cx); free(ua_cy);
}

/* src/ice_velocity_module.f90:1499-1529 */
static void relax(S *s, double *u, double *v, double rel)
{
  for (size_t p = 0; p < s->ncx; ++p) u[p] = rel * u[p] + (1.0 - rel) * s->u_prev[p];
  for (size_t p = 0; p < s->ncy; ++p) v[p] = rel * v[p] + (1.0 - rel) * s->v_prev[p];
}

/* src/ice_velocity_module.f90:1530-1601 (the reference accumulates in
 * (i outer, j inner, u then v) order) */
static double visc_iter_resid(S *s, const double *u, const double *v)
{
  const G *g = &s->g;
  long ncx = 0, ncy = 0;
  double r1 = 0.0, r2 = 0.0;
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny; ++j) {
      if (i < g->nx - 1) {
        size_t p = A2(j, i);
        if (fabs(u[p]) > 1e-6) { ncx++; r1 += pow(u[p] - s->u_prev[p], 2.0); r2 += pow(u[p] + s->u_prev[p], 2.0); }
      }
      if (j < g->ny - 1) {
        size_t p = CY(j, i);
        if (fabs(v[p]) > 1e-6) { ncy++; r1 += pow(v[p] - s->v_prev[p], 2.0); r2 += pow(v[p] + s->v_prev[p], 2.0); }
      }
    }
  if (ncx + ncy > 0) return 2.0 * sqrt(r1) / fmax(sqrt(r2), 1e-8);
  return 0.0;
}

/* Relative FP64 residual of the assembled system at (u,v) (no reference
 * counterpart; the measured stop rule). */
static double csr_rel_residual(S *s, const double *u, const double *v)
{
  ASM a;
  asm_setup(&a, s, u, v);
  set_sico_masks(s);
  assemble(&a);
  /* x must hold the state itself, not the assembly's initial guess (which is 0 on zero rows) */
  const G *g = &s->g;
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny; ++j) {
      if (i < g->nx - 1) s->x[s->ij2n_u[A2(j, i)]] = u[A2(j, i)];
      if (j < g->ny - 1) s->x[s->ij2n_v[CY(j, i)]] = v[CY(j, i)];
    }
  double rr = 0.0, bb = 0.0;
  for (int r = 0; r < s->m; ++r) {
    double ax = 0.0;
    for (int k = s->ptr[r]; k < s->ptr[r + 1]; ++k) ax += s->val[k] * s->x[s->idx[k]];
    double d = ax - s->b[r];
    rr += d * d;
    bb += s->b[r] * s->b[r];
  }
  return bb > 0.0 ? sqrt(rr) / sqrt(bb) : sqrt(rr);
}

/* src/ice_velocity_module.f90:1200-1298 */
static void calc_3D_horizontal_velocities_DIVA(S *s, double *u3, double *v3, double *F1)
{
  const G *g = &s->g;
  const imau_cfg *C = g->C;
  const int nz = g->nz;
  const double *Hi = s->io->Hi_a;
  double f[IMAU_NZ_MAX], F1s[IMAU_NZ_MAX];
  for (size_t p = 0; p < s->na; ++p) {
    for (int k = 0; k < nz; ++k) f[k] = (-Hi[p] / s->visc3[p * nz + k]) * C->zeta[k];
    orc_vertical_integration_from_bottom_to_zeta(nz, C->zeta, f, F1 + p * nz);
  }
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny; ++j) {
      if (i < g->nx - 1) {
        size_t p0 = A2(j, i), p1 = A2(j, i + 1);
        for (int k = 0; k < nz; ++k) {
          if (C->choice_ice_margin == IMAU_MARGIN_BC) {
            if (Hi[p0] > 0 && Hi[p1] == 0) F1s[k] = F1[p0 * nz + k];
            else if (Hi[p0] == 0 && Hi[p1] > 0) F1s[k] = F1[p1 * nz + k];
            else if (Hi[p0] > 0 && Hi[p1] > 0) F1s[k] = 0.5 * (F1[p0 * nz + k] + F1[p1 * nz + k]);
            else F1s[k] = 0.0;
          } else F1s[k] = 0.5 * (F1[p0 * nz + k] + F1[p1 * nz + k]);
          u3[A3I(k, j, i)] = s->u_base_cx[p0] + s->taub_cx[p0] * F1s[k];
        }
      }
      if (j < g->ny - 1) {
        size_t p0 = A2(j, i), p1 = A2(j + 1, i);
        for (int k = 0; k < nz; ++k) {
          if (C->choice_ice_margin == IMAU_MARGIN_BC) {
            if (Hi[p0] > 0 && Hi[p1] == 0) F1s[k] = F1[p0 * nz + k];
            else if (Hi[p0] == 0 && Hi[p1] > 0) F1s[k] = F1[p1 * nz + k];
            else if (Hi[p0] > 0 && Hi[p1] > 0) F1s[k] = 0.5 * (F1[p0 * nz + k] + F1[p1 * nz + k]);
            else F1s[k] = 0.0;
          } else F1s[k] = 0.5 * (F1[p0 * nz + k] + F1[p1 * nz + k]);
          v3[CY3(k, j, i)] = s->v_base_cy[CY(j, i)] + s->taub_cy[CY(j, i)] * F1s[k];
        }
      }
    }
}

/* src/ice_velocity_module.f90:2592-2811 (x direction; 2D when nzz==1) */
static void map_vel_cx_to_a(const G *g, const int32_t *mi, const double *ucx, double *ua, int nzz)
{
  const int nx = g->nx, ny = g->ny;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nzz; ++k) {
        double val;
#define UC(jj, ii) ucx[((size_t)(ii) * ny + (jj)) * nzz + k]
        if (i == 0) val = mi[A2(j, i)] == 1 ? UC(j, i) : 0.0;
        else if (i == nx - 1) val = mi[A2(j, i)] == 1 ? UC(j, i - 1) : 0.0;
        else {
          int a = mi[A2(j, i - 1)], b = mi[A2(j, i)], c = mi[A2(j, i + 1)];
          if (a == 1 && b == 1 && c == 1) val = 0.5 * (UC(j, i - 1) + UC(j, i));
          else if (a == 0 && b == 1 && c == 1) val = UC(j, i);
          else if (a == 1 && b == 1 && c == 0) val = UC(j, i - 1);
          else val = 0.0;
        }
#undef UC
        ua[((size_t)i * ny + j) * nzz + k] = val;
      }
}
static void map_vel_cy_to_a(const G *g, const int32_t *mi, const double *vcy, double *va, int nzz)
{
  const int nx = g->nx, ny = g->ny;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nzz; ++k) {
        double val;
#define VC(jj, ii) vcy[((size_t)(ii) * (ny - 1) + (jj)) * nzz + k]
        if (j == 0) val = mi[A2(j, i)] == 1 ? VC(j, i) : 0.0;
        else if (j == ny - 1) val = mi[A2(j, i)] == 1 ? VC(j - 1, i) : 0.0;
        else {
          int a = mi[A2(j - 1, i)], b = mi[A2(j, i)], c = mi[A2(j + 1, i)];
          if (a == 1 && b == 1 && c == 1) val = 0.5 * (VC(j - 1, i) + VC(j, i));
          else if (a == 0 && b == 1 && c == 1) val = VC(j, i);
          else if (a == 1 && b == 1 && c == 0) val = VC(j - 1, i);
          else val = 0.0;
        }
#undef VC
        va[((size_t)i * ny + j) * nzz + k] = val;
      }
}

/* src/ice_velocity_module.f90:534-661.  u3a/v3a are the a-grid 3D velocities
 * of the PREVIOUS call (they are refreshed only afterwards, :491-492). */
static void calc_3D_vertical_velocities(const G *g, const imau_ice_io *io, const double *u3, const double *v3,
                                        const double *u3a, const double *v3a, double *w3)
{
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const double *z = g->C->zeta;
  const double dx = g->dx;
  size_t na = (size_t)nx * ny;
  double *dHs_dx = dalloc(na), *dHs_dy = dalloc(na), *dHi_dx = dalloc(na), *dHi_dy = dalloc(na);
  ddx_a_to_a(g, io->Hs_a, dHs_dx);
  ddy_a_to_a(g, io->Hs_a, dHs_dy);
  ddx_a_to_a(g, io->Hi_a, dHi_dx);
  ddy_a_to_a(g, io->Hi_a, dHi_dy);
#define U3(k, j, i) u3[((size_t)(i) * ny + (j)) * nz + (k)]
#define V3(k, j, i) v3[((size_t)(i) * (ny - 1) + (j)) * nz + (k)]
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      size_t p = A2(j, i);
      double dHbase_dx = dHs_dx[p] - dHi_dx[p], dHbase_dy = dHs_dy[p] - dHi_dy[p];
      double ub, vb;
      if (i == 0) ub = U3(nz - 1, j, 0);
      else if (i == nx - 1) ub = U3(nz - 1, j, nx - 2);
      else ub = 0.5 * (U3(nz - 1, j, i - 1) + U3(nz - 1, j, i));
      if (j == 0) vb = V3(nz - 1, 0, i);
      else if (j == ny - 1) vb = V3(nz - 1, ny - 2, i);
      else vb = 0.5 * (V3(nz - 1, j - 1, i) + V3(nz - 1, j, i));
      if (io->mask_sheet_a[p] == 1) w3[p * nz + nz - 1] = ub * dHbase_dx + vb * dHbase_dy + (io->dHb_dt_a ? io->dHb_dt_a[p] : 0.0);
      else w3[p * nz + nz - 1] = ub * dHbase_dx + vb * dHbase_dy;
      for (int k = nz - 2; k >= 0; --k) {
        double dudx_k, dudx_k1, dvdy_k, dvdy_k1;
        if (i == 0) { dudx_k = (U3(k, j, 1) - U3(k, j, 0)) / dx; dudx_k1 = (U3(k + 1, j, 1) - U3(k + 1, j, 0)) / dx; }
        else if (i == nx - 1) { dudx_k = (U3(k, j, nx - 2) - U3(k, j, nx - 3)) / dx; dudx_k1 = (U3(k + 1, j, nx - 2) - U3(k + 1, j, nx - 3)) / dx; }
        else { dudx_k = (U3(k, j, i) - U3(k, j, i - 1)) / dx; dudx_k1 = (U3(k + 1, j, i) - U3(k + 1, j, i - 1)) / dx; }
        if (j == 0) { dvdy_k = (V3(k, 1, i) - V3(k, 0, i)) / dx; dvdy_k1 = (V3(k + 1, 1, i) - V3(k + 1, 0, i)) / dx; }
        else if (j == ny - 1) { dvdy_k = (V3(k, ny - 2, i) - V3(k, ny - 3, i)) / dx; dvdy_k1 = (V3(k + 1, ny - 2, i) - V3(k + 1, ny - 3, i)) / dx; }
        else { dvdy_k = (V3(k, j, i) - V3(k, j - 1, i)) / dx; dvdy_k1 = (V3(k + 1, j, i) - V3(k + 1, j - 1, i)) / dx; }
        double w1 = (dudx_k + dudx_k1) / 2.0, w2 = (dvdy_k + dvdy_k1) / 2.0;
        double zm = 0.5 * (z[k + 1] + z[k]), dz = z[k + 1] - z[k];
        double Hm = fmax(0.1, io->Hi_a[p]);
        double w3t = ((io->dHs_dx_a[p] - zm * io->dHi_dx_a[p]) / Hm) * ((u3a[p * nz + k + 1] - u3a[p * nz + k]) / dz);
        double w4t = ((io->dHs_dy_a[p] - zm * io->dHi_dy_a[p]) / Hm) * ((v3a[p * nz + k + 1] - v3a[p * nz + k]) / dz);
        w3[p * nz + k] = w3[p * nz + k + 1] - io->Hi_a[p] * (w1 + w2 + w3t + w4t) * dz;
      }
    }
#undef U3
#undef V3
  free(dHs_dx); free(dHs_dy); free(dHi_dx); free(dHi_dy);
}

/* src/ice_velocity_module.f90:377-533 */
static void calc_secondary_velocities(const G *g, imau_ice_io *io, double *u3, double *v3,
                                      const double *u_in_cx, const double *v_in_cy,
                                      double *u_base_cx, double *v_base_cy, double *u_vav, double *v_vav)
{
  const imau_cfg *C = g->C;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  size_t na = (size_t)nx * ny, ncx = (size_t)ny * (nx - 1), ncy = (size_t)(ny - 1) * nx;
  double *u_surf_cx = dalloc(ncx), *v_surf_cy = dalloc(ncy);
  double *u3a = dalloc(na * nz), *v3a = dalloc(na * nz), *w3 = dalloc(na * nz);
  cp_in(u3a, io->u_3D_a, na * nz);
  cp_in(v3a, io->v_3D_a, na * nz);
  int dyn = C->choice_ice_dynamics;
  if (dyn == IMAU_DYN_SIA) {
    memcpy(u3, io->u_3D_SIA_cx, ncx * nz * sizeof(double));
    memcpy(v3, io->v_3D_SIA_cy, ncy * nz * sizeof(double));
    memset(u_base_cx, 0, ncx * sizeof(double));
    memset(v_base_cy, 0, ncy * sizeof(double));
    calc_3D_vertical_velocities(g, io, u3, v3, u3a, v3a, w3);
    for (size_t p = 0; p < ncx; ++p) { u_surf_cx[p] = u3[p * nz]; u_vav[p] = orc_vertical_average(nz, C->zeta, u3 + p * nz); }
    for (size_t p = 0; p < ncy; ++p) { v_surf_cy[p] = v3[p * nz]; v_vav[p] = orc_vertical_average(nz, C->zeta, v3 + p * nz); }
  } else if (dyn == IMAU_DYN_SSA) {
    memcpy(u_base_cx, u_in_cx, ncx * sizeof(double));
    memcpy(v_base_cy, v_in_cy, ncy * sizeof(double));
    for (size_t p = 0; p < ncx; ++p) { u_vav[p] = u_in_cx[p]; u_surf_cx[p] = u_in_cx[p]; for (int k = 0; k < nz; ++k) u3[p * nz + k] = u_in_cx[p]; }
    for (size_t p = 0; p < ncy; ++p) { v_vav[p] = v_in_cy[p]; v_surf_cy[p] = v_in_cy[p]; for (int k = 0; k < nz; ++k) v3[p * nz + k] = v_in_cy[p]; }
  } else if (dyn == IMAU_DYN_SIA_SSA) {
    for (size_t p = 0; p < ncx; ++p) for (int k = 0; k < nz; ++k) u3[p * nz + k] = io->u_3D_SIA_cx[p * nz + k] + u_in_cx[p];
    for (size_t p = 0; p < ncy; ++p) for (int k = 0; k < nz; ++k) v3[p * nz + k] = io->v_3D_SIA_cy[p * nz + k] + v_in_cy[p];
    memcpy(u_base_cx, u_in_cx, ncx * sizeof(double));
    memcpy(v_base_cy, v_in_cy, ncy * sizeof(double));
    calc_3D_vertical_velocities(g, io, u3, v3, u3a, v3a, w3);
    for (size_t p = 0; p < ncx; ++p) { u_surf_cx[p] = u3[p * nz]; u_vav[p] = orc_vertical_average(nz, C->zeta, u3 + p * nz); }
    for (size_t p = 0; p < ncy; ++p) { v_surf_cy[p] = v3[p * nz]; v_vav[p] = orc_vertical_average(nz, C->zeta, v3 + p * nz); }
  } else { /* DIVA */
    calc_3D_vertical_velocities(g, io, u3, v3, u3a, v3a, w3);
    for (size_t p = 0; p < ncx; ++p) u_surf_cx[p] = u3[p * nz];
    for (size_t p = 0; p < ncy; ++p) v_surf_cy[p] = v3[p * nz];
  }
  const int32_t *mi = io->mask_ice_a;
  double *uva = dalloc(na), *vva = dalloc(na), *usa = dalloc(na), *vsa = dalloc(na), *uba = dalloc(na), *vba = dalloc(na);
  map_vel_cx_to_a(g, mi, u3, u3a, nz);
  map_vel_cy_to_a(g, mi, v3, v3a, nz);
  map_vel_cx_to_a(g, mi, u_vav, uva, 1);
  map_vel_cy_to_a(g, mi, v_vav, vva, 1);
  map_vel_cx_to_a(g, mi, u_surf_cx, usa, 1);
  map_vel_cy_to_a(g, mi, v_surf_cy, vsa, 1);
  map_vel_cx_to_a(g, mi, u_base_cx, uba, 1);
  map_vel_cy_to_a(g, mi, v_base_cy, vba, 1);
  double *R = dalloc(na), *ua_v = dalloc(na), *ua_s = dalloc(na), *ua_b = dalloc(na);
  for (size_t p = 0; p < na; ++p) {
    ua_v[p] = sqrt(uva[p] * uva[p] + vva[p] * vva[p]);
    ua_s[p] = sqrt(usa[p] * usa[p] + vsa[p] * vsa[p]);
    ua_b[p] = sqrt(uba[p] * uba[p] + vba[p] * vba[p]);
    if (dyn == IMAU_DYN_SIA) R[p] = 1.0;
    else if (dyn == IMAU_DYN_SSA) R[p] = 0.0;
    else R[p] = (mi[p] == 1) ? fmax(0.0, fmin(1.0, 1.0 - ua_b[p] / ua_s[p])) : 0.0;
  }
  cp_out(io->u_3D_a, u3a, na * nz); cp_out(io->v_3D_a, v3a, na * nz); cp_out(io->w_3D_a, w3, na * nz);
  cp_out(io->u_vav_a, uva, na); cp_out(io->v_vav_a, vva, na); cp_out(io->u_surf_a, usa, na); cp_out(io->v_surf_a, vsa, na);
  cp_out(io->u_base_a, uba, na); cp_out(io->v_base_a, vba, na);
  cp_out(io->uabs_vav_a, ua_v, na); cp_out(io->uabs_surf_a, ua_s, na); cp_out(io->uabs_base_a, ua_b, na);
  cp_out(io->R_shear, R, na);
  cp_out(io->u_surf_cx, u_surf_cx, ncx); cp_out(io->v_surf_cy, v_surf_cy, ncy);
  if (io->w_surf_a) for (size_t p = 0; p < na; ++p) io->w_surf_a[p] = w3[p * nz];
  if (io->w_base_a) for (size_t p = 0; p < na; ++p) io->w_base_a[p] = w3[p * nz + nz - 1];
  free(u_surf_cx); free(v_surf_cy); free(u3a); free(v3a); free(w3);
  free(uva); free(vva); free(usa); free(vsa); free(uba); free(vba); free(R); free(ua_v); free(ua_s); free(ua_b);
}

static void write_common(S *s, imau_ice_io *io)
{
  const G *g = &s->g;
  size_t nz = g->nz;
  cp_out(io->taub_cx, s->taub_cx, s->ncx); cp_out(io->taub_cy, s->taub_cy, s->ncy);
  cp_out(io->visc_eff_3D_a, s->visc3, nz * s->na);
  cp_out(io->du_dz_3D_cx, s->du_dz, nz * s->ncx); cp_out(io->dv_dz_3D_cy, s->dv_dz, nz * s->ncy);
  cp_out(io->beta_a, s->beta_a, s->na); cp_out(io->beta_eff_a, s->beta_eff_a, s->na);
  cp_out(io->beta_eff_cx, s->beta_eff_cx, s->ncx); cp_out(io->beta_eff_cy, s->beta_eff_cy, s->ncy);
  cp_out(io->taudx_cx, s->taudx, s->ncx); cp_out(io->taudy_cy, s->taudy, s->ncy);
  cp_out(io->N_a, s->N_a, s->na); cp_out(io->N_b, s->N_b, s->nb);
  cp_out(io->N_cx, s->N_cx, s->ncx); cp_out(io->N_cy, s->N_cy, s->ncy);
  cp_out(io->visc_eff_int_a, s->visc_int_a, s->na); cp_out(io->F2_a, s->F2_a, s->na);
  cp_out(io->f_grnd_a, s->fg_a, s->na); cp_out(io->f_grnd_cx, s->fg_cx, s->ncx);
  cp_out(io->f_grnd_cy, s->fg_cy, s->ncy); cp_out(io->f_grnd_b, s->fg_b, s->nb);
  cp_out(io->overburden_pressure_a, s->overb, s->na); cp_out(io->pore_water_pressure_a, s->pore, s->na);
  cp_out(io->Neff_a, s->Neff, s->na);
  cp_out(io->beta_sq_a, s->beta_sq, s->na); cp_out(io->alpha_sq_a, s->alpha_sq, s->na);
  cp_out(io->phi_fric_a, s->phi_fric, s->na); cp_out(io->tauc_a, s->tauc, s->na);
  if (io->DIVA_mask_combi) memcpy(io->DIVA_mask_combi, s->combi, s->na * sizeof(int32_t));
  if (io->DIVA_isfront_inner) memcpy(io->DIVA_isfront_inner, s->fin, s->na * sizeof(int32_t));
  if (io->DIVA_isfront_outer) memcpy(io->DIVA_isfront_outer, s->fout, s->na * sizeof(int32_t));
  cp_out(io->u_cx_prev, s->u_prev, s->ncx); cp_out(io->v_cy_prev, s->v_prev, s->ncy);
  cp_out(io->u_base_cx, s->u_base_cx, s->ncx); cp_out(io->v_base_cy, s->v_base_cy, s->ncy);
}

static void calc_taud(S *s)
{
  /* src/ice_velocity_module.f90:288-294 */
  for (size_t p = 0; p < s->ncx; ++p) s->taudx[p] = -ICE_DENSITY * GRAV * s->io->Hi_cx[p] * s->io->dHs_dx_cx[p];
  for (size_t p = 0; p < s->ncy; ++p) s->taudy[p] = -ICE_DENSITY * GRAV * s->io->Hi_cy[p] * s->io->dHs_dy_cy[p];
}

static int sum_sheet(const S *s)
{
  long t = 0;
  for (size_t p = 0; p < s->na; ++p) t += s->io->mask_sheet_a[p];
  return t > 0;
}

/* the configuration checks of calc_bed_roughness_Martin2011 / _MISMIPplus
 * (src/basal_conditions_and_sliding_module.f90:789-792, 896-898) */
/* ISMIP-HOM E: the slip columns of the 51-line Arolla file, one per grid
 * column (src/basal_conditions_and_sliding_module.f90:1372-1389) */
static int check_ismip_e(const imau_grid *gr, const imau_cfg *c, const imau_ice_io *io)
{
  if (c->choice_sliding_law != IMAU_SLID_IDEALISED || c->choice_idealised_sliding_law != IMAU_IDEAL_ISMIP_HOM_E) return 0;
  if (!io->ISMIP_HOM_E_slides) return fail(IMAU_ERR_ARG, "ISMIP_HOM_E needs ISMIP_HOM_E_slides");
  if (gr->nx != 51) return fail(IMAU_ERR_CONFIG, "ISMIP_HOM_E: the Arolla file has 51 lines, one per grid column");
  return 0;
}

static int check_roughness(const imau_cfg *c)
{
  int law = c->choice_sliding_law;
  if (c->choice_basal_roughness != IMAU_ROUGH_PARAMETERISED || law == IMAU_SLID_NO_SLIDING || law == IMAU_SLID_IDEALISED)
    return 0;
  if (c->choice_param_basal_roughness == IMAU_PROUGH_MARTIN2011 && law != IMAU_SLID_COULOMB &&
      law != IMAU_SLID_COULOMB_REGULARISED && law != IMAU_SLID_ZOET_IVERSON)
    return fail(IMAU_ERR_CONFIG, "Martin2011 roughness needs Coulomb, Coulomb_regularised or Zoet-Iverson sliding");
  if (c->choice_param_basal_roughness == IMAU_PROUGH_MISMIPPLUS && law != IMAU_SLID_WEERTMAN &&
      law != IMAU_SLID_TSAI2015 && law != IMAU_SLID_SCHOOF2005)
    return fail(IMAU_ERR_CONFIG, "MISMIP+ roughness needs Weertman, Tsai2015 or Schoof2005 sliding");
  return 0;
}

/* src/ice_velocity_module.f90:249-374 */
/* estimate_visc_iter_UV_errors, src/ice_velocity_module.f90:1602-1652: the
 * relative change of this iteration where the velocity chk (ice%u_vav_cx)
 * exceeds 1e-6 m/yr */
static void visc_iter_errors(const S *s, const double *chk_u, const double *chk_v, imau_ice_io *io)
{
  if (io->DIVA_err_cx)
    for (size_t p = 0; p < s->ncx; ++p)
      io->DIVA_err_cx[p] = fabs(chk_u ? chk_u[p] : 0.0) > 1e-6
                               ? 2.0 * fabs(s->u[p] - s->u_prev[p]) / fabs(s->u[p] + s->u_prev[p] + 1e-5) : 0.0;
  if (io->DIVA_err_cy)
    for (size_t p = 0; p < s->ncy; ++p)
      io->DIVA_err_cy[p] = fabs(chk_v ? chk_v[p] : 0.0) > 1e-6
                               ? 2.0 * fabs(s->v[p] - s->v_prev[p]) / fabs(s->v[p] + s->v_prev[p] + 1e-5) : 0.0;
}
static void visc_iter_errors_init(const S *s, imau_ice_io *io)  /* :176-177, :303-304 */
{
  if (io->DIVA_err_cx) for (size_t p = 0; p < s->ncx; ++p) io->DIVA_err_cx[p] = 1e5;
  if (io->DIVA_err_cy) for (size_t p = 0; p < s->ncy; ++p) io->DIVA_err_cy[p] = 1e5;
}

int orc_solve_diva(const imau_grid *gr, const imau_cfg *c, imau_ice_io *io, int lin)
{
  if (c->choice_ice_dynamics != IMAU_DYN_DIVA) return fail(IMAU_ERR_CONFIG, "solve_DIVA: choice_ice_dynamics is not DIVA");
  if (check_roughness(c)) return IMAU_ERR_CONFIG;
  { int e = check_ismip_e(gr, c, io); if (e) return e; }
  S s;
  st_init(&s, gr, c, io, 0);
  const G *g = &s.g;
  io->stat_picard_iterations = 0;
  io->stat_linear_sweeps = 0;
  if (!sum_sheet(&s)) {
    memset(io->u_vav_cx, 0, s.ncx * sizeof(double));
    memset(io->v_vav_cy, 0, s.ncy * sizeof(double));
    st_free(&s);
    return 0;
  }
  calc_taud(&s);
  calc_basal_conditions(&s);
  determine_grounded_fractions(g, io->TAF_a, s.fg_a, s.fg_cx, s.fg_cy, s.fg_b);
  visc_iter_errors_init(&s, io);
  int it = 0, rc = 0;
  double resid = 0.0;
  /* under the residual rule the state returned is the one whose residual was
   * evaluated: the velocities, basal stress and the viscosity they were
   * computed with (the viscosity of the stop check is its next iterate) */
  double *visc_keep = (c->stop_rule == IMAU_STOP_RESIDUAL) ? dalloc(g->nz * s.na) : NULL;
  for (;;) {
    it++;
    if (visc_keep) memcpy(visc_keep, s.visc3, g->nz * s.na * sizeof(double));
    calc_vertical_shear_strain_rates(&s);
    calc_effective_viscosity(&s, s.u, s.v);
    calc_sliding_term_beta(&s, s.u, s.v);
    calc_F_integral(&s);
    calc_beta_eff(&s);
    if (c->stop_rule == IMAU_STOP_RESIDUAL) {
      double r = csr_rel_residual(&s, s.u, s.v);
      io->stat_nonlinear_residual = r;
      if (r <= c->nonlinear_rtol || it > c->DIVA_visc_it_nit) {
        memcpy(s.visc3, visc_keep, g->nz * s.na * sizeof(double));
        it--;
        break;
      }
    }
    memcpy(s.u_prev, s.u, s.ncx * sizeof(double));
    memcpy(s.v_prev, s.v, s.ncy * sizeof(double));
    rc = solve_linearised(&s, s.u, s.v, lin);
    if (rc) break;
    apply_velocity_limits(&s, s.u, s.v);
    relax(&s, s.u, s.v, c->DIVA_visc_it_relax);
    resid = visc_iter_resid(&s, s.u, s.v);
    int conv = 0;
    if (c->stop_rule == IMAU_STOP_PICARD) {
      if (resid < c->DIVA_visc_it_norm_dUV_tol) conv = 1;
      else if (it >= c->DIVA_visc_it_nit) conv = 1;
    }
    if (!conv) visc_iter_errors(&s, s.u, s.v, io);  /* :355 */
    calc_basal_stress(&s);
    calc_basal_velocities(&s);
    if (conv) break;
  }
  free(visc_keep);
  io->stat_picard_iterations = it;
  io->stat_resid_UV = resid;
  if (rc) { st_free(&s); return rc; }
  size_t nz = g->nz;
  double *u3 = dalloc(nz * s.ncx), *v3 = dalloc(nz * s.ncy), *F1 = dalloc(nz * s.na);
  calc_3D_horizontal_velocities_DIVA(&s, u3, v3, F1);
  memcpy(io->u_vav_cx, s.u, s.ncx * sizeof(double));
  memcpy(io->v_vav_cy, s.v, s.ncy * sizeof(double));
  write_common(&s, io);
  double *uv = dalloc(s.ncx), *vv = dalloc(s.ncy);
  memcpy(uv, s.u, s.ncx * sizeof(double));
  memcpy(vv, s.v, s.ncy * sizeof(double));
  calc_secondary_velocities(g, io, u3, v3, s.u, s.v, s.u_base_cx, s.v_base_cy, uv, vv);
  cp_out(io->u_3D_cx, u3, nz * s.ncx); cp_out(io->v_3D_cy, v3, nz * s.ncy); cp_out(io->F1_3D_a, F1, nz * s.na);
  if (c->stop_rule != IMAU_STOP_RESIDUAL) io->stat_nonlinear_residual = NAN;
  free(u3); free(v3); free(F1); free(uv); free(vv);
  st_free(&s);
  return 0;
}

/* src/ice_velocity_module.f90:115-248 */
int orc_solve_ssa(const imau_grid *gr, const imau_cfg *c, imau_ice_io *io, int lin)
{
  if (c->choice_ice_dynamics != IMAU_DYN_SSA && c->choice_ice_dynamics != IMAU_DYN_SIA_SSA)
    return fail(IMAU_ERR_CONFIG, "solve_SSA: choice_ice_dynamics is not SSA or SIA/SSA");
  if (check_roughness(c)) return IMAU_ERR_CONFIG;
  { int e = check_ismip_e(gr, c, io); if (e) return e; }
  S s;
  st_init(&s, gr, c, io, 1);
  const G *g = &s.g;
  io->stat_picard_iterations = 0;
  if (!sum_sheet(&s) || c->choice_sliding_law == IMAU_SLID_NO_SLIDING) {
    memset(io->u_SSA_cx, 0, s.ncx * sizeof(double));
    memset(io->v_SSA_cy, 0, s.ncy * sizeof(double));
    st_free(&s);
    return 0;
  }
  calc_taud(&s);
  calc_basal_conditions(&s);
  determine_grounded_fractions(g, io->TAF_a, s.fg_a, s.fg_cx, s.fg_cy, s.fg_b);
  /* SSA: no vertical shear (du_dz stays zero) */
  visc_iter_errors_init(&s, io);
  int it = 0, rc = 0;
  double resid = 0.0;
  double *visc_keep = (c->stop_rule == IMAU_STOP_RESIDUAL) ? dalloc(g->nz * s.na) : NULL;
  for (;;) {
    it++;
    if (visc_keep) memcpy(visc_keep, s.visc3, g->nz * s.na * sizeof(double));
    calc_effective_viscosity(&s, s.u, s.v);
    calc_sliding_term_beta(&s, s.u, s.v);
    memcpy(s.beta_eff_a, s.beta_a, s.na * sizeof(double));
    map_a_to_cx(g, s.beta_eff_a, s.beta_eff_cx);
    map_a_to_cy(g, s.beta_eff_a, s.beta_eff_cy);
    for (size_t p = 0; p < s.ncx; ++p) s.beta_eff_cx[p] *= s.fg_cx[p] * s.fg_cx[p];
    for (size_t p = 0; p < s.ncy; ++p) s.beta_eff_cy[p] *= s.fg_cy[p] * s.fg_cy[p];
    if (c->stop_rule == IMAU_STOP_RESIDUAL) {
      double r = csr_rel_residual(&s, s.u, s.v);
      io->stat_nonlinear_residual = r;
      if (r <= c->nonlinear_rtol || it > c->DIVA_visc_it_nit) {
        memcpy(s.visc3, visc_keep, g->nz * s.na * sizeof(double));
        it--;
        break;
      }
    }
    memcpy(s.u_prev, s.u, s.ncx * sizeof(double));
    memcpy(s.v_prev, s.v, s.ncy * sizeof(double));
    rc = solve_linearised(&s, s.u, s.v, lin);
    if (rc) break;
    apply_velocity_limits(&s, s.u, s.v);
    relax(&s, s.u, s.v, c->DIVA_visc_it_relax);
    resid = visc_iter_resid(&s, s.u, s.v);
    if (c->stop_rule == IMAU_STOP_PICARD && (resid < c->DIVA_visc_it_norm_dUV_tol || it >= c->DIVA_visc_it_nit)) break;
    visc_iter_errors(&s, io->u_vav_cx, io->v_vav_cy, io);  /* :238, the test on the state's u_vav */
  }
  free(visc_keep);
  io->stat_picard_iterations = it;
  io->stat_resid_UV = resid;
  if (rc) { st_free(&s); return rc; }
  size_t nz = g->nz;
  memcpy(io->u_SSA_cx, s.u, s.ncx * sizeof(double));
  memcpy(io->v_SSA_cy, s.v, s.ncy * sizeof(double));
  write_common(&s, io);
  double *u3 = dalloc(nz * s.ncx), *v3 = dalloc(nz * s.ncy);
  double *ub = dalloc(s.ncx), *vb = dalloc(s.ncy), *uv = dalloc(s.ncx), *vv = dalloc(s.ncy);
  calc_secondary_velocities(g, io, u3, v3, s.u, s.v, ub, vb, uv, vv);
  cp_out(io->u_3D_cx, u3, nz * s.ncx); cp_out(io->v_3D_cy, v3, nz * s.ncy);
  cp_out(io->u_base_cx, ub, s.ncx); cp_out(io->v_base_cy, vb, s.ncy);
  cp_out(io->u_vav_cx, uv, s.ncx); cp_out(io->v_vav_cy, vv, s.ncy);
  free(u3); free(v3); free(ub); free(vb); free(uv); free(vv);
  st_free(&s);
  return 0;
}

/* src/ice_velocity_module.f90:36-114 */
int orc_solve_sia(const imau_grid *gr, const imau_cfg *c, imau_ice_io *io)
{
  if (c->choice_ice_dynamics != IMAU_DYN_SIA && c->choice_ice_dynamics != IMAU_DYN_SIA_SSA)
    return fail(IMAU_ERR_CONFIG, "solve_SIA: choice_ice_dynamics is not SIA or SIA/SSA");
  G gg = {gr->nx, gr->ny, c->nz, gr->dx, gr->x, gr->y, c};
  const G *g = &gg;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const double n = c->n_flow;
  double f[IMAU_NZ_MAX], D[IMAU_NZ_MAX];
  const int32_t *ms = io->mask_sheet_a;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      if (i < nx - 1) {
        size_t p = A2(j, i);
        if (ms[A2(j, i)] == 0 && ms[A2(j, i + 1)] == 0) {
          for (int k = 0; k < nz; ++k) io->u_3D_SIA_cx[p * nz + k] = 0.0;
        } else {
          double D0 = pow(ICE_DENSITY * GRAV * io->Hi_cx[p], n) *
                      pow(io->dHs_dx_cx[p] * io->dHs_dx_cx[p] + io->dHs_dy_cx[p] * io->dHs_dy_cx[p], (n - 1.0) / 2.0);
          for (int k = 0; k < nz; ++k) f[k] = io->A_flow_3D_cx[p * nz + k] * pow(c->zeta[k], n);
          orc_vertical_integration_from_bottom_to_zeta(nz, c->zeta, f, D);
          for (int k = 0; k < nz; ++k) {
            double Dd = 2.0 * io->Hi_cx[p] * D[k];
            io->u_3D_SIA_cx[p * nz + k] = fmax(D0 * Dd, -1e5) * io->dHs_dx_cx[p];
          }
        }
      }
      if (j < ny - 1) {
        size_t p = CY(j, i);
        if (ms[A2(j, i)] == 0 && ms[A2(j + 1, i)] == 0) {
          for (int k = 0; k < nz; ++k) io->v_3D_SIA_cy[p * nz + k] = 0.0;
        } else {
          double D0 = pow(ICE_DENSITY * GRAV * io->Hi_cy[p], n) *
                      pow(io->dHs_dx_cy[p] * io->dHs_dx_cy[p] + io->dHs_dy_cy[p] * io->dHs_dy_cy[p], (n - 1.0) / 2.0);
          for (int k = 0; k < nz; ++k) f[k] = io->A_flow_3D_cy[p * nz + k] * pow(c->zeta[k], n);
          orc_vertical_integration_from_bottom_to_zeta(nz, c->zeta, f, D);
          for (int k = 0; k < nz; ++k) {
            double Dd = 2.0 * io->Hi_cy[p] * D[k];
            io->v_3D_SIA_cy[p * nz + k] = fmax(D0 * Dd, -1e5) * io->dHs_dy_cy[p];
          }
        }
      }
    }
  {
    /* solve_SIA always finishes with calc_secondary_velocities (:109); in
     * SIA/SSA mode that combines the new SIA columns with the current u_SSA */
    size_t ncx = (size_t)ny * (nx - 1), ncy = (size_t)(ny - 1) * nx;
    double *u3 = dalloc(ncx * nz), *v3 = dalloc(ncy * nz), *ub = dalloc(ncx), *vb = dalloc(ncy);
    double *uv = dalloc(ncx), *vv = dalloc(ncy);
    calc_secondary_velocities(g, io, u3, v3, io->u_SSA_cx, io->v_SSA_cy, ub, vb, uv, vv);
    cp_out(io->u_3D_cx, u3, nz * ncx); cp_out(io->v_3D_cy, v3, nz * ncy);
    cp_out(io->u_base_cx, ub, ncx); cp_out(io->v_base_cy, vb, ncy);
    cp_out(io->u_vav_cx, uv, ncx); cp_out(io->v_vav_cy, vv, ncy);
    free(u3); free(v3); free(ub); free(vb); free(uv); free(vv);
  }
  return 0;
}

/* FP64 relative nonlinear residual at the io state (see header). */
int orc_nonlinear_residual(const imau_grid *gr, const imau_cfg *c, const imau_ice_io *io, double *res)
{
  int ssa = c->choice_ice_dynamics != IMAU_DYN_DIVA;
  S s;
  st_init(&s, gr, c, io, ssa);
  const G *g = &s.g;
  calc_taud(&s);
  calc_basal_conditions(&s);
  determine_grounded_fractions(g, io->TAF_a, s.fg_a, s.fg_cx, s.fg_cy, s.fg_b);
  if (!ssa) {
    calc_vertical_shear_strain_rates(&s);
    calc_effective_viscosity(&s, s.u, s.v);
    calc_sliding_term_beta(&s, s.u, s.v);
    calc_F_integral(&s);
    calc_beta_eff(&s);
  } else {
    calc_effective_viscosity(&s, s.u, s.v);
    calc_sliding_term_beta(&s, s.u, s.v);
    memcpy(s.beta_eff_a, s.beta_a, s.na * sizeof(double));
    map_a_to_cx(g, s.beta_eff_a, s.beta_eff_cx);
    map_a_to_cy(g, s.beta_eff_a, s.beta_eff_cy);
    for (size_t p = 0; p < s.ncx; ++p) s.beta_eff_cx[p] *= s.fg_cx[p] * s.fg_cx[p];
    for (size_t p = 0; p < s.ncy; ++p) s.beta_eff_cy[p] *= s.fg_cy[p] * s.fg_cy[p];
  }
  *res = csr_rel_residual(&s, s.u, s.v);
  st_free(&s);
  return 0;
}

/* CSR of the system at the coefficients held in io (N_a, N_b, N_cx, N_cy,
 * beta_eff_cx, beta_eff_cy, taudx_cx, taudy_cy), for unit tests. */
int orc_assemble_csr(const imau_grid *gr, const imau_cfg *c, const imau_ice_io *io,
                     const double *u_cx, const double *v_cy,
                     int32_t *ptr, int32_t *idx, double *val, double *b, double *x)
{
  S s;
  st_init(&s, gr, c, io, 0);
  cp_in(s.N_a, io->N_a, s.na); cp_in(s.N_b, io->N_b, s.nb);
  cp_in(s.N_cx, io->N_cx, s.ncx); cp_in(s.N_cy, io->N_cy, s.ncy);
  cp_in(s.beta_eff_cx, io->beta_eff_cx, s.ncx); cp_in(s.beta_eff_cy, io->beta_eff_cy, s.ncy);
  cp_in(s.taudx, io->taudx_cx, s.ncx); cp_in(s.taudy, io->taudy_cy, s.ncy);
  set_sico_masks(&s);
  ASM a;
  asm_setup(&a, &s, u_cx, v_cy);
  a.ptr = ptr; a.idx = idx; a.val = val; a.b = b; a.x = x;
  int nnz = assemble(&a);
  st_free(&s);
  return nnz;
}

/* sizes and a few field offsets of the ABI structs as this C compiler lays
 * them out (lets the tests check the Python mirror of include/imau_vel.h) */
#include <stddef.h>
void orc_abi_layout(int64_t *out)
{
  out[0] = (int64_t)sizeof(imau_grid);
  out[1] = (int64_t)sizeof(imau_cfg);
  out[2] = (int64_t)sizeof(imau_ice_io);
  out[3] = (int64_t)offsetof(imau_cfg, zeta);
  out[4] = (int64_t)offsetof(imau_cfg, check_interval);
  out[5] = (int64_t)offsetof(imau_ice_io, u_vav_cx);
  out[6] = (int64_t)offsetof(imau_ice_io, DIVA_SOR_nit);
  out[7] = (int64_t)offsetof(imau_ice_io, stat_seconds);
}

/* sizes of the other ABI structs and the offset of their last member, in the
 * order thickness cfg / io, producer cfg / io, time-step cfg / state, refgeo io */
void orc_abi_layout2(int64_t *out)
{
  out[0] = (int64_t)sizeof(imau_thickness_cfg);
  out[1] = (int64_t)sizeof(imau_thickness_io);
  out[2] = (int64_t)sizeof(imau_producer_cfg);
  out[3] = (int64_t)sizeof(imau_producer_io);
  out[4] = (int64_t)sizeof(imau_timestep_cfg);
  out[5] = (int64_t)sizeof(imau_timestep_state);
  out[6] = (int64_t)sizeof(imau_refgeo_io);
  out[7] = (int64_t)offsetof(imau_timestep_cfg, continental_shelf_min_height);
  out[8] = (int64_t)offsetof(imau_timestep_state, nsteps);
  out[9] = (int64_t)offsetof(imau_refgeo_io, Hb_GIAeq);
  out[10] = (int64_t)offsetof(imau_timestep_cfg, choice_calving_law);
}

/* ------------------------------------------------------------------------ */
/* ice-thickness update, src/ice_thickness_module.f90:25-909 and            */
/* calc_critical_timestep_adv, src/ice_dynamics_module.f90:726-762          */
/* ------------------------------------------------------------------------ */
static double hi_from_hb_and_hs(double Hb, double Hs, double SL)
{ /* src/utilities_module.f90:160-183 */
  double Hs_float = Hb + (SL - Hb) * SEAWATER_DENSITY / ICE_DENSITY;
  if (Hs > Hs_float) return Hs - Hb;
  return (Hs - SL) / (1.0 - ICE_DENSITY / SEAWATER_DENSITY);
}

/* net mass balance of a cell (m/yr): only the floating fraction at the
 * calving front, none on ice-free ocean (:219-227, :393-401) */
static double mb_net(const imau_thickness_io *io, size_t p)
{
  double mb = io->SMB_year[p] + io->BMB[p];
  if (io->mask_cf_a[p] == 1 && io->mask_shelf_a[p] == 1) return mb * io->float_margin_frac_a[p];
  if (io->mask_ocean_a[p] == 1 && io->mask_shelf_a[p] == 0) return 0.0;
  return mb;
}

/* src/ice_thickness_module.f90:59-312 */
static void dhi_explicit(const imau_grid *g, const imau_thickness_io *io, double dt, double *Qx, double *Qy,
                         double *dHi, double *Ht)
{
  const int nx = g->nx, ny = g->ny;
  const double dx = g->dx;
  size_t na = (size_t)nx * ny;
  for (int i = 0; i < nx - 1; ++i)
    for (int j = 0; j < ny; ++j) {
      double u = io->u_vav_cx[(size_t)i * ny + j];
      Qx[(size_t)i * ny + j] = u * io->Hi_a[(size_t)(u > 0.0 ? i : i + 1) * ny + j] * dx * dt;
    }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny - 1; ++j) {
      double v = io->v_vav_cy[(size_t)i * (ny - 1) + j];
      Qy[(size_t)i * (ny - 1) + j] = v * io->Hi_a[(size_t)i * ny + (v > 0.0 ? j : j + 1)] * dx * dt;
    }
  double *dVmb = dalloc(na);
#define QX(jj, ii) Qx[(size_t)(ii) * ny + (jj)]
#define QY(jj, ii) Qy[(size_t)(ii) * (ny - 1) + (jj)]
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      size_t p = (size_t)i * ny + j;
      dVmb[p] = mb_net(io, p) * dx * dx * dt;
      double Vav = io->Hi_a[p] * dx * dx, out = 0.0;
      if (i > 0 && QX(j, i - 1) < 0.0) out -= QX(j, i - 1);
      if (i < nx - 1 && QX(j, i) > 0.0) out += QX(j, i);
      if (j > 0 && QY(j - 1, i) < 0.0) out -= QY(j - 1, i);
      if (j < ny - 1 && QY(j, i) > 0.0) out += QY(j, i);
      double rescale = 1.0;
      if (-dVmb[p] >= Vav) { dVmb[p] = -Vav; rescale = 0.0; }
      if (out > Vav + dVmb[p]) rescale = (Vav + dVmb[p]) / out;
      if (rescale < 1.0) {
        if (i > 0 && QX(j, i - 1) < 0.0) QX(j, i - 1) *= rescale;
        if (i < nx - 1 && QX(j, i) > 0.0) QX(j, i) *= rescale;
        if (j > 0 && QY(j - 1, i) < 0.0) QY(j - 1, i) *= rescale;
        if (j < ny - 1 && QY(j, i) > 0.0) QY(j, i) *= rescale;
      }
    }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      size_t p = (size_t)i * ny + j;
      double a = dx * dx * dt, d = dVmb[p] / a;
      if (i > 0) d += QX(j, i - 1) / a;
      if (i < nx - 1) d -= QX(j, i) / a;
      if (j > 0) d += QY(j - 1, i) / a;
      if (j < ny - 1) d -= QY(j, i) / a;
      d = fmax(d, -io->Hi_a[p] / dt);
      dHi[p] = d;
      Ht[p] = io->Hi_a[p] + d * dt;
    }
#undef QX
#undef QY
  free(dVmb);
}

/* velocity through a cell face, zero across a calving front whose floating
 * cell is not yet full (:474-540, :574-640) */
static double face_vel(const imau_grid *g, const imau_thickness_io *io, int j, int i, int jn, int in, double vel)
{
  const int ny = g->ny;
  size_t p = (size_t)i * ny + j, q = (size_t)in * ny + jn;
  if (io->mask_shelf_a[p] == 1 && io->mask_ocean_a[q] == 1 && io->mask_ice_a[q] == 0) {
    if (io->float_margin_frac_a[p] < 1.0) return 0.0;
  } else if (io->mask_shelf_a[q] == 1 && io->mask_ocean_a[p] == 1 && io->mask_ice_a[p] == 0) {
    if (io->float_margin_frac_a[q] < 1.0) return 0.0;
  }
  return vel;
}

/* src/ice_thickness_module.f90:313-718: CSR in the (i outer, j inner) order,
 * fs = 2.5; solved with the exact banded LU (standing in for SOR / PETSc) */
static int dhi_semiimplicit(const imau_grid *g, const imau_thickness_io *io, double dt, double *dHi, double *Ht)
{
  const int nx = g->nx, ny = g->ny;
  const double dx = g->dx, fs = 2.5;
  const int m = nx * ny;
  int32_t *ptr = ialloc(m + 1), *idx = ialloc(5 * (size_t)m);
  double *val = dalloc(5 * (size_t)m), *b = dalloc(m), *x = dalloc(m);
  int k = 0;
  const int32_t *mi = io->mask_ice_a, *ms = io->mask_shelf_a, *mo = io->mask_ocean_a;
  const double *H = io->Hi_a;
#define P2(jj, ii) ((size_t)(ii) * ny + (jj))
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      int n = (int)P2(j, i);
      ptr[n] = k;
      int im1 = i > 0 ? i - 1 : 0, ip1 = i < nx - 1 ? i + 1 : nx - 1;
      int jm1 = j > 0 ? j - 1 : 0, jp1 = j < ny - 1 ? j + 1 : ny - 1;
      int scheme; /* 0 MB-only, 1 explicit, 2 semi-implicit (:353-385) */
      if (mi[P2(j, i)] == 0 && mi[P2(j, im1)] == 0 && mi[P2(j, ip1)] == 0 && mi[P2(jm1, i)] == 0 && mi[P2(jp1, i)] == 0)
        scheme = 0;
      else if (ms[P2(j, i)] == 1 &&
               ((mo[P2(j, im1)] == 1 && mi[P2(j, im1)] == 0) || (mo[P2(j, ip1)] == 1 && mi[P2(j, ip1)] == 0) ||
                (mo[P2(jm1, i)] == 1 && mi[P2(jm1, i)] == 0) || (mo[P2(jp1, i)] == 1 && mi[P2(jp1, i)] == 0)))
        scheme = 1;
      else if (mo[P2(j, i)] == 1 && mi[P2(j, i)] == 0 &&
               (ms[P2(j, im1)] == 1 || ms[P2(j, ip1)] == 1 || ms[P2(jm1, i)] == 1 || ms[P2(jp1, i)] == 1))
        scheme = 1;
      else
        scheme = 2;
      double mb = mb_net(io, P2(j, i));
      double uw = 0.0, ue = 0.0, vs = 0.0, vn = 0.0;
      if (i > 0) uw = face_vel(g, io, j, i, j, i - 1, io->u_vav_cx[P2(j, i - 1)]);
      if (i < nx - 1) ue = face_vel(g, io, j, i, j, i + 1, io->u_vav_cx[P2(j, i)]);
      if (j > 0) vs = face_vel(g, io, j, i, j - 1, i, io->v_vav_cy[(size_t)i * (ny - 1) + j - 1]);
      if (j < ny - 1) vn = face_vel(g, io, j, i, j + 1, i, io->v_vav_cy[(size_t)i * (ny - 1) + j]);
      if (scheme == 0) {
        double h = fmax(0.0, H[P2(j, i)] + mb * dt);
        idx[k] = n; val[k++] = 1.0; b[n] = h; x[n] = h;
      } else if (scheme == 1) {
        /* upwind fluxes (m/yr) with the face velocities (:441-560) */
        double Qw = 0.0, Qe = 0.0, Qs = 0.0, Qn = 0.0;
        if (i > 0) Qw = uw * H[P2(j, uw > 0.0 ? i - 1 : i)] / dx;
        if (i < nx - 1) Qe = -ue * H[P2(j, ue > 0.0 ? i : i + 1)] / dx;
        if (j > 0) Qs = vs * H[P2(vs > 0.0 ? j - 1 : j, i)] / dx;
        if (j < ny - 1) Qn = -vn * H[P2(vn > 0.0 ? j : j + 1, i)] / dx;
        double h = fmax(0.0, H[P2(j, i)] + dt * (mb + Qw + Qe + Qs + Qn));
        idx[k] = n; val[k++] = 1.0; b[n] = h; x[n] = h;
      } else {
        double d = 1.0 / dt;
        if (i > 0) d -= (fs / dx) * fmin(uw, 0.0);
        if (i < nx - 1) d += (fs / dx) * fmax(ue, 0.0);
        if (j > 0) d -= (fs / dx) * fmin(vs, 0.0);
        if (j < ny - 1) d += (fs / dx) * fmax(vn, 0.0);
        idx[k] = n; val[k++] = d;
        if (i > 0) { idx[k] = (int)P2(j, i - 1); val[k++] = (-fs / dx) * fmax(uw, 0.0); }
        if (i < nx - 1) { idx[k] = (int)P2(j, i + 1); val[k++] = (fs / dx) * fmin(ue, 0.0); }
        if (j > 0) { idx[k] = (int)P2(j - 1, i); val[k++] = (-fs / dx) * fmax(vs, 0.0); }
        if (j < ny - 1) { idx[k] = (int)P2(j + 1, i); val[k++] = (fs / dx) * fmin(vn, 0.0); }
        double bb = mb + H[P2(j, i)] / dt;
        if (i > 0) bb += H[P2(j, i)] * (1.0 - fs) / dx * fmin(uw, 0.0) + H[P2(j, i - 1)] * (1.0 - fs) / dx * fmax(uw, 0.0);
        if (i < nx - 1) bb -= H[P2(j, i)] * (1.0 - fs) / dx * fmax(ue, 0.0) + H[P2(j, i + 1)] * (1.0 - fs) / dx * fmin(ue, 0.0);
        if (j > 0) bb += H[P2(j, i)] * (1.0 - fs) / dx * fmin(vs, 0.0) + H[P2(j - 1, i)] * (1.0 - fs) / dx * fmax(vs, 0.0);
        if (j < ny - 1) bb -= H[P2(j, i)] * (1.0 - fs) / dx * fmax(vn, 0.0) + H[P2(j + 1, i)] * (1.0 - fs) / dx * fmin(vn, 0.0);
        b[n] = bb;
        x[n] = H[P2(j, i)];
      }
    }
  ptr[m] = k;
#undef P2
  int rc = orc_banded_lu_solve(m, ptr, idx, val, b, x);
  if (!rc)
    for (int n = 0; n < m; ++n) {
      Ht[n] = fmax(0.0, x[n]);
      dHi[n] = (Ht[n] - H[n]) / dt;
    }
  free(ptr); free(idx); free(val); free(b); free(x);
  return rc;
}

/* src/ice_thickness_module.f90:719-909, sides in the reference's order */
static void thickness_bc(const imau_grid *g, const imau_thickness_cfg *c, const imau_thickness_io *io, double dt,
                         double *dHi, double *Ht)
{
  const int nx = g->nx, ny = g->ny;
  const int diva = c->choice_ice_dynamics == IMAU_DYN_DIVA;
  const double *H = io->Hi_a, *Hb = io->Hb_a, *SL = io->SL_a;
#define P2(jj, ii) ((size_t)(ii) * ny + (jj))
  for (int side = 0; side < 4; ++side) {
    int bc = (&c->ice_thickness_west_BC)[side];
    int along = (side < 2) ? ny : nx;
    for (int q = 0; q < along; ++q) {
      /* cell on the boundary, its inner neighbour, the second row in, the periodic source */
      int jb, ib, jn, in, j2, i2, jp, ip;
      if (side == 0) { jb = q; ib = 0; jn = q; in = 1; j2 = q; i2 = 1; jp = q; ip = nx - 2; }
      else if (side == 1) { jb = q; ib = nx - 1; jn = q; in = nx - 2; j2 = q; i2 = nx - 2; jp = q; ip = 1; }
      else if (side == 2) { jb = 0; ib = q; jn = 1; in = q; j2 = 1; i2 = q; jp = ny - 2; ip = q; }
      else { jb = ny - 1; ib = q; jn = ny - 2; in = q; j2 = ny - 2; i2 = q; jp = 1; ip = q; }
      size_t p = P2(jb, ib);
      if (bc == IMAU_HBC_ZERO) {
        Ht[p] = 0.0; dHi[p] = -H[p] / dt;
        if (diva) { size_t p2 = P2(j2, i2); Ht[p2] = 0.0; dHi[p2] = -H[p2] / dt; }
      } else if (bc == IMAU_HBC_INFINITE) {
        size_t pn = P2(jn, in);
        double Hs = surface_elevation(Ht[pn], Hb[pn], SL[pn]);
        Ht[p] = hi_from_hb_and_hs(Hb[p], Hs, SL[p]);
        dHi[p] = (Ht[p] - H[p]) / dt;
      } else if (bc == IMAU_HBC_PERIODIC) {
        Ht[p] = H[P2(jp, ip)];
        dHi[p] = (Ht[p] - H[p]) / dt;
      } else if (bc == IMAU_HBC_FIXED) {
        Ht[p] = H[p]; dHi[p] = 0.0;
      }
    }
  }
#undef P2
}

int orc_thickness_update(const imau_grid *g, const imau_thickness_cfg *c, imau_thickness_io *io, double dt)
{
  const int nx = g->nx, ny = g->ny;
  size_t na = (size_t)nx * ny;
  double *Qx = dalloc((size_t)ny * (nx - 1)), *Qy = dalloc((size_t)(ny - 1) * nx);
  int rc = 0;
  if (c->choice_ice_integration_method == IMAU_DHI_NONE) {
    for (size_t p = 0; p < na; ++p) { io->dHi_dt_a[p] = 0.0; io->Hi_tplusdt_a[p] = io->Hi_a[p]; }
  } else if (c->choice_ice_integration_method == IMAU_DHI_EXPLICIT) {
    dhi_explicit(g, io, dt, Qx, Qy, io->dHi_dt_a, io->Hi_tplusdt_a);
    cp_out(io->Qx_cx, Qx, (size_t)ny * (nx - 1));
    cp_out(io->Qy_cy, Qy, (size_t)(ny - 1) * nx);
  } else if (c->choice_ice_integration_method == IMAU_DHI_SEMIIMPLICIT) {
    rc = dhi_semiimplicit(g, io, dt, io->dHi_dt_a, io->Hi_tplusdt_a);
  } else {
    rc = fail(IMAU_ERR_CONFIG, "unknown choice_ice_integration_method");
  }
  if (!rc) thickness_bc(g, c, io, dt, io->dHi_dt_a, io->Hi_tplusdt_a);
  free(Qx); free(Qy);
  return rc;
}

double orc_critical_timestep_adv(const imau_grid *g, const double *u_cx, const double *v_cy, double dt_max)
{
  const int nx = g->nx, ny = g->ny;
  double dt_crit = 2.0 * dt_max;
  for (int i = 1; i < nx - 1; ++i)
    for (int j = 1; j < ny - 1; ++j) {
      double dt = g->dx / (fabs(u_cx[(size_t)i * ny + j]) + fabs(v_cy[(size_t)i * (ny - 1) + j]));
      dt_crit = fmin(dt_crit, dt);
    }
  return fmin(dt_max, dt_crit * 0.9);
}

/* working arrays of the producer fields (one grid) */
typedef struct {
  double *Hs, *TAF, *dHs_dt, *Hi_cx, *Hi_cy, *Hi_b, *A3cx, *A3cy, *Avcx, *Avcy;
  double *dHi_dx_a, *dHi_dy_a, *dHs_dx_a, *dHs_dy_a, *dHs_dx_cx, *dHs_dy_cx, *dHs_dx_cy, *dHs_dy_cy;
  int32_t *land, *ocean, *ice, *sheet, *shelf, *coast, *margin, *gl, *cf, *total;
} ProducerBufs_;

static void producer_bufs_alloc(ProducerBufs_ *b, size_t na, size_t ncx, size_t ncy, int nz)
{
  b->Hs = dalloc(na); b->TAF = dalloc(na); b->dHs_dt = dalloc(na);
  b->Hi_cx = dalloc(ncx); b->Hi_cy = dalloc(ncy); b->Hi_b = dalloc(na);
  b->A3cx = dalloc(ncx * nz); b->A3cy = dalloc(ncy * nz); b->Avcx = dalloc(ncx); b->Avcy = dalloc(ncy);
  b->dHi_dx_a = dalloc(na); b->dHi_dy_a = dalloc(na); b->dHs_dx_a = dalloc(na); b->dHs_dy_a = dalloc(na);
  b->dHs_dx_cx = dalloc(ncx); b->dHs_dy_cx = dalloc(ncx); b->dHs_dx_cy = dalloc(ncy); b->dHs_dy_cy = dalloc(ncy);
  b->land = ialloc(na); b->ocean = ialloc(na); b->ice = ialloc(na); b->sheet = ialloc(na); b->shelf = ialloc(na);
  b->coast = ialloc(na); b->margin = ialloc(na); b->gl = ialloc(na); b->cf = ialloc(na); b->total = ialloc(na);
}
static void producer_bufs_free(ProducerBufs_ *b)
{
  double *d[] = {b->Hs, b->TAF, b->dHs_dt, b->Hi_cx, b->Hi_cy, b->Hi_b, b->A3cx, b->A3cy, b->Avcx, b->Avcy,
                 b->dHi_dx_a, b->dHi_dy_a, b->dHs_dx_a, b->dHs_dy_a, b->dHs_dx_cx, b->dHs_dy_cx, b->dHs_dx_cy,
                 b->dHs_dy_cy};
  int32_t *m[] = {b->land, b->ocean, b->ice, b->sheet, b->shelf, b->coast, b->margin, b->gl, b->cf, b->total};
  for (size_t q = 0; q < sizeof d / sizeof d[0]; ++q) free(d[q]);
  for (size_t q = 0; q < sizeof m / sizeof m[0]; ++q) free(m[q]);
}
static void producer_io_bind(imau_producer_io *o, const ProducerBufs_ *b, const double *Hi, const double *Hb,
                             const double *SL, double *A3, double *Av)
{
  memset(o, 0, sizeof *o);
  o->Hi_a = Hi; o->Hb_a = Hb; o->SL_a = SL;
  o->Hs_a = b->Hs; o->TAF_a = b->TAF; o->dHs_dt_a = b->dHs_dt;
  o->mask_land_a = b->land; o->mask_ocean_a = b->ocean; o->mask_ice_a = b->ice; o->mask_sheet_a = b->sheet;
  o->mask_shelf_a = b->shelf; o->mask_coast_a = b->coast; o->mask_margin_a = b->margin; o->mask_gl_a = b->gl;
  o->mask_cf_a = b->cf; o->mask_a = b->total;
  o->Hi_cx = b->Hi_cx; o->Hi_cy = b->Hi_cy; o->Hi_b = b->Hi_b;
  o->A_flow_3D_a = A3; o->A_flow_vav_a = Av; o->A_flow_3D_cx = b->A3cx; o->A_flow_3D_cy = b->A3cy;
  o->A_flow_vav_cx = b->Avcx; o->A_flow_vav_cy = b->Avcy;
  o->dHi_dx_a = b->dHi_dx_a; o->dHi_dy_a = b->dHi_dy_a; o->dHs_dx_a = b->dHs_dx_a; o->dHs_dy_a = b->dHs_dy_a;
  o->dHs_dx_cx = b->dHs_dx_cx; o->dHs_dy_cx = b->dHs_dy_cx; o->dHs_dx_cy = b->dHs_dx_cy; o->dHs_dy_cy = b->dHs_dy_cy;
}
/* the fields of geo that are not NULL, from the working set */
static void producer_io_copy_out(imau_producer_io *geo, const imau_producer_io *w, size_t na, size_t ncx, size_t ncy,
                                 size_t nb, int nz)
{
  const size_t d = sizeof(double), q = sizeof(int32_t);
#define CPO(f, n, s) if (geo->f && w->f) memcpy(geo->f, w->f, (n) * (s))
  CPO(Hs_a, na, d); CPO(TAF_a, na, d); CPO(dHs_dt_a, na, d);
  CPO(mask_land_a, na, q); CPO(mask_ocean_a, na, q); CPO(mask_ice_a, na, q); CPO(mask_sheet_a, na, q);
  CPO(mask_shelf_a, na, q); CPO(mask_coast_a, na, q); CPO(mask_margin_a, na, q); CPO(mask_gl_a, na, q);
  CPO(mask_cf_a, na, q); CPO(mask_a, na, q);
  CPO(Hi_cx, ncx, d); CPO(Hi_cy, ncy, d); CPO(Hi_b, nb, d);
  CPO(A_flow_3D_a, na * nz, d); CPO(A_flow_vav_a, na, d); CPO(A_flow_3D_cx, ncx * nz, d);
  CPO(A_flow_3D_cy, ncy * nz, d); CPO(A_flow_vav_cx, ncx, d); CPO(A_flow_vav_cy, ncy, d);
  CPO(dHi_dx_a, na, d); CPO(dHi_dy_a, na, d); CPO(dHs_dx_a, na, d); CPO(dHs_dy_a, na, d);
  CPO(dHs_dx_cx, ncx, d); CPO(dHs_dy_cx, ncx, d); CPO(dHs_dx_cy, ncy, d); CPO(dHs_dy_cy, ncy, d);
#undef CPO
}

/* ------------------------------------------------------------------------ */
/* time stepping with choice_timestepping = 'direct':                       */
/* run_ice_dynamics_direct (src/ice_dynamics_module.f90:63-242) and         */
/* update_ice_thickness (:435-648)                                          */
/* ------------------------------------------------------------------------ */
/* calc_critical_timestep_SIA, :651-725 */
static double crit_sia(const G *g, const double *dt_max, const double *u3sia, const double *v3sia, const double *u3,
                       const double *v3, const double *Hi_cx, const double *Hi_cy, const double *dHs_dx_cx,
                       const double *dHs_dy_cy, const int32_t *sheet, const double *zeta)
{
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  double dtc = *dt_max;
  double *col = dalloc(nz);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      double D = 1e-9;
      if (i < nx - 1 && sheet[A2(j, i)] == 1 && sheet[A2(j, i + 1)] == 1) {
        for (int k = 0; k < nz; ++k) col[k] = u3sia[A3I(k, j, i)];
        const double u = orc_vertical_average(nz, zeta, col);
        D = fmax(D, fabs(u * Hi_cx[A2(j, i)] / dHs_dx_cx[A2(j, i)]));
      }
      if (j < ny - 1 && sheet[A2(j, i)] == 1 && sheet[A2(j + 1, i)] == 1) {
        for (int k = 0; k < nz; ++k) col[k] = v3sia[CY3(k, j, i)];
        const double v = orc_vertical_average(nz, zeta, col);
        D = fmax(D, fabs(v * Hi_cy[CY(j, i)] / dHs_dy_cy[CY(j, i)]));
      }
      dtc = fmin(dtc, g->dx * g->dx / (6.0 * D));
      for (int k = 0; k < nz; ++k) {
        double ua, va;
        if (i == 0) ua = fabs(u3[A3I(k, j, 0)]);
        else if (i == nx - 1) ua = fabs(u3[A3I(k, j, nx - 2)]);
        else ua = fmax(fabs(u3[A3I(k, j, i - 1)]), fabs(u3[A3I(k, j, i)]));
        if (j == 0) va = fabs(v3[CY3(k, 0, i)]);
        else if (j == ny - 1) va = fabs(v3[CY3(k, ny - 2, i)]);
        else va = fmax(fabs(v3[CY3(k, j - 1, i)]), fabs(v3[CY3(k, j, i)]));
        dtc = fmin(dtc, g->dx / (ua + va));
      }
    }
  free(col);
  return dtc * 0.9;
}

/* determine_floating_margin_fraction, src/general_ice_model_data_module.f90:784-865 */
/* (with the effective front thickness Hi_eff_cf_a when heff is not NULL) */
static void floating_margin_fraction(const G *g, const double *Hi, const int32_t *ice, const int32_t *cf,
                                     const int32_t *shelf, double *fmf, double *heff)
{
  const int nx = g->nx, ny = g->ny;
  for (int i = 0; i < nx; ++i)
    for (int j = 1; j < ny - 1; ++j) {
      const size_t p = A2(j, i);
      double f = ice[p] == 1 ? 1.0 : 0.0, h = ice[p] == 1 ? Hi[p] : 0.0;
      if (cf[p] == 1 && shelf[p] == 1) {
        int noncf = 0;
        double hmax = 0.0;
        for (int ii = (i > 0 ? i - 1 : 0); ii <= (i < nx - 1 ? i + 1 : nx - 1); ++ii)
          for (int jj = j - 1; jj <= j + 1; ++jj)
            if (ice[A2(jj, ii)] == 1 && cf[A2(jj, ii)] == 0) {
              noncf = 1;
              hmax = fmax(hmax, Hi[A2(jj, ii)]);
            }
        if (!noncf || hmax < Hi[p]) {
          f = 1.0;
          h = Hi[p];
        } else {
          f = Hi[p] / hmax;
          h = hmax;
        }
      }
      fmf[p] = f;
      if (heff) heff[p] = h;
    }
}

/* the fixed-geometry switches of update_ice_thickness,
 * src/ice_dynamics_module.f90:460-559: H(t+dt) = H(t) on the shelf and its
 * grounding line, on the sheet and its grounding line, or on the grounding
 * line only (masks of the state at t) */
static int any3(const G *g, const int32_t *m, int j, int i)
{
  for (int ii = (i > 0 ? i - 1 : 0); ii <= (i < g->nx - 1 ? i + 1 : g->nx - 1); ++ii)
    for (int jj = (j > 0 ? j - 1 : 0); jj <= (j < g->ny - 1 ? j + 1 : g->ny - 1); ++jj)
      if (m[(size_t)ii * g->ny + jj] == 1) return 1;
  return 0;
}
static void fix_geometry(const G *g, const imau_timestep_cfg *sc, const double *Hi, const int32_t *sheet,
                         const int32_t *shelf, double *Ht)
{
  for (int i = 0; i < g->nx; ++i)
    for (int j = 0; j < g->ny; ++j) {
      const size_t p = (size_t)i * g->ny + j;
      const int gl_sheet = sheet[p] == 1 && any3(g, shelf, j, i);  /* sheet cell next to the shelf */
      const int gl_shelf = shelf[p] == 1 && any3(g, sheet, j, i);  /* shelf cell next to the sheet */
      int fix = 0;
      if (sc->fixed_shelf_geometry && (shelf[p] == 1 || gl_sheet)) fix = 1;
      if (sc->fixed_sheet_geometry && (sheet[p] == 1 || gl_shelf)) fix = 1;
      if (sc->fixed_grounding_line && (gl_sheet || gl_shelf)) fix = 1;
      if (fix) Ht[p] = Hi[p];
    }
}

/* threshold_thickness_calving, src/calving_module.f90:62-91 */
static void threshold_calving(size_t na, double *Hi, const int32_t *cf, const double *heff, double thr)
{
  for (size_t p = 0; p < na; ++p)
    if (cf[p] == 1 && heff[p] < thr) Hi[p] = 0.0;
}

/* the ice removals of update_ice_thickness after the calving,
 * src/ice_dynamics_module.f90:590-635 */
static void remove_ice(size_t na, const imau_timestep_cfg *sc, const imau_refgeo_io *rg, const double *Hb,
                       const double *SL, double *Hi)
{
  for (size_t p = 0; p < na; ++p) {
    if (rg && rg->mask_noice_a && rg->mask_noice_a[p] == 1) Hi[p] = 0.0;
    if (sc->do_remove_shelves && is_floating(Hi[p], Hb[p], SL[p])) Hi[p] = 0.0;
    if (sc->remove_shelves_larger_than_PD && rg->Hi_PD[p] == 0.0 && rg->Hb_PD[p] < 0.0) Hi[p] = 0.0;
    if (sc->continental_shelf_calving && rg->Hi_GIAeq[p] == 0.0 && rg->Hb_GIAeq[p] < sc->continental_shelf_min_height)
      Hi[p] = 0.0;
  }
}

/* remove_unconnected_shelves, src/ice_thickness_module.f90:932-1044: a
 * depth-first fill through the 4 neighbours from the shelf cells next to the
 * sheet; shelf cells it does not reach lose their ice */
static void remove_unconnected_shelves(const G *g, double *Hi, const int32_t *shelf, const int32_t *sheet)
{
  const int nx = g->nx, ny = g->ny;
  int32_t *map = ialloc((size_t)nx * ny);
  int32_t *stack = ialloc(2 * (size_t)nx * ny + 8);
  long n = 0;
  for (int i = 1; i < nx - 1; ++i)
    for (int j = 1; j < ny - 1; ++j)
      if (shelf[A2(j, i)] == 1 && (sheet[A2(j, i + 1)] == 1 || sheet[A2(j, i - 1)] == 1 || sheet[A2(j + 1, i)] == 1 ||
                                   sheet[A2(j - 1, i)] == 1)) {
        stack[2 * n] = i; stack[2 * n + 1] = j; ++n;
        map[A2(j, i)] = 1;
      }
  while (n > 0) {
    --n;
    const int i = stack[2 * n], j = stack[2 * n + 1];
    if (shelf[A2(j, i)] != 1) continue;
    map[A2(j, i)] = 2;
    const int ni[4] = {i - 1, i + 1, i, i}, nj[4] = {j, j, j - 1, j + 1};
    for (int q = 0; q < 4; ++q) {
      if (ni[q] < 0 || ni[q] >= nx || nj[q] < 0 || nj[q] >= ny) continue;
      if (map[A2(nj[q], ni[q])] == 0) {
        map[A2(nj[q], ni[q])] = 1;
        stack[2 * n] = ni[q]; stack[2 * n + 1] = nj[q]; ++n;
      }
    }
  }
  for (size_t p = 0; p < (size_t)nx * ny; ++p)
    if (shelf[p] == 1 && map[p] == 0) Hi[p] = 0.0;
  free(map);
  free(stack);
}

/* the time loop of run_ice_dynamics_direct (pcs == NULL) or run_ice_dynamics_pc
 * (src/ice_dynamics_module.f90:63-242, :243-432) */
static int run_loop(const imau_grid *gr, const imau_cfg *c, const imau_thickness_cfg *tc,
                    const imau_producer_cfg *pc, const imau_timestep_cfg *sc, imau_ice_io *io,
                    const imau_refgeo_io *rg, const double *SMB, const double *BMB, double *Hi_end,
                    imau_producer_io *geo, imau_timestep_state *st, imau_pc_state *pcs, double t_end)
{
  const int dyn = c->choice_ice_dynamics;
  if (!pcs && dyn != IMAU_DYN_NONE && dyn != IMAU_DYN_SIA && dyn != IMAU_DYN_SSA && dyn != IMAU_DYN_SIA_SSA)
    return fail(-2, "direct time stepping works with 'none', 'SIA', 'SSA' or 'SIA/SSA' only");
  if (!(sc->dt_startup_phase > 0.0) || !(sc->dt_min > 0.0) || sc->dt_max < sc->dt_min)
    return fail(-1, "need dt_startup_phase > 0 and 0 < dt_min <= dt_max");
  if (sc->choice_calving_law != IMAU_CALVING_NONE && sc->choice_calving_law != IMAU_CALVING_THRESHOLD_THICKNESS)
    return fail(-2, "unknown choice_calving_law");
  if ((sc->remove_shelves_larger_than_PD && !(rg && rg->Hi_PD && rg->Hb_PD)) ||
      (sc->continental_shelf_calving && !(rg && rg->Hi_GIAeq && rg->Hb_GIAeq)))
    return fail(-1, "the reference geometry of a removal option is missing");
  G gg = {gr->nx, gr->ny, c->nz, gr->dx, gr->x, gr->y, c};
  const G *g = &gg;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const size_t na = (size_t)nx * ny, ncx = (size_t)ny * (nx - 1), ncy = (size_t)(ny - 1) * nx;
  /* working geometry and derived fields */
  double *Hi = dalloc(na), *fmf = dalloc(na), *dHi = dalloc(na), *Ht = dalloc(na), *zero = dalloc(na);
  double *heff = dalloc(na);
  double *A3 = dalloc(na * nz), *Av = dalloc(na);
  memcpy(Hi, io->Hi_a, na * sizeof(double));
  memcpy(A3, io->A_flow_3D_a, na * nz * sizeof(double));
  memcpy(Av, io->A_flow_vav_a, na * sizeof(double));
  ProducerBufs_ PB;
  producer_bufs_alloc(&PB, na, ncx, ncy, nz);
  imau_producer_io pio;
  producer_io_bind(&pio, &PB, Hi, io->Hb_a, io->SL_a, A3, Av);
  /* velocity outputs the loop reads, where the caller did not ask for them */
  double *u_vav = io->u_vav_cx ? io->u_vav_cx : dalloc(ncx), *v_vav = io->v_vav_cy ? io->v_vav_cy : dalloc(ncy);
  double *u3 = io->u_3D_cx ? io->u_3D_cx : dalloc(ncx * nz), *v3 = io->v_3D_cy ? io->v_3D_cy : dalloc(ncy * nz);
  double *u3s = io->u_3D_SIA_cx ? io->u_3D_SIA_cx : dalloc(ncx * nz);
  double *v3s = io->v_3D_SIA_cy ? io->v_3D_SIA_cy : dalloc(ncy * nz);
  double *uS = io->u_SSA_cx ? io->u_SSA_cx : dalloc(ncx), *vS = io->v_SSA_cy ? io->v_SSA_cy : dalloc(ncy);
  imau_ice_io w = *io;
  w.Hi_a = Hi; w.Hs_a = PB.Hs; w.TAF_a = PB.TAF;
  w.dHs_dx_a = PB.dHs_dx_a; w.dHs_dy_a = PB.dHs_dy_a; w.dHi_dx_a = PB.dHi_dx_a; w.dHi_dy_a = PB.dHi_dy_a;
  w.A_flow_vav_a = Av;
  w.mask_land_a = PB.land; w.mask_ocean_a = PB.ocean; w.mask_ice_a = PB.ice; w.mask_sheet_a = PB.sheet;
  w.mask_shelf_a = PB.shelf;
  w.Hi_cx = PB.Hi_cx; w.dHs_dx_cx = PB.dHs_dx_cx; w.dHs_dy_cx = PB.dHs_dy_cx;
  w.Hi_cy = PB.Hi_cy; w.dHs_dx_cy = PB.dHs_dx_cy; w.dHs_dy_cy = PB.dHs_dy_cy; w.Hi_b = PB.Hi_b;
  w.A_flow_3D_a = A3; w.A_flow_3D_cx = PB.A3cx; w.A_flow_3D_cy = PB.A3cy;
  w.u_vav_cx = u_vav; w.v_vav_cy = v_vav; w.u_3D_cx = u3; w.v_3D_cy = v3; w.u_3D_SIA_cx = u3s; w.v_3D_SIA_cy = v3s;
  w.u_SSA_cx = uS; w.v_SSA_cy = vS;
  /* the derived fields of the state at entry, and its floating-margin fraction */
  pio.dHi_dt_a = NULL;
  orc_producers_update(gr, pc, &pio);
  floating_margin_fraction(g, Hi, PB.ice, PB.cf, PB.shelf, fmf, heff);
  /* predictor-corrector fields (type_ice_model :293-299) */
  double *Hn = NULL, *Hnm1 = NULL, *Hstar = NULL, *Hold = NULL, *Hpred = NULL;
  if (pcs) {
    Hn = dalloc(na); Hnm1 = dalloc(na); Hstar = dalloc(na); Hold = dalloc(na); Hpred = dalloc(na);
    memcpy(Hn, pcs->dHidt_Hn_un, na * sizeof(double));
    memcpy(dHi, pcs->dHi_dt_a, na * sizeof(double));
  }
  imau_thickness_io t;  /* calc_dHi_dt at the current geometry and velocities */
  memset(&t, 0, sizeof t);
  t.Hi_a = Hi; t.Hb_a = io->Hb_a; t.SL_a = io->SL_a; t.u_vav_cx = u_vav; t.v_vav_cy = v_vav;
  t.SMB_year = SMB; t.BMB = BMB; t.float_margin_frac_a = fmf;
  t.mask_ice_a = PB.ice; t.mask_shelf_a = PB.shelf; t.mask_ocean_a = PB.ocean; t.mask_cf_a = PB.cf;
  t.Hi_tplusdt_a = Ht;
  const int32_t nit0 = io->DIVA_SOR_nit;
  const double tol0 = io->DIVA_SOR_tol, rtol0 = io->DIVA_PETSc_rtol, atol0 = io->DIVA_PETSc_abstol;
  double time = st->time;
  int rc = 0;
  st->nsteps = 0;
  while (time < t_end) {
    /* :88-99: tighter velocity solves in the start-up phase; the reference's
     * r = 0 at the very start would ask for unbounded iterations, so r is
     * floored at 1e-3 */
    double r = 1.0;
    if (time <= sc->start_time_of_run + sc->dt_startup_phase)
      r = fmax(1e-3, 0.01 * 0.99 * (time - sc->start_time_of_run) / sc->dt_startup_phase);
    w.DIVA_SOR_nit = nit0 * (int32_t)ceil(1.0 / r);
    w.DIVA_SOR_tol = tol0 * r;
    w.DIVA_PETSc_rtol = rtol0 * r;
    w.DIVA_PETSc_abstol = atol0 * r;
    /* :101-108 */
    double dt_max;
    if (time <= sc->start_time_of_run + sc->dt_startup_phase)
      dt_max = sc->dt_min + (sc->dt_max - sc->dt_min) * pow((time - sc->start_time_of_run) / sc->dt_startup_phase, 2);
    else if (time >= sc->end_time_of_run - sc->dt_startup_phase)
      dt_max = sc->dt_min + (sc->dt_max - sc->dt_min) * pow((sc->end_time_of_run - time) / sc->dt_startup_phase, 2);
    else
      dt_max = sc->dt_max;
    if (pcs && dyn != IMAU_DYN_NONE && time == pcs->t_next) {  /* :312-415 */
      const double dt_adv = orc_critical_timestep_adv(gr, u_vav, v_vav, sc->dt_max);
      pcs->dt_crit_ice_prev = pcs->dt_crit_ice;
      const double dt_from_pc = pow(sc->pc_epsilon / pcs->pc_eta, sc->pc_k_I + sc->pc_k_p) *
                                pow(sc->pc_epsilon / pcs->pc_eta_prev, -sc->pc_k_p) * pcs->dt;
      double dc = fmin(fmin(sc->dt_max, 2.0 * pcs->dt_crit_ice_prev), fmin(dt_adv, dt_from_pc));
      dc = fmax(sc->dt_min, fmax(0.5 * pcs->dt_crit_ice_prev, dc));
      dc = fmax(sc->dt_min, fmin(dt_max, dc));
      pcs->dt_crit_ice = dc;
      const double z = dc / pcs->dt_crit_ice_prev;
      pcs->pc_zeta = z;
      memcpy(Hnm1, Hn, na * sizeof(double));
      t.dHi_dt_a = Hn;
      if ((rc = orc_thickness_update(gr, tc, &t, dc))) break;
      for (size_t p = 0; p < na; ++p) {
        Hpred[p] = fmax(0.0, Hi[p] + dc * ((1.0 + z / 2.0) * Hn[p] - (z / 2.0) * Hnm1[p]));
        Hold[p] = Hi[p];
        Hi[p] = Hpred[p];
      }
      pio.dHi_dt_a = Hn;
      pio.dHb_dt_a = io->dHb_dt_a ? io->dHb_dt_a : zero;
      orc_producers_update(gr, pc, &pio);
      if (dyn == IMAU_DYN_SIA || dyn == IMAU_DYN_SIA_SSA)
        if ((rc = orc_solve_sia(gr, c, &w))) break;
      if (dyn == IMAU_DYN_SSA || dyn == IMAU_DYN_SIA_SSA)
        if ((rc = orc_solve_ssa(gr, c, &w, ORC_LIN_BANDED_LU))) break;
      if (dyn == IMAU_DYN_DIVA)
        if ((rc = orc_solve_diva(gr, c, &w, ORC_LIN_BANDED_LU))) break;
      pcs->t_next = time + dc;
      t.dHi_dt_a = Hstar;
      if ((rc = orc_thickness_update(gr, tc, &t, dc))) break;
      memcpy(Hi, Hold, na * sizeof(double));
      pio.dHi_dt_a = Hstar;
      orc_producers_update(gr, pc, &pio);
      pcs->pc_eta_prev = pcs->pc_eta;
      double eta = sc->pc_eta_min;
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
          const size_t p = (size_t)i * ny + j;
          const double hc = fmax(0.0, Hold[p] + 0.5 * dc * (Hn[p] + Hstar[p]));
          dHi[p] = (hc - Hi[p]) / dc;
          if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2 || PB.sheet[p] != 1) continue;  /* :786-808 */
          int gl = 0;
          for (int di = -1; di <= 1; ++di)
            for (int dj = -1; dj <= 1; ++dj) gl |= PB.gl[(size_t)(i + di) * ny + (j + dj)];
          if (!gl) eta = fmax(eta, fabs(z * (hc - Hpred[p]) / ((3.0 * z + 3.0) * dc)));
        }
      pcs->pc_eta = eta;
    }
    if (!pcs && (dyn == IMAU_DYN_SIA || dyn == IMAU_DYN_SIA_SSA) && time == st->t_next_SIA) {
      if ((rc = orc_solve_sia(gr, c, &w))) break;
      double d = crit_sia(g, &sc->dt_max, u3s, v3s, u3, v3, PB.Hi_cx, PB.Hi_cy, PB.dHs_dx_cx, PB.dHs_dy_cy, PB.sheet,
                          c->zeta);
      d = fmax(sc->dt_min, fmin(dt_max, d));
      st->dt_crit_SIA = d;
      st->t_next_SIA = time + d;
    }
    if (!pcs && (dyn == IMAU_DYN_SSA || dyn == IMAU_DYN_SIA_SSA) && time == st->t_next_SSA) {
      if ((rc = orc_solve_ssa(gr, c, &w, ORC_LIN_BANDED_LU))) break;
      double d = orc_critical_timestep_adv(gr, u_vav, v_vav, sc->dt_max);
      d = fmax(sc->dt_min, fmin(dt_max, d));
      st->dt_crit_SSA = d;
      st->t_next_SSA = time + d;
    }
    /* determine_timesteps_and_actions, :838-856 and :933 */
    double t_next = fmin(t_end, time + sc->dt_max);
    if (pcs) {
      if (dyn != IMAU_DYN_NONE) t_next = fmin(t_next, pcs->t_next);
    } else {
      if (dyn == IMAU_DYN_SIA || dyn == IMAU_DYN_SIA_SSA) t_next = fmin(t_next, st->t_next_SIA);
      if (dyn == IMAU_DYN_SSA || dyn == IMAU_DYN_SIA_SSA) t_next = fmin(t_next, st->t_next_SSA);
    }
    const double dt = t_next - time;
    st->dt = dt;
    if (pcs) {  /* :257-260 for 'none'; :428 */
      pcs->dt = dt;
      if (dyn == IMAU_DYN_NONE) memset(dHi, 0, na * sizeof(double));
      for (size_t p = 0; p < na; ++p) Ht[p] = Hi[p] + dt * dHi[p];
    } else if (dyn == IMAU_DYN_NONE) {  /* :224-236 */
      for (size_t p = 0; p < na; ++p) { dHi[p] = 0.0; Ht[p] = Hi[p]; }
    } else {
      t.dHi_dt_a = dHi;
      if ((rc = orc_thickness_update(gr, tc, &t, dt))) break;
      for (size_t p = 0; p < na; ++p) Ht[p] = Hi[p] + dt * dHi[p];
    }
    /* update_ice_thickness, :457-645 */
    if (sc->fixed_shelf_geometry || sc->fixed_sheet_geometry || sc->fixed_grounding_line)
      fix_geometry(g, sc, Hi, PB.sheet, PB.shelf, Ht);
    for (size_t p = 0; p < na; ++p) Hi[p] = fmax(0.0, Ht[p]);
    pio.dHi_dt_a = NULL;
    orc_producers_update(gr, pc, &pio);                         /* determine_masks_ice / _transitions */
    floating_margin_fraction(g, Hi, PB.ice, PB.cf, PB.shelf, fmf, heff);
    if (sc->choice_calving_law == IMAU_CALVING_THRESHOLD_THICKNESS) {  /* :574-588, twice */
      threshold_calving(na, Hi, PB.cf, heff, sc->calving_threshold_thickness);
      orc_producers_update(gr, pc, &pio);
      floating_margin_fraction(g, Hi, PB.ice, PB.cf, PB.shelf, fmf, heff);
      threshold_calving(na, Hi, PB.cf, heff, sc->calving_threshold_thickness);
      orc_producers_update(gr, pc, &pio);
    }
    remove_unconnected_shelves(g, Hi, PB.shelf, PB.sheet);
    remove_ice(na, sc, rg, io->Hb_a, io->SL_a, Hi);
    pio.dHi_dt_a = dHi;                                          /* update_general_ice_model_data */
    pio.dHb_dt_a = io->dHb_dt_a ? io->dHb_dt_a : zero;
    orc_producers_update(gr, pc, &pio);
    time += dt;
    ++st->nsteps;
  }
  st->time = time;
  /* the statistics of the last velocity solve */
  io->stat_picard_iterations = w.stat_picard_iterations;
  io->stat_linear_sweeps = w.stat_linear_sweeps;
  io->stat_resid_UV = w.stat_resid_UV;
  io->stat_nonlinear_residual = w.stat_nonlinear_residual;
  io->stat_seconds = w.stat_seconds;
  if (!rc) {
    memcpy(Hi_end, Hi, na * sizeof(double));
    if (geo) {
      producer_io_copy_out(geo, &pio, na, ncx, ncy, (size_t)(ny - 1) * (nx - 1), nz);
      if (geo->float_margin_frac_a) memcpy(geo->float_margin_frac_a, fmf, na * sizeof(double));
      if (geo->Hi_eff_cf_a) memcpy(geo->Hi_eff_cf_a, heff, na * sizeof(double));
    }
  }
  if (!io->u_vav_cx) free(u_vav);
  if (!io->v_vav_cy) free(v_vav);
  if (!io->u_3D_cx) free(u3);
  if (!io->v_3D_cy) free(v3);
  if (!io->u_3D_SIA_cx) free(u3s);
  if (!io->v_3D_SIA_cy) free(v3s);
  if (!io->u_SSA_cx) free(uS);
  if (!io->v_SSA_cy) free(vS);
  if (pcs) {
    if (!rc) {
      memcpy(pcs->dHidt_Hn_un, Hn, na * sizeof(double));
      memcpy(pcs->dHi_dt_a, dHi, na * sizeof(double));
    }
    free(Hn); free(Hnm1); free(Hstar); free(Hold); free(Hpred);
  }
  producer_bufs_free(&PB);
  free(Hi); free(fmf); free(dHi); free(Ht); free(zero); free(A3); free(Av); free(heff);
  return rc;
}

int orc_run_ice_dynamics_direct(const imau_grid *gr, const imau_cfg *c, const imau_thickness_cfg *tc,
                                const imau_producer_cfg *pc, const imau_timestep_cfg *sc, imau_ice_io *io,
                                const imau_refgeo_io *rg, const double *SMB, const double *BMB, double *Hi_end,
                                imau_producer_io *geo, imau_timestep_state *st, double t_end)
{
  return run_loop(gr, c, tc, pc, sc, io, rg, SMB, BMB, Hi_end, geo, st, NULL, t_end);
}

int orc_run_ice_dynamics_pc(const imau_grid *gr, const imau_cfg *c, const imau_thickness_cfg *tc,
                            const imau_producer_cfg *pc, const imau_timestep_cfg *sc, imau_ice_io *io,
                            const imau_refgeo_io *rg, const double *SMB, const double *BMB, double *Hi_end,
                            imau_producer_io *geo, imau_timestep_state *st, imau_pc_state *pcs, double t_end)
{
  if (!pcs || !pcs->dHidt_Hn_un || !pcs->dHi_dt_a) return fail(-1, "pc state fields are NULL");
  if (pcs->t_next < st->time) return fail(-1, "pc t_next is before the current time");
  return run_loop(gr, c, tc, pc, sc, io, rg, SMB, BMB, Hi_end, geo, st, pcs, t_end);
}

/* building blocks of the loop, exposed for unit tests */
void orc_floating_margin_fraction(const imau_grid *gr, const double *Hi, const int32_t *ice, const int32_t *cf,
                                  const int32_t *shelf, double *fmf, double *heff)
{
  G gg = {gr->nx, gr->ny, 1, gr->dx, gr->x, gr->y, NULL};
  floating_margin_fraction(&gg, Hi, ice, cf, shelf, fmf, heff);
}
void orc_fix_geometry(const imau_grid *gr, const imau_timestep_cfg *sc, const double *Hi, const int32_t *sheet,
                      const int32_t *shelf, double *Ht)
{
  G gg = {gr->nx, gr->ny, 1, gr->dx, gr->x, gr->y, NULL};
  fix_geometry(&gg, sc, Hi, sheet, shelf, Ht);
}
void orc_remove_unconnected_shelves(const imau_grid *gr, double *Hi, const int32_t *shelf, const int32_t *sheet)
{
  G gg = {gr->nx, gr->ny, 1, gr->dx, gr->x, gr->y, NULL};
  remove_unconnected_shelves(&gg, Hi, shelf, sheet);
}
double orc_critical_timestep_sia(const imau_grid *gr, int nz, const double *zeta, double dt_max, const double *u3sia,
                                 const double *v3sia, const double *u3, const double *v3, const double *Hi_cx,
                                 const double *Hi_cy, const double *dHs_dx_cx, const double *dHs_dy_cy,
                                 const int32_t *sheet)
{
  G gg = {gr->nx, gr->ny, nz, gr->dx, gr->x, gr->y, NULL};
  return crit_sia(&gg, &dt_max, u3sia, v3sia, u3, v3, Hi_cx, Hi_cy, dHs_dx_cx, dHs_dy_cy, sheet, zeta);
}
