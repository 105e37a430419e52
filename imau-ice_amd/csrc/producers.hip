// producers.hip -- the fields the velocity solve reads, produced on the
// device from the geometry and the temperature:
// update_general_ice_model_data (src/general_ice_model_data_module.f90:30-84)
// with determine_masks (:87-346) and the ocean flood fill (:868-986), and
// calc_ice_rheology (src/thermodynamics_module.f90:641-753).
//
// One member per call in the solver's slab layout (2D fields of every grid in
// a-grid-shaped slabs, p2(j,i) = off + i*ny + j); 3D fields take explicit
// strides so both the host's (nz,ny,nx) order and the solver's k-plane-major
// order can be read and written.  Every kernel but the flood fill is one
// thread per cell.  The flood fill, a depth-first stack walk in the reference,
// is computed as what it returns -- the floating cells connected to the domain
// edge through the 8 neighbours of floating cells -- by rounds of tile-local
// growth to a fixed point in LDS; growth is monotone, so the rounds end at the
// same set whatever their order.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "dev_common.h"
#include "producers.h"

// Multiply-adds are not contracted in this file: the thickness and mask
// arithmetic rounds as the reference's does, so that a thickness driven to
// zero (H + dt (-H / dt)) is zero and not a residue of a fused multiply-add
// that would count as ice.
#pragma clang fp contract(off)

namespace {

constexpr int TB = 256;
constexpr double RHO_I = 910.0, RHO_W = 1028.0;  // src/parameters_module.f90:20-21

struct Cell {
  int j, i;
  long long p;
};

__device__ __forceinline__ bool cell_of(const Member &m, int q, Cell &c) {
  if (q >= m.nx * m.ny) return false;
  c.i = q / m.ny;
  c.j = q - c.i * m.ny;
  c.p = m.off + q;
  return true;
}

// src/utilities_module.f90:126-158, with the reference's order of operations
__device__ __forceinline__ bool is_floating(double Hi, double Hb, double SL) {
  return Hi < (SL - Hb) * RHO_W / RHO_I;
}
__device__ __forceinline__ double surface_elevation(double Hi, double Hb, double SL) {
  return Hi + fmax(SL - RHO_I / RHO_W * Hi, Hb);
}
__device__ __forceinline__ double thickness_above_floatation(double Hi, double Hb, double SL) {
  return Hi - fmax(0.0, (SL - Hb) * (RHO_W / RHO_I));
}

__global__ void __launch_bounds__(TB) k_prod_point(const Member m, const imau::ProdIn in, const imau::Prod2 o,
                                                   const imau::ProdMasks mk, int floodfill) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const long long p = c.p;
  const double Hi = in.Hi[p], Hb = in.Hb[p], SL = in.SL[p];
  o.Hs[p] = surface_elevation(Hi, Hb, SL);
  if (o.TAF) o.TAF[p] = thickness_above_floatation(Hi, Hb, SL);
  // :47-52 reads the land mask before determine_masks updates it
  if (in.dHi_dt && in.dHb_dt && o.dHs_dt) {
    o.dHs_dt[p] = mk.land[p] == 1 ? in.dHb_dt[p] + in.dHi_dt[p] : in.dHi_dt[p] * (1.0 - RHO_I / RHO_W);
  }
  const bool flt = is_floating(Hi, Hb, SL);
  if (!floodfill) {  // :139-148
    mk.land[p] = flt ? 0 : 1;
    mk.ocean[p] = flt ? 1 : 0;
  } else {  // the flood fill starts from the floating edge cells (:905-924)
    const bool edge = c.i == 0 || c.i == m.nx - 1 || c.j == 0 || c.j == m.ny - 1;
    mk.ocean[p] = (edge && flt) ? 1 : 0;
  }
}

// flood-fill tiles: FI x FJ cells plus a one-cell halo in LDS
constexpr int FI = 16, FJ = 64;

__global__ void __launch_bounds__(TB) k_prod_flood(const Member m, const imau::ProdIn in, int32_t *__restrict__ ocean,
                                                   int *__restrict__ changed) {
  __shared__ uint8_t fl[FI + 2][FJ + 2], oc[FI + 2][FJ + 2];
  __shared__ int any;
  const int ntj = (m.ny + FJ - 1) / FJ;
  const int i0 = (blockIdx.x / ntj) * FI, j0 = (blockIdx.x % ntj) * FJ;
  const int tid = threadIdx.x;
  for (int q = tid; q < (FI + 2) * (FJ + 2); q += TB) {
    const int li = q / (FJ + 2), lj = q - li * (FJ + 2);
    const int i = i0 + li - 1, j = j0 + lj - 1;
    uint8_t f = 0, o = 0;
    if (i >= 0 && j >= 0 && i < m.nx && j < m.ny) {
      const long long p = m.off + (long long)i * m.ny + j;
      f = is_floating(in.Hi[p], in.Hb[p], in.SL[p]) ? 1 : 0;
      o = ocean[p] ? 1 : 0;
    }
    fl[li][lj] = f;
    oc[li][lj] = o;
  }
  bool grew = false;
  for (;;) {
    if (tid == 0) any = 0;
    __syncthreads();
    int mine = 0;
    for (int q = tid; q < FI * FJ; q += TB) {
      const int li = 1 + q / FJ, lj = 1 + q % FJ;
      if (!fl[li][lj] || oc[li][lj]) continue;
      const bool nb = oc[li - 1][lj - 1] | oc[li - 1][lj] | oc[li - 1][lj + 1] | oc[li][lj - 1] | oc[li][lj + 1] |
                      oc[li + 1][lj - 1] | oc[li + 1][lj] | oc[li + 1][lj + 1];
      if (nb) {
        oc[li][lj] = 1;
        mine = 1;
      }
    }
    if (mine) any = 1;
    __syncthreads();
    const int a = any;
    __syncthreads();  // every thread has read `any` before it is cleared again
    if (!a) break;
    grew = true;
  }
  if (!grew) return;
  int wrote = 0;
  for (int q = tid; q < FI * FJ; q += TB) {
    const int li = 1 + q / FJ, lj = 1 + q % FJ;
    const int i = i0 + li - 1, j = j0 + lj - 1;
    if (i >= m.nx || j >= m.ny || !oc[li][lj]) continue;
    const long long p = m.off + (long long)i * m.ny + j;
    if (!ocean[p]) {
      ocean[p] = 1;
      wrote = 1;
    }
  }
  if (wrote) *changed = 1;
}

__global__ void __launch_bounds__(TB) k_prod_land(const Member m, const imau::ProdMasks mk) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  mk.land[c.p] = 1 - mk.ocean[c.p];
}

// ice / sheet / shelf (:178-198), transitions (:219-276), total (:296-341),
// H on cx / cy / b and the slopes (:62-80)
__global__ void __launch_bounds__(TB) k_prod_masks(const Member m, const imau::ProdIn in, const imau::Prod2 o,
                                                   const imau::ProdMasks mk) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const int nx = m.nx, ny = m.ny, i = c.i, j = c.j;
  const long long p = c.p;
  auto at = [&](int jj, int ii) { return m.off + (long long)ii * ny + jj; };
  auto ice = [&](long long q) { return in.Hi[q] > DBL_MIN; };  // :187, Hi > TINY(Hi)
  auto ocn = [&](long long q) { return mk.ocean[q] == 1; };
  const bool ic = ice(p), oc = ocn(p), ld = mk.land[p] == 1;
  const bool sheet = ic && ld, shelf = ic && oc;
  mk.ice[p] = ic;
  mk.sheet[p] = sheet;
  mk.shelf[p] = shelf;
  int coast = 0, margin = 0, gl = 0, cf = 0;
  if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
    const long long n4[4] = {at(j - 1, i), at(j + 1, i), at(j, i - 1), at(j, i + 1)};
    for (int s = 0; s < 4; ++s) {
      const long long q = n4[s];
      const bool iq = ice(q), oq = ocn(q);
      if (ld && !ic && oq) coast = 1;
      if (ic && !iq) margin = 1;
      if (sheet && iq && oq) gl = 1;
      if (ic && oq && !iq) cf = 1;
    }
  }
  mk.coast[p] = coast;
  mk.margin[p] = margin;
  mk.gl[p] = gl;
  mk.cf[p] = cf;
  int t = ld ? IMAU_TYPE_LAND : IMAU_TYPE_OCEAN;
  if (coast) t = IMAU_TYPE_COAST;
  if (sheet) t = IMAU_TYPE_SHEET;
  else if (shelf) t = IMAU_TYPE_SHELF;
  if (margin) t = IMAU_TYPE_MARGIN;
  if (gl) t = IMAU_TYPE_GROUNDINGLINE;
  if (cf) t = IMAU_TYPE_CALVINGFRONT;
  mk.total[p] = t;
  // maps and slopes (src/derivatives_and_grids_module.f90: map_a_to_cx/cy/b,
  // ddx/ddy a->a, a->cx, a->cy, with their one-sided or halved edge rows)
  const double *H = in.Hi, *S = o.Hs;
  const double dx = m.dx;
  if (i < nx - 1) o.Hi_cx[p] = (H[p] + H[at(j, i + 1)]) / 2.0;
  if (j < ny - 1) o.Hi_cy[p] = (H[p] + H[at(j + 1, i)]) / 2.0;
  if (i < nx - 1 && j < ny - 1) o.Hi_b[p] = (H[p] + H[at(j, i + 1)] + H[at(j + 1, i)] + H[at(j + 1, i + 1)]) / 4.0;
  auto ddx_a = [&](const double *d) {
    if (i == 0) return (d[at(j, 1)] - d[at(j, 0)]) / dx;
    if (i == nx - 1) return (d[at(j, nx - 1)] - d[at(j, nx - 2)]) / dx;
    return (d[at(j, i + 1)] - d[at(j, i - 1)]) / (2 * dx);
  };
  auto ddy_a = [&](const double *d) {
    if (j == 0) return (d[at(1, i)] - d[at(0, i)]) / dx;
    if (j == ny - 1) return (d[at(ny - 1, i)] - d[at(ny - 2, i)]) / dx;
    return (d[at(j + 1, i)] - d[at(j - 1, i)]) / (2 * dx);
  };
  o.dHi_dx[p] = ddx_a(H);
  o.dHi_dy[p] = ddy_a(H);
  o.dHs_dx[p] = ddx_a(S);
  o.dHs_dy[p] = ddy_a(S);
  if (i < nx - 1) {
    o.dHs_dx_cx[p] = (S[at(j, i + 1)] - S[p]) / dx;
    const int n = ny - 1;
    double v;
    if (j == 0) v = (S[at(1, i)] + S[at(1, i + 1)] - S[at(0, i)] - S[at(0, i + 1)]) / (2.0 * dx);
    else if (j == n) v = (S[at(n, i)] + S[at(n, i + 1)] - S[at(n - 1, i)] - S[at(n - 1, i + 1)]) / (2.0 * dx);
    else v = (S[at(j + 1, i)] + S[at(j + 1, i + 1)] - S[at(j - 1, i)] - S[at(j - 1, i + 1)]) / (4.0 * dx);
    o.dHs_dy_cx[p] = v;
  }
  if (j < ny - 1) {
    o.dHs_dy_cy[p] = (S[at(j + 1, i)] - S[p]) / dx;
    const int n = nx - 1;
    double v;
    if (i == 0) v = (S[at(j, 1)] + S[at(j + 1, 1)] - S[at(j, 0)] - S[at(j + 1, 0)]) / (2.0 * dx);
    else if (i == n) v = (S[at(j, n)] + S[at(j + 1, n)] - S[at(j, n - 1)] - S[at(j + 1, n - 1)]) / (2.0 * dx);
    else v = (S[at(j, i + 1)] + S[at(j + 1, i + 1)] - S[at(j, i - 1)] - S[at(j + 1, i - 1)]) / (4.0 * dx);
    o.dHs_dx_cy[p] = v;
  }
}

// A on cx / cy (map_a_to_cx/cy_3D and _2D, :66-70 and
// src/thermodynamics_module.f90:739-743); one thread per (cell, level)
__global__ void __launch_bounds__(TB) k_prod_map_A(const Member m, int nz, const double *__restrict__ A3,
                                                   imau::Str3 sa, double *__restrict__ A3cx, imau::Str3 sx,
                                                   double *__restrict__ A3cy, imau::Str3 sy,
                                                   const double *__restrict__ Av, double *__restrict__ Avcx,
                                                   double *__restrict__ Avcy) {
  const long long na = (long long)m.nx * m.ny;
  const long long q = (long long)blockIdx.x * TB + threadIdx.x;
  if (q >= na * (nz + 1)) return;
  const int k = (int)(q / na);  // k == nz: the 2D vertical average
  const int c = (int)(q - (long long)k * na);
  const int i = c / m.ny, j = c - i * m.ny;
  if (k == nz) {
    if (!Av) return;
    const long long p = m.off + c;
    if (Avcx && i < m.nx - 1) Avcx[p] = (Av[p] + Av[p + m.ny]) / 2.0;
    if (Avcy && j < m.ny - 1) Avcy[p] = (Av[p] + Av[p + 1]) / 2.0;
    return;
  }
  if (!A3) return;
  const double a = A3[k * sa.sk + i * sa.si + j * sa.sj];
  if (A3cx && i < m.nx - 1) A3cx[k * sx.sk + i * sx.si + j * sx.sj] = (a + A3[k * sa.sk + (i + 1) * sa.si + j * sa.sj]) / 2.0;
  if (A3cy && j < m.ny - 1) A3cy[k * sy.sk + i * sy.si + j * sy.sj] = (a + A3[k * sa.sk + i * sa.si + (j + 1) * sa.sj]) / 2.0;
}

// calc_ice_rheology (src/thermodynamics_module.f90:662-737): per column, the
// flow factor of every level, the enhancement factor and the vertical average
__global__ void __launch_bounds__(TB) k_prod_rheology(const Member m, const imau_producer_cfg cfg, double A_mismip,
                                                      const double *__restrict__ Ti, imau::Str3 st,
                                                      const imau::ProdMasks mk, double *__restrict__ A3,
                                                      imau::Str3 sa, double *__restrict__ Av) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const int nz = cfg.nz;
  const double A_low = 1.14e-05, A_high = 5.47e+10, Q_low = 6.0e+04, Q_high = 13.9e+04, R_gas = 8.314;  // :654-658, src/parameters_module.f90:23
  const bool ice = mk.ice[c.p] == 1;
  const double enh = mk.sheet[c.p] == 1 ? cfg.m_enh_sheet : (mk.shelf[c.p] == 1 ? cfg.m_enh_shelf : 1.0);
  const bool apply = mk.sheet[c.p] == 1 || mk.shelf[c.p] == 1;
  double s = 0.0, prev = 0.0;
  for (int k = 0; k < nz; ++k) {
    double a;
    if (cfg.choice_ice_rheology == IMAU_RHEO_UNIFORM) {
      a = cfg.uniform_flow_factor;
    } else if (cfg.choice_ice_rheology == IMAU_RHEO_MISMIP_MOD) {
      a = A_mismip;
    } else if (ice) {
      const double T = Ti[k * st.sk + c.i * st.si + c.j * st.sj];
      a = T < 263.15 ? A_low * exp(-Q_low / (R_gas * T)) : A_high * exp(-Q_high / (R_gas * T));
    } else {
      a = cfg.choice_ice_margin == IMAU_MARGIN_BC ? 0.0 : A_low * exp(-Q_low / (R_gas * 263.15));
    }
    if (apply) a *= enh;
    A3[k * sa.sk + c.i * sa.si + c.j * sa.sj] = a;
    if (k > 0) s += 0.5 * (a + prev) * (cfg.zeta[k] - cfg.zeta[k - 1]);
    prev = a;
  }
  Av[c.p] = s;
}

// determine_floating_margin_fraction (src/general_ice_model_data_module.f90:803-858):
// rows 1 .. ny-2 only, as the reference; heff (may be null) = Hi_eff_cf_a
__global__ void __launch_bounds__(TB) k_float_margin(const Member m, const double *__restrict__ Hi,
                                                    const imau::ProdMasks mk, double *__restrict__ fmf,
                                                    double *__restrict__ heff) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const int nx = m.nx, ny = m.ny, i = c.i, j = c.j;
  if (j < 1 || j > ny - 2) return;
  const long long p = c.p;
  const bool ice = mk.ice[p] == 1;
  double f = ice ? 1.0 : 0.0, h = ice ? Hi[p] : 0.0;
  if (mk.cf[p] == 1 && mk.shelf[p] == 1) {
    bool noncf = false;
    double hmax = 0.0;
    for (int ii = max(0, i - 1); ii <= min(nx - 1, i + 1); ++ii)
      for (int jj = j - 1; jj <= j + 1; ++jj) {
        const long long q = m.off + (long long)ii * ny + jj;
        if (mk.ice[q] == 1 && mk.cf[q] == 0) {
          noncf = true;
          hmax = fmax(hmax, Hi[q]);
        }
      }
    const bool whole = !noncf || hmax < Hi[p];
    f = whole ? 1.0 : Hi[p] / hmax;
    h = whole ? Hi[p] : hmax;
  }
  fmf[p] = f;
  if (heff) heff[p] = h;
}

// the fixed-geometry switches of update_ice_thickness
// (src/ice_dynamics_module.f90:460-559), with the masks of the state at t:
// H(t+dt) = H(t) on the shelf and the grounding line next to it (bit 0), the
// sheet and the grounding line next to it (bit 1), the grounding line (bit 2)
__device__ __forceinline__ bool any3(const Member &m, const int32_t *__restrict__ mask, int i, int j) {
  for (int ii = max(0, i - 1); ii <= min(m.nx - 1, i + 1); ++ii)
    for (int jj = max(0, j - 1); jj <= min(m.ny - 1, j + 1); ++jj)
      if (mask[m.off + (long long)ii * m.ny + jj] == 1) return true;
  return false;
}
__global__ void __launch_bounds__(TB) k_fix_geometry(const Member m, int which, const double *__restrict__ Hi,
                                                    const int32_t *__restrict__ sheet,
                                                    const int32_t *__restrict__ shelf, double *__restrict__ Ht) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const bool sh = shelf[c.p] == 1, gr = sheet[c.p] == 1;
  const bool gl_sheet = gr && any3(m, shelf, c.i, c.j), gl_shelf = sh && any3(m, sheet, c.i, c.j);
  const bool fix = ((which & 1) && (sh || gl_sheet)) || ((which & 2) && (gr || gl_shelf)) ||
                   ((which & 4) && (gl_sheet || gl_shelf));
  if (fix) Ht[c.p] = Hi[c.p];
}

// threshold_thickness_calving (src/calving_module.f90:62-91)
__global__ void __launch_bounds__(TB) k_calve(const Member m, const int32_t *__restrict__ cf,
                                             const double *__restrict__ heff, double thr, double *__restrict__ Hi) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  if (cf[c.p] == 1 && heff[c.p] < thr) Hi[c.p] = 0.0;
}

// the ice removals after the calving (src/ice_dynamics_module.f90:590-635)
__global__ void __launch_bounds__(TB) k_remove_ice(const Member m, const imau::Removal r, const double *__restrict__ Hb,
                                                  const double *__restrict__ SL, double *__restrict__ Hi) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const long long p = c.p;
  double h = Hi[p];
  if (r.noice && r.noice[p] == 1) h = 0.0;
  if (r.shelves && is_floating(h, Hb[p], SL[p])) h = 0.0;
  if (r.Hi_PD && r.Hi_PD[p] == 0.0 && r.Hb_PD[p] < 0.0) h = 0.0;
  if (r.Hi_GIA && r.Hi_GIA[p] == 0.0 && r.Hb_GIA[p] < r.min_height) h = 0.0;
  Hi[p] = h;
}

// remove_unconnected_shelves (src/ice_thickness_module.f90:932-1044), what its
// stack walk returns: the shelf cells 4-connected through shelf cells to an
// interior shelf cell that has a 4-neighbour sheet cell keep their ice.
// Seeds, then rounds of tile-local growth (as the ocean flood fill), then the removal.
__global__ void __launch_bounds__(TB) k_shelf_seed(const Member m, const imau::ProdMasks mk, int32_t *__restrict__ keep) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const int nx = m.nx, ny = m.ny, i = c.i, j = c.j;
  const long long p = c.p;
  int k = 0;
  if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && mk.shelf[p] == 1)
    k = (mk.sheet[p + ny] == 1 || mk.sheet[p - ny] == 1 || mk.sheet[p + 1] == 1 || mk.sheet[p - 1] == 1) ? 1 : 0;
  keep[p] = k;
}

__global__ void __launch_bounds__(TB) k_shelf_flood(const Member m, const int32_t *__restrict__ shelf,
                                                   int32_t *__restrict__ keep, int *__restrict__ changed) {
  __shared__ uint8_t al[FI + 2][FJ + 2], kp[FI + 2][FJ + 2];
  __shared__ int any;
  const int ntj = (m.ny + FJ - 1) / FJ;
  const int i0 = (blockIdx.x / ntj) * FI, j0 = (blockIdx.x % ntj) * FJ;
  const int tid = threadIdx.x;
  for (int q = tid; q < (FI + 2) * (FJ + 2); q += TB) {
    const int li = q / (FJ + 2), lj = q - li * (FJ + 2);
    const int i = i0 + li - 1, j = j0 + lj - 1;
    uint8_t a = 0, k = 0;
    if (i >= 0 && j >= 0 && i < m.nx && j < m.ny) {
      const long long p = m.off + (long long)i * m.ny + j;
      a = shelf[p] == 1 ? 1 : 0;
      k = keep[p] ? 1 : 0;
    }
    al[li][lj] = a;
    kp[li][lj] = k;
  }
  bool grew = false;
  for (;;) {
    if (tid == 0) any = 0;
    __syncthreads();
    int mine = 0;
    for (int q = tid; q < FI * FJ; q += TB) {
      const int li = 1 + q / FJ, lj = 1 + q % FJ;
      if (!al[li][lj] || kp[li][lj]) continue;
      if (kp[li - 1][lj] | kp[li + 1][lj] | kp[li][lj - 1] | kp[li][lj + 1]) {
        kp[li][lj] = 1;
        mine = 1;
      }
    }
    if (mine) any = 1;
    __syncthreads();
    const int a = any;
    __syncthreads();  // every thread has read `any` before it is cleared again
    if (!a) break;
    grew = true;
  }
  if (!grew) return;
  int wrote = 0;
  for (int q = tid; q < FI * FJ; q += TB) {
    const int li = 1 + q / FJ, lj = 1 + q % FJ;
    const int i = i0 + li - 1, j = j0 + lj - 1;
    if (i >= m.nx || j >= m.ny || !kp[li][lj]) continue;
    const long long p = m.off + (long long)i * m.ny + j;
    if (!keep[p]) {
      keep[p] = 1;
      wrote = 1;
    }
  }
  if (wrote) *changed = 1;
}

__global__ void __launch_bounds__(TB) k_shelf_remove(const Member m, const int32_t *__restrict__ shelf,
                                                    const int32_t *__restrict__ keep, double *__restrict__ Hi) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  if (shelf[c.p] == 1 && !keep[c.p]) Hi[c.p] = 0.0;
}

int blocks(const Member &m) { return (m.nx * m.ny + TB - 1) / TB; }

}  // namespace

namespace imau {

void prod_point(hipStream_t s, const Member &m, const ProdIn &in, const Prod2 &o, const ProdMasks &mk, int floodfill) {
  hipLaunchKernelGGL(k_prod_point, dim3(blocks(m)), dim3(TB), 0, s, m, in, o, mk, floodfill);
}
void prod_flood_round(hipStream_t s, const Member &m, const ProdIn &in, const ProdMasks &mk, int *changed) {
  const int nt = ((m.nx + FI - 1) / FI) * ((m.ny + FJ - 1) / FJ);
  hipLaunchKernelGGL(k_prod_flood, dim3(nt), dim3(TB), 0, s, m, in, mk.ocean, changed);
}
void prod_land(hipStream_t s, const Member &m, const ProdMasks &mk) {
  hipLaunchKernelGGL(k_prod_land, dim3(blocks(m)), dim3(TB), 0, s, m, mk);
}
void prod_masks_maps(hipStream_t s, const Member &m, const ProdIn &in, const Prod2 &o, const ProdMasks &mk) {
  hipLaunchKernelGGL(k_prod_masks, dim3(blocks(m)), dim3(TB), 0, s, m, in, o, mk);
}
void prod_map_A(hipStream_t s, const Member &m, int nz, const double *A3, Str3 sa, double *A3cx, Str3 sx, double *A3cy,
                Str3 sy, const double *Av, double *Avcx, double *Avcy) {
  const long long n = (long long)m.nx * m.ny * (nz + 1);
  hipLaunchKernelGGL(k_prod_map_A, dim3((unsigned)((n + TB - 1) / TB)), dim3(TB), 0, s, m, nz, A3, sa, A3cx, sx, A3cy,
                     sy, Av, Avcx, Avcy);
}
void prod_float_margin(hipStream_t s, const Member &m, const double *Hi, const ProdMasks &mk, double *fmf,
                       double *heff) {
  hipLaunchKernelGGL(k_float_margin, dim3(blocks(m)), dim3(TB), 0, s, m, Hi, mk, fmf, heff);
}
void prod_fix_geometry(hipStream_t s, const Member &m, int which, const double *Hi, const ProdMasks &mk, double *Ht) {
  hipLaunchKernelGGL(k_fix_geometry, dim3(blocks(m)), dim3(TB), 0, s, m, which, Hi, mk.sheet, mk.shelf, Ht);
}
void prod_calve(hipStream_t s, const Member &m, const ProdMasks &mk, const double *heff, double thr, double *Hi) {
  hipLaunchKernelGGL(k_calve, dim3(blocks(m)), dim3(TB), 0, s, m, mk.cf, heff, thr, Hi);
}
void prod_remove_ice(hipStream_t s, const Member &m, const Removal &r, const double *Hb, const double *SL, double *Hi) {
  hipLaunchKernelGGL(k_remove_ice, dim3(blocks(m)), dim3(TB), 0, s, m, r, Hb, SL, Hi);
}
void prod_shelf_seed(hipStream_t s, const Member &m, const ProdMasks &mk, int32_t *keep) {
  hipLaunchKernelGGL(k_shelf_seed, dim3(blocks(m)), dim3(TB), 0, s, m, mk, keep);
}
void prod_shelf_round(hipStream_t s, const Member &m, const ProdMasks &mk, int32_t *keep, int *changed) {
  const int nt = ((m.nx + FI - 1) / FI) * ((m.ny + FJ - 1) / FJ);
  hipLaunchKernelGGL(k_shelf_flood, dim3(nt), dim3(TB), 0, s, m, mk.shelf, keep, changed);
}
void prod_shelf_remove(hipStream_t s, const Member &m, const ProdMasks &mk, const int32_t *keep, double *Hi) {
  hipLaunchKernelGGL(k_shelf_remove, dim3(blocks(m)), dim3(TB), 0, s, m, mk.shelf, keep, Hi);
}
void prod_rheology(hipStream_t s, const Member &m, const imau_producer_cfg &c, double time, const double *Ti, Str3 st,
                   const ProdMasks &mk, double *A3, Str3 sa, double *Av) {
  // MISMIP_mod (:701-713): the step function of time
  double A = 1.0e-16;
  if (time < 15000.0) A = 1.0e-16;
  else if (time < 30000.0) A = 1.0e-17;
  else if (time < 45000.0) A = 1.0e-16;
  hipLaunchKernelGGL(k_prod_rheology, dim3(blocks(m)), dim3(TB), 0, s, m, c, A, Ti, st, mk, A3, sa, Av);
}

}  // namespace imau
