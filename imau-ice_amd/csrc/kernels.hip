// kernels.hip -- HIP/gfx950 kernels of the ice-velocity hot path.
//
// Every kernel takes the tile work-list: workgroup b processes tile
// tiles[b] = (member g, i0, j0) of IMAU_TI x IMAU_TJ a-grid cells with a
// 64 x 4 thread block (lane = j, wavefront = i).  The same launch therefore
// covers one grid or an ensemble of uneven grids.
#include "dev_common.h"
#include "kernels.h"
#include "stencil.h"

#define TILE_SETUP                                  \
  const int tb = imau_xcd_tile();                   \
  const Tile t = tiles[tb];                         \
  const Member m = mem[t.g];                        \
  const int j = t.j0 + (int)threadIdx.x;

#define FOR_TILE_ROWS for (int ii = threadIdx.y; ii < IMAU_TI; ii += IMAU_WG_I)
#define FOR_VTILE_ROWS _Pragma("unroll 1") for (int ii = threadIdx.y; ii < IMAU_VTI; ii += IMAU_WG_I)

// ---------------------------------------------------------------------------
// block reductions (wave64 shuffles, then LDS across the 4 wavefronts)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}
template <int NV>
__device__ void block_sum_store(double (&v)[NV], double *out) {
  __shared__ double red[IMAU_WG_I][NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = wave_sum(v[q]);
    if (threadIdx.x == 0) red[threadIdx.y][q] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int w = 0; w < IMAU_WG_I; ++w) s += red[w][q];
      out[q] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// sliding laws, src/basal_conditions_and_sliding_module.f90:908-1432
// ---------------------------------------------------------------------------
// `slid` is kp.slid, or the law as a compile-time constant so that a kernel
// specialised on it carries only that law
__device__ __forceinline__ double sliding_beta(const KP &kp, int slid, double ua, double va, double beta_sq,
                                               double alpha_sq, double phi, double tauc, double Neff, double x,
                                               double y, double A_vav, double beta_e) {
  const double uabs = sqrt(kp.delta_v * kp.delta_v + ua * ua + va * va);
  switch (slid) {
    case 0: return 0.0;
    case 4: return beta_sq * pow(uabs, 1.0 / kp.wm - 1.0);                                   // Weertman :984-987
    case 2: return tauc / uabs;                                                               // Coulomb :1039-1041
    case 3: return tauc * pow(uabs, kp.q_plastic - 1.0) / pow(kp.u_threshold, kp.q_plastic);  // :1078-1083
    case 5: return fmin(alpha_sq * Neff, beta_sq * pow(uabs, 1.0 / kp.wm)) * pow(uabs, -1.0);  // Tsai :1128-1131
    case 6: {                                                                                 // Schoof :1175-1183
      const double mm = kp.wm;
      return ((beta_sq * pow(uabs, 1.0 / mm) * alpha_sq * Neff) /
              pow(pow(beta_sq, mm) * uabs + pow(alpha_sq * Neff, mm), 1.0 / mm)) * pow(uabs, -1.0);
    }
    case 7:                                                                                   // Zoet-Iverson :1220-1223
      return Neff * tan((D_PI / 180.0) * phi) * pow(uabs, 1.0 / kp.zi_p - 1.0) * pow(uabs + kp.zi_ut, -1.0 / kp.zi_p);
    case 1:                                                                                   // idealised :1288-1432
      if (kp.ideal == 0) return 1000.0 + 1000.0 * sin(2.0 * D_PI * x / kp.ismip_L) * sin(2.0 * D_PI * y / kp.ismip_L);
      if (kp.ideal == 1) return 1000.0 + 1000.0 * sin(2.0 * D_PI * x / kp.ismip_L);
      if (kp.ideal == 2) return beta_e;                                                      // :1352-1399
      if (kp.ideal == 3) return pow(A_vav * 1000.0, -1.0);
      return __builtin_nan("");
    default: return __builtin_nan("");
  }
}

// Schoof (2006) till yield stress (src/utilities_module.f90:836-877), tauc part
__device__ double schoof_tauc(double tantheta, double h0, double L, double mexp, double y) {
  const double f = D_ICE_DENSITY * D_GRAV * h0 * tantheta;
  return f * pow(fabs(y / L), mexp);
}

// ---------------------------------------------------------------------------
// grounded fraction of a bilinear quad, src/general_ice_model_data_module.f90:521-781
// ---------------------------------------------------------------------------
__device__ __forceinline__ double gf_perturb(double f, double s) {
  const double ftol = 1e-4;
  if (f == 0.0) return ftol * s;
  if (f > 0.0) return fmax(ftol * s, f);
  if (f < 0.0) return fmin(-ftol * s, f);
  return f;
}
__device__ __forceinline__ void gf_rot(double &NW, double &NE, double &SW, double &SE) {
  const double a = NW, b = NE, c = SE, d = SW;
  NW = b; NE = c; SE = d; SW = a;
}
__device__ __forceinline__ double gf_corner(double aa, double bb, double cc, double dd) {
  return ((bb * cc - aa * dd) * log(fabs(1.0 - (aa * dd) / (bb * cc))) + aa * dd) / (dd * dd);
}
__device__ __forceinline__ double gf_strip(double aa, double bb, double cc, double dd, double x) {
  return ((bb * cc - aa * dd) * log(fabs(cc + dd * x)) - bb * dd * x) / (dd * dd);
}
__device__ double fraction_above_zero(double fNW, double fNE, double fSW, double fSE) {
  double NW = gf_perturb(fNW, 1.1), NE = gf_perturb(fNE, 1.21), SW = gf_perturb(fSW, 1.13), SE = gf_perturb(fSE, 0.97);
  double phi;
  if (NW >= 0 && NE >= 0 && SW >= 0 && SE >= 0) return 1.0;
  if (NW <= 0 && NE <= 0 && SW <= 0 && SE <= 0) return 0.0;
  int scen = 0;
  for (int it = 0; it < 5 && scen == 0; ++it) {
    gf_rot(NW, NE, SW, SE);
    if (SW > 0 && SE < 0 && NE < 0 && NW < 0) scen = 1;
    else if (SW < 0 && SE > 0 && NE > 0 && NW > 0) scen = 2;
    else if (SW > 0 && SE > 0 && NE < 0 && NW < 0) scen = 3;
    else if (SW > 0 && SE < 0 && NE > 0 && NW < 0) scen = 4;
  }
  double aa, bb, cc, dd;
  if (scen == 1) {
    aa = SW; bb = SE - SW; cc = NW - SW; dd = NE + SW - NW - SE;
    phi = gf_corner(aa, bb, cc, dd);
    if (fabs(dd) < 1e-4) {
      SW = (SW > 0) ? SW + 0.1 : SW - 0.1;
      aa = SW; bb = SE - SW; cc = NW - SW; dd = NE + SW - NW - SE;
      phi = gf_corner(aa, bb, cc, dd);
    }
  } else if (scen == 2) {
    aa = -SW; bb = -(SE - SW); cc = -(NW - SW); dd = -(NE + SW - NW - SE);
    phi = 1.0 - gf_corner(aa, bb, cc, dd);
    if (fabs(dd) < 1e-4) {
      SW = (SW > 0) ? SW + 0.1 : SW - 0.1;
      aa = -SW; bb = -(SE - SW); cc = -(NW - SW); dd = -(NE + SW - NW - SE);
      phi = 1.0 - gf_corner(aa, bb, cc, dd);
    }
  } else if (scen == 3) {
    if (fabs(1.0 - NW / NE) < 1e-6 && fabs(1.0 - SW / SE) < 1e-6) {
      phi = SW / (SW - NW);
    } else {
      aa = SW; bb = SE - SW; cc = NW - SW; dd = NE + SW - NW - SE;
      phi = gf_strip(aa, bb, cc, dd, 1.0) - gf_strip(aa, bb, cc, dd, 0.0);
      if (fabs(dd) < 1e-4) {
        SW = (SW > 0) ? SW + 0.1 : SW - 0.1;
        aa = SW; bb = SE - SW; cc = NW - SW; dd = NE + SW - NW - SE;
        phi = gf_strip(aa, bb, cc, dd, 1.0) - gf_strip(aa, bb, cc, dd, 0.0);
      }
    }
  } else if (scen == 4) {
    aa = SW; bb = SE - SW; cc = NW - SW; dd = NE + SW - NW - SE;
    phi = gf_corner(aa, bb, cc, dd);
    gf_rot(NW, NE, SW, SE);
    gf_rot(NW, NE, SW, SE);
    aa = SW; bb = SE - SW; cc = NW - SW; dd = NE + SW - NW - SE;
    phi += gf_corner(aa, bb, cc, dd);
  } else {
    phi = __builtin_nan("");
  }
  return fmax(0.0, fmin(1.0, phi));
}

// ---------------------------------------------------------------------------
// per-solve setup 1: driving stress, basal conditions, grounded-fraction
// quadrants, SICOPOLIS masks, calving-front rhs, B = 0.5 A^(-1/n)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(IMAU_WG) k_setup1(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                    const KP kp, const Fields f, const double *__restrict__ yg) {
  TILE_SETUP
  if (j >= m.ny) return;
  const int nx = m.nx, ny = m.ny;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= nx) break;
    const long long p = p2(m, j, i);
    // taudx/taudy, src/ice_velocity_module.f90:288-294
    if (i < nx - 1) f.taudx[p] = -D_ICE_DENSITY * D_GRAV * f.Hi_cx[p] * f.dHs_dx_cx[p];
    if (j < ny - 1) f.taudy[p] = -D_ICE_DENSITY * D_GRAV * f.Hi_cy[p] * f.dHs_dy_cy[p];
    // basal hydrology, src/basal_conditions_and_sliding_module.f90:90-238
    const double Hi = f.Hi[p], Hb = f.Hb[p], SL = f.SL[p];
    double pore;
    if (kp.hydro == 0) pore = 0.0;
    else if (kp.hydro == 1) pore = -D_SEAWATER_DENSITY * D_GRAV * (SL - Hb);
    else {
      const double lp = fmin(1.0, fmax(0.0, 1.0 - (Hb - SL - kp.hydro_min) / (kp.hydro_max - kp.hydro_min)));
      pore = 0.96 * D_ICE_DENSITY * D_GRAV * Hi * lp;
    }
    const double ob = D_ICE_DENSITY * D_GRAV * Hi;
    f.pore[p] = pore;
    f.overb[p] = ob;
    f.Neff[p] = fmax(0.0, ob - pore);
    // bed roughness parameterisations, :243-294, :769-903
    if (kp.slid != 0 && kp.slid != 1 && kp.rough == 1) {
      if (kp.prough == 0) {
        const double w = fmin(1.0, fmax(0.0, (Hb - kp.phi_hb_min) / (kp.phi_hb_max - kp.phi_hb_min)));
        f.phi_fric[p] = (1.0 - w) * kp.phi_min + w * kp.phi_max;
      } else if (kp.prough == 1) {
        f.tauc[p] = schoof_tauc(fabs(f.dHs_dx_a[p]), Hi, kp.ssa_L, kp.ssa_m, yg[m.offy + j]);
      } else {
        // MISMIP+ (:846-899): Weertman sets beta_sq only, Tsai2015 / Schoof2005 both
        f.beta_sq[p] = 1.0e4;
        if (kp.slid != 4) f.alpha_sq[p] = 0.5;
      }
    }
    // grounded-fraction quadrants from the edge-clamped TAF field, :420-520
    {
      const int im = iclamp(i - 1, 0, nx - 1), ip = iclamp(i + 1, 0, nx - 1);
      const int jm = iclamp(j - 1, 0, ny - 1), jp = iclamp(j + 1, 0, ny - 1);
      const double T_nw = f.TAF[p2(m, jp, im)], T_n = f.TAF[p2(m, jp, i)], T_ne = f.TAF[p2(m, jp, ip)];
      const double T_w = f.TAF[p2(m, j, im)], T_m = f.TAF[p], T_e = f.TAF[p2(m, j, ip)];
      const double T_sw = f.TAF[p2(m, jm, im)], T_s = f.TAF[p2(m, jm, i)], T_se = f.TAF[p2(m, jm, ip)];
      const double f_NW = 0.25 * (T_nw + T_n + T_w + T_m);
      const double f_N = 0.5 * (T_n + T_m);
      const double f_NE = 0.25 * (T_n + T_ne + T_m + T_e);
      const double f_W = 0.5 * (T_w + T_m);
      const double f_m = T_m;
      const double f_E = 0.5 * (T_m + T_e);
      const double f_SW = 0.25 * (T_w + T_m + T_sw + T_s);
      const double f_S = 0.5 * (T_m + T_s);
      const double f_SE = 0.25 * (T_m + T_e + T_s + T_se);
      f.qNW[p] = fraction_above_zero(f_NW, f_N, f_W, f_m);
      f.qNE[p] = fraction_above_zero(f_N, f_NE, f_m, f_E);
      f.qSW[p] = fraction_above_zero(f_W, f_m, f_SW, f_S);
      f.qSE[p] = fraction_above_zero(f_m, f_E, f_S, f_SE);
    }
    // set_sico_masks, src/ice_velocity_module.f90:1322-1431
    {
      const bool flt = f.TAF[p] < 0.0;
      const int combi = (Hi == 0.0) ? (flt ? 2 : 1) : (flt ? 3 : 0);
      f.combi[p] = combi;
      const int i1 = i > 0 ? i - 1 : 0, i2 = i < nx - 1 ? i + 1 : nx - 1;
      const int j1 = j > 0 ? j - 1 : 0, j2 = j < ny - 1 ? j + 1 : ny - 1;
      const double Hw = f.Hi[p2(m, j, i1)], He = f.Hi[p2(m, j, i2)], Hs_ = f.Hi[p2(m, j1, i)], Hn = f.Hi[p2(m, j2, i)];
      int fin = (Hi > 0.0 && (Hw == 0.0 || He == 0.0 || Hs_ == 0.0 || Hn == 0.0)) ? 1 : 0;
      const int fout = (Hi == 0.0 && (Hw > 0.0 || He > 0.0 || Hs_ > 0.0 || Hn > 0.0)) ? 1 : 0;
      if (fin == 1 && combi == 0) fin = 0;
      f.fin[p] = fin;
      f.fout[p] = fout;
      // calving-front rhs, :2250-2261
      double H_ocn = 0.0;
      if (SL > Hb) H_ocn = fmin(D_ICE_DENSITY / D_SEAWATER_DENSITY * Hi, SL - Hb);
      f.front_rhs[p] = 0.5 * D_ICE_DENSITY * D_GRAV * Hi * Hi - 0.5 * D_SEAWATER_DENSITY * D_GRAV * H_ocn * H_ocn;
    }
    // B = 0.5 * A^(-1/n) (first factor of src/ice_velocity_module.f90:799)
    for (int k = 0; k < kp.nz; ++k) {
      const long long q = p3(m, k, j, i);
      const double A = f.A3[q];
      f.B3[q] = (kp.n_flow == 3.0) ? 0.5 / cbrt(A) : 0.5 * pow(A, -1.0 / kp.n_flow);
    }
  }
}

__device__ __forceinline__ int corner_bc(int b1, int b2) {
  if (b1 == 1 || b2 == 1) return 1;  // zero
  if (b1 == 0 || b2 == 0) return 0;  // infinite
  return 2;                          // periodic
}
// BC code of a domain-boundary row at (j,i) with last indices imax/jmax, :1954-2004
__device__ __forceinline__ int boundary_bc(const int *bc, int i, int j, int imax, int jmax) {
  const int bw = bc[0], be = bc[1], bs = bc[2], bn = bc[3];
  if (i == 0 && j == 0) return corner_bc(bs, bw);
  if (i == imax && j == 0) return corner_bc(bs, be);
  if (i == 0 && j == jmax) return corner_bc(bn, bw);
  if (i == imax && j == jmax) return corner_bc(bn, be);
  if (i == 0) return bw;
  if (i == imax) return be;
  if (j == 0) return bs;
  return bn;
}
__device__ __forceinline__ uint8_t bc_rowtype(int bc) { return bc == 0 ? RT_BND_INF : (bc == 1 ? RT_BND_ZERO : RT_BND_PER); }

// per-solve setup 2: grounded fractions on all grids, row types
__global__ void __launch_bounds__(IMAU_WG) k_setup2(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                    const KP kp, const Fields f) {
  TILE_SETUP
  if (j >= m.ny) return;
  const int nx = m.nx, ny = m.ny;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= nx) break;
    const long long p = p2(m, j, i);
    // src/general_ice_model_data_module.f90:380-400
    f.fg_a[p] = 0.25 * (f.qNW[p] + f.qNE[p] + f.qSW[p] + f.qSE[p]);
    if (i < nx - 1) f.fg_cx[p] = 0.25 * (f.qNE[p] + f.qSE[p] + f.qNW[p2(m, j, i + 1)] + f.qSW[p2(m, j, i + 1)]);
    if (j < ny - 1) f.fg_cy[p] = 0.25 * (f.qNE[p] + f.qNW[p] + f.qSE[p2(m, j + 1, i)] + f.qSW[p2(m, j + 1, i)]);
    if (i < nx - 1 && j < ny - 1)
      f.fg_b[p] = 0.25 * (f.qNE[p] + f.qNW[p2(m, j, i + 1)] + f.qSE[p2(m, j + 1, i)] + f.qSW[p2(m, j + 1, i + 1)]);
    // row types, src/ice_velocity_module.f90:1762-1931
    if (i < nx - 1) {
      uint8_t rt = RT_FREE;
      if (i == 0 || i == nx - 2 || j == 0 || j == ny - 1) rt = bc_rowtype(boundary_bc(kp.bc_u, i, j, nx - 2, ny - 1));
      else if (kp.margin == 0) {
        const int c0 = f.combi[p], c1 = f.combi[p2(m, j, i + 1)];
        if ((c0 == 3 && c1 == 1) || (c0 == 1 && c1 == 3)) rt = RT_ZERO;
        else if ((f.fin[p] == 1 && f.fout[p2(m, j, i + 1)] == 1) || (f.fout[p] == 1 && f.fin[p2(m, j, i + 1)] == 1)) {
          const int i1 = (f.fin[p] == 1) ? i : i + 1;
          if (f.fout[p2(m, j, i1 - 1)] == 1 && f.fout[p2(m, j, i1 + 1)] == 1) rt = RT_ZERO;
          else rt = (i1 == i) ? RT_MARGIN_I : RT_MARGIN_O;
        }
      }
      f.rt_u[p] = rt;
    }
    if (j < ny - 1) {
      uint8_t rt = RT_FREE;
      if (i == 0 || i == nx - 1 || j == 0 || j == ny - 2) rt = bc_rowtype(boundary_bc(kp.bc_v, i, j, nx - 1, ny - 2));
      else if (kp.margin == 0) {
        const int c0 = f.combi[p], c1 = f.combi[p2(m, j + 1, i)];
        if ((c0 == 3 && c1 == 1) || (c0 == 1 && c1 == 3)) rt = RT_ZERO;
        else if ((f.fin[p] == 1 && f.fout[p2(m, j + 1, i)] == 1) || (f.fout[p] == 1 && f.fin[p2(m, j + 1, i)] == 1)) {
          const int j1 = (f.fin[p] == 1) ? j : j + 1;
          if (f.fout[p2(m, j1 - 1, i)] == 1 && f.fout[p2(m, j1 + 1, i)] == 1) rt = RT_ZERO;
          else rt = (j1 == j) ? RT_MARGIN_I : RT_MARGIN_O;
        }
      }
      f.rt_v[p] = rt;
    }
  }
}

// ---------------------------------------------------------------------------
// Picard step 1: vertical shear strain rates + effective viscosity + N_a +
// F2 + beta + beta_eff on the a-grid (fused A4, A5, A6, A7, A8-a).
// Reads visc (previous iteration), writes visc_out.
// src/ice_velocity_module.f90:664-1086
// ---------------------------------------------------------------------------
// Picard step 1a: the horizontal strain-rate invariant on the a-grid, from the
// staggered velocities (A5, src/ice_velocity_module.f90:722-899, with the
// staggered derivatives of src/derivatives_and_grids_module.f90:755-938).  A
// separate kernel so the many boundary variants do not cost k_visc occupancy.
__global__ void __launch_bounds__(IMAU_WG) k_strain(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                    const KP kp, const Fields f, const int *__restrict__ active) {
  TILE_SETUP
  if (j >= m.ny || !active[t.g]) return;
  const int nx = m.nx, ny = m.ny;
  const double dx = m.dx;
  const double *__restrict__ u = f.u;
  const double *__restrict__ v = f.v;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= nx) break;
    const long long p = p2(m, j, i);
    // horizontal strain rates on the a-grid (ddx_cx_to_a, ddy_cy_to_a, ddy_cx_to_a,
    // ddx_cy_to_a; src/derivatives_and_grids_module.f90:755-938)
#define U(jj, ii_) u[p2(m, (jj), (ii_))]
#define V(jj, ii_) v[p2(m, (jj), (ii_))]
    double dudx, dvdy, dudy, dvdx;
    {
      const int ic = iclamp(i, 1, nx - 2);
      dudx = (U(j, ic) - U(j, ic - 1)) / dx;
      const int jc = iclamp(j, 1, ny - 2);
      dvdy = (V(jc, i) - V(jc - 1, i)) / dx;
    }
    const bool iint = (i > 0 && i < nx - 1), jint = (j > 0 && j < ny - 1);
    if (iint && jint) {
      dudy = (U(j + 1, i - 1) + U(j + 1, i) - U(j - 1, i - 1) - U(j - 1, i)) / (4.0 * dx);
      dvdx = (V(j - 1, i + 1) + V(j, i + 1) - V(j - 1, i - 1) - V(j, i - 1)) / (4.0 * dx);
    } else if (iint) {  // j == 0 or ny-1
      if (j == 0) {
        dudy = (U(1, i - 1) + U(1, i) - U(0, i - 1) - U(0, i)) / (4.0 * dx);
        dvdx = (V(0, i + 1) - V(0, i - 1)) / (2.0 * dx);
      } else {
        dudy = (U(ny - 1, i - 1) + U(ny - 1, i) - U(ny - 2, i - 1) - U(ny - 2, i)) / (4.0 * dx);
        dvdx = (V(ny - 2, i + 1) - V(ny - 2, i - 1)) / (2.0 * dx);
      }
    } else if (jint) {  // i == 0 or nx-1
      if (i == 0) {
        dudy = (U(j + 1, 0) - U(j - 1, 0)) / (2.0 * dx);
        dvdx = (V(j - 1, 1) + V(j, 1) - V(j - 1, 0) - V(j, 0)) / (4.0 * dx);
      } else {
        dudy = (U(j + 1, nx - 2) - U(j - 1, nx - 2)) / (2.0 * dx);
        dvdx = (V(j - 1, nx - 1) + V(j, nx - 1) - V(j - 1, nx - 2) - V(j, nx - 2)) / (4.0 * dx);
      }
    } else {  // corners
      const int ic = (i == 0) ? 0 : nx - 2;
      const int jv = (j == 0) ? 0 : ny - 2;
      dudy = (j == 0) ? (U(1, ic) - U(0, ic)) / dx : (U(ny - 1, ic) - U(ny - 2, ic)) / dx;
      dvdx = (i == 0) ? (V(jv, 1) - V(jv, 0)) / dx : (V(jv, nx - 1) - V(jv, nx - 2)) / dx;
    }
#undef U
#undef V
    const double eps_h = dudx * dudx + dvdy * dvdy + dudx * dvdy + 0.25 * (dudy + dvdx) * (dudy + dvdx) + kp.eps0;

    f.eps_h[p] = eps_h;
  }
}

// Per-column constants of the 3D viscosity, shared by both viscosity kernels:
// the basal stresses and calc_staggered_margin choices of the vertical shear
// terms (:691-692, :704-705, :775-788) and the column's strain invariant.
template <class T>
struct ViscCol {
  T tbl, tbr, tcl, tcr, eps_h, HiT, vmin, expo;
  int sl, sr, sb, st, sxa, sya;
  bool noice_l, noice_r, noice_b, noice_t, use_avg_x, use_avg_y, i0, ie, j0, je;
};

template <bool DIVA, class T>
__device__ __forceinline__ ViscCol<T> visc_col(const Member &m, const KP &kp, const Fields &f, int j, int i,
                                               double Hi, double eps_h) {
  const int nx = m.nx, ny = m.ny;
  ViscCol<T> c;
  const int icl = (i == 0) ? 0 : (i == nx - 1 ? nx - 2 : i - 1);  // left cx index used
  const int icr = (i == 0) ? 0 : (i == nx - 1 ? nx - 2 : i);      // right cx index used
  const int jcl = (j == 0) ? 0 : (j == ny - 1 ? ny - 2 : j - 1);
  const int jcr = (j == 0) ? 0 : (j == ny - 1 ? ny - 2 : j);
  c.use_avg_x = (i > 0 && i < nx - 1);
  c.use_avg_y = (j > 0 && j < ny - 1);
  c.i0 = i == 0; c.ie = i == nx - 1; c.j0 = j == 0; c.je = j == ny - 1;
  c.tbl = c.tbr = c.tcl = c.tcr = (T)0.0;
  double Hl0 = 0, Hl1 = 0, Hr0 = 0, Hr1 = 0, Hb0 = 0, Hb1 = 0, Ht0 = 0, Ht1 = 0;
  c.noice_l = c.noice_r = c.noice_b = c.noice_t = true;
  if (DIVA) {
    c.tbl = (T)f.taub_cx[p2(m, j, icl)]; c.tbr = (T)f.taub_cx[p2(m, j, icr)];
    c.tcl = (T)f.taub_cy[p2(m, jcl, i)]; c.tcr = (T)f.taub_cy[p2(m, jcr, i)];
    Hl0 = f.Hi[p2(m, j, icl)]; Hl1 = f.Hi[p2(m, j, icl + 1)];
    Hr0 = f.Hi[p2(m, j, icr)]; Hr1 = f.Hi[p2(m, j, icr + 1)];
    Hb0 = f.Hi[p2(m, jcl, i)]; Hb1 = f.Hi[p2(m, jcl + 1, i)];
    Ht0 = f.Hi[p2(m, jcr, i)]; Ht1 = f.Hi[p2(m, jcr + 1, i)];
    c.noice_l = f.m_ice[p2(m, j, icl)] == 0 && f.m_ice[p2(m, j, icl + 1)] == 0;
    c.noice_r = f.m_ice[p2(m, j, icr)] == 0 && f.m_ice[p2(m, j, icr + 1)] == 0;
    c.noice_b = f.m_ice[p2(m, jcl, i)] == 0 && f.m_ice[p2(m, jcl + 1, i)] == 0;
    c.noice_t = f.m_ice[p2(m, jcr, i)] == 0 && f.m_ice[p2(m, jcr + 1, i)] == 0;
  }
  const double Hxl = c.use_avg_x ? f.Hi[p2(m, j, i - 1)] : 0.0;
  const double Hyl = c.use_avg_y ? f.Hi[p2(m, j - 1, i)] : 0.0;
  // calc_staggered_margin choices, fixed per column: 0 first, 1 second, 2 mean
  c.sl = stag_sel(Hl0, Hl1); c.sr = stag_sel(Hr0, Hr1); c.sb = stag_sel(Hb0, Hb1); c.st = stag_sel(Ht0, Ht1);
  c.sxa = stag_sel(Hxl, Hi); c.sya = stag_sel(Hyl, Hi);
  c.expo = (T)((1.0 - kp.n_flow) / (2.0 * kp.n_flow));
  c.HiT = (T)Hi; c.eps_h = (T)eps_h; c.vmin = (T)kp.visc_min;
  return c;
}

// The viscosity of one level from the previous viscosity at the centre and
// its four neighbours (out-of-grid neighbours 0) and the rate-factor term B
// (:664-721, :775-806).  Multiply-adds are not contracted, so every kernel
// that calls it rounds alike whatever the surrounding code.
// FP32 viscosity arithmetic (the Picard loop's FP32 viscosity and the warm
// start's column completion) takes the hardware reciprocal, log2 and exp2
// (v_rcp_f32, v_log_f32, v_exp_f32: about 1 ulp each) in place of the
// correctly rounded division and cbrt / pow; FP64 keeps exact arithmetic
__device__ __forceinline__ float wdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double wdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float wpow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
__device__ __forceinline__ double wpow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float wcbrt(float x) { return __builtin_amdgcn_exp2f((1.0f / 3.0f) * __builtin_amdgcn_logf(x)); }
__device__ __forceinline__ double wcbrt(double x) { return cbrt(x); }

template <bool DIVA, bool THIRD, class T>
__device__ __forceinline__ T visc_level(const ViscCol<T> &c, T z, T vc, T vw, T ve, T vs, T vn, T B) {
#pragma clang fp contract(off)
  T du_dz_a = (T)0.0, dv_dz_a = (T)0.0;
  if (DIVA) {
    T dl = (T)0.0, dr = (T)0.0, db = (T)0.0, dt = (T)0.0;
    if (!c.noice_l) dl = wdiv(c.tbl, stag_pick<T>(c.i0 ? vc : vw, c.i0 ? ve : vc, c.sl)) * z;
    if (!c.noice_r) dr = wdiv(c.tbr, stag_pick<T>(c.ie ? vw : vc, c.ie ? vc : ve, c.sr)) * z;
    if (!c.noice_b) db = wdiv(c.tcl, stag_pick<T>(c.j0 ? vc : vs, c.j0 ? vn : vc, c.sb)) * z;
    if (!c.noice_t) dt = wdiv(c.tcr, stag_pick<T>(c.je ? vs : vc, c.je ? vc : vn, c.st)) * z;
    du_dz_a = c.use_avg_x ? stag_pick<T>(dl, dr, c.sxa) : dl;
    dv_dz_a = c.use_avg_y ? stag_pick<T>(db, dt, c.sya) : db;
  }
  const T eps_sq = c.eps_h + (T)0.25 * (du_dz_a * du_dz_a + dv_dz_a * dv_dz_a);
  const T eta = THIRD ? wdiv(B, wcbrt(eps_sq)) : B * wpow(eps_sq, c.expo);
  return fmax(eta, c.vmin);  // NaN (cold start 0/0) -> visc_eff_min
}

// trapezoid steps of the viscosity integral and the F2 integrand (H/eta) z^2
// (vertical_integrate, src/utilities_module.f90:73-96; :960-1013)
template <class T>
__device__ __forceinline__ void visc_trap(int k, T z, T zprev, T eta, T HiT, T &vint, T &F2int, T &eta_prev,
                                          T &g_prev) {
#pragma clang fp contract(off)
  const T gk = wdiv(HiT, eta) * (z * z);
  if (k > 0) {
    const T dz = z - zprev;
    vint += (T)0.5 * (eta + eta_prev) * dz;
    F2int += (T)0.5 * (gk + g_prev) * dz;
  }
  eta_prev = eta;
  g_prev = gk;
}

template <bool DIVA, bool THIRD, class T>
__global__ void __launch_bounds__(IMAU_WG) k_visc(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                  const KP kp, const Fields f, const T *__restrict__ visc_in,
                                                  T *__restrict__ visc_out, const T *__restrict__ B3,
                                                  const int *__restrict__ active) {
  TILE_SETUP
  if (j >= m.ny) return;
  const int nx = m.nx, ny = m.ny, nz = kp.nz;
  if (!active[t.g]) {  // finished member: keep its viscosity in the buffer that becomes current
    FOR_VTILE_ROWS {
      const int i = t.i0 + ii;
      if (i >= nx) break;
      for (int k = 0; k < nz; ++k) visc_out[p3(m, k, j, i)] = visc_in[p3(m, k, j, i)];
    }
    return;
  }
  FOR_VTILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= nx) break;
    const long long p = p2(m, j, i);
    const double Hi = f.Hi[p];
    const int mi = f.m_ice[p];
    const double eps_h = f.eps_h[p];  // k_strain

    const bool bc_skip = (kp.margin == 0 && mi == 0);
    T vint = (T)0.0, F2int = (T)0.0;
    if (bc_skip) {
      for (int k = 0; k < nz; ++k) visc_out[p3(m, k, j, i)] = (T)0.0;
    } else {
      const ViscCol<T> cc = visc_col<DIVA, T>(m, kp, f, j, i, Hi, eps_h);
      T eta_prev_k = (T)0.0, g_prev_k = (T)0.0;
      // the previous viscosity at the centre and its four neighbours and the
      // rate-factor term, loaded IMAU_VISC_PF levels ahead into a ring of
      // registers, so that the loads of the next levels are in flight while
      // level k is computed
      constexpr int PF = IMAU_VISC_PF;
      const long long plane = (long long)nx * ny;
      const long long pc = p3(m, 0, j, i);
      const long long ow = (i > 0) ? -(long long)ny : 0, oe = (i < nx - 1) ? (long long)ny : 0;
      const long long os = (j > 0) ? -1 : 0, on = (j < ny - 1) ? 1 : 0;
      T rc[PF], rw[PF], re[PF], rs[PF], rn[PF], rB[PF];
      auto load = [&](int k, int sl) {
        const long long q = pc + k * plane;
        if (DIVA) {
          rc[sl] = visc_in[q];
          rw[sl] = visc_in[q + ow];
          re[sl] = visc_in[q + oe];
          rs[sl] = visc_in[q + os];
          rn[sl] = visc_in[q + on];
        }
        rB[sl] = B3[q];
      };
#pragma unroll
      for (int sl = 0; sl < PF; ++sl)
        if (sl < nz) load(sl, sl);
      auto level = [&](int k, int sl) {
        const T z = (T)kp.zeta[k];
        // out-of-grid neighbours read the centre (an in-bounds address) and are
        // replaced by 0, as the reference's rows never use them
        T vc = (T)0.0, vw = (T)0.0, ve = (T)0.0, vs = (T)0.0, vn = (T)0.0;
        if (DIVA) {
          vc = rc[sl];
          vw = (i > 0) ? rw[sl] : (T)0.0;
          ve = (i < nx - 1) ? re[sl] : (T)0.0;
          vs = (j > 0) ? rs[sl] : (T)0.0;
          vn = (j < ny - 1) ? rn[sl] : (T)0.0;
        }
        const T B = rB[sl];
        if (k + PF < nz) load(k + PF, sl);
        const T eta = visc_level<DIVA, THIRD, T>(cc, z, vc, vw, ve, vs, vn, B);
        visc_out[pc + k * plane] = eta;
        visc_trap<T>(k, z, k > 0 ? (T)kp.zeta[k - 1] : (T)0.0, eta, cc.HiT, vint, F2int, eta_prev_k, g_prev_k);
      };
#pragma unroll 1
      for (int k0 = 0; k0 < nz; k0 += PF)
#pragma unroll
        for (int sl = 0; sl < PF; ++sl)
          if (k0 + sl < nz) level(k0 + sl, sl);
      f.visc_int[p] = (double)vint;
      f.N_a[p] = (double)vint * ((kp.margin == 0) ? Hi : fmax(0.1, Hi));
    }
    // F2 integral (calc_F_integral, :960-1013)
    if (DIVA) f.F2[p] = (mi == 1) ? fmax(kp.F_int_min, (double)F2int) : ((kp.margin == 0) ? 0.0 : kp.F_int_min);
  }
}

// The DIVA viscosity kernel for nz = IMAU_MNZ, marching along i: one
// wavefront per tile of t.pad rows x 64 columns, lane = column j.  The
// previous viscosity of rows i-1, i and i+1 (all levels, plus the columns
// j0-1 and j0+64 loaded by the edge lanes) sits in an LDS ring of three rows,
// and the tile advances one row per step, loading only row i+1: each level of
// visc_in is read once per row plus two halo rows per tile, instead of five
// times (the centre and four neighbours) as in k_visc.  Both kernels compute
// each level with visc_level and visc_trap, so the results are the same bit
// for bit.
template <bool THIRD, class T>
__global__ void __launch_bounds__(64) k_visc_march(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                   const KP kp, const Fields f, const T *__restrict__ visc_in,
                                                   T *__restrict__ visc_out, const T *__restrict__ B3,
                                                   const int *__restrict__ active) {
  constexpr int NZ = IMAU_MNZ, W = 66;
  __shared__ T ring[3][NZ][W];
  const Tile t = tiles[imau_xcd_tile()];
  const Member m = mem[t.g];
  const int lane = (int)threadIdx.x;
  const int j = t.j0 + lane;
  const int nx = m.nx, ny = m.ny;
  const bool jin = j < ny;
  const int i1 = min(t.i0 + t.pad, nx);  // t.pad: the tile's rows (the autotuned tile height)
  if (!active[t.g]) {  // finished member: keep its viscosity in the buffer that becomes current
    if (!jin) return;
    for (int i = t.i0; i < i1; ++i)
      for (int k = 0; k < NZ; ++k) visc_out[p3(m, k, j, i)] = visc_in[p3(m, k, j, i)];
    return;
  }
  // row r of visc_in into ring slot s: own column (0 beyond the grid) and the halo columns
  auto load_row = [&](int r, int sl) {
    const bool rin = r >= 0 && r < nx;
    for (int k = 0; k < NZ; ++k) {
      ring[sl][k][lane + 1] = (rin && jin) ? visc_in[p3(m, k, j, r)] : (T)0.0;
      if (lane == 0) ring[sl][k][0] = (rin && j > 0) ? visc_in[p3(m, k, j - 1, r)] : (T)0.0;
      if (lane == 63) ring[sl][k][W - 1] = (rin && j + 1 < ny) ? visc_in[p3(m, k, j + 1, r)] : (T)0.0;
    }
  };
  int sp = 0, sc = 1, sn = 2;
  load_row(t.i0 - 1, sp);
  load_row(t.i0, sc);
#pragma unroll 1
  for (int i = t.i0; i < i1; ++i) {
    load_row(i + 1, sn);
    __syncthreads();
    if (jin) {
      const long long p = p2(m, j, i);
      const double Hi = f.Hi[p];
      const int mi = f.m_ice[p];
      const double eps_h = f.eps_h[p];
      const bool bc_skip = (kp.margin == 0 && mi == 0);
      const ViscCol<T> cc = visc_col<true, T>(m, kp, f, j, i, Hi, eps_h);
      T vint = (T)0.0, F2int = (T)0.0, eta_prev_k = (T)0.0, g_prev_k = (T)0.0;
#pragma unroll 1
      for (int k = 0; k < NZ; ++k) {
        const T vc = ring[sc][k][lane + 1], vw = ring[sp][k][lane + 1], ve = ring[sn][k][lane + 1];
        const T vs = ring[sc][k][lane], vn = ring[sc][k][lane + 2];
        const T z = (T)kp.zeta[k];
        T eta = visc_level<true, THIRD, T>(cc, z, vc, vw, ve, vs, vn, B3[p3(m, k, j, i)]);
        if (bc_skip) eta = (T)0.0;
        visc_out[p3(m, k, j, i)] = eta;
        visc_trap<T>(k, z, k > 0 ? (T)kp.zeta[k - 1] : (T)0.0, eta, cc.HiT, vint, F2int, eta_prev_k, g_prev_k);
      }
      if (!bc_skip) {
        f.visc_int[p] = (double)vint;
        f.N_a[p] = (double)vint * ((kp.margin == 0) ? Hi : fmax(0.1, Hi));
      }
      f.F2[p] = (mi == 1) ? fmax(kp.F_int_min, (double)F2int) : ((kp.margin == 0) ? 0.0 : kp.F_int_min);
    }
    __syncthreads();  // slot sp is overwritten by the next step's load
    const int s0 = sp;
    sp = sc;
    sc = sn;
    sn = s0;
  }
}

// ---------------------------------------------------------------------------
// Picard step 1b: basal friction on the a-grid (calc_sliding_term_beta,
// :900-959) and beta_eff (calc_beta_eff, :1014-1062), after k_visc has written
// F2.  A separate kernel so the register-hungry sliding laws do not limit the
// occupancy of the 3D viscosity kernel.
// ---------------------------------------------------------------------------
// SLID >= 0: the kernel carries only that sliding law (kp.slid == SLID)
template <bool DIVA, int SLID>
__global__ void __launch_bounds__(IMAU_WG) k_beta(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                  const KP kp, const Fields f, const double *__restrict__ xg,
                                                  const double *__restrict__ yg, const int *__restrict__ active) {
  TILE_SETUP
  if (j >= m.ny || !active[t.g]) return;
  const int nx = m.nx, ny = m.ny, nz = kp.nz;
  const int slid = SLID >= 0 ? SLID : kp.slid;
  const double *__restrict__ u = f.u;
  const double *__restrict__ v = f.v;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= nx) break;
    const long long p = p2(m, j, i);
    // velocities on the a-grid for the sliding law (map_cx_to_a / map_cy_to_a, :1579-1644)
#define U(jj, ii_) u[p2(m, (jj), (ii_))]
#define V(jj, ii_) v[p2(m, (jj), (ii_))]
    const double ua = (i == 0) ? U(j, 0) : (i == nx - 1 ? U(j, nx - 2) : 0.5 * (U(j, i - 1) + U(j, i)));
    const double va = (j == 0) ? V(0, i) : (j == ny - 1 ? V(ny - 2, i) : 0.5 * (V(j - 1, i) + V(j, i)));
#undef U
#undef V
    const double F2 = DIVA ? f.F2[p] : 0.0;
    // basal friction (calc_sliding_term_beta, :900-959)
    double tauc = f.tauc[p];
    if (slid == 3 || (slid == 2 && !(kp.rough == 1 && kp.prough == 1))) {
      tauc = tan((D_PI / 180.0) * f.phi_fric[p]) * f.Neff[p];
      f.tauc[p] = tauc;
    }
    double beta = sliding_beta(kp, slid, ua, va, f.beta_sq[p], f.alpha_sq[p], f.phi_fric[p], tauc, f.Neff[p],
                               xg[m.offx + i], yg[m.offy + j], f.A_vav ? f.A_vav[p] : 0.0,
                               (slid == 1 && kp.ideal == 2) ? f.ismip_e[m.offx + i] : 0.0);
    beta = fmin(beta, kp.beta_max);
    if (!kp.glsub && f.m_ocean[p] == 1) beta = 0.0;
    f.beta_a[p] = beta;
    // beta_eff (calc_beta_eff, :1014-1062); SSA uses beta itself (:193)
    double be;
    if (DIVA) {
      if (slid == 0) be = 1.0 / F2;
      else {
        be = beta / (1.0 + beta * F2);
        if (kp.freeze) {
          const long long qb = p3(m, nz - 1, j, i);
          if (f.Tpmp3[qb] - f.Ti3[qb] > kp.dT_freeze) be = 1.0 / F2;
        }
      }
    } else {
      be = beta;
    }
    f.beta_eff_a[p] = be;
  }
}

// ---------------------------------------------------------------------------
// Column-local completion of a velocity-only warm start (DIVA).  Per a-grid
// column, at the given velocities: viscosity with the vertical shear of the
// column's own basal stress (A4, A5), F2 (A7), the sliding law and beta_eff
// (A6, A8) and |tau_b| = beta_eff |u_a| (A9), iterated `iters` times in
// registers from the shear-free viscosity.  The Picard loop's staggered forms
// (tau_b and viscosity averaged onto cx / cy, :664-721) are replaced by the
// column's own values; its first iteration restores them.  Writes the
// viscosity (in place), visc_int, N_a, F2, beta_a and beta_eff_a.
// ---------------------------------------------------------------------------

// SLID >= 0: the kernel carries only that sliding law.  NZ > 0: the level
// count is fixed at compile time and the column's viscosity is kept in LDS
// across the iterations (one slot per level and thread, level-major, so the
// lanes of a wavefront touch consecutive words); NZ = 0 reads kp.nz and keeps
// it in a private array.
template <bool THIRD, class T, int SLID, int NZ>
__global__ void __launch_bounds__(IMAU_WG) k_visc_column(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                         const KP kp, const Fields f, T *__restrict__ visc,
                                                         const T *__restrict__ B3, const double *__restrict__ xg,
                                                         const double *__restrict__ yg, int iters,
                                                         const int *__restrict__ active) {
  TILE_SETUP
  if (j >= m.ny || !active[t.g]) return;
  const int nx = m.nx, ny = m.ny;
  const int nz = NZ > 0 ? NZ : kp.nz;
  const int slid = SLID >= 0 ? SLID : kp.slid;
  __shared__ T s_vk[NZ > 0 ? NZ : 1][IMAU_WG];
  T p_vk[NZ > 0 ? 1 : IMAU_NZMAX];
  const int tid = threadIdx.y * IMAU_TJ + threadIdx.x;
  auto vk = [&](int k) -> T & { return NZ > 0 ? s_vk[NZ > 0 ? k : 0][tid] : p_vk[NZ > 0 ? 0 : k]; };
  FOR_VTILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= nx) break;
    const long long p = p2(m, j, i);
    const double Hi = f.Hi[p];
    const int mi = f.m_ice[p];
    const double ua = (i == 0) ? f.u[p2(m, j, 0)]
                      : (i == nx - 1 ? f.u[p2(m, j, nx - 2)] : 0.5 * (f.u[p2(m, j, i - 1)] + f.u[p]));
    const double va = (j == 0) ? f.v[p2(m, 0, i)]
                      : (j == ny - 1 ? f.v[p2(m, ny - 2, i)] : 0.5 * (f.v[p2(m, j - 1, i)] + f.v[p]));
    double tauc = f.tauc[p];
    if (slid == 3 || (slid == 2 && !(kp.rough == 1 && kp.prough == 1)))
      tauc = tan((D_PI / 180.0) * f.phi_fric[p]) * f.Neff[p];
    double beta = sliding_beta(kp, slid, ua, va, f.beta_sq[p], f.alpha_sq[p], f.phi_fric[p], tauc, f.Neff[p],
                               xg[m.offx + i], yg[m.offy + j], f.A_vav ? f.A_vav[p] : 0.0,
                               (slid == 1 && kp.ideal == 2) ? f.ismip_e[m.offx + i] : 0.0);
    beta = fmin(beta, kp.beta_max);
    if (!kp.glsub && f.m_ocean[p] == 1) beta = 0.0;
    bool frozen = false;
    if (kp.freeze) {
      const long long qb = p3(m, nz - 1, j, i);
      frozen = f.Tpmp3[qb] - f.Ti3[qb] > kp.dT_freeze;
    }
    const double speed = sqrt(ua * ua + va * va);
    const bool bc_skip = (kp.margin == 0 && mi == 0);
    const T expo = (T)((1.0 - kp.n_flow) / (2.0 * kp.n_flow));
    const T HiT = (T)Hi, eps_h = (T)f.eps_h[p], vmin = (T)kp.visc_min;
    double F2 = 0.0, be = 0.0, vint_out = 0.0;
    T tb = (T)0.0;
#pragma unroll 1
    for (int it = 0; it < iters; ++it) {
      T vint = (T)0.0, F2int = (T)0.0, eta_prev = (T)0.0, g_prev = (T)0.0;
#pragma unroll 1
      for (int k = 0; k < nz; ++k) {
        const T z = (T)kp.zeta[k];
        if (bc_skip) { vk(k) = (T)0.0; continue; }
        const T du = (it == 0) ? (T)0.0 : wdiv(tb * z, vk(k));
        const T eps_sq = eps_h + (T)0.25 * du * du;
        const T B = B3[p3(m, k, j, i)];
        T eta = B * wpow(eps_sq, THIRD ? (T)(-1.0 / 3.0) : expo);
        eta = fmax(eta, vmin);
        const T gk = wdiv(HiT, eta) * (z * z);
        if (k > 0) {
          const T dz = z - (T)kp.zeta[k - 1];
          vint += (T)0.5 * (eta + eta_prev) * dz;
          F2int += (T)0.5 * (gk + g_prev) * dz;
        }
        eta_prev = eta;
        g_prev = gk;
        vk(k) = eta;
      }
      vint_out = (double)vint;
      F2 = (mi == 1) ? fmax(kp.F_int_min, (double)F2int) : ((kp.margin == 0) ? 0.0 : kp.F_int_min);
      be = (slid == 0 || frozen) ? 1.0 / F2 : beta / (1.0 + beta * F2);
      tb = (T)(be * speed);
    }
    for (int k = 0; k < nz; ++k) visc[p3(m, k, j, i)] = vk(k);
    if (!bc_skip) {
      f.visc_int[p] = vint_out;
      f.N_a[p] = vint_out * ((kp.margin == 0) ? Hi : fmax(0.1, Hi));
    }
    f.F2[p] = F2;
    f.beta_a[p] = beta;
    f.beta_eff_a[p] = be;
  }
}

// ---------------------------------------------------------------------------
// Picard step 2: staggered coefficients and right-hand sides
// N_b (:820-874), beta_eff on cx/cy with f_grnd^2 (:1063-1076, SSA :195-206),
// N_cx/N_cy for the sans-cross-terms form (:876-880), rhs of every row.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(IMAU_WG) k_coef(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                  const KP kp, const Fields f, int diva,
                                                  const int *__restrict__ active) {
  TILE_SETUP
  if (j >= m.ny || !active[t.g]) return;
  const int nx = m.nx, ny = m.ny;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= nx) break;
    const long long p = p2(m, j, i);
    if (i < nx - 1 && j < ny - 1) {
      const long long pe = p2(m, j, i + 1), pn = p2(m, j + 1, i), pne = p2(m, j + 1, i + 1);
      double vb;
      if (kp.margin == 0) {
        double s = 0.0, w = 0.0;
        if (f.m_ice[p] == 1) { s += f.visc_int[p]; w += 1.0; }
        if (f.m_ice[pe] == 1) { s += f.visc_int[pe]; w += 1.0; }
        if (f.m_ice[pn] == 1) { s += f.visc_int[pn]; w += 1.0; }
        if (f.m_ice[pne] == 1) { s += f.visc_int[pne]; w += 1.0; }
        vb = (w > 0.0) ? s / w : 0.0;
        f.N_b[p] = vb * f.Hi_b[p];
      } else {
        vb = (f.visc_int[p] + f.visc_int[pe] + f.visc_int[pn] + f.visc_int[pne]) / 4.0;
        f.N_b[p] = vb * fmax(0.1, f.Hi_b[p]);
      }
    }
    const bool use_fg = diva ? (kp.glsub != 0) : true;
    if (i < nx - 1) {
      const long long pe = p2(m, j, i + 1);
      double b = (f.beta_eff_a[p] + f.beta_eff_a[pe]) / 2.0;
      if (use_fg) b *= f.fg_cx[p] * f.fg_cx[p];
      f.beta_eff_cx[p] = b;
      if (!kp.cross) f.N_cx[p] = (f.N_a[p] + f.N_a[pe]) / 2.0;
    }
    if (j < ny - 1) {
      const long long pn = p2(m, j + 1, i);
      double b = (f.beta_eff_a[p] + f.beta_eff_a[pn]) / 2.0;
      if (use_fg) b *= f.fg_cy[p] * f.fg_cy[p];
      f.beta_eff_cy[p] = b;
      if (!kp.cross) f.N_cy[p] = (f.N_a[p] + f.N_a[pn]) / 2.0;
    }
  }
}

// ---------------------------------------------------------------------------
// The linearised staggered DIVA/SSA rows (src/ice_velocity_module.f90:1762-2589),
// matrix-free.  For the unknown at (j,i) they return the off-diagonal sum S,
// the diagonal D and the right-hand side b, so that the row reads S + D x = b.
// ---------------------------------------------------------------------------
using CoefP = CoefT<double>;
using GAcc = GAccT<double>;

struct Vec2 {
  double *u, *v;
};
struct CVec2 {
  const double *u, *v;
};

// LDS tile of the sweep (u in [0], v in [1]) with a one-cell halo
__shared__ double g_sx[2][IMAU_TI + 2][IMAU_TJ + 2];

// LDS-tile accessor (the rows evaluated through it never leave the tile halo;
// periodic boundary rows, which do, are handled separately in the sweep)
template <int COMP>
struct LAcc {
  int i0, j0;
  __device__ __forceinline__ double operator()(int jj, int ii) const { return g_sx[COMP][ii - i0 + 1][jj - j0 + 1]; }
};

// CROSS: 1 / 0 fixes the cross-term choice at compile time, -1 reads kp.cross
template <int CROSS, class AU, class AV>
__device__ __forceinline__ void u_row_x(const Member &m, const KP &kp, const CoefP &c, int j, int i, uint8_t rt,
                                        const AU &U, const AV &V, double &S, double &D, double &b) {
  u_sd_x<CROSS, double>(m, kp, c, j, i, rt, U, V, S, D);
  if (rt == RT_FREE) {
    const double dx2 = m.dx * m.dx;
    const bool cross = CROSS >= 0 ? CROSS != 0 : kp.cross != 0;
    b = cross ? -c.taudx[p2(m, j, i)] * dx2 : -c.taudx[p2(m, j, i)] * dx2 / c.Ncx[p2(m, j, i)];
  } else if (rt == RT_MARGIN_I || rt == RT_MARGIN_O) {
    b = c.frhs[p2(m, j, (rt == RT_MARGIN_I) ? i : i + 1)];
  } else {
    b = 0.0;
  }
}

template <int CROSS, class AU, class AV>
__device__ __forceinline__ void v_row_x(const Member &m, const KP &kp, const CoefP &c, int j, int i, uint8_t rt,
                                        const AU &U, const AV &V, double &S, double &D, double &b) {
  v_sd_x<CROSS, double>(m, kp, c, j, i, rt, U, V, S, D);
  if (rt == RT_FREE) {
    const double dx2 = m.dx * m.dx;
    const bool cross = CROSS >= 0 ? CROSS != 0 : kp.cross != 0;
    b = cross ? -c.taudy[p2(m, j, i)] * dx2 : -c.taudy[p2(m, j, i)] * dx2 / c.Ncy[p2(m, j, i)];
  } else if (rt == RT_MARGIN_I || rt == RT_MARGIN_O) {
    b = c.frhs[p2(m, (rt == RT_MARGIN_I) ? j : j + 1, i)];
  } else {
    b = 0.0;
  }
}

template <class AU, class AV>
__device__ __forceinline__ void u_row(const Member &m, const KP &kp, const CoefP &c, int j, int i, uint8_t rt,
                                      const AU &U, const AV &V, double &S, double &D, double &b) {
  u_row_x<-1>(m, kp, c, j, i, rt, U, V, S, D, b);
}
template <class AU, class AV>
__device__ __forceinline__ void v_row(const Member &m, const KP &kp, const CoefP &c, int j, int i, uint8_t rt,
                                      const AU &U, const AV &V, double &S, double &D, double &b) {
  v_row_x<-1>(m, kp, c, j, i, rt, U, V, S, D, b);
}



#undef su
#undef sv

// ---------------------------------------------------------------------------
// Picard step 3 (after the linear solve): velocity limits (:1432-1498),
// relaxation (:1499-1529), Picard residual partial sums (:1530-1601),
// basal stress (:1087-1120) and basal velocity (:1121-1199).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(IMAU_WG) k_post(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                  const KP kp, const Fields f, const double *__restrict__ xu,
                                                  const double *__restrict__ xv, const int *__restrict__ active,
                                                  int diva, double *__restrict__ partial,
                                                  const double *__restrict__ relax, AMix mix) {
  TILE_SETUP
  double acc[3] = {0.0, 0.0, 0.0};
  const int nx = m.nx, ny = m.ny;
  const double rel = relax ? relax[t.g] : kp.relax, vmax = kp.vel_max;
  // Anderson mixing: u_new = sum_s (a_s U_s + b_s F_s) over the stored slots
  double wa[IMAU_AND_SLOTS], wb[IMAU_AND_SLOTS];
  if (mix.w)
#pragma unroll
    for (int q = 0; q < IMAU_AND_SLOTS; ++q) {
      wa[q] = mix.w[t.g * 2 * IMAU_AND_SLOTS + q];
      wb[q] = mix.w[t.g * 2 * IMAU_AND_SLOTS + IMAU_AND_SLOTS + q];
    }
  if (j < ny && active[t.g]) {
    FOR_TILE_ROWS {
      const int i = t.i0 + ii;
      if (i >= nx) break;
      const long long p = p2(m, j, i);
      if (i < nx - 1) {
        double x = xu[p];
        // map_cy_to_cx of the solved v, :1757-1797
        double vcx;
        if (j == 0) vcx = (xv[p2(m, 0, i)] + xv[p2(m, 0, i + 1)]) / 2.0;
        else if (j == ny - 1) vcx = (xv[p2(m, ny - 2, i)] + xv[p2(m, ny - 2, i + 1)]) / 2.0;
        else vcx = (xv[p2(m, j - 1, i)] + xv[p2(m, j - 1, i + 1)] + xv[p2(m, j, i)] + xv[p2(m, j, i + 1)]) / 4.0;
        const double ua = sqrt(x * x + vcx * vcx);
        if (ua > vmax) x = x * vmax / ua;
        const double uo = f.u[p];
        double un = rel * x + (1.0 - rel) * uo;
        if (mix.w) {
          un = 0.0;
#pragma unroll
          for (int q = 0; q < IMAU_AND_SLOTS; ++q)
            if (q < mix.ns) un += wa[q] * mix.U.u[q][p] + wb[q] * mix.F.u[q][p];
        }
        f.u[p] = un;
        f.u_prev[p] = uo;
        if (f.errt_u)  // estimate_visc_iter_UV_errors, :1626-1631
          f.errt_u[p] = fabs(diva ? un : f.chk_u[p]) > 1e-6 ? 2.0 * fabs(un - uo) / fabs(un + uo + 1e-5) : 0.0;
        if (fabs(un) > 1e-6) {
          acc[0] += (un - uo) * (un - uo);
          acc[1] += (un + uo) * (un + uo);
          acc[2] += 1.0;
        }
        if (diva) {
          const double tb = f.beta_eff_cx[p] * un;
          f.taub_cx[p] = tb;
          double ub;
          if (kp.slid == 0) ub = 0.0;
          else {
            const double F2s = stag_margin(f.F2[p], f.F2[p2(m, j, i + 1)], f.Hi[p], f.Hi[p2(m, j, i + 1)]);
            ub = un - tb * F2s;
            if (kp.freeze) {
              const long long qb = p3(m, kp.nz - 1, j, i);
              if (f.Tpmp3[qb] - f.Ti3[qb] > kp.dT_freeze) ub = 0.0;
            }
          }
          f.u_base[p] = ub;
        }
      }
      if (j < ny - 1) {
        double x = xv[p];
        // map_cx_to_cy of the solved u, :1718-1756
        double ucy;
        if (i == 0) ucy = (xu[p2(m, j, 0)] + xu[p2(m, j + 1, 0)]) / 2.0;
        else if (i == nx - 1) ucy = (xu[p2(m, j, nx - 2)] + xu[p2(m, j + 1, nx - 2)]) / 2.0;
        else ucy = (xu[p2(m, j, i - 1)] + xu[p2(m, j, i)] + xu[p2(m, j + 1, i - 1)] + xu[p2(m, j + 1, i)]) / 4.0;
        const double ua = sqrt(ucy * ucy + x * x);
        if (ua > vmax) x = x * vmax / ua;
        const double vo = f.v[p];
        double vn = rel * x + (1.0 - rel) * vo;
        if (mix.w) {
          vn = 0.0;
#pragma unroll
          for (int q = 0; q < IMAU_AND_SLOTS; ++q)
            if (q < mix.ns) vn += wa[q] * mix.U.v[q][p] + wb[q] * mix.F.v[q][p];
        }
        f.v[p] = vn;
        f.v_prev[p] = vo;
        if (f.errt_v)  // :1633-1638
          f.errt_v[p] = fabs(diva ? vn : f.chk_v[p]) > 1e-6 ? 2.0 * fabs(vn - vo) / fabs(vn + vo + 1e-5) : 0.0;
        if (fabs(vn) > 1e-6) {
          acc[0] += (vn - vo) * (vn - vo);
          acc[1] += (vn + vo) * (vn + vo);
          acc[2] += 1.0;
        }
        if (diva) {
          const double tb = f.beta_eff_cy[p] * vn;
          f.taub_cy[p] = tb;
          double vb;
          if (kp.slid == 0) vb = 0.0;
          else {
            const double F2s = stag_margin(f.F2[p], f.F2[p2(m, j + 1, i)], f.Hi[p], f.Hi[p2(m, j + 1, i)]);
            vb = vn - tb * F2s;
            if (kp.freeze) {
              const long long qb = p3(m, kp.nz - 1, j, i);
              if (f.Tpmp3[qb] - f.Ti3[qb] > kp.dT_freeze) vb = 0.0;
            }
          }
          f.v_base[p] = vb;
        }
      }
    }
  }
  block_sum_store<3>(acc, partial + 3 * (long long)tb);
}

// Basal stress at the current velocities, tau_b = beta_eff u (:1087-1120),
// without a linear solve: the viscosity passes that complete a warm start
// (given velocities, cold viscosity and basal stress) before the Picard loop.
// Reads beta_eff on cx / cy, which k_coef writes.
__global__ void __launch_bounds__(IMAU_WG) k_taub(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                  const Fields f, const int *__restrict__ active) {
  TILE_SETUP
  if (j >= m.ny || !active[t.g]) return;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    if (i < m.nx - 1) f.taub_cx[p] = f.beta_eff_cx[p] * f.u[p];
    if (j < m.ny - 1) f.taub_cy[p] = f.beta_eff_cy[p] * f.v[p];
  }
}

// reset the flagged members to the state the problem was created with (the
// stall restart of a warm-started Picard loop): velocities, basal stress and
// the 3D viscosity (also its FP32 copy when the viscosity runs in FP32)
__global__ void __launch_bounds__(IMAU_WG) k_reset_members(const Member *__restrict__ mem,
                                                           const Tile *__restrict__ tiles, const Fields f,
                                                           imau::ResetSrc src, double *__restrict__ visc,
                                                           float *__restrict__ visc32, int nz,
                                                           const int *__restrict__ flag) {
  TILE_SETUP
  if (j >= m.ny || !flag[t.g]) return;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    f.u[p] = src.u[p];
    f.v[p] = src.v[p];
    f.taub_cx[p] = src.tbx[p];
    f.taub_cy[p] = src.tby[p];
    for (int k = 0; k < nz; ++k) {
      const long long q = p3(m, k, j, i);
      visc[q] = src.visc[q];
      if (visc32) visc32[q] = (float)src.visc[q];
    }
  }
}

// dst = src on the 3D slabs of the members with flag[g] != 0
template <class T>
__global__ void __launch_bounds__(IMAU_WG) k_copy3_members(const Member *__restrict__ mem,
                                                           const Tile *__restrict__ tiles, const T *__restrict__ src,
                                                           T *__restrict__ dst, int nz, const int *__restrict__ flag) {
  TILE_SETUP
  if (j >= m.ny || !flag[t.g]) return;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    for (int k = 0; k < nz; ++k) dst[p3(m, k, j, i)] = src[p3(m, k, j, i)];
  }
}

// dst = src (u and v slabs) on the members with flag[g] != 0
__global__ void __launch_bounds__(IMAU_WG) k_copy2_members(const Member *__restrict__ mem,
                                                           const Tile *__restrict__ tiles, CVec2 src, Vec2 dst,
                                                           const int *__restrict__ flag) {
  TILE_SETUP
  if (j >= m.ny || !flag[t.g]) return;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    dst.u[p] = src.u[p];
    dst.v[p] = src.v[p];
  }
}

// ---------------------------------------------------------------------------
// warm-start network I/O: image g of the network is member g's a-grid with
// rows i (x) and columns j (y), pixel off[g] + i ny + j
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned short f2bf_rne(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// input features, BF16 [pix][16]: H/1e3, Hs/1e3, Hb/1e3, 100 ds/dx, 100 ds/dy,
// log10(beta^2) - 4, log10(A) + 16.5, ice mask, and asinh(u_est / 10 m/yr) for
// both components of the local estimate u_est = -(u_b + u_d) grad s / |grad s|
// with Weertman sliding u_b = (tau_d / beta^2)^m and the SIA deformation
// velocity u_d = 2 A tau_d^n H / (n + 2) under tau_d = rho g H |grad s|
__global__ void __launch_bounds__(IMAU_WG) k_net_features(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                          const KP kp, const Fields f,
                                                          const long long *__restrict__ poff,
                                                          unsigned short *__restrict__ x) {
  TILE_SETUP
  if (j >= m.ny) return;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    const long long q = poff[t.g] + (long long)i * m.ny + j;
    const double H = f.Hi[p], sx = f.dHs_dx_a[p], sy = f.dHs_dy_a[p];
    const double gs = sqrt(sx * sx + sy * sy);
    const double taud = D_ICE_DENSITY * D_GRAV * H * gs;
    const double b2 = f.beta_sq ? f.beta_sq[p] : 0.0, A = f.A_vav[p];
    double speed = 0.0;
    if (H > 0.0 && gs > 0.0) {
      if (b2 > 0.0) speed += pow(taud / b2, kp.wm);
      if (A > 0.0) speed += 2.0 * A * pow(taud, kp.n_flow) * H / (kp.n_flow + 2.0);
    }
    speed = fmin(speed, kp.vel_max);
    const double ex = gs > 0.0 ? -speed * sx / gs : 0.0, ey = gs > 0.0 ? -speed * sy / gs : 0.0;
    float v[16];
    v[0] = (float)(H * 1e-3);
    v[1] = (float)(f.Hs[p] * 1e-3);
    v[2] = (float)(f.Hb[p] * 1e-3);
    v[3] = (float)(100.0 * sx);
    v[4] = (float)(100.0 * sy);
    v[5] = b2 > 0.0 ? (float)(log10(b2) - 4.0) : 0.0f;
    v[6] = A > 0.0 ? (float)(log10(A) + 16.5) : 0.0f;
    v[7] = H > 0.0 ? 1.0f : 0.0f;
    v[8] = (float)asinh(ex / 10.0);
    v[9] = (float)asinh(ey / 10.0);
#pragma unroll
    for (int k = 10; k < 16; ++k) v[k] = 0.0f;
    // four 8-byte stores per pixel (the 32 bytes of its channels)
    typedef unsigned short us4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int k = 0; k < 16; k += 4)
      *reinterpret_cast<us4 *>(x + q * 16 + k) = us4{f2bf_rne(v[k]), f2bf_rne(v[k + 1]), f2bf_rne(v[k + 2]),
                                                     f2bf_rne(v[k + 3])};
  }
}

// training targets: asinh(u_a / 10 m/yr), asinh(v_a / 10 m/yr) of the a-grid
// velocities (mean of the two staggered neighbours, the one inside at the
// domain edge), weight 1 on ice and 0 elsewhere
__global__ void __launch_bounds__(IMAU_WG) k_net_targets(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                         const Fields f, const long long *__restrict__ poff,
                                                         float *__restrict__ tg, float *__restrict__ wt) {
  TILE_SETUP
  if (j >= m.ny) return;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    const long long q = poff[t.g] + (long long)i * m.ny + j;
    const double ua = (i == 0) ? f.u[p2(m, j, 0)]
                      : (i == m.nx - 1) ? f.u[p2(m, j, m.nx - 2)] : 0.5 * (f.u[p2(m, j, i - 1)] + f.u[p]);
    const double va = (j == 0) ? f.v[p2(m, 0, i)]
                      : (j == m.ny - 1) ? f.v[p2(m, m.ny - 2, i)] : 0.5 * (f.v[p2(m, j - 1, i)] + f.v[p]);
    tg[2 * q] = (float)asinh(ua / 10.0);
    tg[2 * q + 1] = (float)asinh(va / 10.0);
    wt[q] = f.Hi[p] > 0.0 ? 1.0f : 0.0f;
  }
}

// warm start from the network output o[pix][2] (asinh scale): a-grid
// velocities 10 sinh(o), mapped to cx / cy by the mean of the two neighbours
// and limited to vel_max
__global__ void __launch_bounds__(IMAU_WG) k_net_warm(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                      const KP kp, const Fields f, const long long *__restrict__ poff,
                                                      const float *__restrict__ o, const int *__restrict__ active) {
  TILE_SETUP
  if (j >= m.ny || !active[t.g]) return;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    const long long q = poff[t.g] + (long long)i * m.ny + j;
    const double vm = kp.vel_max;
    if (i < m.nx - 1) {
      const double u = 5.0 * (sinh((double)o[2 * q]) + sinh((double)o[2 * (q + m.ny)]));
      f.u[p] = fmax(-vm, fmin(vm, u));
    }
    if (j < m.ny - 1) {
      const double v = 5.0 * (sinh((double)o[2 * q + 1]) + sinh((double)o[2 * (q + 1) + 1]));
      f.v[p] = fmax(-vm, fmin(vm, v));
    }
  }
}

// FP64 residual partial sums: sum (S + D x - b)^2 and sum b^2 over every row
// of the system assembled at the current coefficients.
__global__ void __launch_bounds__(IMAU_WG) k_resid(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                   const KP kp, const CoefP c, const double *__restrict__ xu,
                                                   const double *__restrict__ xv, double *__restrict__ partial) {
  TILE_SETUP
  double acc[2] = {0.0, 0.0};
  const int nx = m.nx, ny = m.ny;
  const GAcc U{xu, m}, V{xv, m};
  if (j < ny) {
    FOR_TILE_ROWS {
      const int i = t.i0 + ii;
      if (i >= nx) break;
      const long long p = p2(m, j, i);
      double S, D, b;
      if (i < nx - 1) {
        u_row(m, kp, c, j, i, c.rt_u[p], U, V, S, D, b);
        const double r = S + D * xu[p] - b;
        acc[0] += r * r;
        acc[1] += b * b;
      }
      if (j < ny - 1) {
        v_row(m, kp, c, j, i, c.rt_v[p], U, V, S, D, b);
        const double r = S + D * xv[p] - b;
        acc[0] += r * r;
        acc[1] += b * b;
      }
    }
  }
  block_sum_store<2>(acc, partial + 2 * (long long)tb);
}

// The state seen through a row.  Pass 0 reads the state and notes whether the
// row read a zero row's unknown (where the start vector x0 differs from the
// state); pass 1 reads x0, which is the state with the zero rows set to 0.
struct RAcc {
  const double *x;
  const uint8_t *rt;
  Member m;
  bool *hit;
  bool x0;
  __device__ __forceinline__ double operator()(int jj, int ii) const {
    const long long p = p2(m, jj, ii);
    const uint8_t t = rt[p];
    const bool z = (t == RT_ZERO || t == RT_BND_ZERO);
    if (z) *hit = true;
    return (z && x0) ? 0.0 : x[p];
  }
};

// one row of k_resid_init: adds |row(u) - b|^2, b^2 and |b - row(x0)|^2 to acc,
// writes x0 and r = b - row(x0); the row is evaluated a second time only when
// it reads a zero row
template <int CROSS, bool ISU>
__device__ __forceinline__ void resid_init_row(const Member &m, const KP &kp, const CoefP &c, int j, int i,
                                               long long p, const double *xs, RAcc &U, RAcc &V, double *x0,
                                               double *r, double acc[3]) {
  const uint8_t rt = ISU ? c.rt_u[p] : c.rt_v[p];
  const double x = (rt == RT_ZERO || rt == RT_BND_ZERO) ? 0.0 : xs[p];
  bool hit = false;
  U.hit = V.hit = &hit;
  double rl = 0.0;
#pragma unroll 1
  for (int pass = 0; pass < 2; ++pass) {
    U.x0 = V.x0 = pass == 1;
    double S, D, b;
    if (ISU) u_row_x<CROSS>(m, kp, c, j, i, rt, U, V, S, D, b);
    else v_row_x<CROSS>(m, kp, c, j, i, rt, U, V, S, D, b);
    if (pass == 0) {
      const double rn = S + D * xs[p] - b;
      acc[0] += rn * rn;
      acc[1] += b * b;
    }
    rl = b - (S + D * x);
    if (!hit) break;
  }
  x0[p] = x;
  r[p] = rl;
  acc[2] += rl * rl;
}

// The FP64 nonlinear residual at the current state (as k_resid) fused with the
// start of the linear solve: x0 = u, v with the zero rows set to 0 (k_init_x)
// and r = b - A x0.  Partial sums per tile: |A u - b|^2, |b|^2, |r|^2.
template <int CROSS>
__global__ void __launch_bounds__(IMAU_WG) k_resid_init(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                        const KP kp, const CoefP c, const double *__restrict__ xu,
                                                        const double *__restrict__ xv, Vec2 x0, Vec2 r,
                                                        const int *__restrict__ active, double *__restrict__ partial) {
  TILE_SETUP
  double acc[3] = {0.0, 0.0, 0.0};
  const int nx = m.nx, ny = m.ny;
  bool hit = false;
  RAcc U{xu, c.rt_u, m, &hit, false}, V{xv, c.rt_v, m, &hit, false};
  if (j < ny && active[t.g]) {
    FOR_TILE_ROWS {
      const int i = t.i0 + ii;
      if (i >= nx) break;
      const long long p = p2(m, j, i);
      if (i < nx - 1) resid_init_row<CROSS, true>(m, kp, c, j, i, p, xu, U, V, x0.u, r.u, acc);
      if (j < ny - 1) resid_init_row<CROSS, false>(m, kp, c, j, i, p, xv, U, V, x0.v, r.v, acc);
    }
  }
  block_sum_store<3>(acc, partial + 3 * (long long)tb);
}

// Sum NV partials per member over its tile range [tstart[g], tstart[g+1])
// in a fixed order (deterministic).  One workgroup per member.
template <int NV>
__global__ void __launch_bounds__(256) k_member_sum(const int *__restrict__ tstart, const double *__restrict__ partial,
                                                    double *__restrict__ out) {
  const int g = blockIdx.x;
  const int a = tstart[g], b = tstart[g + 1];
  __shared__ double red[NV][256];
  double s[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) s[q] = 0.0;
  for (int k = a + (int)threadIdx.x; k < b; k += 256)
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q] += partial[(long long)NV * k + q];
#pragma unroll
  for (int q = 0; q < NV; ++q) red[q][threadIdx.x] = s[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int q = 0; q < NV; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) out[NV * g + q] = red[q][0];
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = v;
}
__global__ void k_fill_i32(int *p, int n, int v) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = v;
}
// relative residual per member from the member sums (|Au-b|^2, |b|^2)
__global__ void k_resid_ratio(int ngrids, const double *__restrict__ sums, double *__restrict__ r) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngrids) return;
  const double rr = sums[2 * g], bb = sums[2 * g + 1];
  r[g] = bb > 0.0 ? sqrt(rr) / sqrt(bb) : sqrt(rr);
}

// copy the Picard state into the linear-solve iterate (initial guess, the
// reference's x(n) = u(j,i); zero rows start from 0)
__global__ void __launch_bounds__(IMAU_WG) k_init_x(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                    const Fields f, double *__restrict__ xu, double *__restrict__ xv) {
  TILE_SETUP
  if (j >= m.ny) return;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    const uint8_t ru = (i < m.nx - 1) ? f.rt_u[p] : RT_ZERO;
    const uint8_t rv = (j < m.ny - 1) ? f.rt_v[p] : RT_ZERO;
    xu[p] = (ru == RT_ZERO || ru == RT_BND_ZERO) ? 0.0 : f.u[p];
    xv[p] = (rv == RT_ZERO || rv == RT_BND_ZERO) ? 0.0 : f.v[p];
  }
}

// ---------------------------------------------------------------------------
// final 3D velocities and secondary fields
// ---------------------------------------------------------------------------
// F1 column (src/ice_velocity_module.f90:1222-1227)
__global__ void __launch_bounds__(IMAU_WG) k_F1(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                const KP kp, const Fields f, const double *__restrict__ visc,
                                                OutF o) {
  TILE_SETUP
  if (j >= m.ny) return;
  const int nz = kp.nz;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const double H = f.Hi[p2(m, j, i)];
    double Fk1 = 0.0;
    double gk1 = (-H / visc[p3(m, nz - 1, j, i)]) * kp.zeta[nz - 1];
    o.F1[p3(m, nz - 1, j, i)] = 0.0;
    for (int k = nz - 2; k >= 0; --k) {
      const double gk = (-H / visc[p3(m, k, j, i)]) * kp.zeta[k];
      const double Fk = Fk1 - 0.5 * (gk1 + gk) * (kp.zeta[k + 1] - kp.zeta[k]);
      o.F1[p3(m, k, j, i)] = Fk;
      Fk1 = Fk;
      gk1 = gk;
    }
  }
}

// u_3D_cx / v_3D_cy of the DIVA (:1230-1289)
__global__ void __launch_bounds__(IMAU_WG) k_uv3d(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                  const KP kp, const Fields f, OutF o) {
  TILE_SETUP
  if (j >= m.ny) return;
  const int nz = kp.nz;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    const double H0 = f.Hi[p];
    for (int dir = 0; dir < 2; ++dir) {
      if (dir == 0 && i >= m.nx - 1) continue;
      if (dir == 1 && j >= m.ny - 1) continue;
      const int j1 = dir ? j + 1 : j, i1 = dir ? i : i + 1;
      const double H1 = f.Hi[p2(m, j1, i1)];
      const double base = dir ? f.v_base[p] : f.u_base[p];
      const double tb = dir ? f.taub_cy[p] : f.taub_cx[p];
      double *out = dir ? o.v3 : o.u3;
      for (int k = 0; k < nz; ++k) {
        const double F0 = o.F1[p3(m, k, j, i)], Fn = o.F1[p3(m, k, j1, i1)];
        double Fs;
        if (kp.margin == 0) {
          if (H0 > 0.0 && H1 == 0.0) Fs = F0;
          else if (H0 == 0.0 && H1 > 0.0) Fs = Fn;
          else if (H0 > 0.0 && H1 > 0.0) Fs = 0.5 * (F0 + Fn);
          else Fs = 0.0;
        } else {
          Fs = 0.5 * (F0 + Fn);
        }
        out[p3(m, k, j, i)] = base + tb * Fs;
      }
    }
  }
}

__device__ __forceinline__ double vavg_col(const double *a, const Member &m, const KP &kp, int j, int i) {
  double s = 0.0;
  for (int k = 0; k < kp.nz - 1; ++k)
    s += 0.5 * (a[p3(m, k + 1, j, i)] + a[p3(m, k, j, i)]) * (kp.zeta[k + 1] - kp.zeta[k]);
  return s;
}

// 3D columns, basal and depth-averaged velocity for the non-DIVA modes
// (calc_secondary_velocities branches, :393-475).  mode: 1 SIA, 2 SSA, 3 SIA/SSA
__global__ void __launch_bounds__(IMAU_WG) k_compose(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                     const KP kp, const Fields f, OutF o, int mode) {
  TILE_SETUP
  if (j >= m.ny) return;
  const int nz = kp.nz;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    for (int dir = 0; dir < 2; ++dir) {
      if (dir == 0 && i >= m.nx - 1) continue;
      if (dir == 1 && j >= m.ny - 1) continue;
      double *u3 = dir ? o.v3 : o.u3;
      const double *s3 = dir ? o.v3sia : o.u3sia;
      const double us = dir ? f.v[p] : f.u[p];
      double *ub = dir ? f.v_base : f.u_base;
      double *uv = dir ? o.vvav : o.uvav;
      for (int k = 0; k < nz; ++k) {
        const long long q = p3(m, k, j, i);
        u3[q] = (mode == 1) ? s3[q] : (mode == 2 ? us : s3[q] + us);
      }
      ub[p] = (mode == 1) ? 0.0 : us;
      uv[p] = (mode == 2) ? us : vavg_col(u3, m, kp, j, i);
    }
  }
}

// selection of map_velocity_from_cx_to_a / _cy_to_a (:2592-2811) at a-cell
// (j,i) from the ice mask: 0 -> 0, 1 -> the mean of the two staggered values,
// 2 -> the upper one (index i or j), 3 -> the lower one (index i-1 or j-1);
// at the first / last a cell only the edge value exists (2 / 3)
__device__ __forceinline__ int mv_sel(const Member &m, const Fields &f, int j, int i, bool ydir) {
  const int n = ydir ? m.ny : m.nx, c = ydir ? j : i;
  const int mi = f.m_ice[p2(m, j, i)];
  if (c == 0) return mi == 1 ? 2 : 0;
  if (c == n - 1) return mi == 1 ? 3 : 0;
  const int a0 = ydir ? f.m_ice[p2(m, j - 1, i)] : f.m_ice[p2(m, j, i - 1)];
  const int a2 = ydir ? f.m_ice[p2(m, j + 1, i)] : f.m_ice[p2(m, j, i + 1)];
  if (mi != 1) return 0;
  if (a0 == 1 && a2 == 1) return 1;
  if (a0 == 0 && a2 == 1) return 2;
  if (a0 == 1 && a2 == 0) return 3;
  return 0;
}
__device__ __forceinline__ double mv_pick(int sel, double lo, double hi) {
  return sel == 1 ? 0.5 * (lo + hi) : (sel == 2 ? hi : (sel == 3 ? lo : 0.0));
}
// the same map of a 2D staggered field (n_stag = n - 1 values along the direction)
__device__ __forceinline__ double mv_2d(const double *a, const Member &m, int sel, int j, int i, bool ydir) {
  const int n = ydir ? m.ny : m.nx, c = ydir ? j : i;
  const double lo = c > 0 ? a[ydir ? p2(m, j - 1, i) : p2(m, j, i - 1)] : 0.0;
  const double hi = c < n - 1 ? a[p2(m, j, i)] : 0.0;
  return mv_pick(sel, lo, hi);
}

// One pass over each column, base to surface: w_3D_a by mass conservation
// (:534-661, when W; u3a/v3a hold the a-grid 3D velocities of the previous
// call as the reference reads them, each read before it is replaced), the
// a-grid copies of the 3D velocities, the depth-averaged, surface and basal
// a-grid velocities, |u|, the shear ratio and the surface and basal w
// (:490-533, :642-650).  The staggered values a level needs are loaded once
// and serve both the divergence and the a-grid map.
template <bool W>
__global__ void __launch_bounds__(IMAU_WG) k_final_a(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                     const KP kp, const Fields f, OutF o, int dyn,
                                                     const double *__restrict__ uvav,
                                                     const double *__restrict__ vvav) {
  TILE_SETUP
  if (j >= m.ny) return;
  const int nx = m.nx, ny = m.ny, nz = kp.nz;
  const double dx = m.dx;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= nx) break;
    const long long p = p2(m, j, i);
    // the staggered columns read: iL/iR (jL/jR) of the divergence, which
    // also hold the a-grid map's lower / upper values
    const int iL = (i == 0) ? 0 : (i == nx - 1 ? nx - 3 : i - 1);
    const int iR = (i == 0) ? 1 : (i == nx - 1 ? nx - 2 : i);
    const int jL = (j == 0) ? 0 : (j == ny - 1 ? ny - 3 : j - 1);
    const int jR = (j == 0) ? 1 : (j == ny - 1 ? ny - 2 : j);
    const int sx = mv_sel(m, f, j, i, false), sy = mv_sel(m, f, j, i, true);
    auto ulo = [&](double uL, double uR) { return i == nx - 1 ? uR : uL; };
    auto uhi = [&](double uL, double uR) { return i == 0 ? uL : uR; };
    auto vlo = [&](double vL, double vR) { return j == ny - 1 ? vR : vL; };
    auto vhi = [&](double vL, double vR) { return j == 0 ? vL : vR; };
    const int kb = nz - 1;
    double uL1 = o.u3[p3(m, kb, j, iL)], uR1 = o.u3[p3(m, kb, j, iR)];
    double vL1 = o.v3[p3(m, kb, jL, i)], vR1 = o.v3[p3(m, kb, jR, i)];
    double w = 0.0, wb = 0.0, ua1 = 0.0, va1 = 0.0;
    double dHsx = 0.0, dHix = 0.0, dHsy = 0.0, dHiy = 0.0, Hm = 0.0, Hi = 0.0;
    if (W) {
      // local ddx_a_to_a / ddy_a_to_a of Hs and Hi (:563-566)
      double dHs_dx, dHs_dy, dHi_dx, dHi_dy;
      if (i == 0) { dHs_dx = (f.Hs[p2(m, j, 1)] - f.Hs[p]) / dx; dHi_dx = (f.Hi[p2(m, j, 1)] - f.Hi[p]) / dx; }
      else if (i == nx - 1) { dHs_dx = (f.Hs[p] - f.Hs[p2(m, j, nx - 2)]) / dx; dHi_dx = (f.Hi[p] - f.Hi[p2(m, j, nx - 2)]) / dx; }
      else { dHs_dx = (f.Hs[p2(m, j, i + 1)] - f.Hs[p2(m, j, i - 1)]) / (2 * dx); dHi_dx = (f.Hi[p2(m, j, i + 1)] - f.Hi[p2(m, j, i - 1)]) / (2 * dx); }
      if (j == 0) { dHs_dy = (f.Hs[p2(m, 1, i)] - f.Hs[p]) / dx; dHi_dy = (f.Hi[p2(m, 1, i)] - f.Hi[p]) / dx; }
      else if (j == ny - 1) { dHs_dy = (f.Hs[p] - f.Hs[p2(m, ny - 2, i)]) / dx; dHi_dy = (f.Hi[p] - f.Hi[p2(m, ny - 2, i)]) / dx; }
      else { dHs_dy = (f.Hs[p2(m, j + 1, i)] - f.Hs[p2(m, j - 1, i)]) / (2 * dx); dHi_dy = (f.Hi[p2(m, j + 1, i)] - f.Hi[p2(m, j - 1, i)]) / (2 * dx); }
      const double dHbase_dx = dHs_dx - dHi_dx, dHbase_dy = dHs_dy - dHi_dy;
      const double ub = (i == 0) ? uL1 : (i == nx - 1 ? uR1 : 0.5 * (uL1 + uR1));
      const double vb = (j == 0) ? vL1 : (j == ny - 1 ? vR1 : 0.5 * (vL1 + vR1));
      w = ub * dHbase_dx + vb * dHbase_dy;
      if (f.m_sheet[p] == 1 && f.dHb_dt) w += f.dHb_dt[p];
      o.w3[p3(m, kb, j, i)] = w;
      wb = w;
      Hi = f.Hi[p];
      Hm = fmax(0.1, Hi);
      dHsx = f.dHs_dx_a[p]; dHix = f.dHi_dx_a[p]; dHsy = f.dHs_dy_a[p]; dHiy = f.dHi_dy_a[p];
      ua1 = o.u3a[p3(m, kb, j, i)];
      va1 = o.v3a[p3(m, kb, j, i)];
    }
    double usa = mv_pick(sx, ulo(uL1, uR1), uhi(uL1, uR1)), vsa = mv_pick(sy, vlo(vL1, vR1), vhi(vL1, vR1));
    o.u3a[p3(m, kb, j, i)] = usa;
    o.v3a[p3(m, kb, j, i)] = vsa;
    for (int k = nz - 2; k >= 0; --k) {
      const double uL = o.u3[p3(m, k, j, iL)], uR = o.u3[p3(m, k, j, iR)];
      const double vL = o.v3[p3(m, k, jL, i)], vR = o.v3[p3(m, k, jR, i)];
      if (W) {
        const double ua = o.u3a[p3(m, k, j, i)], va = o.v3a[p3(m, k, j, i)];
        const double dudx_k = (uR - uL) / dx, dudx_k1 = (uR1 - uL1) / dx;
        const double dvdy_k = (vR - vL) / dx, dvdy_k1 = (vR1 - vL1) / dx;
        const double w1 = (dudx_k + dudx_k1) / 2.0, w2 = (dvdy_k + dvdy_k1) / 2.0;
        const double zm = 0.5 * (kp.zeta[k + 1] + kp.zeta[k]), dz = kp.zeta[k + 1] - kp.zeta[k];
        const double w3 = ((dHsx - zm * dHix) / Hm) * ((ua1 - ua) / dz);
        const double w4 = ((dHsy - zm * dHiy) / Hm) * ((va1 - va) / dz);
        w = w - Hi * (w1 + w2 + w3 + w4) * dz;
        o.w3[p3(m, k, j, i)] = w;
        ua1 = ua;
        va1 = va;
      }
      usa = mv_pick(sx, ulo(uL, uR), uhi(uL, uR));
      vsa = mv_pick(sy, vlo(vL, vR), vhi(vL, vR));
      o.u3a[p3(m, k, j, i)] = usa;
      o.v3a[p3(m, k, j, i)] = vsa;
      uL1 = uL; uR1 = uR; vL1 = vL; vR1 = vR;
    }
    // usa / vsa now hold the surface (k = 0) values
    const double uva = mv_2d(uvav, m, sx, j, i, false), vva = mv_2d(vvav, m, sy, j, i, true);
    const double uba = mv_2d(f.u_base, m, sx, j, i, false), vba = mv_2d(f.v_base, m, sy, j, i, true);
    o.u_vav_a[p] = uva; o.v_vav_a[p] = vva;
    o.u_surf_a[p] = usa; o.v_surf_a[p] = vsa;
    o.u_base_a[p] = uba; o.v_base_a[p] = vba;
    const double av = sqrt(uva * uva + vva * vva), as = sqrt(usa * usa + vsa * vsa), ab = sqrt(uba * uba + vba * vba);
    o.uabs_vav[p] = av; o.uabs_surf[p] = as; o.uabs_base[p] = ab;
    double R;
    if (dyn == 1) R = 1.0;
    else if (dyn == 2) R = 0.0;
    else R = (f.m_ice[p] == 1) ? fmax(0.0, fmin(1.0, 1.0 - ab / as)) : 0.0;
    o.R_shear[p] = R;
    o.w_surf[p] = W ? w : o.w3[p3(m, 0, j, i)];
    o.w_base[p] = W ? wb : o.w3[p3(m, kb, j, i)];
  }
}

// SIA columns (solve_SIA, src/ice_velocity_module.f90:61-105)
__global__ void __launch_bounds__(IMAU_WG) k_sia(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                 const KP kp, const Fields f, const double *__restrict__ Acx,
                                                 const double *__restrict__ Acy, OutF o) {
  TILE_SETUP
  if (j >= m.ny) return;
  const int nz = kp.nz;
  const double n = kp.n_flow;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    for (int dir = 0; dir < 2; ++dir) {
      if (dir == 0 && i >= m.nx - 1) continue;
      if (dir == 1 && j >= m.ny - 1) continue;
      const int j1 = dir ? j + 1 : j, i1 = dir ? i : i + 1;
      double *out = dir ? o.v3sia : o.u3sia;
      if (f.m_sheet[p] == 0 && f.m_sheet[p2(m, j1, i1)] == 0) {
        for (int k = 0; k < nz; ++k) out[p3(m, k, j, i)] = 0.0;
        continue;
      }
      const double H = dir ? f.Hi_cy[p] : f.Hi_cx[p];
      const double sx = dir ? f.dHs_dx_cy[p] : f.dHs_dx_cx[p];
      const double sy = dir ? f.dHs_dy_cy[p] : f.dHs_dy_cx[p];
      const double slope = dir ? sy : sx;
      const double *A = dir ? Acy : Acx;
      const double D0 = pow(D_ICE_DENSITY * D_GRAV * H, n) * pow(sx * sx + sy * sy, (n - 1.0) / 2.0);
      // integral from the bed to zeta of A zeta^n (vertical_integration_from_bottom_to_zeta)
      double Dk1 = 0.0, gk1 = A[p3(m, nz - 1, j, i)] * pow(kp.zeta[nz - 1], n);
      out[p3(m, nz - 1, j, i)] = fmax(D0 * (2.0 * H * 0.0), -1e5) * slope;
      for (int k = nz - 2; k >= 0; --k) {
        const double gk = A[p3(m, k, j, i)] * pow(kp.zeta[k], n);
        const double Dk = Dk1 - 0.5 * (gk1 + gk) * (kp.zeta[k + 1] - kp.zeta[k]);
        out[p3(m, k, j, i)] = fmax(D0 * (2.0 * H * Dk), -1e5) * slope;
        Dk1 = Dk;
        gk1 = gk;
      }
    }
  }
}

// (du, dv) = v on the 2D slabs of the members with flag[g] != 0
__global__ void __launch_bounds__(IMAU_WG) k_fill2_members(const Member *__restrict__ mem,
                                                           const Tile *__restrict__ tiles, Vec2 dst, double v,
                                                           const int *__restrict__ flag) {
  TILE_SETUP
  if (j >= m.ny || !flag[t.g]) return;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    dst.u[p] = v;
    dst.v[p] = v;
  }
}

// layout transposes between the host (k fastest, leading dimension ldj)
// and the device (k-plane major, leading dimension ny) for ni x nj columns
__global__ void k_h2d_3d(const Member *__restrict__ mem, int g, const double *__restrict__ src, double *__restrict__ dst,
                         int ni, int nj, int ldj, int nz) {
  const Member m = mem[g];
  const long long n = (long long)ni * nj * nz;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(q % nz);
    const long long col = q / nz;
    const int j = (int)(col % nj), i = (int)(col / nj);
    dst[p3(m, k, j, i)] = src[((long long)i * ldj + j) * nz + k];
  }
}
__global__ void k_d2h_3d(const Member *__restrict__ mem, int g, const double *__restrict__ src, double *__restrict__ dst,
                         int ni, int nj, int ldj, int nz) {
  const Member m = mem[g];
  const long long n = (long long)ni * nj * nz;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(q % nz);
    const long long col = q / nz;
    const int j = (int)(col % nj), i = (int)(col / nj);
    dst[((long long)i * ldj + j) * nz + k] = src[p3(m, k, j, i)];
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace imau {
static inline dim3 wg() { return dim3(IMAU_TJ, IMAU_WG_I, 1); }
static inline CoefP to_p(const Coef &c) {
  CoefP p;
  p.Na = c.Na; p.Nb = c.Nb; p.Ncx = c.Ncx; p.Ncy = c.Ncy; p.bcx = c.bcx; p.bcy = c.bcy;
  p.taudx = c.taudx; p.taudy = c.taudy; p.frhs = c.frhs; p.rt_u = c.rt_u; p.rt_v = c.rt_v;
  return p;
}
void setup1(const Launch &L, const KP &kp, const Fields &f, const double *yg) {
  hipLaunchKernelGGL(k_setup1, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, yg);
}
void setup2(const Launch &L, const KP &kp, const Fields &f) {
  hipLaunchKernelGGL(k_setup2, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f);
}
void visc3d(const Launch &L, const Launch &LV, const KP &kp, const Fields &f, bool diva, const double *visc_in,
            double *visc_out, const int *active) {
  // n = 3 (every shipped configuration) takes the cube-root form of the viscosity
  const bool third = kp.n_flow == 3.0;
  if (diva && kp.nz == IMAU_MNZ && LV.nmtiles > 0) {
    hipLaunchKernelGGL(k_strain, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, active);
    auto km = third ? k_visc_march<true, double> : k_visc_march<false, double>;
    hipLaunchKernelGGL(km, dim3(LV.nmtiles), dim3(64), 0, L.s, L.mem, LV.mtiles, kp, f, visc_in, visc_out, f.B3,
                       active);
    return;
  }
  auto kv = diva ? (third ? k_visc<true, true, double> : k_visc<true, false, double>)
                 : (third ? k_visc<false, true, double> : k_visc<false, false, double>);
  hipLaunchKernelGGL(k_strain, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, active);
  hipLaunchKernelGGL(kv, dim3(LV.ntiles), wg(), 0, L.s, L.mem, LV.tiles, kp, f, visc_in, visc_out, f.B3, active);
}
void visc3d32(const Launch &L, const Launch &LV, const KP &kp, const Fields &f, bool diva, const float *visc_in,
              float *visc_out, const float *B3, const int *active) {
  const bool third = kp.n_flow == 3.0;
  if (diva && kp.nz == IMAU_MNZ && LV.nmtiles > 0) {
    hipLaunchKernelGGL(k_strain, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, active);
    auto km = third ? k_visc_march<true, float> : k_visc_march<false, float>;
    hipLaunchKernelGGL(km, dim3(LV.nmtiles), dim3(64), 0, L.s, L.mem, LV.mtiles, kp, f, visc_in, visc_out, B3, active);
    return;
  }
  auto kv = diva ? (third ? k_visc<true, true, float> : k_visc<true, false, float>)
                 : (third ? k_visc<false, true, float> : k_visc<false, false, float>);
  hipLaunchKernelGGL(k_strain, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, active);
  hipLaunchKernelGGL(kv, dim3(LV.ntiles), wg(), 0, L.s, L.mem, LV.tiles, kp, f, visc_in, visc_out, B3, active);
}
__global__ void k_d2f(const double *__restrict__ a, float *__restrict__ b, long long n) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    b[q] = (float)a[q];
}
__global__ void k_f2d(const float *__restrict__ a, double *__restrict__ b, long long n) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    b[q] = (double)a[q];
}
void to_f32(hipStream_t s, const double *a, float *b, long long n) {
  hipLaunchKernelGGL(k_d2f, dim3((int)std::min<long long>((n + 255) / 256, 8192)), dim3(256), 0, s, a, b, n);
}
void to_f64(hipStream_t s, const float *a, double *b, long long n) {
  hipLaunchKernelGGL(k_f2d, dim3((int)std::min<long long>((n + 255) / 256, 8192)), dim3(256), 0, s, a, b, n);
}
void beta(const Launch &L, const KP &kp, const Fields &f, bool diva, const double *xg, const double *yg,
          const int *active) {
  // Weertman (the common law) gets its own specialisation
  auto k = diva ? (kp.slid == 4 ? k_beta<true, 4> : k_beta<true, -1>)
                : (kp.slid == 4 ? k_beta<false, 4> : k_beta<false, -1>);
  hipLaunchKernelGGL(k, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, xg, yg, active);
}
void coef(const Launch &L, const KP &kp, const Fields &f, bool diva, const int *active) {
  hipLaunchKernelGGL(k_coef, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, diva ? 1 : 0, active);
}
void net_features(const Launch &L, const KP &kp, const Fields &f, const long long *poff, unsigned short *x) {
  hipLaunchKernelGGL(k_net_features, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, poff, x);
}
void net_targets(const Launch &L, const Fields &f, const long long *poff, float *tg, float *wt) {
  hipLaunchKernelGGL(k_net_targets, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, f, poff, tg, wt);
}
void net_warm(const Launch &L, const KP &kp, const Fields &f, const long long *poff, const float *o,
              const int *active) {
  hipLaunchKernelGGL(k_net_warm, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, poff, o, active);
}
// the k_visc_column specialisation for kp: n = 3, Weertman sliding and
// nz = IMAU_MNZ each get their own
template <class T>
static auto column_kernel(const KP &kp) {
  const bool third = kp.n_flow == 3.0, weert = kp.slid == 4, fixed = kp.nz == IMAU_MNZ;
  if (third) {
    if (weert) return fixed ? k_visc_column<true, T, 4, IMAU_MNZ> : k_visc_column<true, T, 4, 0>;
    return fixed ? k_visc_column<true, T, -1, IMAU_MNZ> : k_visc_column<true, T, -1, 0>;
  }
  if (weert) return fixed ? k_visc_column<false, T, 4, IMAU_MNZ> : k_visc_column<false, T, 4, 0>;
  return fixed ? k_visc_column<false, T, -1, IMAU_MNZ> : k_visc_column<false, T, -1, 0>;
}
void visc_column(const Launch &L, const Launch &LV, const KP &kp, const Fields &f, double *visc, const double *xg,
                 const double *yg, int iters, const int *active) {
  hipLaunchKernelGGL(k_strain, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, active);
  hipLaunchKernelGGL(column_kernel<double>(kp), dim3(LV.ntiles), wg(), 0, L.s, L.mem, LV.tiles, kp, f, visc, f.B3, xg,
                     yg, iters, active);
}
void visc_column32(const Launch &L, const Launch &LV, const KP &kp, const Fields &f, float *visc, const float *B3,
                   const double *xg, const double *yg, int iters, const int *active) {
  hipLaunchKernelGGL(k_strain, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, active);
  hipLaunchKernelGGL(column_kernel<float>(kp), dim3(LV.ntiles), wg(), 0, L.s, L.mem, LV.tiles, kp, f, visc, B3, xg, yg,
                     iters, active);
}
void taub(const Launch &L, const Fields &f, const int *active) {
  hipLaunchKernelGGL(k_taub, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, f, active);
}
void init_x(const Launch &L, const Fields &f, double *xu, double *xv) {
  hipLaunchKernelGGL(k_init_x, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, f, xu, xv);
}
void post(const Launch &L, const KP &kp, const Fields &f, const double *xu, const double *xv, const int *active,
          bool diva, double *partial, const double *relax, const AMix &mix) {
  hipLaunchKernelGGL(k_post, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, xu, xv, active, diva ? 1 : 0, partial,
                     relax, mix);
}
void resid(const Launch &L, const KP &kp, const Coef &c, const double *xu, const double *xv, double *partial) {
  hipLaunchKernelGGL(k_resid, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, to_p(c), xu, xv, partial);
}
void member_sum3(hipStream_t s, int ngrids, const int *tstart, const double *partial, double *out) {
  hipLaunchKernelGGL(k_member_sum<3>, dim3(ngrids), dim3(256), 0, s, tstart, partial, out);
}
void resid_init(const Launch &L, const KP &kp, const Coef &c, const double *xu, const double *xv, double *x0u,
                double *x0v, double *ru, double *rv, const int *active, double *partial) {
  auto k = kp.cross ? k_resid_init<1> : k_resid_init<0>;
  hipLaunchKernelGGL(k, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, to_p(c), xu, xv, Vec2{x0u, x0v},
                     Vec2{ru, rv}, active, partial);
}
void member_sum2(hipStream_t s, int ngrids, const int *tstart, const double *partial, double *out) {
  hipLaunchKernelGGL(k_member_sum<2>, dim3(ngrids), dim3(256), 0, s, tstart, partial, out);
}
void fill_u64(hipStream_t s, unsigned long long *p, int n, unsigned long long v) {
  hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}
void reset_members(const Launch &L, const KP &kp, const Fields &f, const ResetSrc &src, double *visc,
                   float *visc32, const int *flag) {
  hipLaunchKernelGGL(k_reset_members, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, f, src, visc, visc32, kp.nz, flag);
}
void copy2_members(const Launch &L, const double *su, const double *sv, double *du, double *dv, const int *flag) {
  hipLaunchKernelGGL(k_copy2_members, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, CVec2{su, sv}, Vec2{du, dv}, flag);
}
void copy3_members(const Launch &L, int nz, const double *src, double *dst, const int *flag) {
  hipLaunchKernelGGL(k_copy3_members<double>, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, src, dst, nz, flag);
}
void copy3_members(const Launch &L, int nz, const float *src, float *dst, const int *flag) {
  hipLaunchKernelGGL(k_copy3_members<float>, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, src, dst, nz, flag);
}
void fill2_members(const Launch &L, double *du, double *dv, double v, const int *flag) {
  hipLaunchKernelGGL(k_fill2_members, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, Vec2{du, dv}, v, flag);
}
void fill_i32(hipStream_t s, int *p, int n, int v) {
  hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}
void resid_ratio(hipStream_t s, int ngrids, const double *sums, double *r) {
  hipLaunchKernelGGL(k_resid_ratio, dim3((ngrids + 255) / 256), dim3(256), 0, s, ngrids, sums, r);
}
void F1(const Launch &L, const KP &kp, const Fields &f, const double *visc, const OutF &o) {
  hipLaunchKernelGGL(k_F1, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, visc, o);
}
void uv3d(const Launch &L, const KP &kp, const Fields &f, const OutF &o) {
  hipLaunchKernelGGL(k_uv3d, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, o);
}
void compose(const Launch &L, const KP &kp, const Fields &f, const OutF &o, int mode) {
  hipLaunchKernelGGL(k_compose, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, o, mode);
}
void final_a(const Launch &L, const KP &kp, const Fields &f, const OutF &o, int dyn, const double *uvav,
             const double *vvav, bool with_w) {
  if (with_w)
    hipLaunchKernelGGL(k_final_a<true>, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, o, dyn, uvav, vvav);
  else
    hipLaunchKernelGGL(k_final_a<false>, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, o, dyn, uvav, vvav);
}
void sia(const Launch &L, const KP &kp, const Fields &f, const double *Acx, const double *Acy, const OutF &o) {
  hipLaunchKernelGGL(k_sia, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, Acx, Acy, o);
}
void h2d_3d(hipStream_t s, const Member *mem, int g, const double *src, double *dst, int ni, int nj, int ldj, int nz) {
  long long n = (long long)ni * nj * nz;
  int blocks = (int)((n + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_h2d_3d, dim3(blocks), dim3(256), 0, s, mem, g, src, dst, ni, nj, ldj, nz);
}
void d2h_3d(hipStream_t s, const Member *mem, int g, const double *src, double *dst, int ni, int nj, int ldj, int nz) {
  long long n = (long long)ni * nj * nz;
  int blocks = (int)((n + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_d2h_3d, dim3(blocks), dim3(256), 0, s, mem, g, src, dst, ni, nj, ldj, nz);
}
}  // namespace imau

// ===========================================================================
// Krylov linear solver (BiCGStab, right-preconditioned) and global
// multicolour SOR.  Vectors are (u-slab, v-slab) pairs in the member slabs;
// per-member scalars live on the device (sc[g*NSC + k]) so that an
// iteration needs no host round trip.
// ===========================================================================
namespace bicg {
enum { RHO = 0, ALPHA = 1, OMEGA = 2, BETA = 3, RR = 4, NSC = 8 };
}


// y = A x over every row; with w != null also partial sums (w, y) and (y, y)
template <int CROSS>
__global__ void __launch_bounds__(IMAU_WG) k_apply(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                   const KP kp, const CoefP c, CVec2 x, Vec2 y, CVec2 w,
                                                   const int *__restrict__ active, double *__restrict__ partial) {
  TILE_SETUP
  double acc[2] = {0.0, 0.0};
  const GAcc U{x.u, m}, V{x.v, m};
  if (j < m.ny && active[t.g]) {
    FOR_TILE_ROWS {
      const int i = t.i0 + ii;
      if (i >= m.nx) break;
      const long long p = p2(m, j, i);
      double S, D;
      if (i < m.nx - 1) {
        u_sd_x<CROSS, double>(m, kp, c, j, i, c.rt_u[p], U, V, S, D);
        const double yv = S + D * x.u[p];
        y.u[p] = yv;
        if (w.u) { acc[0] += w.u[p] * yv; acc[1] += yv * yv; }
      }
      if (j < m.ny - 1) {
        v_sd_x<CROSS, double>(m, kp, c, j, i, c.rt_v[p], U, V, S, D);
        const double yv = S + D * x.v[p];
        y.v[p] = yv;
        if (w.u) { acc[0] += w.v[p] * yv; acc[1] += yv * yv; }
      }
    }
  }
  if (partial) block_sum_store<2>(acc, partial + 2 * (long long)tb);
}

// r = b - A x, rhat = r, p = v = 0; partials (r, r), (b, b)
__global__ void __launch_bounds__(IMAU_WG) k_bicg_init(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                       const KP kp, const CoefP c, CVec2 x, Vec2 r, Vec2 rh, Vec2 pp,
                                                       Vec2 vv, const int *__restrict__ active,
                                                       double *__restrict__ partial) {
  TILE_SETUP
  double acc[2] = {0.0, 0.0};
  const GAcc U{x.u, m}, V{x.v, m};
  if (j < m.ny && active[t.g]) {
    FOR_TILE_ROWS {
      const int i = t.i0 + ii;
      if (i >= m.nx) break;
      const long long p = p2(m, j, i);
      double S, D, b;
      if (i < m.nx - 1) {
        u_row(m, kp, c, j, i, c.rt_u[p], U, V, S, D, b);
        const double rv = b - (S + D * x.u[p]);
        r.u[p] = rv;
        if (rh.u) { rh.u[p] = rv; pp.u[p] = 0.0; vv.u[p] = 0.0; }
        acc[0] += rv * rv; acc[1] += b * b;
      }
      if (j < m.ny - 1) {
        v_row(m, kp, c, j, i, c.rt_v[p], U, V, S, D, b);
        const double rv = b - (S + D * x.v[p]);
        r.v[p] = rv;
        if (rh.u) { rh.v[p] = rv; pp.v[p] = 0.0; vv.v[p] = 0.0; }
        acc[0] += rv * rv; acc[1] += b * b;
      }
    }
  }
  block_sum_store<2>(acc, partial + 2 * (long long)tb);
}

#define VEC_LOOP(...)                                               \
  TILE_SETUP                                                       \
  if (j < m.ny && active[t.g]) {                                   \
    FOR_TILE_ROWS {                                                \
      const int i = t.i0 + ii;                                     \
      if (i >= m.nx) break;                                        \
      const long long p = p2(m, j, i);                             \
      const bool hu = i < m.nx - 1, hv = j < m.ny - 1;             \
      __VA_ARGS__                                                  \
    }                                                              \
  }

// p = r + beta (p - omega v)
__global__ void __launch_bounds__(IMAU_WG) k_bicg_p(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                    CVec2 r, Vec2 pp, CVec2 vv, const double *__restrict__ sc,
                                                    const int *__restrict__ active) {
  VEC_LOOP({
    const double beta = sc[t.g * bicg::NSC + bicg::BETA], om = sc[t.g * bicg::NSC + bicg::OMEGA];
    if (hu) pp.u[p] = r.u[p] + beta * (pp.u[p] - om * vv.u[p]);
    if (hv) pp.v[p] = r.v[p] + beta * (pp.v[p] - om * vv.v[p]);
  })
}

// s = r - alpha v (s overwrites r)
__global__ void __launch_bounds__(IMAU_WG) k_bicg_s(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                    Vec2 r, CVec2 vv, const double *__restrict__ sc,
                                                    const int *__restrict__ active) {
  VEC_LOOP({
    const double al = sc[t.g * bicg::NSC + bicg::ALPHA];
    if (hu) r.u[p] -= al * vv.u[p];
    if (hv) r.v[p] -= al * vv.v[p];
  })
}

// x += alpha phat + omega shat ; r = s - omega t ; partials (rhat, r), (r, r)
__global__ void __launch_bounds__(IMAU_WG) k_bicg_xr(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                     Vec2 x, Vec2 r, CVec2 ph, CVec2 sh, CVec2 tt, CVec2 rh,
                                                     const double *__restrict__ sc, const int *__restrict__ active,
                                                     double *__restrict__ partial) {
  double acc[2] = {0.0, 0.0};
  VEC_LOOP({
    const double al = sc[t.g * bicg::NSC + bicg::ALPHA], om = sc[t.g * bicg::NSC + bicg::OMEGA];
    if (hu) {
      x.u[p] += al * ph.u[p] + om * sh.u[p];
      const double rv = r.u[p] - om * tt.u[p];
      r.u[p] = rv;
      acc[0] += rh.u[p] * rv; acc[1] += rv * rv;
    }
    if (hv) {
      x.v[p] += al * ph.v[p] + om * sh.v[p];
      const double rv = r.v[p] - om * tt.v[p];
      r.v[p] = rv;
      acc[0] += rh.v[p] * rv; acc[1] += rv * rv;
    }
  })
  block_sum_store<2>(acc, partial + 2 * (long long)tb);
}

// per-member reduction of two partials and the BiCGStab scalar update.
// stage 0 (after init):  rho = (r,r), rr = (r,r), alpha = omega = 1, beta = 0
// stage 1 (after v=Ap):  alpha = rho / (rhat, v)
// stage 2 (after t=As):  omega = (t,s) / (t,t)
// stage 3 (after x,r):   beta = (rho_new/rho)(alpha/omega), rho = rho_new, rr = (r,r)
// out[2g], out[2g+1] receive the two sums (for the host's convergence test)
__global__ void __launch_bounds__(256) k_bicg_scalar(const int *__restrict__ tstart, const double *__restrict__ partial,
                                                     double *__restrict__ sc, double *__restrict__ out,
                                                     const int *__restrict__ active, int stage) {
  const int g = blockIdx.x;
  const int a = tstart[g], b = tstart[g + 1];
  __shared__ double red[2][256];
  double s0 = 0.0, s1 = 0.0;
  for (int k = a + (int)threadIdx.x; k < b; k += 256) {
    s0 += partial[2 * (long long)k];
    s1 += partial[2 * (long long)k + 1];
  }
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  s0 = red[0][0];
  s1 = red[1][0];
  out[2 * g] = s0;
  out[2 * g + 1] = s1;
  if (!active[g]) return;
  double *S = sc + g * bicg::NSC;
  if (stage == 0) {
    S[bicg::RHO] = s0; S[bicg::RR] = s0; S[bicg::ALPHA] = 1.0; S[bicg::OMEGA] = 1.0; S[bicg::BETA] = 0.0;
  } else if (stage == 1) {
    S[bicg::ALPHA] = (s0 != 0.0) ? S[bicg::RHO] / s0 : 0.0;
  } else if (stage == 2) {
    S[bicg::OMEGA] = (s1 != 0.0) ? s0 / s1 : 0.0;
  } else {
    const double rho = S[bicg::RHO], om = S[bicg::OMEGA];
    S[bicg::BETA] = (rho != 0.0 && om != 0.0) ? (s0 / rho) * (S[bicg::ALPHA] / om) : 0.0;
    S[bicg::RHO] = s0;
    S[bicg::RR] = s1;
  }
}

// ---------------------------------------------------------------------------
// Preconditioner z = M^-1 y: symmetric Gauss-Seidel on every tile (forward
// colours u-red, u-black, v-red, v-black, boundary rows; then backward), from
// a zero start, with the tile halo held at zero (block Jacobi across tiles).
// nsweeps symmetric sweeps per application.  The row right-hand side is y.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(IMAU_WG) k_prec_tile(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                       const KP kp, const CoefP c, CVec2 y, Vec2 z,
                                                       const int *__restrict__ active, int nsweeps) {
  __shared__ double red_unused;
  (void)red_unused;
  const Tile t = tiles[imau_xcd_tile()];
  const Member m = mem[t.g];
  const int nx = m.nx, ny = m.ny;
  const int tid = threadIdx.y * IMAU_TJ + threadIdx.x;
  if (!active[t.g]) return;
  for (int q = tid; q < (IMAU_TI + 2) * (IMAU_TJ + 2); q += IMAU_WG) {
    g_sx[0][q / (IMAU_TJ + 2)][q % (IMAU_TJ + 2)] = 0.0;
    g_sx[1][q / (IMAU_TJ + 2)][q % (IMAU_TJ + 2)] = 0.0;
  }
  __syncthreads();
  const LAcc<0> U{t.i0, t.j0};
  const LAcc<1> V{t.i0, t.j0};
  const int j = t.j0 + (int)threadIdx.x;
  const int lj = threadIdx.x + 1;
  for (int sw = 0; sw < 2 * nsweeps; ++sw) {
    const bool fwd = (sw & 1) == 0;
    for (int q = 0; q < 5; ++q) {
      // forward: 0 1 2 3 B ; backward: 3 2 1 0 B
      const int ph = (q == 4) ? 4 : (fwd ? q : 3 - q);
      FOR_TILE_ROWS {
        const int i = t.i0 + ii;
        if (i >= nx || j >= ny) continue;
        const long long p = p2(m, j, i);
        const int li = ii + 1;
        if (ph < 4) {
          const bool isu = ph < 2;
          if (((i + j) & 1) != (ph & 1)) continue;
          if (isu && i >= nx - 1) continue;
          if (!isu && j >= ny - 1) continue;
          const uint8_t rt = isu ? c.rt_u[p] : c.rt_v[p];
          if (is_bnd(rt)) continue;
          double S, D, b;
          if (isu) {
            u_row(m, kp, c, j, i, rt, U, V, S, D, b);
            g_sx[0][li][lj] = (y.u[p] - S) / D;
          } else {
            v_row(m, kp, c, j, i, rt, U, V, S, D, b);
            g_sx[1][li][lj] = (y.v[p] - S) / D;
          }
        } else {
          if (i < nx - 1) {
            const uint8_t rt = c.rt_u[p];
            if (is_bnd(rt)) {
              double val = y.u[p];
              if (rt == RT_BND_INF) {
                int jo, io;
                bnd_src(rt, i, j, nx - 2, ny - 1, jo, io);
                val += g_sx[0][io - t.i0 + 1][jo - t.j0 + 1];
              }
              g_sx[0][li][lj] = val;
            }
          }
          if (j < ny - 1) {
            const uint8_t rt = c.rt_v[p];
            if (is_bnd(rt)) {
              double val = y.v[p];
              if (rt == RT_BND_INF) {
                int jo, io;
                bnd_src(rt, i, j, nx - 1, ny - 2, jo, io);
                val += g_sx[1][io - t.i0 + 1][jo - t.j0 + 1];
              }
              g_sx[1][li][lj] = val;
            }
          }
        }
      }
      __syncthreads();
    }
  }
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= nx || j >= ny) continue;
    const long long p = p2(m, j, i);
    if (i < nx - 1) z.u[p] = g_sx[0][ii + 1][lj];
    if (j < ny - 1) z.v[p] = g_sx[1][ii + 1][lj];
  }
}

// ---------------------------------------------------------------------------
// Global multicolour SOR (reference SOR semantics, src/utilities_module.f90:644-718,
// in a parallel ordering): one launch updates one colour in place.
// colour 0 u-red, 1 u-black, 2 v-red, 3 v-black, 4 boundary rows.
// With rhs.u != nullptr the right-hand side is rhs (boundary rows included)
// instead of the assembled b: the multigrid smoother on A z = y.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(IMAU_WG) k_sor_color(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                       const KP kp, const CoefP c, Vec2 x, int color, double omega,
                                                       const int *__restrict__ active, int track,
                                                       unsigned long long *__restrict__ resmax, CVec2 rhs) {
  __shared__ double red[IMAU_WG_I];
  TILE_SETUP
  double rmax = 0.0;
  const GAcc U{x.u, m}, V{x.v, m};
  const int nx = m.nx, ny = m.ny;
  if (j < ny && active[t.g]) {
    FOR_TILE_ROWS {
      const int i = t.i0 + ii;
      if (i >= nx) break;
      const long long p = p2(m, j, i);
      if (color < 4) {
        const bool isu = color < 2;
        if (((i + j) & 1) != (color & 1)) continue;
        if (isu && i >= nx - 1) continue;
        if (!isu && j >= ny - 1) continue;
        const uint8_t rt = isu ? c.rt_u[p] : c.rt_v[p];
        if (is_bnd(rt)) continue;
        double S, D, b;
        if (isu) u_row(m, kp, c, j, i, rt, U, V, S, D, b);
        else v_row(m, kp, c, j, i, rt, U, V, S, D, b);
        if (rhs.u) b = isu ? rhs.u[p] : rhs.v[p];
        double *xx = isu ? x.u : x.v;
        const double xo = xx[p];
        const double res = (S + D * xo - b) / D;
        rmax = fmax(rmax, fabs(res));
        xx[p] = (rt == RT_FREE) ? xo - omega * res : xo - res;
      } else {
        if (i < nx - 1 && is_bnd(c.rt_u[p])) {
          const uint8_t rt = c.rt_u[p];
          double val = rhs.u ? rhs.u[p] : 0.0;
          if (rt != RT_BND_ZERO) {
            int jo, io;
            bnd_src(rt, i, j, nx - 2, ny - 1, jo, io);
            val += x.u[p2(m, jo, io)];
          }
          rmax = fmax(rmax, fabs(x.u[p] - val));
          x.u[p] = val;
        }
        if (j < ny - 1 && is_bnd(c.rt_v[p])) {
          const uint8_t rt = c.rt_v[p];
          double val = rhs.v ? rhs.v[p] : 0.0;
          if (rt != RT_BND_ZERO) {
            int jo, io;
            bnd_src(rt, i, j, nx - 1, ny - 2, jo, io);
            val += x.v[p2(m, jo, io)];
          }
          rmax = fmax(rmax, fabs(x.v[p] - val));
          x.v[p] = val;
        }
      }
    }
  }
  if (track) {
    const double w = wave_max(rmax);
    if (threadIdx.x == 0) red[threadIdx.y] = w;
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      double r = red[0];
      for (int q = 1; q < IMAU_WG_I; ++q) r = fmax(r, red[q]);
      typedef __attribute__((address_space(1))) unsigned long long gu64;
      __hip_atomic_fetch_max((gu64 *)(resmax + t.g), (unsigned long long)__double_as_longlong(r), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

namespace imau {
void apply(const Launch &L, const KP &kp, const Coef &c, const double *xu, const double *xv, double *yu, double *yv,
           const double *wu, const double *wv, const int *active, double *partial) {
  auto k = kp.cross ? k_apply<1> : k_apply<0>;
  hipLaunchKernelGGL(k, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, to_p(c), CVec2{xu, xv}, Vec2{yu, yv},
                     CVec2{wu, wv}, active, partial);
}
void bicg_init(const Launch &L, const KP &kp, const Coef &c, const double *xu, const double *xv, double *ru, double *rv,
               double *rhu, double *rhv, double *pu, double *pv, double *vu, double *vv, const int *active,
               double *partial) {
  hipLaunchKernelGGL(k_bicg_init, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, to_p(c), CVec2{xu, xv}, Vec2{ru, rv},
                     Vec2{rhu, rhv}, Vec2{pu, pv}, Vec2{vu, vv}, active, partial);
}
void bicg_p(const Launch &L, const double *ru, const double *rv, double *pu, double *pv, const double *vu,
            const double *vv, const double *sc, const int *active) {
  hipLaunchKernelGGL(k_bicg_p, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, CVec2{ru, rv}, Vec2{pu, pv}, CVec2{vu, vv},
                     sc, active);
}
void bicg_s(const Launch &L, double *ru, double *rv, const double *vu, const double *vv, const double *sc,
            const int *active) {
  hipLaunchKernelGGL(k_bicg_s, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, Vec2{ru, rv}, CVec2{vu, vv}, sc, active);
}
void bicg_xr(const Launch &L, double *xu, double *xv, double *ru, double *rv, const double *phu, const double *phv,
             const double *shu, const double *shv, const double *tu, const double *tv, const double *rhu,
             const double *rhv, const double *sc, const int *active, double *partial) {
  hipLaunchKernelGGL(k_bicg_xr, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, Vec2{xu, xv}, Vec2{ru, rv},
                     CVec2{phu, phv}, CVec2{shu, shv}, CVec2{tu, tv}, CVec2{rhu, rhv}, sc, active, partial);
}
void bicg_scalar(hipStream_t s, int ngrids, const int *tstart, const double *partial, double *sc, double *out,
                 const int *active, int stage) {
  hipLaunchKernelGGL(k_bicg_scalar, dim3(ngrids), dim3(256), 0, s, tstart, partial, sc, out, active, stage);
}
void prec_tile(const Launch &L, const KP &kp, const Coef &c, const double *yu, const double *yv, double *zu, double *zv,
               const int *active, int nsweeps) {
  hipLaunchKernelGGL(k_prec_tile, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, to_p(c), CVec2{yu, yv},
                     Vec2{zu, zv}, active, nsweeps);
}
void sor_color(const Launch &L, const KP &kp, const Coef &c, double *xu, double *xv, int color, double omega,
               const int *active, bool track, unsigned long long *resmax, const double *ru, const double *rv) {
  hipLaunchKernelGGL(k_sor_color, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, to_p(c), Vec2{xu, xv}, color, omega,
                     active, track ? 1 : 0, resmax, CVec2{ru, rv});
}
}  // namespace imau

// ---------------------------------------------------------------------------
// GCR (generalised conjugate residual, restarted every IMAU_GCR_M directions),
// right-preconditioned: one preconditioner application and one operator
// application per iteration.  For each member the search directions Q_i = A Z_i
// are kept orthonormal; scalars live on the device (sc[g*8 + ...]).
// ---------------------------------------------------------------------------
namespace gcr {
enum { BETA = 0, INVN = IMAU_GCR_M, ALPHA = IMAU_GCR_M + 1, NV = 2 * IMAU_GCR_M + 2 };
}

// partial sums per tile: (q,Q_i) and (r,Q_i) for i < nj, (q,q), (r,q)
__global__ void __launch_bounds__(IMAU_WG) k_gcr_dots(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                      CVec2 q, CVec2 r, GList Q, int nj,
                                                      const int *__restrict__ active, double *__restrict__ partial) {
  TILE_SETUP
  double acc[gcr::NV];
#pragma unroll
  for (int k = 0; k < gcr::NV; ++k) acc[k] = 0.0;
  if (j < m.ny && active[t.g]) {
    FOR_TILE_ROWS {
      const int i = t.i0 + ii;
      if (i >= m.nx) break;
      const long long p = p2(m, j, i);
      for (int comp = 0; comp < 2; ++comp) {
        if (comp == 0 ? (i >= m.nx - 1) : (j >= m.ny - 1)) continue;
        const double qv = (comp == 0 ? q.u : q.v)[p], rv = (comp == 0 ? r.u : r.v)[p];
#pragma unroll
        for (int k = 0; k < IMAU_GCR_M; ++k)
          if (k < nj) {
            const double w = (comp == 0 ? Q.u[k] : Q.v[k])[p];
            acc[k] += qv * w;
            acc[IMAU_GCR_M + k] += rv * w;
          }
        acc[2 * IMAU_GCR_M] += qv * qv;
        acc[2 * IMAU_GCR_M + 1] += rv * qv;
      }
    }
  }
  block_sum_store<gcr::NV>(acc, partial + gcr::NV * (long long)tb);
}

// per member: beta_i = (q,Q_i), q' = q - sum beta_i Q_i, |q'| from the Gram
// identity, alpha = (r, q'/|q'|)
__global__ void __launch_bounds__(256) k_gcr_scalar(const int *__restrict__ tstart, const double *__restrict__ partial,
                                                    double *__restrict__ sc, const int *__restrict__ active, int nj) {
  const int g = blockIdx.x;
  if (!active[g]) return;
  const int a = tstart[g], b = tstart[g + 1];
  __shared__ double red[gcr::NV][256];
  double s[gcr::NV];
#pragma unroll
  for (int q = 0; q < gcr::NV; ++q) s[q] = 0.0;
  for (int k = a + (int)threadIdx.x; k < b; k += 256)
#pragma unroll
    for (int q = 0; q < gcr::NV; ++q) s[q] += partial[(long long)gcr::NV * k + q];
#pragma unroll
  for (int q = 0; q < gcr::NV; ++q) red[q][threadIdx.x] = s[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int q = 0; q < gcr::NV; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double *S = sc + g * bicg::NSC;
  double qq = red[2 * IMAU_GCR_M][0], rq = red[2 * IMAU_GCR_M + 1][0];
  for (int k = 0; k < IMAU_GCR_M; ++k) {
    const double beta = (k < nj) ? red[k][0] : 0.0;
    S[gcr::BETA + k] = beta;
    qq -= beta * beta;
    rq -= beta * ((k < nj) ? red[IMAU_GCR_M + k][0] : 0.0);
  }
  const bool ok = qq > 0.0 && isfinite(qq);
  const double inv = ok ? 1.0 / sqrt(qq) : 0.0;
  S[gcr::INVN] = inv;
  S[gcr::ALPHA] = ok ? rq * inv : 0.0;
}

// Q_nj = (q - sum beta_i Q_i)/|q'|, Z_nj likewise, x += alpha Z_nj, r -= alpha Q_nj;
// q and z are read from, and replaced in, the slots Q_nj and Z_nj; partial (r,r)
__global__ void __launch_bounds__(IMAU_WG) k_gcr_update(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                        GList Q, GList Z, int nj, Vec2 x, Vec2 r,
                                                        const double *__restrict__ sc, const int *__restrict__ active,
                                                        double *__restrict__ partial) {
  TILE_SETUP
  double acc[2] = {0.0, 0.0};
  if (j < m.ny && active[t.g]) {
    const double *S = sc + t.g * bicg::NSC;
    const double inv = S[gcr::INVN], al = S[gcr::ALPHA];
    double be[IMAU_GCR_M];
#pragma unroll
    for (int k = 0; k < IMAU_GCR_M; ++k) be[k] = S[gcr::BETA + k];
    FOR_TILE_ROWS {
      const int i = t.i0 + ii;
      if (i >= m.nx) break;
      const long long p = p2(m, j, i);
      for (int comp = 0; comp < 2; ++comp) {
        if (comp == 0 ? (i >= m.nx - 1) : (j >= m.ny - 1)) continue;
        double *qn = const_cast<double *>(comp == 0 ? Q.u[nj] : Q.v[nj]);
        double *zn = const_cast<double *>(comp == 0 ? Z.u[nj] : Z.v[nj]);
        double qv = qn[p], zv = zn[p];
#pragma unroll
        for (int k = 0; k < IMAU_GCR_M; ++k)
          if (k < nj) {
            qv -= be[k] * (comp == 0 ? Q.u[k] : Q.v[k])[p];
            zv -= be[k] * (comp == 0 ? Z.u[k] : Z.v[k])[p];
          }
        qv *= inv;
        zv *= inv;
        qn[p] = qv;
        zn[p] = zv;
        (comp == 0 ? x.u : x.v)[p] += al * zv;
        double &rr = (comp == 0 ? r.u : r.v)[p];
        rr -= al * qv;
        acc[0] += rr * rr;
      }
    }
  }
  block_sum_store<2>(acc, partial + 2 * (long long)tb);
}

namespace imau {
void gcr_dots(const Launch &L, const double *qu, const double *qv, const double *ru, const double *rv, const GList &Q,
              int nj, const int *active, double *partial) {
  hipLaunchKernelGGL(k_gcr_dots, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, CVec2{qu, qv}, CVec2{ru, rv}, Q, nj,
                     active, partial);
}
void gcr_scalar(hipStream_t s, int ngrids, const int *tstart, const double *partial, double *sc, const int *active,
                int nj) {
  hipLaunchKernelGGL(k_gcr_scalar, dim3(ngrids), dim3(256), 0, s, tstart, partial, sc, active, nj);
}
void gcr_update(const Launch &L, const GList &Q, const GList &Z, int nj, double *xu, double *xv, double *ru, double *rv,
                const double *sc, const int *active, double *partial) {
  hipLaunchKernelGGL(k_gcr_update, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, Q, Z, nj, Vec2{xu, xv},
                     Vec2{ru, rv}, sc, active, partial);
}
}  // namespace imau

// ---------------------------------------------------------------------------
// Anderson acceleration of the Picard map u -> G(u) (the limited solution of
// the linearised system): store U_k = u and F_k = G(u) - u in slot k and
// accumulate (F_k, F_s) for every stored slot s; the per-member mixing
// weights are solved on the host and applied by k_post.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(IMAU_WG) k_and_prep(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                      const KP kp, const Fields f, const double *__restrict__ xu,
                                                      const double *__restrict__ xv, AMix mix, int slot,
                                                      const int *__restrict__ active, double *__restrict__ partial) {
  TILE_SETUP
  double acc[IMAU_AND_SLOTS];
#pragma unroll
  for (int q = 0; q < IMAU_AND_SLOTS; ++q) acc[q] = 0.0;
  const int nx = m.nx, ny = m.ny;
  const double vmax = kp.vel_max;
  if (j < ny && active[t.g]) {
    double *Uu = const_cast<double *>(mix.U.u[slot]), *Uv = const_cast<double *>(mix.U.v[slot]);
    double *Fu = const_cast<double *>(mix.F.u[slot]), *Fv = const_cast<double *>(mix.F.v[slot]);
    FOR_TILE_ROWS {
      const int i = t.i0 + ii;
      if (i >= nx) break;
      const long long p = p2(m, j, i);
      if (i < nx - 1) {  // the velocity limit of k_post (:1432-1498)
        double x = xu[p];
        double vcx;
        if (j == 0) vcx = (xv[p2(m, 0, i)] + xv[p2(m, 0, i + 1)]) / 2.0;
        else if (j == ny - 1) vcx = (xv[p2(m, ny - 2, i)] + xv[p2(m, ny - 2, i + 1)]) / 2.0;
        else vcx = (xv[p2(m, j - 1, i)] + xv[p2(m, j - 1, i + 1)] + xv[p2(m, j, i)] + xv[p2(m, j, i + 1)]) / 4.0;
        const double ua = sqrt(x * x + vcx * vcx);
        if (ua > vmax) x = x * vmax / ua;
        const double uo = f.u[p], fk = x - uo;
        Uu[p] = uo;
        Fu[p] = fk;
#pragma unroll
        for (int q = 0; q < IMAU_AND_SLOTS; ++q)
          if (q < mix.ns) acc[q] += fk * (q == slot ? fk : mix.F.u[q][p]);
      }
      if (j < ny - 1) {
        double x = xv[p];
        double ucy;
        if (i == 0) ucy = (xu[p2(m, j, 0)] + xu[p2(m, j + 1, 0)]) / 2.0;
        else if (i == nx - 1) ucy = (xu[p2(m, j, nx - 2)] + xu[p2(m, j + 1, nx - 2)]) / 2.0;
        else ucy = (xu[p2(m, j, i - 1)] + xu[p2(m, j, i)] + xu[p2(m, j + 1, i - 1)] + xu[p2(m, j + 1, i)]) / 4.0;
        const double ua = sqrt(ucy * ucy + x * x);
        if (ua > vmax) x = x * vmax / ua;
        const double vo = f.v[p], fk = x - vo;
        Uv[p] = vo;
        Fv[p] = fk;
#pragma unroll
        for (int q = 0; q < IMAU_AND_SLOTS; ++q)
          if (q < mix.ns) acc[q] += fk * (q == slot ? fk : mix.F.v[q][p]);
      }
    }
  }
  block_sum_store<IMAU_AND_SLOTS>(acc, partial + IMAU_AND_SLOTS * (long long)tb);
}

namespace imau {
void and_prep(const Launch &L, const KP &kp, const Fields &f, const double *xu, const double *xv, const AMix &mix,
              int slot, const int *active, double *partial) {
  hipLaunchKernelGGL(k_and_prep, dim3(L.ntiles), wg(), 0, L.s, L.mem, L.tiles, kp, f, xu, xv, mix, slot, active,
                     partial);
}
void member_sum_and(hipStream_t s, int ngrids, const int *tstart, const double *partial, double *out) {
  hipLaunchKernelGGL(k_member_sum<IMAU_AND_SLOTS>, dim3(ngrids), dim3(256), 0, s, tstart, partial, out);
}
}  // namespace imau
