// dev_common.h -- shared device-side definitions of the MI355X velocity path.
//
// Device layout (chosen for coalescing, not the host's):
//   * every 2D field of member g (a, cx, cy and b grids alike) lives in an
//     a-grid-shaped slab with leading dimension ny:  p2(j,i) = off + i*ny + j.
//     cx/cy/b fields simply leave their last column / row unused.
//   * every 3D field is stored k-plane major: p3(k,j,i) = off3 + k*nx*ny + i*ny + j,
//     so a wavefront reading one k level of 64 neighbouring columns is one
//     contiguous 512-byte access.
//   * members of an ensemble are concatenated; a tile work-list maps each
//     workgroup to (member, i0, j0), so one launch sweeps uneven grids.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define IMAU_TJ 64          // tile extent along j (one wavefront)
#define IMAU_TI 8           // tile extent along i
#define IMAU_WG_I 4         // wavefronts per workgroup along i
#define IMAU_WG (IMAU_TJ * IMAU_WG_I)
// fused smoother sweep: tiles of IMAU_FTI x IMAU_FTJ cells staged in LDS with a
// recomputed halo of IMAU_FH cells (four colour phases + the boundary rows)
#define IMAU_FTI 16
#define IMAU_FTJ 64
#define IMAU_FH 5
// tile height of the 3D viscosity kernel (16 and 32 measured slower than 8 on
// the 64-member ensemble: fewer workgroups in flight)
#define IMAU_VTI 8
// levels of the previous viscosity the 3D viscosity kernel loads ahead (1, 2,
// 3, 4 measured 561, 610, 700, 692 us on the 64-member ensemble: deeper
// prefetch costs occupancy)
#define IMAU_VISC_PF 1
// rows marched by one wavefront of the row-marching 3D viscosity kernel
// (nz = 15): each lane keeps its column's viscosity at rows i-1, i, i+1 in an
// LDS ring, so every level of the previous viscosity is read once, plus two
// halo rows per tile.  0 selects the IMAU_VTI-row tiled kernel instead, the
// default (measured 3x faster than 32-row marching on the 64-member ensemble:
// one 64-lane wavefront per tile leaves too few waves in flight).
// imau_vel_problem_autotune times the candidates IMAU_MTI_CANDIDATES on the
// problem itself and keeps the fastest.
#define IMAU_MTI 0
#define IMAU_MTI_CANDIDATES {0, 4, 8, 16, 32}
#define IMAU_MNZ 15

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b on XCD
// b % 8), each with its own L2.  Tile kernels run tile imau_xcd_tile() rather
// than tile b: XCD x takes the x-th contiguous eighth of the spatially ordered
// tile list, so the halo rows and columns that neighbouring tiles share are
// fetched into one L2 rather than two.  A bijection of [0, gridDim.x); per-tile
// partial sums are stored at the tile's index, so each member's partials stay
// contiguous.
__device__ __forceinline__ int imau_xcd_tile() {
  const int n = (int)gridDim.x, b = (int)blockIdx.x;
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

// Anderson acceleration of the Picard loop: up to IMAU_AND_SLOTS stored
// iterates (IMAU_AND_SLOTS - 1 differences)
#define IMAU_AND_SLOTS 4

// restart length of the GCR Krylov solver
#define IMAU_GCR_M 4
// pointers to the stored GCR directions (u and v components)
struct GList {
  const double *u[IMAU_GCR_M];
  const double *v[IMAU_GCR_M];
};
struct SList {
  const double *u[IMAU_AND_SLOTS];
  const double *v[IMAU_AND_SLOTS];
};
// Anderson state for k_and_prep / k_post: slots U_s, F_s; ns valid slots;
// w = per-member weights a_s then b_s (2 * IMAU_AND_SLOTS each) of
// u_new = sum_s a_s U_s + b_s F_s; null = the configured relaxation
struct AMix {
  SList U, F;
  int ns;
  const double *w;
};
#define IMAU_NZMAX 32

// physical constants, src/parameters_module.f90:11-24
#define D_ICE_DENSITY      910.0
#define D_SEAWATER_DENSITY 1028.0
#define D_GRAV             9.81
#define D_PI               3.141592653589793

struct Member {
  int nx, ny;
  double dx;
  long long off;   // 2D base offset
  long long off3;  // 3D base offset (= nz * off)
  long long offx;  // offset of the member's x(nx) coordinates
  long long offy;  // offset of the member's y(ny) coordinates
};

struct Tile {
  int g, i0, j0, pad;
};

// row types of the linearised system (one per u and per v unknown)
enum : uint8_t {
  RT_FREE = 0,       // 9-point interior row (src/ice_velocity_module.f90:2342-2589)
  RT_BND_INF = 1,    // domain boundary 'infinite'  (:1932-2191)
  RT_BND_ZERO = 2,   // domain boundary 'zero'
  RT_BND_PER = 3,    // domain boundary 'periodic'
  RT_ZERO = 4,       // shelf next to land / single-pixel ridge (BC margin mode, :1787-1796, :2224-2230)
  RT_MARGIN_I = 5,   // calving-front row, ice on the row's own side (i1 = i)
  RT_MARGIN_O = 6    // calving-front row, ice on the far side (i1 = i+1)
};

// kernel parameters: the configuration scalars, passed by value
struct KP {
  int nz;
  double zeta[IMAU_NZMAX];
  int dyn, margin, slid, ideal, hydro, rough, prough;
  int bc_u[4], bc_v[4];  // west, east, south, north
  int cross, glsub, freeze;
  double n_flow, eps0, visc_min, beta_max, vel_max, relax;
  double delta_v, wm, q_plastic, u_threshold, zi_ut, zi_p, dT_freeze;
  double hydro_min, hydro_max, phi_hb_min, phi_hb_max, phi_min, phi_max;
  double ssa_L, ssa_m, ismip_L;
  double F_int_min;
};

// device pointers of every field (ld ny slabs, see above)
struct Fields {
  // static inputs
  const double *Hi, *Hb, *SL, *Hs, *TAF, *dHs_dx_a, *dHs_dy_a, *dHi_dx_a, *dHi_dy_a, *dHb_dt, *A_vav;
  const int32_t *m_ice, *m_sheet, *m_ocean;
  const double *Hi_cx, *dHs_dx_cx, *dHs_dy_cx, *Hi_cy, *dHs_dx_cy, *dHs_dy_cy, *Hi_b;
  const double *Ti3, *Tpmp3;           // ice / pressure-melting temperature (3D), may be null
  const double *A3;                    // A_flow_3D_a (3D)
  double *B3;                          // 0.5*A^(-1/n) (3D), derived from A3 once per solve
  double *beta_sq, *alpha_sq, *phi_fric, *tauc;
  // state
  double *u, *v;                       // cx / cy velocity (vav for DIVA, SSA field for SSA)
  double *taub_cx, *taub_cy;
  double *visc_a, *visc_b;             // 3D viscosity, double buffer (a: current, b: other)
  // per-iteration coefficients
  double *N_a, *N_b, *N_cx, *N_cy, *visc_int, *F2, *beta_a, *beta_eff_a, *beta_eff_cx, *beta_eff_cy;
  // per-solve data
  double *taudx, *taudy, *fg_a, *fg_cx, *fg_cy, *fg_b, *overb, *pore, *Neff, *front_rhs;
  double *qNW, *qNE, *qSW, *qSE;       // grounded-fraction quadrants (scratch)
  int32_t *combi, *fin, *fout;
  uint8_t *rt_u, *rt_v;
  // linear solve
  double *xu0, *xv0, *xu1, *xv1;       // double-buffered iterate
  double *u_prev, *v_prev, *u_base, *v_base;
  const double *ismip_e;               // ISMIP-HOM E beta per global column (offx + i): 0 or 1e30
  double *errt_u, *errt_v;             // this iteration's error estimate (null: not requested)
  const double *chk_u, *chk_v;         // the velocities its |u| > 1e-6 test reads (SSA: u_vav of the state)
  double *eps_h;                       // horizontal strain-rate invariant on the a-grid (k_strain -> k_visc)
};

// one member of a coarse multigrid level: u unknowns on (nyu, nxu), v on
// (nyv, nxv), stored in a slab of ny = max(nyu,nyv) by nx = max(nxu,nxv)
// coarsest multigrid level: every member at most IMAU_CMAX x IMAU_CMAX (one workgroup, LDS)
#define IMAU_CMAX 32

struct CMember {
  int nxu, nyu, nxv, nyv, nx, ny;
  long long off;
};

// a coarse level: explicit staggered 9-point stencils.  u-row slots:
// 0 C, 1 U(J,I-1), 2 U(J,I+1), 3 U(J-1,I), 4 U(J+1,I), 5 V(J-1,I), 6 V(J-1,I+1),
// 7 V(J,I), 8 V(J,I+1).  v-row slots: 0 C, 1 V(J,I-1), 2 V(J,I+1), 3 V(J-1,I),
// 4 V(J+1,I), 5 U(J,I-1), 6 U(J,I), 7 U(J+1,I-1), 8 U(J+1,I).  cu[k*na + p].
template <class T>
struct CLevelT {
  const CMember *mem;
  const Tile *tiles;
  const int *depth;  // per member: index of its coarsest level (>= 1)
  int lev;           // index of this level (1 = first coarse level)
  int ntiles;
  long long na;
  T *cu, *cv;        // 9 coefficient planes per component: c[k * na + p]
  T *xu, *xv, *bu, *bv;
};

// buffers of the final (once per solve) phase
struct OutF {
  double *F1, *u3, *v3, *u3sia, *v3sia, *w3, *u3a, *v3a;        // 3D
  double *uvav, *vvav;                                          // cx / cy depth-averaged velocity
  double *u_vav_a, *v_vav_a, *u_surf_a, *v_surf_a, *u_base_a, *v_base_a;
  double *uabs_vav, *uabs_surf, *uabs_base, *R_shear, *w_surf, *w_base;
};

__device__ __forceinline__ long long p2(const Member &m, int j, int i) { return m.off + (long long)i * m.ny + j; }
__device__ __forceinline__ long long p3(const Member &m, int k, int j, int i) {
  return m.off3 + (long long)k * m.nx * m.ny + (long long)i * m.ny + j;
}

// calc_staggered_margin, src/ice_velocity_module.f90:1300-1319
__device__ __forceinline__ double stag_margin(double v0, double v1, double H0, double H1) {
  if (H0 > 0.0 && H1 == 0.0) return v0;
  if (H0 == 0.0 && H1 > 0.0) return v1;
  return 0.5 * (v0 + v1);
}

// the same choice made once from the thicknesses, then applied to values
__device__ __forceinline__ int stag_sel(double H0, double H1) {
  if (H0 > 0.0 && H1 == 0.0) return 0;
  if (H0 == 0.0 && H1 > 0.0) return 1;
  return 2;
}
template <class T = double>
__device__ __forceinline__ T stag_pick(T v0, T v1, int sel) {
  return sel == 0 ? v0 : (sel == 1 ? v1 : (T)0.5 * (v0 + v1));
}

__device__ __forceinline__ int iclamp(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }
