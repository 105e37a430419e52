// mg.hip -- aggregation multigrid preconditioner for the linearised staggered
// DIVA/SSA system (src/ice_velocity_module.f90:1762-2589), in FP64 or FP32.
//
// Level 0 is the matrix-free fine operator (stencil.h).  Coarse levels hold
// explicit staggered 9-point stencils built on the device as Galerkin products
// R A P, with P piecewise constant on 2 x 2 aggregates of each velocity
// component ('infinite' boundary unknowns follow the aggregate of their
// interior source; 'zero' and 'periodic' boundary unknowns get no coarse
// counterpart) and R = P^T over the interior and calving-front rows.  A 2 x 2
// aggregation of the staggered u (cx) and v (cy) points keeps the u-v coupling
// pattern of the fine rows, so every coarse level has the same 9-point form.
// Smoothing is one four-colour Gauss-Seidel sweep (u red, u black, v red,
// v black, then the boundary rows at level 0) before and after the coarse
// correction; each member's coarsest level (at most IMAU_CMAX x IMAU_CMAX)
// is solved by symmetric Gauss-Seidel sweeps in LDS by one workgroup.
#include <type_traits>

#include "mg.h"
#include "stencil.h"

// the FP32 V-cycle (a preconditioner) divides by the diagonal with the
// hardware reciprocal (v_rcp_f32, about 1 ulp); the FP64 one divides exactly
__device__ __forceinline__ float mdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double mdiv(double a, double b) { return a / b; }

#define TILE_SETUP                                  \
  const int tb = imau_xcd_tile();                   \
  const Tile t = tiles[tb];                         \
  const Member m = mem[t.g];                        \
  const int j = t.j0 + (int)threadIdx.x;

#define CTILE_SETUP                                  \
  const int tb = imau_xcd_tile();                    \
  const Tile t = L.tiles[tb];                        \
  const CMember cm = L.mem[t.g];                     \
  const int J = t.j0 + (int)threadIdx.x;

#define FOR_TILE_ROWS for (int ii = threadIdx.y; ii < IMAU_TI; ii += IMAU_WG_I)

namespace {

template <class T>
struct VecT {
  T *u, *v;
};
template <class T>
struct CVecT {
  const T *u, *v;
};

__device__ __forceinline__ long long cp2(const CMember &m, int J, int I) { return m.off + (long long)I * m.ny + J; }

__device__ __forceinline__ bool cvalid(const CMember &m, int comp, int J, int I) {
  return comp == 0 ? (J >= 0 && I >= 0 && J < m.nyu && I < m.nxu) : (J >= 0 && I >= 0 && J < m.nyv && I < m.nxv);
}

// coarse slot of the coupling from a row of component rowc to an unknown of
// component colc at aggregate offset (dJ, dI); -1 outside the 9-point pattern
__device__ __forceinline__ int agg_slot(int rowc, int colc, int dJ, int dI) {
  if (rowc == colc) {
    if (dJ == 0 && dI == 0) return 0;
    if (dJ == 0 && dI == -1) return 1;
    if (dJ == 0 && dI == 1) return 2;
    if (dJ == -1 && dI == 0) return 3;
    if (dJ == 1 && dI == 0) return 4;
    return -1;
  }
  if (rowc == 0) {  // u row, v column
    if (dJ == -1 && dI == 0) return 5;
    if (dJ == -1 && dI == 1) return 6;
    if (dJ == 0 && dI == 0) return 7;
    if (dJ == 0 && dI == 1) return 8;
    return -1;
  }
  if (dJ == 0 && dI == -1) return 5;  // v row, u column
  if (dJ == 0 && dI == 0) return 6;
  if (dJ == 1 && dI == -1) return 7;
  if (dJ == 1 && dI == 0) return 8;
  return -1;
}

// unknown (component cc, JJ, II) addressed by slot k of a row of component rowc at (J, I)
__device__ __forceinline__ void slot_pos(int rowc, int k, int J, int I, int &cc, int &JJ, int &II) {
  if (k < 5) {
    cc = rowc;
    JJ = J + (k == 3 ? -1 : (k == 4 ? 1 : 0));
    II = I + (k == 1 ? -1 : (k == 2 ? 1 : 0));
  } else if (rowc == 0) {
    cc = 1;
    JJ = J + (k < 7 ? -1 : 0);
    II = I + ((k & 1) ? 0 : 1);
  } else {
    cc = 0;
    JJ = J + (k < 7 ? 0 : 1);
    II = I + ((k & 1) ? -1 : 0);
  }
}

// nonzeros of the fine u row (interior or calving front) at (j,i):
// f(comp, dj, di, value) for the unknown (comp, j + dj, i + di).  The offsets
// of an interior row are constants, so a caller that unrolls over the rows of
// an aggregate resolves every coarse slot at compile time.
// CROSS: 1 / 0 fixes the cross-term choice at compile time, -1 reads kp.cross
template <int CROSS, class F>
__device__ __forceinline__ void u_couplings(const Member &m, const KP &kp, const CoefT<double> &c, int j, int i,
                                            uint8_t rt, F &&f) {
  const double dx2 = m.dx * m.dx;
  if (rt == RT_FREE) {
    if (CROSS >= 0 ? CROSS != 0 : kp.cross != 0) {
      const double Na0 = c.Na[p2(m, j, i)], Na1 = c.Na[p2(m, j, i + 1)];
      const double Nb0 = c.Nb[p2(m, j, i)], Nbm = c.Nb[p2(m, j - 1, i)];
      f(0, 0, 1, 4.0 * Na1); f(0, 0, -1, 4.0 * Na0); f(0, 1, 0, Nb0); f(0, -1, 0, Nbm);
      f(0, 0, 0, -4.0 * (Na1 + Na0) + -1.0 * (Nb0 + Nbm) - c.bcx[p2(m, j, i)] * dx2);
      f(1, 0, 0, -2.0 * Na0 - Nb0); f(1, 0, 1, 2.0 * Na1 + Nb0);
      f(1, -1, 0, 2.0 * Na0 + Nbm); f(1, -1, 1, -2.0 * Na1 - Nbm);
    } else {
      const double Ncx = c.Ncx[p2(m, j, i)];
      f(0, 0, 1, 4.0); f(0, 0, -1, 4.0); f(0, 1, 0, 1.0); f(0, -1, 0, 1.0);
      f(0, 0, 0, -10.0 - dx2 * c.bcx[p2(m, j, i)] / Ncx);
      f(1, 0, 0, -3.0); f(1, 0, 1, 3.0); f(1, -1, 0, 3.0); f(1, -1, 1, -3.0);
    }
  } else {  // RT_MARGIN_I / RT_MARGIN_O
    const int d1 = (rt == RT_MARGIN_I) ? 0 : 1;  // column i1 = i + d1
    const double Na = c.Na[p2(m, j, i + d1)] / m.dx;
    f(0, 0, d1 - 1, -4.0 * Na); f(0, 0, d1, 4.0 * Na); f(1, -1, d1, -2.0 * Na); f(1, 0, d1, 2.0 * Na);
  }
}
template <int CROSS, class F>
__device__ __forceinline__ void v_couplings(const Member &m, const KP &kp, const CoefT<double> &c, int j, int i,
                                            uint8_t rt, F &&f) {
  const double dx2 = m.dx * m.dx;
  if (rt == RT_FREE) {
    if (CROSS >= 0 ? CROSS != 0 : kp.cross != 0) {
      const double Na0 = c.Na[p2(m, j, i)], Na1 = c.Na[p2(m, j + 1, i)];
      const double Nb0 = c.Nb[p2(m, j, i)], Nbm = c.Nb[p2(m, j, i - 1)];
      f(1, 1, 0, 4.0 * Na1); f(1, -1, 0, 4.0 * Na0); f(1, 0, 1, Nb0); f(1, 0, -1, Nbm);
      f(1, 0, 0, -4.0 * (Na1 + Na0) + -1.0 * (Nb0 + Nbm) - c.bcy[p2(m, j, i)] * dx2);
      f(0, 1, -1, -2.0 * Na1 - Nbm); f(0, 1, 0, 2.0 * Na1 + Nb0);
      f(0, 0, -1, 2.0 * Na0 + Nbm); f(0, 0, 0, -2.0 * Na0 - Nb0);
    } else {
      const double Ncy = c.Ncy[p2(m, j, i)];
      f(1, 1, 0, 4.0); f(1, -1, 0, 4.0); f(1, 0, 1, 1.0); f(1, 0, -1, 1.0);
      f(1, 0, 0, -10.0 - dx2 * c.bcy[p2(m, j, i)] / Ncy);
      f(0, 1, -1, -3.0); f(0, 1, 0, 3.0); f(0, 0, -1, 3.0); f(0, 0, 0, -3.0);
    }
  } else {
    const int d1 = (rt == RT_MARGIN_I) ? 0 : 1;  // row j1 = j + d1
    const double Na = c.Na[p2(m, j + d1, i)] / m.dx;
    f(1, d1 - 1, 0, -4.0 * Na); f(1, d1, 0, 4.0 * Na); f(0, d1, -1, -2.0 * Na); f(0, d1, 0, 2.0 * Na);
  }
}

// level-1 aggregate of the fine unknown (comp, jj, ii); false if it has none
__device__ __forceinline__ bool fine_agg(const Member &m, const uint8_t *rt_u, const uint8_t *rt_v, int comp, int jj,
                                         int ii, int &J, int &I) {
  uint8_t rt = comp == 0 ? rt_u[p2(m, jj, ii)] : rt_v[p2(m, jj, ii)];
  if (rt == RT_BND_INF) {
    int jo, io;
    bnd_src(rt, ii, jj, comp == 0 ? m.nx - 2 : m.nx - 1, comp == 0 ? m.ny - 1 : m.ny - 2, jo, io);
    jj = jo;
    ii = io;
    rt = comp == 0 ? rt_u[p2(m, jj, ii)] : rt_v[p2(m, jj, ii)];
  }
  if (!in_row(rt)) return false;
  J = jj >> 1;
  I = ii >> 1;
  return true;
}

__device__ __forceinline__ bool fine_valid(const Member &m, int comp, int j, int i) {
  return comp == 0 ? (i <= m.nx - 2 && j <= m.ny - 1) : (i <= m.nx - 1 && j <= m.ny - 2);
}

// ---------------------------------------------------------------------------
// Galerkin products
// ---------------------------------------------------------------------------
// fine (level 0, FP64 coefficients) -> level 1
// One level-1 cell, the components not marked plain: every coupling's
// unknown is mapped to its aggregate (fine_agg) and slot at run time.
template <class T, int CROSS>
__device__ __forceinline__ void gal0_cell(const Member &m, const KP &kp, const CoefT<double> &c, const CLevelT<T> &L,
                                          const CMember &cm, int J, int I, long long pc, int fl) {
  for (int comp = 0; comp < 2; ++comp) {
    if (!cvalid(cm, comp, J, I) || ((fl >> comp) & 1)) continue;
    double cf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int dj = 0; dj < 2; ++dj)
      for (int di = 0; di < 2; ++di) {
        const int j = 2 * J + dj, i = 2 * I + di;
        if (!fine_valid(m, comp, j, i)) continue;
        const uint8_t rt = comp == 0 ? c.rt_u[p2(m, j, i)] : c.rt_v[p2(m, j, i)];
        if (!in_row(rt)) continue;
        auto acc = [&](int cc, int oj, int oi, double val) {
          int Jc, Ic;
          if (!fine_agg(m, c.rt_u, c.rt_v, cc, j + oj, i + oi, Jc, Ic)) return;
          const int s = agg_slot(comp, cc, Jc - J, Ic - I);
          if (s >= 0) cf[s] += val;
        };
        if (comp == 0) u_couplings<CROSS>(m, kp, c, j, i, rt, acc);
        else v_couplings<CROSS>(m, kp, c, j, i, rt, acc);
      }
    T *dst = comp == 0 ? L.cu : L.cv;
#pragma unroll
    for (int k = 0; k < 9; ++k) dst[k * L.na + pc] = (T)cf[k];
  }
}

// fine -> level 1 on the cells that are not plain (boundary rows, calving
// fronts, zero rows and their neighbours): a compact list built with the
// flags, so no wavefront of the interior waits for one such lane
template <class T, int CROSS>
__global__ void __launch_bounds__(256) k_gal0_list(const Member *__restrict__ mem, const KP kp, const CoefT<double> c,
                                                   CLevelT<T> L, const int *__restrict__ active,
                                                   const uint8_t *__restrict__ plain, const int2 *__restrict__ list,
                                                   int n) {
  const int q = (int)(blockIdx.x * 256 + threadIdx.x);
  if (q >= n) return;
  const int2 e = list[q];
  if (!active[e.x]) return;
  const CMember cm = L.mem[e.x];
  const int I = e.y / cm.ny, J = e.y - I * cm.ny;
  const long long pc = cm.off + e.y;
  gal0_cell<T, CROSS>(mem[e.x], kp, c, L, cm, J, I, pc, plain[pc]);
}

// Plain aggregates: all four fine rows of a level-1 aggregate of component
// comp are interior rows (RT_FREE) and every unknown their couplings reach is
// an interior or calving-front unknown.  The aggregate of each such unknown
// is then (j >> 1, i >> 1), so a coupling's coarse slot follows from its
// offsets alone.  Bit comp of plain[level-1 cell]; once per solve (the row
// types are fixed for a solve).
__device__ __forceinline__ bool plain_row(const Member &m, const uint8_t *rt_u, const uint8_t *rt_v, int comp, int j,
                                          int i) {
  if (!fine_valid(m, comp, j, i)) return false;
  if ((comp == 0 ? rt_u[p2(m, j, i)] : rt_v[p2(m, j, i)]) != RT_FREE) return false;
  // the unknowns of u_couplings / v_couplings for an interior row
  bool ok;
  if (comp == 0)
    ok = in_row(rt_u[p2(m, j, i + 1)]) && in_row(rt_u[p2(m, j, i - 1)]) && in_row(rt_u[p2(m, j + 1, i)]) &&
         in_row(rt_u[p2(m, j - 1, i)]) && in_row(rt_v[p2(m, j, i)]) && in_row(rt_v[p2(m, j, i + 1)]) &&
         in_row(rt_v[p2(m, j - 1, i)]) && in_row(rt_v[p2(m, j - 1, i + 1)]);
  else
    ok = in_row(rt_v[p2(m, j + 1, i)]) && in_row(rt_v[p2(m, j - 1, i)]) && in_row(rt_v[p2(m, j, i + 1)]) &&
         in_row(rt_v[p2(m, j, i - 1)]) && in_row(rt_u[p2(m, j + 1, i - 1)]) && in_row(rt_u[p2(m, j + 1, i)]) &&
         in_row(rt_u[p2(m, j, i - 1)]) && in_row(rt_u[p2(m, j, i)]);
  return ok;
}
// The cells with a valid component that is not plain are appended to list
// (member, index within the member's level), counted in *n (zeroed first).
__global__ void __launch_bounds__(IMAU_WG) k_gal_plain_flags(const Member *__restrict__ mem,
                                                             const uint8_t *__restrict__ rt_u,
                                                             const uint8_t *__restrict__ rt_v, CLevelT<float> L,
                                                             uint8_t *__restrict__ plain, int2 *__restrict__ list,
                                                             int *__restrict__ n) {
  CTILE_SETUP
  if (J >= cm.ny) return;
  const Member m = mem[t.g];
  FOR_TILE_ROWS {
    const int I = t.i0 + ii;
    if (I >= cm.nx) break;
    int fl = 0, valid = 0;
    for (int comp = 0; comp < 2; ++comp) {
      bool ok = cvalid(cm, comp, J, I);
      if (ok) valid |= 1 << comp;
      for (int q = 0; q < 4 && ok; ++q) ok = plain_row(m, rt_u, rt_v, comp, 2 * J + (q >> 1), 2 * I + (q & 1));
      if (ok) fl |= 1 << comp;
    }
    plain[cp2(cm, J, I)] = (uint8_t)fl;
    // one atomic per wavefront: the lanes with a cell to append take
    // consecutive slots after the wavefront's base
    const bool add = (valid & ~fl) != 0;
    const unsigned long long mask = __ballot(add);
    if (mask) {
      const int lane = (int)__lane_id(), lead = __ffsll((long long)mask) - 1;
      int base = 0;
      if (lane == lead) base = atomicAdd(n, __popcll(mask));
      base = __shfl(base, lead);
      if (add) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = make_int2(t.g, I * cm.ny + J);
    }
  }
}

// one fine row (2J+DJ, 2I+DI) of a plain aggregate: every coupling offset is a
// literal after inlining, so the slot is a constant and cf stays in registers
template <int CROSS, int COMP, int DJ, int DI>
__device__ __forceinline__ void gal_plain_row(const Member &m, const KP &kp, const CoefT<double> &c, int J, int I,
                                              double (&cf)[9]) {
  auto acc = [&](int cc, int oj, int oi, double val) {
    const int s = agg_slot(COMP, cc, (DJ + oj) >> 1, (DI + oi) >> 1);
    if (s >= 0) cf[s] += val;
  };
  if (COMP == 0) u_couplings<CROSS>(m, kp, c, 2 * J + DJ, 2 * I + DI, RT_FREE, acc);
  else v_couplings<CROSS>(m, kp, c, 2 * J + DJ, 2 * I + DI, RT_FREE, acc);
}

// fine -> level 1 on the plain aggregates: the same sums in the same order as
// k_gal0, with straight-line loads and no aggregate look-ups
template <class T, int CROSS, int COMP>
__device__ __forceinline__ void gal_plain(const Member &m, const KP &kp, const CoefT<double> &c, const CLevelT<T> &L,
                                          int J, int I, long long pc) {
  double cf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  gal_plain_row<CROSS, COMP, 0, 0>(m, kp, c, J, I, cf);
  gal_plain_row<CROSS, COMP, 0, 1>(m, kp, c, J, I, cf);
  gal_plain_row<CROSS, COMP, 1, 0>(m, kp, c, J, I, cf);
  gal_plain_row<CROSS, COMP, 1, 1>(m, kp, c, J, I, cf);
  T *dst = COMP == 0 ? L.cu : L.cv;
#pragma unroll
  for (int k = 0; k < 9; ++k) dst[k * L.na + pc] = (T)cf[k];
}
template <class T, int CROSS>
__global__ void __launch_bounds__(IMAU_WG) k_gal0_plain(const Member *__restrict__ mem, const KP kp,
                                                        const CoefT<double> c, CLevelT<T> L,
                                                        const int *__restrict__ active,
                                                        const uint8_t *__restrict__ plain) {
  CTILE_SETUP
  if (J >= cm.ny || !active[t.g]) return;
  const Member m = mem[t.g];
  FOR_TILE_ROWS {
    const int I = t.i0 + ii;
    if (I >= cm.nx) break;
    const long long pc = cp2(cm, J, I);
    const int fl = plain[pc];
    if (fl & 1) gal_plain<T, CROSS, 0>(m, kp, c, L, J, I, pc);
    if (fl & 2) gal_plain<T, CROSS, 1>(m, kp, c, L, J, I, pc);
  }
}

// one fine cell (2J+DJ, 2I+DI) of a coarse cell whose fine cells and
// couplings all lie inside the level: slot and target are constants
template <class T, int COMP, int DJ, int DI>
__device__ __forceinline__ void galc_plain(const CLevelT<T> &Fl, const CMember &fm, int J, int I, double (&cf)[9]) {
  const long long pf = cp2(fm, 2 * J + DJ, 2 * I + DI);
  const T *src = COMP == 0 ? Fl.cu : Fl.cv;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    int cc, jj, i2;
    slot_pos(COMP, k, DJ, DI, cc, jj, i2);  // offsets from the coarse cell's first fine cell
    const int s = agg_slot(COMP, cc, jj >> 1, i2 >> 1);
    if (s >= 0) cf[s] += (double)src[k * Fl.na + pf];
  }
}

// level l (Fl) -> level l+1 (L), accumulated in FP64.  Away from the
// level's edges (the test below: every fine cell and every unknown its row
// reaches is valid) the sums run with constant slots, in the same order as
// the general loop.
template <class T>
__global__ void __launch_bounds__(IMAU_WG) k_galc(CLevelT<T> Fl, CLevelT<T> L, const int *__restrict__ active) {
  CTILE_SETUP
  if (J >= cm.ny || !active[t.g] || L.depth[t.g] < L.lev) return;
  const CMember fm = Fl.mem[t.g];
  FOR_TILE_ROWS {
    const int I = t.i0 + ii;
    if (I >= cm.nx) break;
    const long long pc = cp2(cm, J, I);
    const bool inner = J >= 1 && I >= 1;
    const bool plain_u = inner && 2 * J + 2 < fm.nyu && 2 * I + 2 < fm.nxu && 2 * J + 1 < fm.nyv && 2 * I + 2 < fm.nxv;
    const bool plain_v = inner && 2 * J + 2 < fm.nyv && 2 * I + 2 < fm.nxv && 2 * J + 2 < fm.nyu && 2 * I + 1 < fm.nxu;
    if (plain_u && cvalid(cm, 0, J, I)) {
      double cf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      galc_plain<T, 0, 0, 0>(Fl, fm, J, I, cf);
      galc_plain<T, 0, 0, 1>(Fl, fm, J, I, cf);
      galc_plain<T, 0, 1, 0>(Fl, fm, J, I, cf);
      galc_plain<T, 0, 1, 1>(Fl, fm, J, I, cf);
#pragma unroll
      for (int k = 0; k < 9; ++k) L.cu[k * L.na + pc] = (T)cf[k];
    }
    if (plain_v && cvalid(cm, 1, J, I)) {
      double cf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      galc_plain<T, 1, 0, 0>(Fl, fm, J, I, cf);
      galc_plain<T, 1, 0, 1>(Fl, fm, J, I, cf);
      galc_plain<T, 1, 1, 0>(Fl, fm, J, I, cf);
      galc_plain<T, 1, 1, 1>(Fl, fm, J, I, cf);
#pragma unroll
      for (int k = 0; k < 9; ++k) L.cv[k * L.na + pc] = (T)cf[k];
    }
    for (int comp = 0; comp < 2; ++comp) {
      if (!cvalid(cm, comp, J, I) || (comp == 0 ? plain_u : plain_v)) continue;
      double cf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      const T *src = comp == 0 ? Fl.cu : Fl.cv;
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const int j = 2 * J + dj, i = 2 * I + di;
          if (!cvalid(fm, comp, j, i)) continue;
          const long long pf = cp2(fm, j, i);
#pragma unroll
          for (int k = 0; k < 9; ++k) {
            int cc, jj, i2;
            slot_pos(comp, k, j, i, cc, jj, i2);
            if (!cvalid(fm, cc, jj, i2)) continue;
            const int s = agg_slot(comp, cc, (jj >> 1) - J, (i2 >> 1) - I);
            if (s >= 0) cf[s] += (double)src[k * Fl.na + pf];
          }
        }
      T *dst = comp == 0 ? L.cu : L.cv;
#pragma unroll
      for (int k = 0; k < 9; ++k) dst[k * L.na + pc] = (T)cf[k];
    }
  }
}

// ---------------------------------------------------------------------------
// coarse-level smoothing and transfers
// ---------------------------------------------------------------------------
// off-diagonal sum S and diagonal D of a coarse row at (J,I)
template <class T>
__device__ __forceinline__ void crow(const CLevelT<T> &L, const CMember &m, int comp, int J, int I, const T *xu,
                                     const T *xv, T &S, T &D) {
  const long long p = cp2(m, J, I);
  const T *cf = comp == 0 ? L.cu : L.cv;
  D = cf[p];
  S = (T)0.0;
#pragma unroll
  for (int k = 1; k < 9; ++k) {
    int cc, JJ, II;
    slot_pos(comp, k, J, I, cc, JJ, II);
    if (!cvalid(m, cc, JJ, II)) continue;
    S += cf[k * L.na + p] * (cc == 0 ? xu : xv)[cp2(m, JJ, II)];
  }
}

// one colour (0 u red, 1 u black, 2 v red, 3 v black) of Gauss-Seidel, in place
template <class T>
__global__ void __launch_bounds__(IMAU_WG) k_cgs(CLevelT<T> L, int color, const int *__restrict__ active) {
  CTILE_SETUP
  if (J >= cm.ny || !active[t.g] || L.depth[t.g] <= L.lev) return;
  const int comp = color >> 1;
  FOR_TILE_ROWS {
    const int I = t.i0 + ii;
    if (I >= cm.nx) break;
    if (((I + J) & 1) != (color & 1) || !cvalid(cm, comp, J, I)) continue;
    T S, D;
    crow(L, cm, comp, J, I, L.xu, L.xv, S, D);
    const long long p = cp2(cm, J, I);
    const T b = (comp == 0 ? L.bu : L.bv)[p];
    (comp == 0 ? L.xu : L.xv)[p] = (D != (T)0.0) ? mdiv(b - S, D) : (T)0.0;
  }
}

// One fused Gauss-Seidel sweep of a coarse level: the tile (IMAU_TI x IMAU_TJ)
// plus a halo of 4 cells is staged in LDS from (xiu, xiv) (zero when xiu is
// null), the four colour phases update shrinking regions (exact up to 3 - p
// cells beyond the tile after phase p) and the tile is written to (xou, xov),
// a different buffer, so the result equals four in-place colour launches.
#define CH 4
#define CRI (IMAU_TI + 2 * CH)
#define CRJ (IMAU_TJ + 2 * CH)
template <class T>
__global__ void __launch_bounds__(IMAU_WG) k_cgs_fused(CLevelT<T> L, const T *__restrict__ xiu,
                                                       const T *__restrict__ xiv, T *__restrict__ xou,
                                                       T *__restrict__ xov, int reverse,
                                                       const int *__restrict__ active) {
  __shared__ T sx[2][CRI * CRJ];
  const Tile t = L.tiles[imau_xcd_tile()];
  if (!active[t.g] || L.depth[t.g] <= L.lev) return;
  const CMember cm = L.mem[t.g];
  const int Ib = t.i0 - CH, Jb = t.j0 - CH;
  const int tid = threadIdx.y * IMAU_TJ + threadIdx.x;
  for (int q = tid; q < CRI * CRJ; q += IMAU_WG) {
    const int li = q / CRJ, lj = q - li * CRJ;
    const int I = Ib + li, J = Jb + lj;
    T a = (T)0.0, b = (T)0.0;
    if (xiu) {
      if (cvalid(cm, 0, J, I)) a = xiu[cp2(cm, J, I)];
      if (cvalid(cm, 1, J, I)) b = xiv[cp2(cm, J, I)];
    }
    sx[0][q] = a;
    sx[1][q] = b;
  }
  __syncthreads();
  for (int ph = 0; ph < 4; ++ph) {
    const int color = reverse ? 3 - ph : ph;
    const int comp = color >> 1;
    const int r = CH - 1 - ph;
    // lanes take the RW / 2 cells of this colour in each row of the region
    // (same-component couplings are 5-point, so cells of one colour do not couple)
    const int RW = IMAU_TJ + 2 * r, RH = RW / 2, n = (IMAU_TI + 2 * r) * RH;
    const T *cf = comp == 0 ? L.cu : L.cv;
    for (int q = tid; q < n; q += IMAU_WG) {
      const int li = CH - r + q / RH;
      const int lj = CH - r + 2 * (q % RH) + ((color ^ (Ib + li + Jb + CH - r)) & 1);
      const int I = Ib + li, J = Jb + lj;
      if (!cvalid(cm, comp, J, I)) continue;
      const long long p = cp2(cm, J, I);
      const T D = cf[p];
      T S = (T)0.0;
#pragma unroll
      for (int k = 1; k < 9; ++k) {
        int cc, JJ, II;
        slot_pos(comp, k, J, I, cc, JJ, II);
        if (!cvalid(cm, cc, JJ, II)) continue;
        S += cf[k * L.na + p] * sx[cc][(II - Ib) * CRJ + (JJ - Jb)];
      }
      const T bb = (comp == 0 ? L.bu : L.bv)[p];
      sx[comp][li * CRJ + lj] = (D != (T)0.0) ? mdiv(bb - S, D) : (T)0.0;
    }
    __syncthreads();
  }
  for (int q = tid; q < IMAU_TI * IMAU_TJ; q += IMAU_WG) {
    const int li = CH + q / IMAU_TJ, lj = CH + q % IMAU_TJ;
    const int I = Ib + li, J = Jb + lj;
    if (cvalid(cm, 0, J, I)) xou[cp2(cm, J, I)] = sx[0][li * CRJ + lj];
    if (cvalid(cm, 1, J, I)) xov[cp2(cm, J, I)] = sx[1][li * CRJ + lj];
  }
}

// b(l+1) = R (b - A x)(l); one thread per coarse cell
template <class T>
__global__ void __launch_bounds__(IMAU_WG) k_crestrict(CLevelT<T> Fl, CLevelT<T> L, const int *__restrict__ active) {
  CTILE_SETUP
  if (J >= cm.ny || !active[t.g] || Fl.depth[t.g] <= Fl.lev) return;
  const CMember fm = Fl.mem[t.g];
  FOR_TILE_ROWS {
    const int I = t.i0 + ii;
    if (I >= cm.nx) break;
    for (int comp = 0; comp < 2; ++comp) {
      if (!cvalid(cm, comp, J, I)) continue;
      T s = (T)0.0;
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const int j = 2 * J + dj, i = 2 * I + di;
          if (!cvalid(fm, comp, j, i)) continue;
          T S, D;
          crow(Fl, fm, comp, j, i, Fl.xu, Fl.xv, S, D);
          const long long pf = cp2(fm, j, i);
          s += (comp == 0 ? Fl.bu : Fl.bv)[pf] - S - D * (comp == 0 ? Fl.xu : Fl.xv)[pf];
        }
      (comp == 0 ? L.bu : L.bv)[cp2(cm, J, I)] = s;
    }
  }
}

// x(l) += P x(l+1); one thread per fine cell
template <class T>
__global__ void __launch_bounds__(IMAU_WG) k_cprolong(CLevelT<T> Fl, CLevelT<T> L, const int *__restrict__ active) {
  const Tile t = Fl.tiles[imau_xcd_tile()];
  const CMember fm = Fl.mem[t.g];
  const int j = t.j0 + (int)threadIdx.x;
  if (j >= fm.ny || !active[t.g] || Fl.depth[t.g] <= Fl.lev) return;
  const CMember cm = L.mem[t.g];
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= fm.nx) break;
    const long long pf = cp2(fm, j, i), pc = cp2(cm, j >> 1, i >> 1);
    if (cvalid(fm, 0, j, i)) Fl.xu[pf] += L.xu[pc];
    if (cvalid(fm, 1, j, i)) Fl.xv[pf] += L.xv[pc];
  }
}

// the coarsest level of each member: one 1024-thread workgroup, thread q owns
// cell (J, I) = (q % ny, q / ny) with its u and v rows in registers and the
// iterate in LDS; nsweeps symmetric four-colour Gauss-Seidel sweeps from zero
template <class T>
__device__ __forceinline__ void coarse_row_load(const CLevelT<T> &L, const CMember &m, int comp, int J, int I,
                                                bool valid, T (&cf)[9], int (&nb)[9], T &b) {
  const long long p = cp2(m, J, I);
  const T *src = comp == 0 ? L.cu : L.cv;
  b = valid ? (comp == 0 ? L.bu : L.bv)[p] : (T)0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    int cc, JJ, II;
    slot_pos(comp, k, J, I, cc, JJ, II);
    const bool ok = valid && (k == 0 || cvalid(m, cc, JJ, II));
    cf[k] = ok ? src[k * L.na + p] : (T)0.0;
    nb[k] = ok ? cc * (IMAU_CMAX * IMAU_CMAX) + II * m.ny + JJ : 0;
  }
}

template <class T>
__global__ void __launch_bounds__(1024) k_coarsest(CLevelT<T> L, const int *__restrict__ active, int nsweeps) {
  __shared__ T sx[2 * IMAU_CMAX * IMAU_CMAX];
  const int g = blockIdx.x;
  if (!active[g] || L.depth[g] != L.lev) return;
  const CMember cm = L.mem[g];
  const int n = cm.nx * cm.ny;
  const int q = threadIdx.x;
  const int J = q % cm.ny, I = q / cm.ny;
  const bool in = q < n;
  const bool vu = in && cvalid(cm, 0, J, I), vv = in && cvalid(cm, 1, J, I);
  T cu[9], cv[9], bu, bv;
  int nu[9], nv[9];
  coarse_row_load(L, cm, 0, J, I, vu, cu, nu, bu);
  coarse_row_load(L, cm, 1, J, I, vv, cv, nv, bv);
  sx[q] = (T)0.0;
  sx[IMAU_CMAX * IMAU_CMAX + q] = (T)0.0;
  __syncthreads();
  const int par = (I + J) & 1;
  for (int sw = 0; sw < 2 * nsweeps; ++sw) {
#pragma unroll
    for (int q4 = 0; q4 < 4; ++q4) {
      const int color = (sw & 1) ? 3 - q4 : q4;
      if (par == (color & 1)) {
        if (color < 2 && vu) {
          T S = (T)0.0;
#pragma unroll
          for (int k = 1; k < 9; ++k) S += cu[k] * sx[nu[k]];
          sx[q] = (cu[0] != (T)0.0) ? mdiv(bu - S, cu[0]) : (T)0.0;
        } else if (color >= 2 && vv) {
          T S = (T)0.0;
#pragma unroll
          for (int k = 1; k < 9; ++k) S += cv[k] * sx[nv[k]];
          sx[IMAU_CMAX * IMAU_CMAX + q] = (cv[0] != (T)0.0) ? mdiv(bv - S, cv[0]) : (T)0.0;
        }
      }
      __syncthreads();
    }
  }
  if (in) {
    const long long p = cp2(cm, J, I);
    if (vu) L.xu[p] = sx[q];
    if (vv) L.xv[p] = sx[IMAU_CMAX * IMAU_CMAX + q];
  }
}

// ---------------------------------------------------------------------------
// level 0 (matrix-free fine operator)
// ---------------------------------------------------------------------------
// one colour of Gauss-Seidel on A z = y (colour 4: the domain-boundary rows)
template <class T>
__global__ void __launch_bounds__(IMAU_WG) k_smooth0(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                     const KP kp, const CoefT<T> c, VecT<T> x, CVecT<T> y, int color,
                                                     const int *__restrict__ active) {
  TILE_SETUP
  if (j >= m.ny || !active[t.g]) return;
  const GAccT<T> U{x.u, m}, V{x.v, m};
  const int nx = m.nx, ny = m.ny;
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
      T S, D;
      if (isu) u_sd<T>(m, kp, c, j, i, rt, U, V, S, D);
      else v_sd<T>(m, kp, c, j, i, rt, U, V, S, D);
      (isu ? x.u : x.v)[p] = mdiv((isu ? y.u : y.v)[p] - S, D);
    } else {
      if (i < nx - 1 && is_bnd(c.rt_u[p])) {
        const uint8_t rt = c.rt_u[p];
        T val = y.u[p];
        if (rt != RT_BND_ZERO) {
          int jo, io;
          bnd_src(rt, i, j, nx - 2, ny - 1, jo, io);
          val += x.u[p2(m, jo, io)];
        }
        x.u[p] = val;
      }
      if (j < ny - 1 && is_bnd(c.rt_v[p])) {
        const uint8_t rt = c.rt_v[p];
        T val = y.v[p];
        if (rt != RT_BND_ZERO) {
          int jo, io;
          bnd_src(rt, i, j, nx - 1, ny - 2, jo, io);
          val += x.v[p2(m, jo, io)];
        }
        x.v[p] = val;
      }
    }
  }
}

// One fused fine-level Gauss-Seidel sweep: the tile plus a halo of IMAU_FH
// cells is staged in LDS from zin (zero when zin is null), the four colour
// phases (u red, u black, v red, v black; reversed when reverse != 0) update
// shrinking regions -- after phase p the values are exact up to IMAU_FH-1-p
// cells beyond the tile, because each row couples only neighbours at distance
// one -- and then the domain-boundary rows of the tile are set from their
// sources (exact one cell beyond the tile).  The tile is written to zout, a
// different buffer, so the result equals the one-launch-per-colour sweep
// exactly.  Periodic boundary rows, whose sources lie across the domain, are
// completed by a separate colour-4 pass of k_smooth0.
#define FRI (IMAU_FTI + 2 * IMAU_FH)
#define FRJ (IMAU_FTJ + 2 * IMAU_FH)

template <class T>
struct LAccT {
  const T *base;  // LDS region, element (0,0) at global (ib, jb)
  int ib, jb;
  __device__ __forceinline__ T operator()(int jj, int ii) const { return base[(ii - ib) * FRJ + (jj - jb)]; }
};

template <class T>
__global__ void __launch_bounds__(IMAU_WG) k_smooth_fused(const Member *__restrict__ mem, const Tile *__restrict__ ftiles,
                                                          const KP kp, const CoefT<T> c, CVecT<T> zin, VecT<T> zout,
                                                          CVecT<T> y, int reverse, const int *__restrict__ active) {
  __shared__ T su[FRI * FRJ], sv[FRI * FRJ];
  const Tile t = ftiles[imau_xcd_tile()];
  if (!active[t.g]) return;
  const Member m = mem[t.g];
  const int nx = m.nx, ny = m.ny;
  const int ib = t.i0 - IMAU_FH, jb = t.j0 - IMAU_FH;
  const int tid = threadIdx.y * IMAU_TJ + threadIdx.x;
  for (int q = tid; q < FRI * FRJ; q += IMAU_WG) {
    const int li = q / FRJ, lj = q - li * FRJ;
    const int i = ib + li, j = jb + lj;
    T u = (T)0.0, v = (T)0.0;
    if (zin.u && i >= 0 && j >= 0 && i < nx && j < ny) {
      const long long p = p2(m, j, i);
      if (i < nx - 1) u = zin.u[p];
      if (j < ny - 1) v = zin.v[p];
    }
    su[q] = u;
    sv[q] = v;
  }
  __syncthreads();
  const LAccT<T> U{su, ib, jb}, V{sv, ib, jb};
  for (int ph = 0; ph < 4; ++ph) {
    const int color = reverse ? 3 - ph : ph;
    const bool isu = color < 2;
    const int r = IMAU_FH - 1 - ph;
    const int li0 = IMAU_FH - r, lj0 = IMAU_FH - r;
    // the region has an even width, so every row holds RW / 2 cells of this
    // colour: lanes take those cells only (cells of one colour do not couple,
    // so the order is free)
    const int RW = IMAU_FTJ + 2 * r, RH = RW / 2, n = (IMAU_FTI + 2 * r) * RH;
    for (int q = tid; q < n; q += IMAU_WG) {
      const int li = li0 + q / RH;
      const int lj = lj0 + 2 * (q % RH) + ((color ^ (ib + li + jb + lj0)) & 1);
      const int i = ib + li, j = jb + lj;
      if (i < 0 || j < 0 || i >= nx || j >= ny) continue;
      if (isu && i >= nx - 1) continue;
      if (!isu && j >= ny - 1) continue;
      const long long p = p2(m, j, i);
      const uint8_t rt = isu ? c.rt_u[p] : c.rt_v[p];
      if (is_bnd(rt)) continue;
      T S, D;
      if (isu) u_sd<T>(m, kp, c, j, i, rt, U, V, S, D);
      else v_sd<T>(m, kp, c, j, i, rt, U, V, S, D);
      (isu ? su : sv)[li * FRJ + lj] = mdiv((isu ? y.u : y.v)[p] - S, D);
    }
    __syncthreads();
  }
  // domain-boundary rows of the tile ('infinite' and 'zero'; sources are interior rows)
  for (int q = tid; q < IMAU_FTI * IMAU_FTJ; q += IMAU_WG) {
    const int li = IMAU_FH + q / IMAU_FTJ, lj = IMAU_FH + q % IMAU_FTJ;
    const int i = ib + li, j = jb + lj;
    if (i >= nx || j >= ny) continue;
    const long long p = p2(m, j, i);
    if (i < nx - 1) {
      const uint8_t rt = c.rt_u[p];
      if (rt == RT_BND_ZERO || rt == RT_BND_INF) {
        T val = y.u[p];
        if (rt == RT_BND_INF) {
          int jo, io;
          bnd_src(rt, i, j, nx - 2, ny - 1, jo, io);
          val += U(jo, io);
        }
        su[li * FRJ + lj] = val;
      }
    }
    if (j < ny - 1) {
      const uint8_t rt = c.rt_v[p];
      if (rt == RT_BND_ZERO || rt == RT_BND_INF) {
        T val = y.v[p];
        if (rt == RT_BND_INF) {
          int jo, io;
          bnd_src(rt, i, j, nx - 1, ny - 2, jo, io);
          val += V(jo, io);
        }
        sv[li * FRJ + lj] = val;
      }
    }
  }
  __syncthreads();
  for (int q = tid; q < IMAU_FTI * IMAU_FTJ; q += IMAU_WG) {
    const int li = IMAU_FH + q / IMAU_FTJ, lj = IMAU_FH + q % IMAU_FTJ;
    const int i = ib + li, j = jb + lj;
    if (i >= nx || j >= ny) continue;
    const long long p = p2(m, j, i);
    if (i < nx - 1) zout.u[p] = su[li * FRJ + lj];
    if (j < ny - 1) zout.v[p] = sv[li * FRJ + lj];
  }
}

// level-1 right-hand side: R (y - A z) over the interior and calving-front rows
template <class T, int CROSS>
__global__ void __launch_bounds__(IMAU_WG) k_restrict0(const Member *__restrict__ mem, const KP kp, const CoefT<T> c,
                                                       CVecT<T> y, CVecT<T> z, CLevelT<T> L,
                                                       const int *__restrict__ active) {
  CTILE_SETUP
  if (J >= cm.ny || !active[t.g]) return;
  const Member m = mem[t.g];
  const GAccT<T> U{z.u, m}, V{z.v, m};
  FOR_TILE_ROWS {
    const int I = t.i0 + ii;
    if (I >= cm.nx) break;
    for (int comp = 0; comp < 2; ++comp) {
      if (!cvalid(cm, comp, J, I)) continue;
      T s = (T)0.0;
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const int j = 2 * J + dj, i = 2 * I + di;
          if (!fine_valid(m, comp, j, i)) continue;
          const long long p = p2(m, j, i);
          const uint8_t rt = comp == 0 ? c.rt_u[p] : c.rt_v[p];
          if (!in_row(rt)) continue;
          T S, D;
          if (comp == 0) {
            u_sd_x<CROSS, T>(m, kp, c, j, i, rt, U, V, S, D);
            s += y.u[p] - S - D * z.u[p];
          } else {
            v_sd_x<CROSS, T>(m, kp, c, j, i, rt, U, V, S, D);
            s += y.v[p] - S - D * z.v[p];
          }
        }
      (comp == 0 ? L.bu : L.bv)[cp2(cm, J, I)] = s;
    }
  }
}

// z += P x(1) on the interior, calving-front and 'infinite' boundary unknowns
template <class T>
__global__ void __launch_bounds__(IMAU_WG) k_prolong0(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                      const uint8_t *__restrict__ rt_u, const uint8_t *__restrict__ rt_v,
                                                      VecT<T> z, CLevelT<T> L, const int *__restrict__ active) {
  TILE_SETUP
  if (j >= m.ny || !active[t.g]) return;
  const CMember cm = L.mem[t.g];
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    int Jc, Ic;
    if (i < m.nx - 1 && fine_agg(m, rt_u, rt_v, 0, j, i, Jc, Ic)) z.u[p] += L.xu[cp2(cm, Jc, Ic)];
    if (j < m.ny - 1 && fine_agg(m, rt_u, rt_v, 1, j, i, Jc, Ic)) z.v[p] += L.xv[cp2(cm, Jc, Ic)];
  }
}

// z = 0 (and, for FP32, y32 = y) on every cell of the active members
template <class T>
__global__ void __launch_bounds__(IMAU_WG) k_prep(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                  CVecT<double> y64, VecT<T> y, VecT<T> z,
                                                  const int *__restrict__ active) {
  TILE_SETUP
  if (j >= m.ny || !active[t.g]) return;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    z.u[p] = (T)0.0;
    z.v[p] = (T)0.0;
    if constexpr (!std::is_same<T, double>::value) {
      y.u[p] = (T)y64.u[p];
      y.v[p] = (T)y64.v[p];
    }
  }
}

// FP64 result of the FP32 cycle
__global__ void __launch_bounds__(IMAU_WG) k_fin(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                 CVecT<float> z32, VecT<double> z, const int *__restrict__ active) {
  TILE_SETUP
  if (j >= m.ny || !active[t.g]) return;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    z.u[p] = (double)z32.u[p];
    z.v[p] = (double)z32.v[p];
  }
}

// FP32 copies of the fine coefficients the smoother reads
__global__ void __launch_bounds__(IMAU_WG) k_coef32(const Member *__restrict__ mem, const Tile *__restrict__ tiles,
                                                    const CoefT<double> c, imau::MGFine32 o, int sans,
                                                    const int *__restrict__ active) {
  TILE_SETUP
  if (j >= m.ny || !active[t.g]) return;
  FOR_TILE_ROWS {
    const int i = t.i0 + ii;
    if (i >= m.nx) break;
    const long long p = p2(m, j, i);
    o.Na[p] = (float)c.Na[p];
    o.Nb[p] = (float)c.Nb[p];
    o.bcx[p] = (float)c.bcx[p];
    o.bcy[p] = (float)c.bcy[p];
    if (sans) {
      o.Ncx[p] = (float)c.Ncx[p];
      o.Ncy[p] = (float)c.Ncy[p];
    }
  }
}

template <class T>
__global__ void k_fill(T *p, long long n) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    p[q] = (T)0.0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
inline dim3 wg() { return dim3(IMAU_TJ, IMAU_WG_I, 1); }

inline CoefT<double> coef64(const imau::Coef &c) {
  return CoefT<double>{c.Na, c.Nb, c.Ncx, c.Ncy, c.bcx, c.bcy, c.taudx, c.taudy, c.frhs, c.rt_u, c.rt_v};
}
inline CoefT<float> coef32(const imau::MGFine32 &f, const imau::Coef &c) {
  return CoefT<float>{f.Na, f.Nb, f.Ncx, f.Ncy, f.bcx, f.bcy, nullptr, nullptr, nullptr, c.rt_u, c.rt_v};
}

template <class T>
CLevelT<T> view(const imau::MGPlan &P, int l) {
  const imau::MGLevelDev &d = P.levels[l];
  CLevelT<T> L{};
  L.mem = d.mem;
  L.tiles = d.tiles;
  L.depth = P.depth;
  L.lev = d.lev;
  L.ntiles = d.ntiles;
  L.na = d.na;
  L.cu = (T *)d.data;
  L.cv = L.cu + 9 * d.na;
  L.xu = L.cv + 9 * d.na;
  L.xv = L.xu + d.na;
  L.bu = L.xv + d.na;
  L.bv = L.bu + d.na;
  return L;
}

// the level with its iterate pointers on the second buffer (after the fused post-smoothing)
template <class T>
CLevelT<T> view2(const imau::MGPlan &P, int l) {
  CLevelT<T> L = view<T>(P, l);
  L.xu = L.bv + L.na;
  L.xv = L.xu + L.na;
  return L;
}

template <class T>
void fill(hipStream_t s, T *p, long long n) {
  int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, s, p, n);
}

template <class T>
void setup_t(const imau::Launch &F, const KP &kp, const imau::Coef &c, const imau::MGPlan &P, const int *act) {
  const CLevelT<T> L1 = view<T>(P, 0);
  auto kp0 = kp.cross ? k_gal0_plain<T, 1> : k_gal0_plain<T, 0>;
  hipLaunchKernelGGL(kp0, dim3(L1.ntiles), wg(), 0, F.s, F.mem, kp, coef64(c), L1, act, P.plain);
  if (P.nlist > 0) {
    auto kg = kp.cross ? k_gal0_list<T, 1> : k_gal0_list<T, 0>;
    hipLaunchKernelGGL(kg, dim3((P.nlist + 255) / 256), dim3(256), 0, F.s, F.mem, kp, coef64(c), L1, act, P.plain,
                       P.list, P.nlist);
  }
  for (size_t l = 1; l < P.levels.size(); ++l) {
    const CLevelT<T> Fl = view<T>(P, (int)l - 1), L = view<T>(P, (int)l);
    hipLaunchKernelGGL(k_galc<T>, dim3(L.ntiles), wg(), 0, F.s, Fl, L, act);
  }
}

// coarse part of the cycle: b(1) is set; x(1) is returned
template <class T>
void coarse_t(hipStream_t s, const imau::MGPlan &P, const int *act) {
  const int nl = (int)P.levels.size();
  for (int l = 0; l < nl; ++l) {
    const CLevelT<T> C = view<T>(P, l);
    fill<T>(s, C.xu, 2 * C.na);
    if (P.levels[l].ends)
      hipLaunchKernelGGL(k_coarsest<T>, dim3(P.ngrids), dim3(1024), 0, s, C, act, P.coarsest_sweeps);
    if (l + 1 == nl) break;
    for (int col = 0; col < 4; ++col) hipLaunchKernelGGL(k_cgs<T>, dim3(C.ntiles), wg(), 0, s, C, col, act);
    const CLevelT<T> N = view<T>(P, l + 1);
    hipLaunchKernelGGL(k_crestrict<T>, dim3(N.ntiles), wg(), 0, s, C, N, act);
  }
  for (int l = nl - 2; l >= 0; --l) {
    const CLevelT<T> C = view<T>(P, l), N = view<T>(P, l + 1);
    hipLaunchKernelGGL(k_cprolong<T>, dim3(C.ntiles), wg(), 0, s, C, N, act);
    for (int col = 3; col >= 0; --col) hipLaunchKernelGGL(k_cgs<T>, dim3(C.ntiles), wg(), 0, s, C, col, act);
  }
}

// coarse part with fused sweeps: pre-smoothing from zero into x, the post-smoothed
// (or coarsest-level) result of every level in the second buffer x2
template <class T>
void coarse_fused_t(hipStream_t s, const imau::MGPlan &P, const int *act) {
  const int nl = (int)P.levels.size();
  for (int l = 0; l < nl; ++l) {
    const CLevelT<T> C = view<T>(P, l), C2 = view2<T>(P, l);
    if (P.levels[l].ends)
      hipLaunchKernelGGL(k_coarsest<T>, dim3(P.ngrids), dim3(1024), 0, s, C2, act, P.coarsest_sweeps);
    if (l + 1 == nl) break;
    hipLaunchKernelGGL(k_cgs_fused<T>, dim3(C.ntiles), wg(), 0, s, C, (const T *)nullptr, (const T *)nullptr, C.xu,
                       C.xv, 0, act);
    const CLevelT<T> N = view<T>(P, l + 1);
    hipLaunchKernelGGL(k_crestrict<T>, dim3(N.ntiles), wg(), 0, s, C, N, act);
  }
  for (int l = nl - 2; l >= 0; --l) {
    const CLevelT<T> C = view<T>(P, l), N2 = view2<T>(P, l + 1);
    hipLaunchKernelGGL(k_cprolong<T>, dim3(C.ntiles), wg(), 0, s, C, N2, act);
    hipLaunchKernelGGL(k_cgs_fused<T>, dim3(C.ntiles), wg(), 0, s, C, (const T *)C.xu, (const T *)C.xv,
                       C.bv + C.na, C.bv + 2 * C.na, 1, act);
  }
}

// one launch per colour; the cycle runs in place on z (zeroed by the caller)
template <class T>
void vcycle_t(const imau::Launch &F, const KP &kp, const CoefT<T> &cf, const imau::MGPlan &P, CVecT<T> y, VecT<T> z,
              const int *act, imau::MGMark mark, void *mctx) {
  hipStream_t s = F.s;
  if (mark) mark(mctx, 0);
  for (int col = 0; col < 5; ++col)
    hipLaunchKernelGGL(k_smooth0<T>, dim3(F.ntiles), wg(), 0, s, F.mem, F.tiles, kp, cf, z, y, col, act);
  if (mark) mark(mctx, 1);
  const CLevelT<T> L1 = view<T>(P, 0);
  auto kr = kp.cross ? k_restrict0<T, 1> : k_restrict0<T, 0>;
  hipLaunchKernelGGL(kr, dim3(L1.ntiles), wg(), 0, s, F.mem, kp, cf, y, CVecT<T>{z.u, z.v}, L1, act);
  coarse_t<T>(s, P, act);
  hipLaunchKernelGGL(k_prolong0<T>, dim3(F.ntiles), wg(), 0, s, F.mem, F.tiles, cf.rt_u, cf.rt_v, z, L1, act);
  for (int col = 3; col >= 0; --col)
    hipLaunchKernelGGL(k_smooth0<T>, dim3(F.ntiles), wg(), 0, s, F.mem, F.tiles, kp, cf, z, y, col, act);
  hipLaunchKernelGGL(k_smooth0<T>, dim3(F.ntiles), wg(), 0, s, F.mem, F.tiles, kp, cf, z, y, 4, act);
}

// fused sweeps: pre-smoothing from zero into za, post-smoothing from za into zb
template <class T>
void vcycle_fused_t(const imau::Launch &F, const KP &kp, const CoefT<T> &cf, const imau::MGPlan &P, CVecT<T> y,
                    VecT<T> za, VecT<T> zb, const int *act, imau::MGMark mark, void *mctx) {
  hipStream_t s = F.s;
  if (mark) mark(mctx, 0);
  hipLaunchKernelGGL(k_smooth_fused<T>, dim3(P.nftiles), wg(), 0, s, F.mem, P.ftiles, kp, cf, CVecT<T>{nullptr, nullptr},
                     za, y, 0, act);
  if (P.periodic)
    hipLaunchKernelGGL(k_smooth0<T>, dim3(F.ntiles), wg(), 0, s, F.mem, F.tiles, kp, cf, za, y, 4, act);
  if (mark) mark(mctx, 1);
  const CLevelT<T> L1 = view<T>(P, 0);
  auto kr = kp.cross ? k_restrict0<T, 1> : k_restrict0<T, 0>;
  hipLaunchKernelGGL(kr, dim3(L1.ntiles), wg(), 0, s, F.mem, kp, cf, y, CVecT<T>{za.u, za.v}, L1, act);
  coarse_fused_t<T>(s, P, act);
  hipLaunchKernelGGL(k_prolong0<T>, dim3(F.ntiles), wg(), 0, s, F.mem, F.tiles, cf.rt_u, cf.rt_v, za, view2<T>(P, 0),
                     act);
  hipLaunchKernelGGL(k_smooth_fused<T>, dim3(P.nftiles), wg(), 0, s, F.mem, P.ftiles, kp, cf, CVecT<T>{za.u, za.v},
                     zb, y, 1, act);
  if (P.periodic)
    hipLaunchKernelGGL(k_smooth0<T>, dim3(F.ntiles), wg(), 0, s, F.mem, F.tiles, kp, cf, zb, y, 4, act);
}

}  // namespace

namespace imau {

void mg_plain_flags(const Launch &F, const uint8_t *rt_u, const uint8_t *rt_v, const MGPlan &P) {
  if (P.levels.empty()) return;
  const CLevelT<float> L1 = view<float>(P, 0);
  fill_i32(F.s, P.nlist_d, 1, 0);
  hipLaunchKernelGGL(k_gal_plain_flags, dim3(L1.ntiles), wg(), 0, F.s, F.mem, rt_u, rt_v, L1, P.plain, P.list,
                     P.nlist_d);
}

void mg_setup(const Launch &F, const KP &kp, const Coef &c, const MGPlan &P, bool fp32, bool galerkin,
              const int *active) {
  if (fp32) {
    hipLaunchKernelGGL(k_coef32, dim3(F.ntiles), wg(), 0, F.s, F.mem, F.tiles, coef64(c), P.f32, kp.cross ? 0 : 1,
                       active);
    if (galerkin) setup_t<float>(F, kp, c, P, active);
  } else if (galerkin) {
    setup_t<double>(F, kp, c, P, active);
  }
}

void mg_vcycle(const Launch &F, const KP &kp, const Coef &c, const MGPlan &P, bool fp32, const double *yu,
               const double *yv, double *zu, double *zv, const int *active, MGMark mark, void *mark_ctx) {
  if (fp32) {
    VecT<float> y32{P.f32.yu, P.f32.yv}, z32{P.f32.zu, P.f32.zv}, z32b{P.f32.zu2, P.f32.zv2};
    hipLaunchKernelGGL(k_prep<float>, dim3(F.ntiles), wg(), 0, F.s, F.mem, F.tiles, CVecT<double>{yu, yv}, y32, z32,
                       active);
    const CoefT<float> cf = coef32(P.f32, c);
    if (P.fused) {
      vcycle_fused_t<float>(F, kp, cf, P, CVecT<float>{y32.u, y32.v}, z32, z32b, active, mark, mark_ctx);
      z32 = z32b;
    } else {
      vcycle_t<float>(F, kp, cf, P, CVecT<float>{y32.u, y32.v}, z32, active, mark, mark_ctx);
    }
    hipLaunchKernelGGL(k_fin, dim3(F.ntiles), wg(), 0, F.s, F.mem, F.tiles, CVecT<float>{z32.u, z32.v},
                       VecT<double>{zu, zv}, active);
  } else if (P.fused) {
    vcycle_fused_t<double>(F, kp, coef64(c), P, CVecT<double>{yu, yv}, VecT<double>{P.zu64, P.zv64},
                           VecT<double>{zu, zv}, active, mark, mark_ctx);
  } else {
    VecT<double> z{zu, zv};
    hipLaunchKernelGGL(k_prep<double>, dim3(F.ntiles), wg(), 0, F.s, F.mem, F.tiles, CVecT<double>{yu, yv},
                       VecT<double>{nullptr, nullptr}, z, active);
    vcycle_t<double>(F, kp, coef64(c), P, CVecT<double>{yu, yv}, z, active, mark, mark_ctx);
  }
}

}  // namespace imau
