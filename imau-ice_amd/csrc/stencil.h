// stencil.h -- the linearised staggered DIVA/SSA rows, matrix-free, shared by
// the FP64 solver kernels (kernels.hip) and the precision-templated multigrid
// (mg.hip).  src/ice_velocity_module.f90:1762-2589.
//
// For the unknown at (j,i) of row type rt, u_sd / v_sd return the off-diagonal
// sum S and the diagonal D so that the row reads S + D x = b; the right-hand
// side b is added by the FP64 callers (kernels.hip: u_row / v_row).
#pragma once
#include "dev_common.h"

template <class T>
struct CoefT {
  const T *Na, *Nb, *Ncx, *Ncy, *bcx, *bcy, *taudx, *taudy, *frhs;
  const uint8_t *rt_u, *rt_v;
  // coefficient accessors used by the rows (an LDS-staged variant provides the same)
  __device__ __forceinline__ T na(const Member &m, int j, int i) const { return Na[p2(m, j, i)]; }
  __device__ __forceinline__ T nb(const Member &m, int j, int i) const { return Nb[p2(m, j, i)]; }
  __device__ __forceinline__ T bx(const Member &m, int j, int i) const { return bcx[p2(m, j, i)]; }
  __device__ __forceinline__ T by(const Member &m, int j, int i) const { return bcy[p2(m, j, i)]; }
  __device__ __forceinline__ T ncx(const Member &m, int j, int i) const { return Ncx[p2(m, j, i)]; }
  __device__ __forceinline__ T ncy(const Member &m, int j, int i) const { return Ncy[p2(m, j, i)]; }
};

template <class T>
struct GAccT {  // global-memory accessor
  const T *x;
  Member m;
  __device__ __forceinline__ T operator()(int jj, int ii) const { return x[p2(m, jj, ii)]; }
};

__device__ __forceinline__ bool is_bnd(uint8_t rt) { return rt == RT_BND_INF || rt == RT_BND_ZERO || rt == RT_BND_PER; }
// rows that take part in the multigrid restriction (interior and calving-front rows)
__device__ __forceinline__ bool in_row(uint8_t rt) { return rt == RT_FREE || rt == RT_MARGIN_I || rt == RT_MARGIN_O; }

// source cell of a domain-boundary row (src/ice_velocity_module.f90:2006-2030)
__device__ __forceinline__ void bnd_src(uint8_t rt, int i, int j, int imax, int jmax, int &jo, int &io) {
  if (rt == RT_BND_INF) {
    io = (i == 0) ? 1 : (i == imax ? imax - 1 : i);
    jo = (j == 0) ? 1 : (j == jmax ? jmax - 1 : j);
  } else {
    io = (i == 0) ? imax - 1 : (i == imax ? 1 : i);
    jo = (j == 0) ? jmax - 1 : (j == jmax ? 1 : j);
  }
}

// u row: free rows with cross terms (:2342-2467) or in the normalised 'sans'
// form (:2468-2589); zero rows (:1787-1796, :2224-2230); domain boundary rows
// (:1932-2191); calving-front stress rows (:2192-2266)
// CROSS: 1 / 0 fixes the cross-term choice at compile time, -1 reads kp.cross
template <int CROSS, class T, class AU, class AV, class C>
__device__ __forceinline__ void u_sd_x(const Member &m, const KP &kp, const C &c, int j, int i, uint8_t rt,
                                       const AU &U, const AV &V, T &S, T &D) {
  const int nx = m.nx, ny = m.ny;
  if (rt == RT_FREE) {
    const T dx2 = (T)(m.dx * m.dx);
    if (CROSS >= 0 ? CROSS != 0 : kp.cross != 0) {
      const T Na0 = c.na(m, j, i), Na1 = c.na(m, j, i + 1);
      const T Nb0 = c.nb(m, j, i), Nbm = c.nb(m, j - 1, i);
      S = (T)4.0 * Na1 * U(j, i + 1) + (T)4.0 * Na0 * U(j, i - 1) + Nb0 * U(j + 1, i) + Nbm * U(j - 1, i) +
          ((T)-2.0 * Na0 - Nb0) * V(j, i) + ((T)2.0 * Na1 + Nb0) * V(j, i + 1) + ((T)2.0 * Na0 + Nbm) * V(j - 1, i) +
          ((T)-2.0 * Na1 - Nbm) * V(j - 1, i + 1);
      D = (T)-4.0 * (Na1 + Na0) + (T)-1.0 * (Nb0 + Nbm) - c.bx(m, j, i) * dx2;
    } else {
      const T Ncx = c.ncx(m, j, i);
      S = (T)4.0 * U(j, i + 1) + (T)4.0 * U(j, i - 1) + U(j + 1, i) + U(j - 1, i) - (T)3.0 * V(j, i) +
          (T)3.0 * V(j, i + 1) + (T)3.0 * V(j - 1, i) - (T)3.0 * V(j - 1, i + 1);
      D = (T)-10.0 - dx2 * c.bx(m, j, i) / Ncx;
    }
    return;
  }
  if (rt == RT_ZERO || rt == RT_BND_ZERO) { S = (T)0.0; D = (T)1.0; return; }
  if (rt == RT_BND_INF || rt == RT_BND_PER) {
    int io_, jo_;
    bnd_src(rt, i, j, nx - 2, ny - 1, jo_, io_);
    S = -U(jo_, io_); D = (T)1.0;
    return;
  }
  const int i1 = (rt == RT_MARGIN_I) ? i : i + 1;
  const T Na = c.na(m, j, i1) / (T)m.dx;
  const T cl = (T)-4.0 * Na, cr = (T)4.0 * Na;
  const T sv = (T)-2.0 * Na * V(j - 1, i1) + (T)2.0 * Na * V(j, i1);
  if (i1 == i) { S = cl * U(j, i - 1) + sv; D = cr; }
  else { S = cr * U(j, i + 1) + sv; D = cl; }
}

template <int CROSS, class T, class AU, class AV, class C>
__device__ __forceinline__ void v_sd_x(const Member &m, const KP &kp, const C &c, int j, int i, uint8_t rt,
                                       const AU &U, const AV &V, T &S, T &D) {
  const int nx = m.nx, ny = m.ny;
  if (rt == RT_FREE) {
    const T dx2 = (T)(m.dx * m.dx);
    if (CROSS >= 0 ? CROSS != 0 : kp.cross != 0) {
      const T Na0 = c.na(m, j, i), Na1 = c.na(m, j + 1, i);
      const T Nb0 = c.nb(m, j, i), Nbm = c.nb(m, j, i - 1);
      S = (T)4.0 * Na1 * V(j + 1, i) + (T)4.0 * Na0 * V(j - 1, i) + Nb0 * V(j, i + 1) + Nbm * V(j, i - 1) +
          ((T)-2.0 * Na1 - Nbm) * U(j + 1, i - 1) + ((T)2.0 * Na1 + Nb0) * U(j + 1, i) +
          ((T)2.0 * Na0 + Nbm) * U(j, i - 1) + ((T)-2.0 * Na0 - Nb0) * U(j, i);
      D = (T)-4.0 * (Na1 + Na0) + (T)-1.0 * (Nb0 + Nbm) - c.by(m, j, i) * dx2;
    } else {
      const T Ncy = c.ncy(m, j, i);
      S = (T)4.0 * V(j + 1, i) + (T)4.0 * V(j - 1, i) + V(j, i + 1) + V(j, i - 1) - (T)3.0 * U(j + 1, i - 1) +
          (T)3.0 * U(j + 1, i) + (T)3.0 * U(j, i - 1) - (T)3.0 * U(j, i);
      D = (T)-10.0 - dx2 * c.by(m, j, i) / Ncy;
    }
    return;
  }
  if (rt == RT_ZERO || rt == RT_BND_ZERO) { S = (T)0.0; D = (T)1.0; return; }
  if (rt == RT_BND_INF || rt == RT_BND_PER) {
    int io_, jo_;
    bnd_src(rt, i, j, nx - 1, ny - 2, jo_, io_);
    S = -V(jo_, io_); D = (T)1.0;
    return;
  }
  const int j1 = (rt == RT_MARGIN_I) ? j : j + 1;
  const T Na = c.na(m, j1, i) / (T)m.dx;
  const T cl = (T)-4.0 * Na, cr = (T)4.0 * Na;
  const T su = (T)-2.0 * Na * U(j1, i - 1) + (T)2.0 * Na * U(j1, i);
  if (j1 == j) { S = cl * V(j - 1, i) + su; D = cr; }
  else { S = cr * V(j + 1, i) + su; D = cl; }
}

template <class T, class AU, class AV, class C = CoefT<T>>
__device__ __forceinline__ void u_sd(const Member &m, const KP &kp, const C &c, int j, int i, uint8_t rt,
                                     const AU &U, const AV &V, T &S, T &D) {
  u_sd_x<-1, T>(m, kp, c, j, i, rt, U, V, S, D);
}
template <class T, class AU, class AV, class C = CoefT<T>>
__device__ __forceinline__ void v_sd(const Member &m, const KP &kp, const C &c, int j, int i, uint8_t rt,
                                     const AU &U, const AV &V, T &S, T &D) {
  v_sd_x<-1, T>(m, kp, c, j, i, rt, U, V, S, D);
}
