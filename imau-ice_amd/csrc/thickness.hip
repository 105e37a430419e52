// thickness.hip -- the ice-thickness update on the device: the consumer of
// u_vav_cx / v_vav_cy (calc_dHi_dt, src/ice_thickness_module.f90:25-909) and
// the advective critical time step (calc_critical_timestep_adv,
// src/ice_dynamics_module.f90:726-762).
//
// One member (grid) per call, in the a-grid-shaped slab layout of the solver
// (p2(j,i) = off + i*ny + j for a, cx and cy fields alike), so the velocities
// of a device-resident problem are read where the solve left them.  Every
// kernel is one thread per cell; the explicit scheme's outflux rescaling is
// per source cell (each face flux is the outflux of exactly one cell), so it
// runs in parallel with the reference's results.
#include <hip/hip_runtime.h>

#include <cmath>

#include "dev_common.h"
#include "thickness.h"

// Multiply-adds are not contracted in this file: the thickness and mask
// arithmetic rounds as the reference's does, so that a thickness driven to
// zero (H + dt (-H / dt)) is zero and not a residue of a fused multiply-add
// that would count as ice.
#pragma clang fp contract(off)

namespace {

constexpr int TB = 256;

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

// net mass balance (m/yr) of a cell (:219-227, :393-401)
__device__ __forceinline__ double mb_net(const imau::ThkIn &in, long long p) {
  const double mb = in.SMB[p] + in.BMB[p];
  if (in.m_cf[p] == 1 && in.m_shelf[p] == 1) return mb * in.fmf[p];
  if (in.m_ocean[p] == 1 && in.m_shelf[p] == 0) return 0.0;
  return mb;
}

// explicit scheme, step 1: upwind face volumes of the step, the mass-balance
// volume and the outflux rescaling factor of every cell (:92-279)
__global__ void __launch_bounds__(TB) k_thk_explicit1(const Member m, const imau::ThkIn in, double dt,
                                                     double *__restrict__ Qx, double *__restrict__ Qy,
                                                     double *__restrict__ dVmb, double *__restrict__ rescale) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const int nx = m.nx, ny = m.ny, i = c.i, j = c.j;
  const double dx = m.dx;
  auto qx = [&](int jj, int ii) {  // flux through the face east of cell (jj, ii)
    const double u = in.u[m.off + (long long)ii * ny + jj];
    return u * in.Hi[m.off + (long long)(u > 0.0 ? ii : ii + 1) * ny + jj] * dx * dt;
  };
  auto qy = [&](int jj, int ii) {  // north of cell (jj, ii)
    const double v = in.v[m.off + (long long)ii * ny + jj];
    return v * in.Hi[m.off + (long long)ii * ny + (v > 0.0 ? jj : jj + 1)] * dx * dt;
  };
  if (i < nx - 1) Qx[c.p] = qx(j, i);
  if (j < ny - 1) Qy[c.p] = qy(j, i);
  double vmb = mb_net(in, c.p) * dx * dx * dt;
  const double Vav = in.Hi[c.p] * dx * dx;
  double out = 0.0;
  if (i > 0) { const double q = qx(j, i - 1); if (!(q > 0.0)) out -= q; }
  if (i < nx - 1) { const double q = qx(j, i); if (!(q < 0.0)) out += q; }
  if (j > 0) { const double q = qy(j - 1, i); if (!(q > 0.0)) out -= q; }
  if (j < ny - 1) { const double q = qy(j, i); if (!(q < 0.0)) out += q; }
  double r = 1.0;
  if (-vmb >= Vav) { vmb = -Vav; r = 0.0; }
  if (out > Vav + vmb) r = (Vav + vmb) / out;
  dVmb[c.p] = vmb;
  rescale[c.p] = r;
}

// explicit scheme, step 2: rescaled fluxes (each by its source cell) and the
// thickness change (:281-295)
__global__ void __launch_bounds__(TB) k_thk_explicit2(const Member m, const imau::ThkIn in, double dt,
                                                     double *__restrict__ Qx, double *__restrict__ Qy,
                                                     const double *__restrict__ dVmb,
                                                     const double *__restrict__ rescale, double *__restrict__ dHi,
                                                     double *__restrict__ Ht) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const int nx = m.nx, ny = m.ny, i = c.i, j = c.j;
  const double dx = m.dx;
  auto rs = [&](int jj, int ii) { return rescale[m.off + (long long)ii * ny + jj]; };
  // a face flux is scaled by the factor of the cell it leaves, if below 1
  auto fx = [&](int jj, int ii) {
    double q = Qx[m.off + (long long)ii * ny + jj];
    const double r = (q > 0.0) ? rs(jj, ii) : (q < 0.0 ? rs(jj, ii + 1) : 1.0);
    if (r < 1.0) q *= r;
    return q;
  };
  auto fy = [&](int jj, int ii) {
    double q = Qy[m.off + (long long)ii * ny + jj];
    const double r = (q > 0.0) ? rs(jj, ii) : (q < 0.0 ? rs(jj + 1, ii) : 1.0);
    if (r < 1.0) q *= r;
    return q;
  };
  const double a = dx * dx * dt;
  double d = dVmb[c.p] / a;
  if (i > 0) d += fx(j, i - 1) / a;
  if (i < nx - 1) d -= fx(j, i) / a;
  if (j > 0) d += fy(j - 1, i) / a;
  if (j < ny - 1) d -= fy(j, i) / a;
  d = fmax(d, -in.Hi[c.p] / dt);
  dHi[c.p] = d;
  Ht[c.p] = in.Hi[c.p] + d * dt;
}

// the rescaled face fluxes (the explicit scheme's output Qx_cx / Qy_cy), after
// k_thk_explicit2 has read the unscaled ones
__global__ void __launch_bounds__(TB) k_thk_store_flux(const Member m, double *__restrict__ Q,
                                                       const double *__restrict__ rescale, int east) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const int ny = m.ny;
  if (east ? c.i >= m.nx - 1 : c.j >= ny - 1) return;
  double q = Q[c.p];
  const long long down = east ? c.p + ny : c.p + 1;
  const double r = (q > 0.0) ? rescale[c.p] : (q < 0.0 ? rescale[down] : 1.0);
  if (r < 1.0) q *= r;
  Q[c.p] = q;
}

// face velocity, zero across a calving front whose floating cell is not yet
// full (:474-540, :574-640)
__device__ __forceinline__ double face_vel(const imau::ThkIn &in, long long p, long long q, double vel) {
  if (in.m_shelf[p] == 1 && in.m_ocean[q] == 1 && in.m_ice[q] == 0) {
    if (in.fmf[p] < 1.0) return 0.0;
  } else if (in.m_shelf[q] == 1 && in.m_ocean[p] == 1 && in.m_ice[p] == 0) {
    if (in.fmf[q] < 1.0) return 0.0;
  }
  return vel;
}

// semi-implicit scheme: one row per cell (:339-430, :441-718), fs = 2.5.
// Identity rows (MB-only and the explicit scheme around calving fronts) get
// cw = ce = cs = cn = 0, d = 1, b = the new thickness.
__global__ void __launch_bounds__(TB) k_thk_rows(const Member m, const imau::ThkIn in, double dt, imau::ThkRows R) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const int nx = m.nx, ny = m.ny, i = c.i, j = c.j;
  const double dx = m.dx, fs = 2.5;
  auto P = [&](int jj, int ii) { return m.off + (long long)ii * ny + jj; };
  const int im1 = i > 0 ? i - 1 : 0, ip1 = i < nx - 1 ? i + 1 : nx - 1;
  const int jm1 = j > 0 ? j - 1 : 0, jp1 = j < ny - 1 ? j + 1 : ny - 1;
  const int32_t *mi = in.m_ice, *ms = in.m_shelf, *mo = in.m_ocean;
  int scheme;
  if (mi[P(j, i)] == 0 && mi[P(j, im1)] == 0 && mi[P(j, ip1)] == 0 && mi[P(jm1, i)] == 0 && mi[P(jp1, i)] == 0)
    scheme = 0;
  else if (ms[P(j, i)] == 1 && ((mo[P(j, im1)] == 1 && mi[P(j, im1)] == 0) || (mo[P(j, ip1)] == 1 && mi[P(j, ip1)] == 0) ||
                                (mo[P(jm1, i)] == 1 && mi[P(jm1, i)] == 0) || (mo[P(jp1, i)] == 1 && mi[P(jp1, i)] == 0)))
    scheme = 1;
  else if (mo[P(j, i)] == 1 && mi[P(j, i)] == 0 &&
           (ms[P(j, im1)] == 1 || ms[P(j, ip1)] == 1 || ms[P(jm1, i)] == 1 || ms[P(jp1, i)] == 1))
    scheme = 1;
  else
    scheme = 2;
  const double mb = mb_net(in, c.p);
  const double *H = in.Hi;
  double uw = 0.0, ue = 0.0, vs = 0.0, vn = 0.0;
  if (i > 0) uw = face_vel(in, c.p, P(j, i - 1), in.u[P(j, i - 1)]);
  if (i < nx - 1) ue = face_vel(in, c.p, P(j, i + 1), in.u[P(j, i)]);
  if (j > 0) vs = face_vel(in, c.p, P(j - 1, i), in.v[P(j - 1, i)]);
  if (j < ny - 1) vn = face_vel(in, c.p, P(j + 1, i), in.v[P(j, i)]);
  double d = 1.0, cw = 0.0, ce = 0.0, cs = 0.0, cn = 0.0, b;
  if (scheme == 0) {
    b = fmax(0.0, H[c.p] + mb * dt);
  } else if (scheme == 1) {
    double Qw = 0.0, Qe = 0.0, Qs = 0.0, Qn = 0.0;
    if (i > 0) Qw = uw * H[P(j, uw > 0.0 ? i - 1 : i)] / dx;
    if (i < nx - 1) Qe = -ue * H[P(j, ue > 0.0 ? i : i + 1)] / dx;
    if (j > 0) Qs = vs * H[P(vs > 0.0 ? j - 1 : j, i)] / dx;
    if (j < ny - 1) Qn = -vn * H[P(vn > 0.0 ? j : j + 1, i)] / dx;
    b = fmax(0.0, H[c.p] + dt * (mb + Qw + Qe + Qs + Qn));
  } else {
    d = 1.0 / dt;
    if (i > 0) d -= (fs / dx) * fmin(uw, 0.0);
    if (i < nx - 1) d += (fs / dx) * fmax(ue, 0.0);
    if (j > 0) d -= (fs / dx) * fmin(vs, 0.0);
    if (j < ny - 1) d += (fs / dx) * fmax(vn, 0.0);
    if (i > 0) cw = (-fs / dx) * fmax(uw, 0.0);
    if (i < nx - 1) ce = (fs / dx) * fmin(ue, 0.0);
    if (j > 0) cs = (-fs / dx) * fmax(vs, 0.0);
    if (j < ny - 1) cn = (fs / dx) * fmin(vn, 0.0);
    b = mb + H[c.p] / dt;
    if (i > 0) b += H[c.p] * (1.0 - fs) / dx * fmin(uw, 0.0) + H[P(j, i - 1)] * (1.0 - fs) / dx * fmax(uw, 0.0);
    if (i < nx - 1) b -= H[c.p] * (1.0 - fs) / dx * fmax(ue, 0.0) + H[P(j, i + 1)] * (1.0 - fs) / dx * fmin(ue, 0.0);
    if (j > 0) b += H[c.p] * (1.0 - fs) / dx * fmin(vs, 0.0) + H[P(j - 1, i)] * (1.0 - fs) / dx * fmax(vs, 0.0);
    if (j < ny - 1) b -= H[c.p] * (1.0 - fs) / dx * fmax(vn, 0.0) + H[P(j + 1, i)] * (1.0 - fs) / dx * fmin(vn, 0.0);
  }
  R.d[c.p] = d;
  R.cw[c.p] = cw;
  R.ce[c.p] = ce;
  R.cs[c.p] = cs;
  R.cn[c.p] = cn;
  R.b[c.p] = b;
}

// y = A x for the 5-point rows
__device__ __forceinline__ double row_apply(const Member &m, const imau::ThkRows &R, const double *x, const Cell &c) {
  const int nx = m.nx, ny = m.ny;
  double s = R.d[c.p] * x[c.p];
  if (c.i > 0) s += R.cw[c.p] * x[c.p - ny];
  if (c.i < nx - 1) s += R.ce[c.p] * x[c.p + ny];
  if (c.j > 0) s += R.cs[c.p] * x[c.p - 1];
  if (c.j < ny - 1) s += R.cn[c.p] * x[c.p + 1];
  return s;
}

// block partial sums, finished by k_thk_sum (fixed order: deterministic)
__device__ __forceinline__ void block_sum(double v, double *partial, int slot, int nslots) {
  __shared__ double red[TB];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(long long)blockIdx.x * nslots + slot] = red[0];
  __syncthreads();
}

// r = b - A x, x = the start (the current thickness); partial: |r|^2, |b|^2
__global__ void __launch_bounds__(TB) k_thk_r0(const Member m, const imau::ThkRows R, const double *__restrict__ x,
                                              double *__restrict__ r, double *__restrict__ rhat,
                                              double *__restrict__ p, double *__restrict__ v,
                                              double *__restrict__ partial) {
  Cell c;
  double rr = 0.0, bb = 0.0;
  if (cell_of(m, blockIdx.x * TB + threadIdx.x, c)) {
    const double res = R.b[c.p] - row_apply(m, R, x, c);
    r[c.p] = res;
    rhat[c.p] = res;
    p[c.p] = 0.0;
    v[c.p] = 0.0;
    rr = res * res;
    bb = R.b[c.p] * R.b[c.p];
  }
  block_sum(rr, partial, 0, 2);
  block_sum(bb, partial, 1, 2);
}

__global__ void __launch_bounds__(TB) k_thk_sum(const double *__restrict__ partial, int nblocks, int nslots,
                                               double *__restrict__ out) {
  __shared__ double red[TB];
  for (int s = 0; s < nslots; ++s) {
    double acc = 0.0;
    for (int q = threadIdx.x; q < nblocks; q += TB) acc += partial[(long long)q * nslots + s];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = TB / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = red[0];
    __syncthreads();
  }
}

// BiCGStab with Jacobi preconditioning; the scalars live in sc[]:
// sc[0] rho_old, sc[1] alpha, sc[2] omega, sc[3] rho (latest dot), ...
// p = r + beta (p - omega v); phat = p / d; v = A phat; partial: (rhat, v)
__global__ void __launch_bounds__(TB) k_thk_bicg1(const Member m, const imau::ThkRows R, const double *__restrict__ r,
                                                 const double *__restrict__ rhat, double *__restrict__ p,
                                                 double *__restrict__ phat, const double *__restrict__ v,
                                                 const double *__restrict__ sc) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const double beta = (sc[3] / sc[0]) * (sc[1] / sc[2]);
  const double pn = r[c.p] + beta * (p[c.p] - sc[2] * v[c.p]);
  p[c.p] = pn;
  phat[c.p] = pn / R.d[c.p];
}
__global__ void __launch_bounds__(TB) k_thk_apply(const Member m, const imau::ThkRows R, const double *__restrict__ x,
                                                 double *__restrict__ y, const double *__restrict__ w,
                                                 double *__restrict__ partial) {
  Cell c;
  double acc = 0.0;
  if (cell_of(m, blockIdx.x * TB + threadIdx.x, c)) {
    const double s = row_apply(m, R, x, c);
    y[c.p] = s;
    acc = w[c.p] * s;
  }
  block_sum(acc, partial, 0, 1);
}
// s = r - alpha v (stored in r); shat = s / d
__global__ void __launch_bounds__(TB) k_thk_bicg2(const Member m, const imau::ThkRows R, double *__restrict__ r,
                                                 const double *__restrict__ v, double *__restrict__ shat,
                                                 const double *__restrict__ sc) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const double s = r[c.p] - sc[1] * v[c.p];
  r[c.p] = s;
  shat[c.p] = s / R.d[c.p];
}
// t = A shat; partial: (t, s), (t, t)
__global__ void __launch_bounds__(TB) k_thk_bicg3(const Member m, const imau::ThkRows R, const double *__restrict__ shat,
                                                 double *__restrict__ t, const double *__restrict__ s,
                                                 double *__restrict__ partial) {
  Cell c;
  double ts = 0.0, tt = 0.0;
  if (cell_of(m, blockIdx.x * TB + threadIdx.x, c)) {
    const double tv = row_apply(m, R, shat, c);
    t[c.p] = tv;
    ts = tv * s[c.p];
    tt = tv * tv;
  }
  block_sum(ts, partial, 0, 2);
  block_sum(tt, partial, 1, 2);
}
// x += alpha phat + omega shat; r = s - omega t; partial: |r|^2, (rhat, r);
// dxmax: the largest update of each block
__global__ void __launch_bounds__(TB) k_thk_bicg4(const Member m, double *__restrict__ x, double *__restrict__ r,
                                                 const double *__restrict__ rhat, const double *__restrict__ phat,
                                                 const double *__restrict__ shat, const double *__restrict__ t,
                                                 const double *__restrict__ sc, double *__restrict__ partial,
                                                 double *__restrict__ dxmax) {
  Cell c;
  double rr = 0.0, rh = 0.0, dm = 0.0;
  if (cell_of(m, blockIdx.x * TB + threadIdx.x, c)) {
    const double dx = sc[1] * phat[c.p] + sc[2] * shat[c.p];
    x[c.p] += dx;
    const double rn = r[c.p] - sc[2] * t[c.p];
    r[c.p] = rn;
    rr = rn * rn;
    rh = rhat[c.p] * rn;
    dm = fabs(dx);
  }
  block_sum(rr, partial, 0, 2);
  block_sum(rh, partial, 1, 2);
  // the largest update of the block, for the SOR-style stop rule
  __shared__ double mx[TB];
  mx[threadIdx.x] = dm;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) dxmax[blockIdx.x] = mx[0];
}
__global__ void __launch_bounds__(TB) k_thk_max(const double *__restrict__ v, int n, double *__restrict__ out) {
  __shared__ double mx[TB];
  double acc = 0.0;
  for (int q = threadIdx.x; q < n; q += TB) acc = fmax(acc, v[q]);
  mx[threadIdx.x] = acc;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = mx[0];
}

// H(t+dt) = max(0, x), dH/dt from it (:420-428)
__global__ void __launch_bounds__(TB) k_thk_finish(const Member m, const double *__restrict__ Hi,
                                                  const double *__restrict__ x, double dt, double *__restrict__ dHi,
                                                  double *__restrict__ Ht) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const double h = fmax(0.0, x[c.p]);
  Ht[c.p] = h;
  dHi[c.p] = (h - Hi[c.p]) / dt;
}

__device__ __forceinline__ double surface_elevation(double Hi, double Hb, double SL) {
  return Hi + fmax(SL - D_ICE_DENSITY / D_SEAWATER_DENSITY * Hi, Hb);
}
__device__ __forceinline__ double hi_from_hb_and_hs(double Hb, double Hs, double SL) {
  const double Hs_float = Hb + (SL - Hb) * D_SEAWATER_DENSITY / D_ICE_DENSITY;
  if (Hs > Hs_float) return Hs - Hb;
  return (Hs - SL) / (1.0 - D_ICE_DENSITY / D_SEAWATER_DENSITY);
}

// one side of apply_ice_thickness_BC (:719-909); side 0 west, 1 east, 2 south, 3 north
__global__ void __launch_bounds__(TB) k_thk_bc(const Member m, const imau::ThkIn in, int side, int bc, int diva,
                                              double dt, double *__restrict__ dHi, double *__restrict__ Ht) {
  const int q = blockIdx.x * TB + threadIdx.x;
  const int nx = m.nx, ny = m.ny;
  const int along = side < 2 ? ny : nx;
  if (q >= along) return;
  int jb, ib, jn, in_, jp, ip;
  if (side == 0) { jb = q; ib = 0; jn = q; in_ = 1; jp = q; ip = nx - 2; }
  else if (side == 1) { jb = q; ib = nx - 1; jn = q; in_ = nx - 2; jp = q; ip = 1; }
  else if (side == 2) { jb = 0; ib = q; jn = 1; in_ = q; jp = ny - 2; ip = q; }
  else { jb = ny - 1; ib = q; jn = ny - 2; in_ = q; jp = 1; ip = q; }
  const long long p = m.off + (long long)ib * ny + jb, pn = m.off + (long long)in_ * ny + jn;
  const double *H = in.Hi;
  if (bc == IMAU_HBC_ZERO) {
    Ht[p] = 0.0;
    dHi[p] = -H[p] / dt;
    if (diva) {
      Ht[pn] = 0.0;
      dHi[pn] = -H[pn] / dt;
    }
  } else if (bc == IMAU_HBC_INFINITE) {
    const double Hs = surface_elevation(Ht[pn], in.Hb[pn], in.SL[pn]);
    Ht[p] = hi_from_hb_and_hs(in.Hb[p], Hs, in.SL[p]);
    dHi[p] = (Ht[p] - H[p]) / dt;
  } else if (bc == IMAU_HBC_PERIODIC) {
    Ht[p] = H[m.off + (long long)ip * ny + jp];
    dHi[p] = (Ht[p] - H[p]) / dt;
  } else if (bc == IMAU_HBC_FIXED) {
    Ht[p] = H[p];
    dHi[p] = 0.0;
  }
}

// partial minima of dx / (|u| + |v|) over the interior (:746-752)
__global__ void __launch_bounds__(TB) k_cfl(const Member m, const double *__restrict__ u, const double *__restrict__ v,
                                           double *__restrict__ partial) {
  Cell c;
  double acc = INFINITY;
  if (cell_of(m, blockIdx.x * TB + threadIdx.x, c) && c.i > 0 && c.i < m.nx - 1 && c.j > 0 && c.j < m.ny - 1)
    acc = m.dx / (fabs(u[c.p]) + fabs(v[c.p]));
  __shared__ double mn[TB];
  mn[threadIdx.x] = acc;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = mn[0];
}
__global__ void __launch_bounds__(TB) k_min(const double *__restrict__ v, int n, double *__restrict__ out) {
  __shared__ double mn[TB];
  double acc = INFINITY;
  for (int q = threadIdx.x; q < n; q += TB) acc = fmin(acc, v[q]);
  mn[threadIdx.x] = acc;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = mn[0];
}

// calc_critical_timestep_SIA (src/ice_dynamics_module.f90:670-721): per cell
// dx^2 / (6 D_SIA) with D_SIA = max(1e-9, |u_vav H / dHs_dx| on the sheet's cx
// and cy faces), and dx / (u_a + v_a) of every level of the 3D velocity;
// block minima into partial (the 0.9 and the dt_max cap are the caller's)
__global__ void __launch_bounds__(TB) k_crit_sia(const Member m, int nz, const imau::Zeta zeta,
                                                const double *__restrict__ u3s, const double *__restrict__ v3s,
                                                const double *__restrict__ u3, const double *__restrict__ v3,
                                                const double *__restrict__ Hi_cx, const double *__restrict__ Hi_cy,
                                                const double *__restrict__ dHs_dx_cx,
                                                const double *__restrict__ dHs_dy_cy,
                                                const int32_t *__restrict__ sheet, double *__restrict__ partial) {
  Cell c;
  double acc = INFINITY;
  if (cell_of(m, blockIdx.x * TB + threadIdx.x, c)) {
    const int nx = m.nx, ny = m.ny, i = c.i, j = c.j;
    const long long plane = (long long)nx * ny;
    auto q3 = [&](int k, int jj, int ii) { return m.off3 + k * plane + (long long)ii * ny + jj; };
    auto vav = [&](const double *f, int jj, int ii) {  // vertical_average, src/utilities_module.f90:97-123
      double s = 0.0;
      for (int k = 0; k < nz - 1; ++k) s += 0.5 * (f[q3(k + 1, jj, ii)] + f[q3(k, jj, ii)]) * (zeta.z[k + 1] - zeta.z[k]);
      return s;
    };
    double D = 1e-9;
    if (i < nx - 1 && sheet[c.p] == 1 && sheet[c.p + ny] == 1)
      D = fmax(D, fabs(vav(u3s, j, i) * Hi_cx[c.p] / dHs_dx_cx[c.p]));
    if (j < ny - 1 && sheet[c.p] == 1 && sheet[c.p + 1] == 1)
      D = fmax(D, fabs(vav(v3s, j, i) * Hi_cy[c.p] / dHs_dy_cy[c.p]));
    acc = m.dx * m.dx / (6.0 * D);
    for (int k = 0; k < nz; ++k) {
      double ua, va;
      if (i == 0) ua = fabs(u3[q3(k, j, 0)]);
      else if (i == nx - 1) ua = fabs(u3[q3(k, j, nx - 2)]);
      else ua = fmax(fabs(u3[q3(k, j, i - 1)]), fabs(u3[q3(k, j, i)]));
      if (j == 0) va = fabs(v3[q3(k, 0, i)]);
      else if (j == ny - 1) va = fabs(v3[q3(k, ny - 2, i)]);
      else va = fmax(fabs(v3[q3(k, j - 1, i)]), fabs(v3[q3(k, j, i)]));
      acc = fmin(acc, m.dx / (ua + va));
    }
  }
  __shared__ double mn[TB];
  mn[threadIdx.x] = acc;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = mn[0];
}

// Hi_tplusdt = Hi + dt dHi_dt (src/ice_dynamics_module.f90:233)
__global__ void __launch_bounds__(TB) k_thk_advance(const Member m, const double *__restrict__ Hi,
                                                   const double *__restrict__ dHi, double dt, double *__restrict__ Ht) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  Ht[c.p] = Hi[c.p] + dt * dHi[c.p];
}

// Hi = max(0, Hi_tplusdt) (src/ice_dynamics_module.f90:559)
__global__ void __launch_bounds__(TB) k_thk_clip(const Member m, const double *__restrict__ Ht, double *__restrict__ Hi) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  Hi[c.p] = fmax(0.0, Ht[c.p]);
}

}  // namespace

namespace imau {

int thk_blocks(const Member &m) { return (m.nx * m.ny + TB - 1) / TB; }

void thk_explicit(hipStream_t s, const Member &m, const ThkIn &in, double dt, const ThkWork &w, double *dHi,
                  double *Ht) {
  const int nb = thk_blocks(m);
  hipLaunchKernelGGL(k_thk_explicit1, dim3(nb), dim3(TB), 0, s, m, in, dt, w.Qx, w.Qy, w.t0, w.t1);
  hipLaunchKernelGGL(k_thk_explicit2, dim3(nb), dim3(TB), 0, s, m, in, dt, w.Qx, w.Qy, w.t0, w.t1, dHi, Ht);
  hipLaunchKernelGGL(k_thk_store_flux, dim3(nb), dim3(TB), 0, s, m, w.Qx, w.t1, 1);
  hipLaunchKernelGGL(k_thk_store_flux, dim3(nb), dim3(TB), 0, s, m, w.Qy, w.t1, 0);
}

void thk_rows(hipStream_t s, const Member &m, const ThkIn &in, double dt, const ThkRows &R) {
  hipLaunchKernelGGL(k_thk_rows, dim3(thk_blocks(m)), dim3(TB), 0, s, m, in, dt, R);
}

void thk_r0(hipStream_t s, const Member &m, const ThkRows &R, const double *x, const ThkWork &w, double *sums) {
  const int nb = thk_blocks(m);
  hipLaunchKernelGGL(k_thk_r0, dim3(nb), dim3(TB), 0, s, m, R, x, w.r, w.rhat, w.p, w.v, w.partial);
  hipLaunchKernelGGL(k_thk_sum, dim3(1), dim3(TB), 0, s, w.partial, nb, 2, sums);
}

// one BiCGStab iteration; sc (device) holds rho_old, alpha, omega, rho; the
// host computes the scalars from the sums it reads back (this path runs once
// per time step, not in the Picard loop)
void thk_bicg_p(hipStream_t s, const Member &m, const ThkRows &R, const ThkWork &w, const double *sc,
                double *sums) {
  const int nb = thk_blocks(m);
  hipLaunchKernelGGL(k_thk_bicg1, dim3(nb), dim3(TB), 0, s, m, R, w.r, w.rhat, w.p, w.phat, w.v, sc);
  hipLaunchKernelGGL(k_thk_apply, dim3(nb), dim3(TB), 0, s, m, R, w.phat, w.v, w.rhat, w.partial);
  hipLaunchKernelGGL(k_thk_sum, dim3(1), dim3(TB), 0, s, w.partial, nb, 1, sums);
}
void thk_bicg_s(hipStream_t s, const Member &m, const ThkRows &R, const ThkWork &w, const double *sc, double *sums) {
  const int nb = thk_blocks(m);
  hipLaunchKernelGGL(k_thk_bicg2, dim3(nb), dim3(TB), 0, s, m, R, w.r, w.v, w.shat, sc);
  hipLaunchKernelGGL(k_thk_bicg3, dim3(nb), dim3(TB), 0, s, m, R, w.shat, w.t, w.r, w.partial);
  hipLaunchKernelGGL(k_thk_sum, dim3(1), dim3(TB), 0, s, w.partial, nb, 2, sums);
}
void thk_bicg_x(hipStream_t s, const Member &m, double *x, const ThkWork &w, const double *sc, double *sums) {
  const int nb = thk_blocks(m);
  hipLaunchKernelGGL(k_thk_bicg4, dim3(nb), dim3(TB), 0, s, m, x, w.r, w.rhat, w.phat, w.shat, w.t, sc, w.partial,
                     w.t0);
  hipLaunchKernelGGL(k_thk_sum, dim3(1), dim3(TB), 0, s, w.partial, nb, 2, sums);
  hipLaunchKernelGGL(k_thk_max, dim3(1), dim3(TB), 0, s, w.t0, nb, sums + 2);
}

void thk_finish(hipStream_t s, const Member &m, const double *Hi, const double *x, double dt, double *dHi,
                double *Ht) {
  hipLaunchKernelGGL(k_thk_finish, dim3(thk_blocks(m)), dim3(TB), 0, s, m, Hi, x, dt, dHi, Ht);
}

void thk_bc(hipStream_t s, const Member &m, const ThkIn &in, const int *bcs, int diva, double dt, double *dHi,
            double *Ht) {
  for (int side = 0; side < 4; ++side) {
    if (bcs[side] == IMAU_HBC_NONE) continue;
    const int along = side < 2 ? m.ny : m.nx;
    hipLaunchKernelGGL(k_thk_bc, dim3((along + TB - 1) / TB), dim3(TB), 0, s, m, in, side, bcs[side], diva, dt, dHi,
                       Ht);
  }
}

void crit_sia(hipStream_t s, const Member &m, int nz, const Zeta &zeta, const double *u3s, const double *v3s,
              const double *u3, const double *v3, const double *Hi_cx, const double *Hi_cy, const double *dHs_dx_cx,
              const double *dHs_dy_cy, const int32_t *sheet, double *partial, double *out) {
  const int nb = thk_blocks(m);
  hipLaunchKernelGGL(k_crit_sia, dim3(nb), dim3(TB), 0, s, m, nz, zeta, u3s, v3s, u3, v3, Hi_cx, Hi_cy, dHs_dx_cx,
                     dHs_dy_cy, sheet, partial);
  hipLaunchKernelGGL(k_min, dim3(1), dim3(TB), 0, s, partial, nb, out);
}
// predictor of run_ice_dynamics_pc (src/ice_dynamics_module.f90:333-340):
// Hi_pred = max(0, Hi + dt ((1 + zeta/2) dHn - (zeta/2) dHnm1)), Hi_old = Hi, Hi = Hi_pred
__global__ void __launch_bounds__(TB) k_pc_pred(const Member m, double *__restrict__ Hi, double *__restrict__ Hold,
                                                double *__restrict__ Hpred, const double *__restrict__ dHn,
                                                const double *__restrict__ dHnm1, double dt, double zeta) {
  Cell c;
  if (!cell_of(m, blockIdx.x * TB + threadIdx.x, c)) return;
  const double h = Hi[c.p];
  const double hp = fmax(0.0, h + dt * ((1.0 + zeta / 2.0) * dHn[c.p] - (zeta / 2.0) * dHnm1[c.p]));
  Hold[c.p] = h;
  Hpred[c.p] = hp;
  Hi[c.p] = hp;
}
// corrector (:404-414) at the restored Hi = Hi_old: Hi_corr = max(0, Hi + dt/2 (dHn +
// dHstar)), dHi = (Hi_corr - Hi) / dt; and calc_pc_truncation_error (:763-816):
// the block maximum of tau = |zeta (Hi_corr - Hi_pred) / ((3 zeta + 3) dt)| over the
// interior sheet cells with no grounding line in their 3 x 3 neighbourhood
__global__ void __launch_bounds__(TB) k_pc_corr(const Member m, const double *__restrict__ Hi,
                                                const double *__restrict__ Hpred, const double *__restrict__ dHn,
                                                const double *__restrict__ dHstar, double dt, double zeta,
                                                const int32_t *__restrict__ sheet, const int32_t *__restrict__ gl,
                                                double eta_min, double *__restrict__ dHi, double *__restrict__ partial) {
  Cell c;
  double acc = eta_min;
  if (cell_of(m, blockIdx.x * TB + threadIdx.x, c)) {
    const double hc = fmax(0.0, Hi[c.p] + 0.5 * dt * (dHn[c.p] + dHstar[c.p]));
    dHi[c.p] = (hc - Hi[c.p]) / dt;
    if (c.i > 0 && c.i < m.nx - 1 && c.j > 0 && c.j < m.ny - 1 && sheet[c.p] == 1) {
      bool no_gl = true;
      for (int di = -1; di <= 1; ++di)
        for (int dj = -1; dj <= 1; ++dj) no_gl = no_gl && gl[c.p + (long long)di * m.ny + dj] == 0;
      if (no_gl) acc = fmax(acc, fabs(zeta * (hc - Hpred[c.p]) / ((3.0 * zeta + 3.0) * dt)));
    }
  }
  __shared__ double mx[TB];
  mx[threadIdx.x] = acc;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = mx[0];
}
__global__ void __launch_bounds__(TB) k_max(const double *__restrict__ v, int n, double *__restrict__ out) {
  __shared__ double mx[TB];
  double acc = -INFINITY;
  for (int q = threadIdx.x; q < n; q += TB) acc = fmax(acc, v[q]);
  mx[threadIdx.x] = acc;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = mx[0];
}

void pc_pred(hipStream_t s, const Member &m, double *Hi, double *Hold, double *Hpred, const double *dHn,
             const double *dHnm1, double dt, double zeta) {
  hipLaunchKernelGGL(k_pc_pred, dim3(thk_blocks(m)), dim3(TB), 0, s, m, Hi, Hold, Hpred, dHn, dHnm1, dt, zeta);
}
void pc_corr(hipStream_t s, const Member &m, const double *Hi, const double *Hpred, const double *dHn,
             const double *dHstar, double dt, double zeta, const int32_t *sheet, const int32_t *gl, double eta_min,
             double *dHi, double *partial, double *eta) {
  const int nb = thk_blocks(m);
  hipLaunchKernelGGL(k_pc_corr, dim3(nb), dim3(TB), 0, s, m, Hi, Hpred, dHn, dHstar, dt, zeta, sheet, gl, eta_min,
                     dHi, partial);
  hipLaunchKernelGGL(k_max, dim3(1), dim3(TB), 0, s, partial, nb, eta);
}
void thk_advance(hipStream_t s, const Member &m, const double *Hi, const double *dHi, double dt, double *Ht) {
  hipLaunchKernelGGL(k_thk_advance, dim3(thk_blocks(m)), dim3(TB), 0, s, m, Hi, dHi, dt, Ht);
}
void thk_clip(hipStream_t s, const Member &m, const double *Ht, double *Hi) {
  hipLaunchKernelGGL(k_thk_clip, dim3(thk_blocks(m)), dim3(TB), 0, s, m, Ht, Hi);
}
void cfl(hipStream_t s, const Member &m, const double *u, const double *v, double *partial, double *out) {
  const int nb = thk_blocks(m);
  hipLaunchKernelGGL(k_cfl, dim3(nb), dim3(TB), 0, s, m, u, v, partial);
  hipLaunchKernelGGL(k_min, dim3(1), dim3(TB), 0, s, partial, nb, out);
}

}  // namespace imau
