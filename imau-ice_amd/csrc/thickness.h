// thickness.h -- launchers of the ice-thickness kernels (thickness.hip).
#pragma once
#include "../../include/imau_vel.h"
#include "dev_common.h"

namespace imau {

// inputs of one member, in the solver's slab layout (u, v: cx / cy in a-grid-shaped slabs)
struct ThkIn {
  const double *Hi, *Hb, *SL, *u, *v, *SMB, *BMB, *fmf;
  const int32_t *m_ice, *m_shelf, *m_ocean, *m_cf;
};
// the semi-implicit scheme's 5-point rows
struct ThkRows {
  double *d, *cw, *ce, *cs, *cn, *b;
};
// work arrays (one slab each) and block partials
struct ThkWork {
  double *Qx, *Qy, *t0, *t1;                       // explicit: fluxes, mass-balance volume, rescale factors
  double *r, *rhat, *p, *phat, *v, *shat, *t;      // BiCGStab vectors
  double *partial;                                 // 2 doubles per block
};

int thk_blocks(const Member &m);
void thk_explicit(hipStream_t s, const Member &m, const ThkIn &in, double dt, const ThkWork &w, double *dHi,
                  double *Ht);
void thk_rows(hipStream_t s, const Member &m, const ThkIn &in, double dt, const ThkRows &R);
// r = b - A x (x the start), rhat = r, p = v = 0; sums[0] = |r|^2, sums[1] = |b|^2
void thk_r0(hipStream_t s, const Member &m, const ThkRows &R, const double *x, const ThkWork &w, double *sums);
// BiCGStab steps with Jacobi preconditioning; sc = {rho_old, alpha, omega, rho} on the device:
// p and phat = p / d, v = A phat, sums[0] = (rhat, v)
void thk_bicg_p(hipStream_t s, const Member &m, const ThkRows &R, const ThkWork &w, const double *sc, double *sums);
// s = r - alpha v (in r), shat = s / d, t = A shat; sums[0] = (t, s), sums[1] = (t, t)
void thk_bicg_s(hipStream_t s, const Member &m, const ThkRows &R, const ThkWork &w, const double *sc, double *sums);
// x += alpha phat + omega shat, r = s - omega t; sums[0] = |r|^2, sums[1] = (rhat, r), sums[2] = max |dx|
void thk_bicg_x(hipStream_t s, const Member &m, double *x, const ThkWork &w, const double *sc, double *sums);
void thk_finish(hipStream_t s, const Member &m, const double *Hi, const double *x, double dt, double *dHi, double *Ht);
// apply_ice_thickness_BC, sides west, east, south, north in that order
void thk_bc(hipStream_t s, const Member &m, const ThkIn &in, const int *bcs, int diva, double dt, double *dHi,
            double *Ht);
// the zeta levels of a column (a kernel argument)
struct Zeta {
  double z[IMAU_NZ_MAX];
};
// calc_critical_timestep_SIA's minimum over cells (before the 0.9 and the
// dt_max cap) into *out (device); 3D fields k-plane major in the member's slab
void crit_sia(hipStream_t s, const Member &m, int nz, const Zeta &zeta, const double *u3s, const double *v3s,
              const double *u3, const double *v3, const double *Hi_cx, const double *Hi_cy, const double *dHs_dx_cx,
              const double *dHs_dy_cy, const int32_t *sheet, double *partial, double *out);
// Ht = Hi + dt dHi; Hi = max(0, Ht)
void thk_advance(hipStream_t s, const Member &m, const double *Hi, const double *dHi, double dt, double *Ht);
void thk_clip(hipStream_t s, const Member &m, const double *Ht, double *Hi);
// predictor-corrector steps of run_ice_dynamics_pc; pc_corr writes dHi and the
// truncation error pc_eta into *eta (device); partial holds thk_blocks(m) doubles
void pc_pred(hipStream_t s, const Member &m, double *Hi, double *Hold, double *Hpred, const double *dHn,
             const double *dHnm1, double dt, double zeta);
void pc_corr(hipStream_t s, const Member &m, const double *Hi, const double *Hpred, const double *dHn,
             const double *dHstar, double dt, double zeta, const int32_t *sheet, const int32_t *gl, double eta_min,
             double *dHi, double *partial, double *eta);
// min over interior cells of dx / (|u| + |v|) into *out (device)
void cfl(hipStream_t s, const Member &m, const double *u, const double *v, double *partial, double *out);

}  // namespace imau
