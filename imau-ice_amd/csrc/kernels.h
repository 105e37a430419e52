// kernels.h -- host-side launchers of the device kernels (kernels.hip).
#pragma once
#include "dev_common.h"

namespace imau {

struct Launch {
  hipStream_t s;
  int ntiles;
  const Member *mem;
  const Tile *tiles;
  // the 3D viscosity launch also carries the tiles of its row-marching
  // kernel (one wavefront per tile); none: the tiled kernel runs
  int nmtiles = 0;
  const Tile *mtiles = nullptr;
};

// coefficient pointers of the matrix-free operator
struct Coef {
  const double *Na, *Nb, *Ncx, *Ncy, *bcx, *bcy, *taudx, *taudy, *frhs;
  const uint8_t *rt_u, *rt_v;
};

void setup1(const Launch &L, const KP &kp, const Fields &f, const double *yg);
void setup2(const Launch &L, const KP &kp, const Fields &f);
// 3D viscosity, N, F2 (k_visc) and then basal friction / beta_eff (k_beta)
// (L: the standard tiles, for k_strain; LV: IMAU_VTI-row tiles, for k_visc)
void visc3d(const Launch &L, const Launch &LV, const KP &kp, const Fields &f, bool diva, const double *visc_in,
            double *visc_out, const int *active);
// the same in FP32 (viscosity storage and arithmetic; N, F2 written in FP64)
void visc3d32(const Launch &L, const Launch &LV, const KP &kp, const Fields &f, bool diva, const float *visc_in,
              float *visc_out, const float *B3, const int *active);
void beta(const Launch &L, const KP &kp, const Fields &f, bool diva, const double *xg, const double *yg,
          const int *active);
// element-wise precision conversion of n values
void to_f32(hipStream_t s, const double *a, float *b, long long n);
void to_f64(hipStream_t s, const float *a, double *b, long long n);
void coef(const Launch &L, const KP &kp, const Fields &f, bool diva, const int *active);
// warm-start network I/O (image g = member g's a-grid, pixel poff[g] + i ny + j)
void net_features(const Launch &L, const KP &kp, const Fields &f, const long long *poff, unsigned short *x);
void net_targets(const Launch &L, const Fields &f, const long long *poff, float *tg, float *wt);
void net_warm(const Launch &L, const KP &kp, const Fields &f, const long long *poff, const float *o,
              const int *active);
// column-local completion of a velocity-only warm start (k_strain + k_visc_column)
void visc_column(const Launch &L, const Launch &LV, const KP &kp, const Fields &f, double *visc, const double *xg,
                 const double *yg, int iters, const int *active);
void visc_column32(const Launch &L, const Launch &LV, const KP &kp, const Fields &f, float *visc, const float *B3,
                   const double *xg, const double *yg, int iters, const int *active);
// tau_b = beta_eff u on cx / cy (warm-start completion)
void taub(const Launch &L, const Fields &f, const int *active);
void init_x(const Launch &L, const Fields &f, double *xu, double *xv);
// relax: per-member relaxation factors, or nullptr for the configured DIVA_visc_it_relax
void post(const Launch &L, const KP &kp, const Fields &f, const double *xu, const double *xv, const int *active,
          bool diva, double *partial, const double *relax, const AMix &mix);
// Anderson: store the slot and accumulate (F_k, F_s); member sums of the IMAU_AND_SLOTS dots
void and_prep(const Launch &L, const KP &kp, const Fields &f, const double *xu, const double *xv, const AMix &mix,
              int slot, const int *active, double *partial);
void member_sum_and(hipStream_t s, int ngrids, const int *tstart, const double *partial, double *out);
void resid(const Launch &L, const KP &kp, const Coef &c, const double *xu, const double *xv, double *partial);
void member_sum3(hipStream_t s, int ngrids, const int *tstart, const double *partial, double *out);
// the nonlinear residual at the state fused with x0 (init_x) and r = b - A x0; partials
// (|Au-b|^2, |b|^2, |r|^2) per tile, reduced with member_sum3
void resid_init(const Launch &L, const KP &kp, const Coef &c, const double *xu, const double *xv, double *x0u,
                double *x0v, double *ru, double *rv, const int *active, double *partial);
void member_sum2(hipStream_t s, int ngrids, const int *tstart, const double *partial, double *out);
void fill_u64(hipStream_t s, unsigned long long *p, int n, unsigned long long v);
// dst = src on the 3D slabs of the members with flag[g] != 0
void copy3_members(const Launch &L, int nz, const double *src, double *dst, const int *flag);
// (du, dv) = (su, sv) on the 2D slabs of the members with flag[g] != 0
void copy2_members(const Launch &L, const double *su, const double *sv, double *du, double *dv, const int *flag);
void copy3_members(const Launch &L, int nz, const float *src, float *dst, const int *flag);
// the state a problem was created with (velocities, basal stress, 3D viscosity)
struct ResetSrc {
  const double *u, *v, *tbx, *tby, *visc;
};
// reset the members with flag[g] != 0 to that state (visc32: also the FP32 viscosity, or null)
void reset_members(const Launch &L, const KP &kp, const Fields &f, const ResetSrc &src, double *visc, float *visc32,
                   const int *flag);
void fill_i32(hipStream_t s, int *p, int n, int v);
// (du, dv) = v on the 2D slabs of the members with flag[g] != 0
void fill2_members(const Launch &L, double *du, double *dv, double v, const int *flag);
// r[g] = sqrt(sums[2g]) / sqrt(sums[2g+1]) (or sqrt(sums[2g]) when the second is 0)
void resid_ratio(hipStream_t s, int ngrids, const double *sums, double *r);
void F1(const Launch &L, const KP &kp, const Fields &f, const double *visc, const OutF &o);
void uv3d(const Launch &L, const KP &kp, const Fields &f, const OutF &o);
void compose(const Launch &L, const KP &kp, const Fields &f, const OutF &o, int mode);
// w_3D_a (with_w), the a-grid 3D velocities and the secondary 2D fields in one pass
void final_a(const Launch &L, const KP &kp, const Fields &f, const OutF &o, int dyn, const double *uvav,
             const double *vvav, bool with_w);
void sia(const Launch &L, const KP &kp, const Fields &f, const double *Acx, const double *Acy, const OutF &o);
void h2d_3d(hipStream_t s, const Member *mem, int g, const double *src, double *dst, int ni, int nj, int ldj, int nz);
void d2h_3d(hipStream_t s, const Member *mem, int g, const double *src, double *dst, int ni, int nj, int ldj, int nz);

// Krylov solver and multicolour SOR
void apply(const Launch &L, const KP &kp, const Coef &c, const double *xu, const double *xv, double *yu, double *yv,
           const double *wu, const double *wv, const int *active, double *partial);
void bicg_init(const Launch &L, const KP &kp, const Coef &c, const double *xu, const double *xv, double *ru, double *rv,
               double *rhu, double *rhv, double *pu, double *pv, double *vu, double *vv, const int *active,
               double *partial);
void bicg_p(const Launch &L, const double *ru, const double *rv, double *pu, double *pv, const double *vu,
            const double *vv, const double *sc, const int *active);
void bicg_s(const Launch &L, double *ru, double *rv, const double *vu, const double *vv, const double *sc,
            const int *active);
void bicg_xr(const Launch &L, double *xu, double *xv, double *ru, double *rv, const double *phu, const double *phv,
             const double *shu, const double *shv, const double *tu, const double *tv, const double *rhu,
             const double *rhv, const double *sc, const int *active, double *partial);
void bicg_scalar(hipStream_t s, int ngrids, const int *tstart, const double *partial, double *sc, double *out,
                 const int *active, int stage);
void prec_tile(const Launch &L, const KP &kp, const Coef &c, const double *yu, const double *yv, double *zu, double *zv,
               const int *active, int nsweeps);
void sor_color(const Launch &L, const KP &kp, const Coef &c, double *xu, double *xv, int color, double omega,
               const int *active, bool track, unsigned long long *resmax, const double *ru = nullptr,
               const double *rv = nullptr);
// GCR: partial dots, per-member scalars, orthogonalising update
void gcr_dots(const Launch &L, const double *qu, const double *qv, const double *ru, const double *rv, const GList &Q,
              int nj, const int *active, double *partial);
void gcr_scalar(hipStream_t s, int ngrids, const int *tstart, const double *partial, double *sc, const int *active,
                int nj);
void gcr_update(const Launch &L, const GList &Q, const GList &Z, int nj, double *xu, double *xv, double *ru, double *rv,
                const double *sc, const int *active, double *partial);
}  // namespace imau
