// kernels.h -- host-side launchers of the device kernels (kernels.hip).
#pragma once
#include "dev_common.h"

struct CoefP;
namespace imau {

struct Launch {
  hipStream_t s;
  int ntiles;
  const Member *mem;
  const Tile *tiles;
};

// coefficient pointers of the matrix-free operator
struct Coef {
  const double *Na, *Nb, *Ncx, *Ncy, *bcx, *bcy, *taudx, *taudy, *frhs;
  const uint8_t *rt_u, *rt_v;
};

void setup1(const Launch &L, const KP &kp, const Fields &f, const double *yg);
void setup2(const Launch &L, const KP &kp, const Fields &f);
void visc(const Launch &L, const KP &kp, const Fields &f, bool diva, const double *visc_in, double *visc_out,
          const double *xg, const double *yg, const int *active);
void coef(const Launch &L, const KP &kp, const Fields &f, bool diva, const int *active);
void init_x(const Launch &L, const Fields &f, double *xu, double *xv);
void sweep(const Launch &L, const KP &kp, const Coef &c, const double *xu_in, const double *xv_in, double *xu_out,
           double *xv_out, const int *active, double omega, bool track, unsigned long long *resmax);
void post(const Launch &L, const KP &kp, const Fields &f, const double *xu, const double *xv, const int *active,
          bool diva, double *partial);
void resid(const Launch &L, const KP &kp, const Coef &c, const double *xu, const double *xv, double *partial);
void member_sum3(hipStream_t s, int ngrids, const int *tstart, const double *partial, double *out);
void member_sum2(hipStream_t s, int ngrids, const int *tstart, const double *partial, double *out);
void fill_u64(hipStream_t s, unsigned long long *p, int n, unsigned long long v);
void F1(const Launch &L, const KP &kp, const Fields &f, const double *visc, const OutF &o);
void uv3d(const Launch &L, const KP &kp, const Fields &f, const OutF &o);
void compose(const Launch &L, const KP &kp, const Fields &f, const OutF &o, int mode);
void w3d(const Launch &L, const KP &kp, const Fields &f, const OutF &o);
void secondary(const Launch &L, const KP &kp, const Fields &f, const OutF &o, int dyn, const double *uvav,
               const double *vvav);
void sia(const Launch &L, const KP &kp, const Fields &f, const double *Acx, const double *Acy, const OutF &o);
void h2d_3d(hipStream_t s, const Member *mem, int g, const double *src, double *dst, int ni, int nj, int ldj, int nz);
void d2h_3d(hipStream_t s, const Member *mem, int g, const double *src, double *dst, int ni, int nj, int ldj, int nz);

}  // namespace imau
