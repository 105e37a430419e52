// mg.h -- host interface of the aggregation multigrid preconditioner (mg.hip).
//
// The preconditioner approximates A^-1 y for the linearised staggered DIVA/SSA
// operator by one V-cycle.  It runs either in FP64 or, for the mixed-precision
// mode, entirely in FP32 (FP32 copies of the fine coefficients, FP32 coarse
// stencils, FP32 iterate) while the Krylov loop around it keeps its vectors,
// operator applications and residual norms in FP64.
#pragma once
#include <vector>

#include "dev_common.h"
#include "kernels.h"

namespace imau {

// one coarse level; data holds [cu 9na][cv 9na][xu][xv][bu][bv][x2u][x2v] (na each)
// elements of the working precision (sized for FP64, FP32 uses the first half)
struct MGLevelDev {
  const CMember *mem;
  const Tile *tiles;
  int ntiles;
  int lev;
  long long na;
  void *data;
  bool ends;  // some member's coarsest level is this one
};

// FP32 copies of the fine-level coefficients and the FP32 right-hand side / iterate
struct MGFine32 {
  float *Na, *Nb, *Ncx, *Ncy, *bcx, *bcy, *yu, *yv, *zu, *zv, *zu2, *zv2;
};

struct MGPlan {
  std::vector<MGLevelDev> levels;
  const int *depth;  // per member: its coarsest level
  int ngrids;
  MGFine32 f32;
  double *zu64, *zv64;   // FP64 scratch iterate of the fused smoother
  const Tile *ftiles;    // IMAU_FTI x IMAU_FTJ tiles of the fused smoother
  int nftiles;
  int coarsest_sweeps;
  uint8_t *plain;        // per level-1 cell: bit c when the component-c aggregate is plain (mg.hip)
  int2 *list;            // the level-1 cells that are not plain: (member, index in the member's level)
  int *nlist_d;          // their number on the device, and on the host after mg_plain_flags
  int nlist;
  bool fused;            // one fused launch per fine-level sweep
  bool periodic;         // some domain boundary is periodic (needs a separate boundary-row pass)
};

// marker callback around the first fine-level smoother sweep (which = 0 before, 1 after)
typedef void (*MGMark)(void *ctx, int which);

// FP32 fine coefficients for the current operator and, with galerkin, the
// Galerkin coarse operators (once per linear solve, or every few solves)
void mg_setup(const Launch &F, const KP &kp, const Coef &c, const MGPlan &plan, bool fp32, bool galerkin,
              const int *active);
// once per solve, after the row types are set: marks the level-1 aggregates
// whose Galerkin sums take the straight-line path
// (the caller copies *plan.nlist_d into plan.nlist before the next mg_setup)
void mg_plain_flags(const Launch &F, const uint8_t *rt_u, const uint8_t *rt_v, const MGPlan &plan);
// z = V-cycle(y) for every active member
void mg_vcycle(const Launch &F, const KP &kp, const Coef &c, const MGPlan &plan, bool fp32, const double *yu,
               const double *yv, double *zu, double *zv, const int *active, MGMark mark, void *mark_ctx);

}  // namespace imau
