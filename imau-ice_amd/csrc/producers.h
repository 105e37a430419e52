// producers.h -- launchers of the input-producer kernels (producers.hip):
// update_general_ice_model_data (src/general_ice_model_data_module.f90:30-346,
// ocean flood fill :868-986) and calc_ice_rheology
// (src/thermodynamics_module.f90:641-753).
#pragma once
#include "../../include/imau_vel.h"
#include "dev_common.h"

namespace imau {

// strides of a 3D field: element (k, j, i) at k*sk + i*si + j*sj
struct Str3 {
  long long sk, si, sj;
};

// 2D fields of one member in a-grid-shaped slabs (leading dimension ny, offset m.off)
struct ProdIn {
  const double *Hi, *Hb, *SL, *dHi_dt, *dHb_dt;  // the rates may be null
};
struct ProdMasks {
  int32_t *land, *ocean, *ice, *sheet, *shelf, *coast, *margin, *gl, *cf, *total;
};
struct Prod2 {
  double *Hs, *TAF, *dHs_dt, *Hi_cx, *Hi_cy, *Hi_b;
  double *dHi_dx, *dHi_dy, *dHs_dx, *dHs_dy, *dHs_dx_cx, *dHs_dy_cx, *dHs_dx_cy, *dHs_dy_cy;
};

// Hs, TAF, dHs/dt (reading the land mask as it is on entry), then the
// land/ocean masks from flotation (floodfill = 0) or the edge cells of the
// flood fill (floodfill = 1)
void prod_point(hipStream_t s, const Member &m, const ProdIn &in, const Prod2 &o, const ProdMasks &mk, int floodfill);
// one round of the flood fill: grows the ocean through the 8 neighbours of
// floating cells, tile by tile to a local fixed point; *changed (device) is
// set to 1 when a cell changed
void prod_flood_round(hipStream_t s, const Member &m, const ProdIn &in, const ProdMasks &mk, int *changed);
// land = 1 - ocean after the flood fill
void prod_land(hipStream_t s, const Member &m, const ProdMasks &mk);
// ice / sheet / shelf, the transition masks, the total mask, H on cx / cy / b
// and the slopes of H and Hs
void prod_masks_maps(hipStream_t s, const Member &m, const ProdIn &in, const Prod2 &o, const ProdMasks &mk);
// A on cx / cy: the vertical average (2D, may be null) and the 3D field
// (null: skipped); nz levels with the given strides
void prod_map_A(hipStream_t s, const Member &m, int nz, const double *A3, Str3 sa, double *A3cx, Str3 sx, double *A3cy,
                Str3 sy, const double *Av, double *Avcx, double *Avcy);
// determine_floating_margin_fraction (reads mk.ice, mk.cf, mk.shelf); heff
// (may be null) receives the effective calving-front thickness Hi_eff_cf_a
void prod_float_margin(hipStream_t s, const Member &m, const double *Hi, const ProdMasks &mk, double *fmf,
                       double *heff);
// update_ice_thickness's fixed-geometry switches on H(t+dt): which = bit 0
// fixed_shelf_geometry, bit 1 fixed_sheet_geometry, bit 2 fixed_grounding_line
void prod_fix_geometry(hipStream_t s, const Member &m, int which, const double *Hi, const ProdMasks &mk, double *Ht);
// threshold_thickness_calving: calving-front cells with heff < thr lose their ice
void prod_calve(hipStream_t s, const Member &m, const ProdMasks &mk, const double *heff, double thr, double *Hi);
// the removals after the calving; a null field switches its rule off
struct Removal {
  const int32_t *noice;                 // mask_noice
  int shelves;                          // do_remove_shelves
  const double *Hi_PD, *Hb_PD;          // remove_shelves_larger_than_PD
  const double *Hi_GIA, *Hb_GIA;        // continental_shelf_calving
  double min_height;                    // continental_shelf_min_height
};
void prod_remove_ice(hipStream_t s, const Member &m, const Removal &r, const double *Hb, const double *SL, double *Hi);
// remove_unconnected_shelves: seeds into keep, rounds of growth through the
// 4 neighbours of shelf cells (*changed set when a cell changed), removal
void prod_shelf_seed(hipStream_t s, const Member &m, const ProdMasks &mk, int32_t *keep);
void prod_shelf_round(hipStream_t s, const Member &m, const ProdMasks &mk, int32_t *keep, int *changed);
void prod_shelf_remove(hipStream_t s, const Member &m, const ProdMasks &mk, const int32_t *keep, double *Hi);
// calc_ice_rheology's flow factor, enhancement factors and vertical average
void prod_rheology(hipStream_t s, const Member &m, const imau_producer_cfg &c, double time, const double *Ti, Str3 st,
                   const ProdMasks &mk, double *A3, Str3 sa, double *Av);

}  // namespace imau
