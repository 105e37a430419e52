/*
 * imau_oracle.h -- CPU FP64 restatement of IMAU-ICE's ice-velocity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU library is compared
 * against; nothing in the product (libimauvel.so, the Python package) links,
 * loads or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it.
 *
 * It implements the same contract as include/imau_vel.h (same structs, same
 * array layouts) with the reference's own algorithm: CSR assembly of the
 * linearised staggered DIVA/SSA and either the reference SOR
 * (src/utilities_module.f90:644-718) or an exact banded LU solve standing in
 * for PETSc's KSP (src/petsc_module.F90:24-88).
 *
 * Pinning (tests/test_oracle_pinning.py): closed forms, invariants and
 * library routines, since the reference has no test suite, no golden vectors
 * and cannot be built here.  PARITY UNPINNED (no closed form, invariant or
 * fixture reaches these parts; they follow the cited reference lines only):
 *   - calving-front and shelf-land rows of choice_ice_margin = 'BC'
 *     (src/ice_velocity_module.f90:1762-1865, 2192-2341);
 *   - the corner logic of 'periodic' / 'zero' domain boundaries (:1932-2191);
 *   - the CISM quadrant rotations with their ftol perturbations
 *     (src/general_ice_model_data_module.f90:539-574, 608-680);
 *   - the cold start (0/0 in du/dz resolved to visc_eff_min through fmax);
 *   - Picard trajectories and SOR iteration counts at reference tolerances;
 *   - PETSc's own Krylov iterations (replaced by the exact banded LU).
 */
#ifndef IMAU_ORACLE_H
#define IMAU_ORACLE_H
#include "../include/imau_vel.h"
#ifdef __cplusplus
extern "C" {
#endif

/* linear-solver choice of the oracle */
enum { ORC_LIN_SOR = 0, ORC_LIN_BANDED_LU = 1 };

int    orc_solve_sia (const imau_grid *g, const imau_cfg *c, imau_ice_io *io);
int    orc_solve_ssa (const imau_grid *g, const imau_cfg *c, imau_ice_io *io, int lin_solver);
int    orc_solve_diva(const imau_grid *g, const imau_cfg *c, imau_ice_io *io, int lin_solver);
int    orc_nonlinear_residual(const imau_grid *g, const imau_cfg *c, const imau_ice_io *io, double *res);
const char *orc_last_error(void);
void   orc_abi_layout(int64_t *out);

/* building blocks, exposed for unit tests */
void   orc_vertical_integration_from_bottom_to_zeta(int nz, const double *zeta, const double *f, double *out);
double orc_vertical_integrate(int nz, const double *zeta, const double *f);
double orc_vertical_average(int nz, const double *zeta, const double *f);
double orc_calc_fraction_above_zero(double f_NW, double f_NE, double f_SW, double f_SE);
double orc_sliding_beta(const imau_cfg *c, double u_a, double v_a, double beta_sq, double alpha_sq,
                        double phi_fric, double tauc, double Neff, double x, double y, double A_vav);
void   orc_update_general_ice_model_data(const imau_grid *g, const imau_cfg *c,
                                          const double *Hi, const double *Hb, const double *SL,
                                          const double *A3,
                                          double *Hs, double *TAF, int32_t *mask_land, int32_t *mask_ocean,
                                          int32_t *mask_ice, int32_t *mask_sheet, int32_t *mask_shelf,
                                          double *Hi_cx, double *Hi_cy, double *Hi_b,
                                          double *A3_cx, double *A3_cy, double *A_vav,
                                          double *dHi_dx_a, double *dHi_dy_a, double *dHs_dx_a, double *dHs_dy_a,
                                          double *dHs_dx_cx, double *dHs_dy_cx, double *dHs_dx_cy, double *dHs_dy_cy);
/* Assemble the CSR of solve_DIVA_stag_linearised for the coefficients in io
 * (N_a, N_b, beta_eff_c*, taud*, masks) and initial guess u,v.  Arrays sized
 * by the caller: ptr (m+1), idx/val (9m), b/x (m); returns nnz. */
int    orc_assemble_csr(const imau_grid *g, const imau_cfg *c, const imau_ice_io *io,
                        const double *u_cx, const double *v_cy,
                        int32_t *ptr, int32_t *idx, double *val, double *b, double *x);
int    orc_sor(int m, const int32_t *ptr, const int32_t *idx, const double *val, const double *b,
               double *x, int nit, double tol, double omega, int *iterations);
int    orc_banded_lu_solve(int m, const int32_t *ptr, const int32_t *idx, const double *val,
                           const double *b, double *x);
/* ice-thickness update (src/ice_thickness_module.f90:25-909) and the
 * advective critical time step (src/ice_dynamics_module.f90:726-762) */
/* update_general_ice_model_data with the full masks (and the ocean flood
 * fill) and calc_ice_rheology, with the library's producer structs */
int    orc_producers_update(const imau_grid *g, const imau_producer_cfg *c, imau_producer_io *io);
int    orc_calc_ice_rheology(const imau_grid *g, const imau_producer_cfg *c, imau_producer_io *io, double time);
/* the 'direct' time loop with update_ice_thickness, same arguments as
 * imau_run_ice_dynamics_direct (include/imau_vel.h) without the context */
int    orc_run_ice_dynamics_direct(const imau_grid *g, const imau_cfg *c, const imau_thickness_cfg *tc,
                                   const imau_producer_cfg *pc, const imau_timestep_cfg *sc, imau_ice_io *io,
                                   const imau_refgeo_io *rg, const double *SMB, const double *BMB, double *Hi_end,
                                   imau_producer_io *geo, imau_timestep_state *st, double t_end);
int    orc_run_ice_dynamics_pc    (const imau_grid *g, const imau_cfg *c, const imau_thickness_cfg *tc,
                                   const imau_producer_cfg *pc, const imau_timestep_cfg *sc, imau_ice_io *io,
                                   const imau_refgeo_io *rg, const double *SMB, const double *BMB, double *Hi_end,
                                   imau_producer_io *geo, imau_timestep_state *st, imau_pc_state *pcs, double t_end);
int    orc_calc_basal_conditions(const imau_grid *g, const imau_cfg *c, imau_ice_io *io, double *phi, double *tauc);
void   orc_floating_margin_fraction(const imau_grid *g, const double *Hi, const int32_t *ice, const int32_t *cf,
                                    const int32_t *shelf, double *fmf, double *heff);
void   orc_fix_geometry(const imau_grid *g, const imau_timestep_cfg *sc, const double *Hi, const int32_t *sheet,
                        const int32_t *shelf, double *Ht);
void   orc_remove_unconnected_shelves(const imau_grid *g, double *Hi, const int32_t *shelf, const int32_t *sheet);
double orc_critical_timestep_sia(const imau_grid *g, int nz, const double *zeta, double dt_max, const double *u3sia,
                                 const double *v3sia, const double *u3, const double *v3, const double *Hi_cx,
                                 const double *Hi_cy, const double *dHs_dx_cx, const double *dHs_dy_cy,
                                 const int32_t *sheet);
int    orc_thickness_update(const imau_grid *g, const imau_thickness_cfg *c, imau_thickness_io *io, double dt);
double orc_critical_timestep_adv(const imau_grid *g, const double *u_cx, const double *v_cy, double dt_max);
/* Schoof (2006) ice-stream closed form, src/utilities_module.f90:836-877 */
void   orc_ssa_schoof2006(double n_flow, double L, double m, double tantheta, double h0,
                          double A_flow, double y, double *U, double *tauc);

#ifdef __cplusplus
}
#endif
#endif
