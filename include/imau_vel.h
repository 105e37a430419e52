/*
 * imau_vel.h -- C ABI of the MI355X ice-velocity library (libimauvel.so).
 *
 * Drop-in replacement for the velocity solvers of IMAU-ICE's
 * ice_velocity_module (src/ice_velocity_module.f90).  A Fortran host keeps its
 * own module of that name whose routines solve_SIA / solve_SSA / solve_DIVA
 * (grid, ice) call the functions below through ISO_C_BINDING.
 *
 * Conventions (identical to the reference data types,
 * src/data_types_module.f90:13-325):
 *   - every array is the host's own storage, borrowed for the duration of the
 *     call, Fortran column-major with 1-based (j,i) mapped to 0-based offsets:
 *       a-grid  (ny  ,nx  )  offset = i*ny     + j
 *       cx-grid (ny  ,nx-1)  offset = i*ny     + j      (u between a(j,i),a(j,i+1))
 *       cy-grid (ny-1,nx  )  offset = i*(ny-1) + j      (v between a(j,i),a(j+1,i))
 *       b-grid  (ny-1,nx-1)  offset = i*(ny-1) + j
 *       3D      (nz,ny,nx)   offset = (i*ny + j)*nz + k (k=0 surface, k=nz-1 base)
 *   - reals are double, masks are int32 0/1;
 *   - a NULL output pointer means "not requested"; a NULL input pointer is an
 *     error unless the configuration does not read that field.
 *   - units as in the reference: m, yr, Pa, Pa^-n yr^-1 (src/data_types_module.f90:93,151,210).
 *
 * Error behaviour: every entry point returns IMAU_OK (0) or a negative status;
 * imau_vel_last_error() gives the message.  The reference calls crash() in the
 * same situations (src/configuration_module.f90:3566-3638); the Fortran shim
 * maps a non-zero status onto crash().
 */
#ifndef IMAU_VEL_H
#define IMAU_VEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMAU_VEL_ABI_VERSION 1
#define IMAU_NZ_MAX 32

/* status codes */
enum {
  IMAU_OK              =  0,
  IMAU_ERR_CONFIG      = -1,  /* unknown choice / inconsistent configuration   */
  IMAU_ERR_ARG         = -2,  /* NULL required pointer, bad grid size          */
  IMAU_ERR_NAN         = -3,  /* NaN / Inf detected in a solver field          */
  IMAU_ERR_DIVERGED    = -4,  /* linear solver diverged (reference: SOR guard) */
  IMAU_ERR_STALLED     = -5,  /* FP64 residual stalled, fallback also failed   */
  IMAU_ERR_DEVICE      = -6,  /* HIP runtime error                             */
  IMAU_ERR_NO_DEVICE   = -7   /* library built without / cannot see a GPU      */
};

/* choice_ice_dynamics (src/configuration_module.f90:271) */
enum { IMAU_DYN_NONE = 0, IMAU_DYN_SIA = 1, IMAU_DYN_SSA = 2, IMAU_DYN_SIA_SSA = 3, IMAU_DYN_DIVA = 4 };
/* choice_ice_margin (src/configuration_module.f90:275) */
enum { IMAU_MARGIN_BC = 0, IMAU_MARGIN_INFINITE_SLAB = 1 };
/* choice_sliding_law (src/configuration_module.f90:349) */
enum {
  IMAU_SLID_NO_SLIDING = 0, IMAU_SLID_IDEALISED = 1, IMAU_SLID_COULOMB = 2,
  IMAU_SLID_COULOMB_REGULARISED = 3, IMAU_SLID_WEERTMAN = 4, IMAU_SLID_TSAI2015 = 5,
  IMAU_SLID_SCHOOF2005 = 6, IMAU_SLID_ZOET_IVERSON = 7
};
/* choice_idealised_sliding_law (src/configuration_module.f90:350) */
enum { IMAU_IDEAL_ISMIP_HOM_C = 0, IMAU_IDEAL_ISMIP_HOM_D = 1, IMAU_IDEAL_ISMIP_HOM_E = 2, IMAU_IDEAL_ISMIP_HOM_F = 3 };
/* choice_basal_hydrology (src/configuration_module.f90:361) */
enum { IMAU_HYDRO_DRY = 0, IMAU_HYDRO_SATURATED = 1, IMAU_HYDRO_MARTIN2011 = 2 };
/* choice_basal_roughness / choice_param_basal_roughness (src/configuration_module.f90:366,373) */
enum { IMAU_ROUGH_UNIFORM = 0, IMAU_ROUGH_PARAMETERISED = 1, IMAU_ROUGH_PRESCRIBED = 2 };
enum { IMAU_PROUGH_MARTIN2011 = 0, IMAU_PROUGH_SSA_ICESTREAM = 1, IMAU_PROUGH_MISMIPPLUS = 2 };
/* DIVA_boundary_BC_* (src/configuration_module.f90:288-295) */
enum { IMAU_BC_INFINITE = 0, IMAU_BC_ZERO = 1, IMAU_BC_PERIODIC = 2 };
/* DIVA_choice_matrix_solver (src/configuration_module.f90:296).  The GPU
 * replaces both reference solvers; IMAU_SOLVER_SOR keeps the SOR stop rule
 * (max-norm of the Gauss-Seidel update <= DIVA_SOR_tol), IMAU_SOLVER_PETSC
 * keeps the KSP stop rule (||r|| <= max(rtol*||b||, abstol)). */
enum { IMAU_SOLVER_SOR = 0, IMAU_SOLVER_PETSC = 1 };
/* arithmetic of the linearised solve (no reference counterpart: the
 * reference is FP64 throughout) */
/* precision (GPU-only): FP64 throughout; MIXED = FP32 multigrid preconditioner
 * inside the FP64 Krylov loop; MIXED_VISC = MIXED plus FP32 viscosity storage
 * and arithmetic (N, F2, the operator, residual and update stay FP64) */
enum { IMAU_PREC_FP64 = 0, IMAU_PREC_MIXED = 1, IMAU_PREC_MIXED_VISC = 2 };
/* Picard stop rule: reference rule (relative velocity change, ice_velocity_module.f90:346-351)
 * or the FP64 relative nonlinear residual ||A(u)u-b(u)||/||b(u)|| <= nonlinear_rtol */
enum { IMAU_STOP_PICARD = 0, IMAU_STOP_RESIDUAL = 1 };

/* type_grid (src/data_types_module.f90:13-29) -- the fields the path reads */
typedef struct imau_grid {
  int32_t nx, ny;
  double  dx;
  const double *x;   /* (nx) a-grid x coordinates [m] */
  const double *y;   /* (ny) a-grid y coordinates [m] */
} imau_grid;

/* The velocity-solver part of the reference's global config struct C
 * (src/configuration_module.f90:271-302, 349-377, 168-171, 425-426, 70-85),
 * with every CHARACTER choice mapped onto an enum. */
typedef struct imau_cfg {
  int32_t choice_ice_dynamics;
  int32_t choice_ice_margin;
  int32_t choice_sliding_law;
  int32_t choice_idealised_sliding_law;
  int32_t choice_basal_hydrology;
  int32_t choice_basal_roughness;
  int32_t choice_param_basal_roughness;
  int32_t DIVA_boundary_BC_u_west, DIVA_boundary_BC_u_east, DIVA_boundary_BC_u_south, DIVA_boundary_BC_u_north;
  int32_t DIVA_boundary_BC_v_west, DIVA_boundary_BC_v_east, DIVA_boundary_BC_v_south, DIVA_boundary_BC_v_north;
  int32_t DIVA_choice_matrix_solver;
  int32_t include_SSADIVA_crossterms;
  int32_t do_GL_subgrid_friction;
  int32_t include_basal_freezing;
  int32_t nz;
  double  zeta[IMAU_NZ_MAX];
  double  n_flow;
  double  DIVA_visc_it_norm_dUV_tol;
  int32_t DIVA_visc_it_nit;
  double  DIVA_visc_it_relax;
  double  DIVA_epsilon_sq_0;
  double  DIVA_visc_eff_min;
  double  DIVA_beta_max;
  double  DIVA_vel_max;
  double  slid_delta_v;
  double  slid_Weertman_m;
  double  slid_Coulomb_reg_q_plastic;
  double  slid_Coulomb_reg_u_threshold;
  double  slid_ZI_ut;
  double  slid_ZI_p;
  double  deltaT_basal_freezing;
  double  Martin2011_hydro_Hb_min, Martin2011_hydro_Hb_max;
  double  Martin2011till_phi_Hb_min, Martin2011till_phi_Hb_max;
  double  Martin2011till_phi_min, Martin2011till_phi_max;
  double  SSA_icestream_L, SSA_icestream_m;
  double  ISMIP_HOM_L;
  /* GPU-only knobs (no reference counterpart) */
  int32_t precision;       /* IMAU_PREC_*                                     */
  int32_t stop_rule;       /* IMAU_STOP_*                                     */
  double  nonlinear_rtol;  /* target of IMAU_STOP_RESIDUAL                    */
  int32_t mg_levels;       /* KSP mode preconditioner: 0 = aggregation multigrid
                              (automatic depth), 1 = tile Gauss-Seidel only     */
  int32_t check_interval;  /* sweeps between linear-solver convergence checks */
  double  picard_relax_max; /* 0: the reference's fixed DIVA_visc_it_relax.  > 0:
                              adaptive relaxation, per member: start at
                              DIVA_visc_it_relax, grow x1.25 after every Picard
                              iteration whose FP64 residual fell, back to
                              DIVA_visc_it_relax after one that rose, capped at
                              this value (needs IMAU_STOP_RESIDUAL)          */
  int32_t mg_smoother;     /* multigrid fine-level smoother: 0 = one fused launch
                              per sweep (tiles with a recomputed halo in LDS),
                              1 = one launch per colour; identical results     */
  int32_t ksp_type;        /* KSP-mode Krylov method: 0 = BiCGStab, 1 = GCR
                              restarted every 4 directions (one preconditioner
                              application per iteration)                       */
  int32_t mg_rebuild;      /* rebuild the Galerkin coarse operators every
                              mg_rebuild-th linear solve (1 = every solve; in
                              between, the previous coarse levels precondition
                              the current fine operator)                      */
  int32_t anderson_m;      /* 0: relaxation (fixed or adaptive).  1..3: Anderson
                              acceleration of the Picard map with m stored
                              differences, restarted for a member whose FP64
                              residual grows by anderson_restart (needs
                              IMAU_STOP_RESIDUAL)                            */
  int32_t warm_visc_iters; /* DIVA: before the Picard loop, this many passes of
                              viscosity, beta_eff and tau_b = beta_eff u at the
                              given velocities, without a linear solve (0 = none,
                              the reference's start).  Completes a warm start
                              that supplies velocities only.                  */
  double  anderson_beta;   /* Anderson step u = sum_i w_i (U_i + beta F_i)
                              (F = G(U) - U); 1 = undamped, in (0, 1]      */
  double  anderson_restart; /* restart a member's Anderson history (and take
                              one relaxation step) when its FP64 residual
                              exceeds this factor times the previous one;
                              >= 1                                           */
  int32_t warm_column_iters; /* DIVA: before the Picard loop (and before the
                              warm_visc_iters passes), complete a
                              velocity-only warm start column by column:
                              viscosity, F2, beta_eff and |tau_b| iterated
                              this many times in one kernel with each
                              column's own basal stress (0 = none)          */
  int32_t stall_window;    /* IMAU_STOP_RESIDUAL: a member whose FP64 residual
                              did not fall by the factor stall_drop over the
                              last stall_window Picard iterations, or whose
                              first residual is not below 1 (the residual of
                              zero velocities), restarts once from the state
                              the problem was created with (the cold start)
                              inside the same Picard loop; its iteration count
                              keeps running.  For a warm-started solve.  0 =
                              off                                           */
  double  stall_drop;      /* > 1                                            */
  double  stall_relax;     /* DIVA_visc_it_relax of a restarted member's
                              unmixed Picard steps (0 = DIVA_visc_it_relax)  */
  int32_t mg_graph;        /* 1: each multigrid V-cycle is replayed from a HIP
                              graph captured on its first use in a solve (one
                              launch instead of one per kernel); 0: launched
                              kernel by kernel.  Identical results          */
} imau_cfg;

/* The fields of type_ice_model (src/data_types_module.f90:56-325) that the
 * velocity path reads or writes. */
typedef struct imau_ice_io {
  /* ---- inputs: geometry and derived data (update_general_ice_model_data) ---- */
  const double  *Hi_a, *Hb_a, *SL_a, *Hs_a, *TAF_a;       /* (ny,nx) */
  const double  *dHs_dx_a, *dHs_dy_a, *dHi_dx_a, *dHi_dy_a, *dHb_dt_a;
  const double  *A_flow_vav_a;                             /* (ny,nx) */
  const int32_t *mask_land_a, *mask_ocean_a, *mask_ice_a, *mask_sheet_a, *mask_shelf_a;
  const double  *Hi_cx, *dHs_dx_cx, *dHs_dy_cx;            /* (ny,nx-1) */
  const double  *Hi_cy, *dHs_dx_cy, *dHs_dy_cy;            /* (ny-1,nx) */
  const double  *Hi_b;                                     /* (ny-1,nx-1) */
  const double  *A_flow_3D_a;                              /* (nz,ny,nx) */
  const double  *A_flow_3D_cx, *A_flow_3D_cy;              /* SIA only */
  const double  *Ti_a, *Ti_pmp_a;                          /* only if include_basal_freezing */
  const int32_t *ISMIP_HOM_E_slides;  /* (nx) ISMIP-HOM E: the fourth column ("slides") of the
                                         C%ISMIP_HOM_E_Arolla_filename file, line i for grid column i
                                         (src/basal_conditions_and_sliding_module.f90:1352-1399): beta is
                                         0 where it is 1, 1e30 elsewhere.  The file has 51 lines, so the
                                         grid has nx = 51.  Read once when the problem is created */
  /* ---- bed roughness (in; updated when choice_basal_roughness=parameterised) ---- */
  double *beta_sq_a, *alpha_sq_a, *phi_fric_a, *tauc_a;    /* (ny,nx) */
  /* ---- in/out state (warm start and result) ---- */
  double *u_vav_cx, *v_vav_cy;                             /* DIVA */
  double *u_SSA_cx, *v_SSA_cy;                             /* SSA, SIA/SSA */
  double *taub_cx, *taub_cy;
  double *visc_eff_3D_a;                                   /* (nz,ny,nx) */
  /* ---- outputs (NULL = not requested) ---- */
  double *u_3D_cx, *v_3D_cy, *u_3D_SIA_cx, *v_3D_SIA_cy;   /* (nz,..) */
  double *u_3D_a, *v_3D_a, *w_3D_a, *F1_3D_a;              /* (nz,ny,nx) */
  double *du_dz_3D_cx, *dv_dz_3D_cy;
  double *u_vav_a, *v_vav_a, *u_surf_a, *v_surf_a, *u_base_a, *v_base_a;
  double *uabs_vav_a, *uabs_surf_a, *uabs_base_a, *w_surf_a, *w_base_a, *R_shear;
  double *u_surf_cx, *v_surf_cy, *u_base_cx, *v_base_cy;
  double *beta_a, *beta_eff_a, *beta_eff_cx, *beta_eff_cy;
  double *taudx_cx, *taudy_cy;
  double *N_a, *N_b, *N_cx, *N_cy, *visc_eff_int_a, *F2_a;
  double *f_grnd_a, *f_grnd_cx, *f_grnd_cy, *f_grnd_b;
  double *overburden_pressure_a, *pore_water_pressure_a, *Neff_a;
  int32_t *DIVA_mask_combi, *DIVA_isfront_inner, *DIVA_isfront_outer;
  double *u_cx_prev, *v_cy_prev;
  double *DIVA_err_cx, *DIVA_err_cy;   /* estimate_visc_iter_UV_errors (src/ice_velocity_module.f90:1602-1652):
                                          1e5 at the start of a solve, then after every Picard iteration
                                          that does not end the loop 2|u-u_prev|/|u+u_prev+1e-5| where
                                          |u_vav| > 1e-6, else 0.  Computed only when requested at create;
                                          a member whose solve returns before its Picard loop (no grounded
                                          ice) keeps the values it had, the caller's at create */
  /* ---- per-call solver tolerances (ice%DIVA_SOR_*, ice%DIVA_PETSc_*,
   *      src/data_types_module.f90:256-260; set by the caller as in
   *      src/ice_dynamics_module.f90:94-99) ---- */
  int32_t DIVA_SOR_nit;
  double  DIVA_SOR_tol, DIVA_SOR_omega;
  double  DIVA_PETSc_rtol, DIVA_PETSc_abstol;
  /* ---- statistics written by the solve (no reference counterpart) ---- */
  int32_t stat_picard_iterations;
  int64_t stat_linear_sweeps;
  double  stat_resid_UV;            /* last Picard relative change           */
  double  stat_nonlinear_residual;  /* FP64 ||A(u)u-b(u)||/||b(u)|| at return */
  double  stat_seconds;             /* wall time of the device solve          */
} imau_ice_io;

typedef struct imau_ctx imau_ctx;

/* Replaces initialise_SSADIVA_solution_matrix (src/ice_velocity_module.f90:2814-2860,
 * called at src/ice_dynamics_module.f90:1197): creates the per-process
 * context with device workspaces for this grid. */
int  imau_vel_init(const imau_grid *grid, const imau_cfg *cfg, imau_ctx **ctx);

/* Replaces solve_SIA (src/ice_velocity_module.f90:36-114). */
int  imau_vel_solve_sia(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg, imau_ice_io *io);

/* Replaces solve_SSA (src/ice_velocity_module.f90:115-248). */
int  imau_vel_solve_ssa(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg, imau_ice_io *io);

/* Replaces solve_DIVA (src/ice_velocity_module.f90:249-374). */
int  imau_vel_solve_diva(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg, imau_ice_io *io);

/* FP64 relative nonlinear residual ||A(u)u-b(u)||_2/||b(u)||_2 of the DIVA
 * (or SSA) system at the state held in io (u, taub, visc_eff_3D_a), with A
 * assembled as solve_DIVA_stag_linearised does (src/ice_velocity_module.f90:1656-2589).
 * No reference counterpart: it is the stop rule of the GPU solves. */
int  imau_vel_nonlinear_residual(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg,
                                 imau_ice_io *io, double *residual);

/* Replaces finalise (no reference counterpart beyond freeing the CSR windows). */
void imau_vel_finalize(imau_ctx *ctx);

/* Text of the last error on this context (or of the last failed init). */
const char *imau_vel_last_error(const imau_ctx *ctx);

/* Kernel timing (no reference counterpart): returns the number of timed
 * launches, their summed duration, measured with HIP events on the context's
 * stream, and the summed a-grid cells of the members each timed launch
 * processed (members that have converged are skipped), since the previous
 * call; then resets the counters and
 * selects what is timed next: IMAU_TIME_OFF, IMAU_TIME_SMOOTHER (one
 * fine-level smoother sweep: the five colour launches of a multigrid sweep,
 * a tile Gauss-Seidel application, or one SOR sweep) or IMAU_TIME_VISC (the
 * 3D viscosity kernel of a Picard iteration). */
enum { IMAU_TIME_OFF = 0, IMAU_TIME_SMOOTHER = 1, IMAU_TIME_VISC = 2 };
int  imau_vel_kernel_timing(imau_ctx *ctx, int32_t what, int64_t *launches, double *ms, int64_t *cells);

/* Fills cfg with the reference defaults (src/configuration_module.f90). */
void imau_vel_default_config(imau_cfg *cfg);

/* ---- device-resident problems (no reference counterpart) ----
 * A problem holds one or more grids (an ensemble) in device memory: create
 * uploads inputs and state, solve runs the velocity solve on the device only,
 * download copies the requested outputs back, reset_state restores the state
 * that was uploaded (for repeated cold/warm-start timing).  The convenience
 * entry points imau_vel_solve_* are create + solve + download. */
typedef struct imau_problem imau_problem;
int  imau_vel_problem_create(imau_ctx *ctx, int32_t ngrids, const imau_grid *grids, const imau_cfg *cfg,
                             const imau_ice_io *ios, imau_problem **prob);
int  imau_vel_problem_solve(imau_ctx *ctx, imau_problem *prob, const imau_cfg *cfg);
int  imau_vel_problem_reset_state(imau_ctx *ctx, imau_problem *prob);
int  imau_vel_problem_download(imau_ctx *ctx, imau_problem *prob, imau_ice_io *ios);
/* per member of the last solve: Picard iterations and the last FP64 relative
 * nonlinear residual (NaN under IMAU_STOP_PICARD); either pointer may be NULL */
int  imau_vel_problem_stats(imau_ctx *ctx, const imau_problem *prob, int32_t *picard_iterations,
                            double *nonlinear_residual);
/* per member of the last solve: the number of stall restarts (0 or 1, see
 * imau_cfg.stall_window) */
int  imau_vel_problem_restarts(imau_ctx *ctx, const imau_problem *prob, int32_t *restarts);
void imau_vel_problem_destroy(imau_problem *prob);

/* No reference counterpart: the project's tile autotuner.  Times the 3D
 * viscosity kernel of DIVA at nz = 15 on this problem's own grids and state
 * with HIP events, over its tile heights (0: the tiled kernel, 8 rows per
 * 256-thread workgroup; 4, 8, 16, 32: the row-marching kernel, that many rows
 * per 64-lane wavefront), keeps the fastest for the problem's later solves and
 * returns it in *tile_rows (may be NULL).  The state is not changed; other
 * configurations keep their fixed tiles.  Every height gives the same
 * viscosity bit for bit. */
int  imau_vel_problem_autotune(imau_ctx *ctx, imau_problem *prob, const imau_cfg *cfg, int32_t *tile_rows);
/* No reference counterpart: sets the tile height the autotuner would choose
 * (one of 0, 4, 8, 16, 32; IMAU_ERR_ARG otherwise). */
int  imau_vel_problem_set_tile_rows(imau_ctx *ctx, imau_problem *prob, int32_t tile_rows);

/* No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the FP64 relative nonlinear residual of every member of a
 * device-resident problem at its current state (u, taub, visc_eff_3D_a), the
 * same value imau_vel_nonlinear_residual gives for that member's state, written
 * to residual_dev[ngrids] in DEVICE memory, asynchronously on the context's
 * stream (no host synchronisation, no allocation, no memset).  Its reduction
 * partials, member sums and flags live in a static device array of
 * IMAU_NOALLOC_SCRATCH_BYTES per GPU: problems needing more (2 doubles per
 * 8 x 64 tile plus 2.5 per member, about 33 M cells) are rejected with
 * IMAU_ERR_ARG.  All calls on one GPU share that array: calls that overlap on
 * one GPU (several contexts or threads) race.  The problem's coefficient
 * diagnostics (N, beta, beta_eff, the second viscosity buffer) are
 * overwritten; its state is not. */
#define IMAU_NOALLOC_SCRATCH_BYTES (1 << 20)
int  imau_vel_problem_residual_noalloc(imau_ctx *ctx, imau_problem *prob, const imau_cfg *cfg,
                                       double *residual_dev);

/* ---- ice-thickness update: the consumer of u_vav_cx / v_vav_cy ----
 * choice_ice_integration_method (src/configuration_module.f90:308) */
enum { IMAU_DHI_NONE = 0, IMAU_DHI_EXPLICIT = 1, IMAU_DHI_SEMIIMPLICIT = 2 };
/* ice_thickness_*_BC (src/configuration_module.f90:325-328) */
enum { IMAU_HBC_NONE = 0, IMAU_HBC_ZERO = 1, IMAU_HBC_INFINITE = 2, IMAU_HBC_PERIODIC = 3, IMAU_HBC_FIXED = 4 };

/* The thickness part of the reference's config struct C
 * (src/configuration_module.f90:308-314, 325-328, dt_max) */
typedef struct imau_thickness_cfg {
  int32_t choice_ice_integration_method;
  int32_t choice_ice_dynamics;      /* 'zero' boundaries clear two rows under DIVA */
  int32_t ice_thickness_west_BC, ice_thickness_east_BC, ice_thickness_south_BC, ice_thickness_north_BC;
  int32_t dHi_choice_matrix_solver; /* IMAU_SOLVER_SOR or IMAU_SOLVER_PETSC (semi-implicit scheme) */
  int32_t dHi_SOR_nit;
  double  dHi_SOR_tol, dHi_SOR_omega;
  double  dHi_PETSc_rtol, dHi_PETSc_abstol;
  double  dt_max;
} imau_thickness_cfg;

/* The fields calc_dHi_dt reads and writes (type_ice_model, type_SMB_model,
 * type_BMB_model; src/data_types_module.f90), host arrays in the layouts above. */
typedef struct imau_thickness_io {
  const double  *Hi_a, *Hb_a, *SL_a;                        /* (ny,nx) */
  const double  *u_vav_cx, *v_vav_cy;                       /* (ny,nx-1), (ny-1,nx) */
  const double  *SMB_year, *BMB, *float_margin_frac_a;      /* (ny,nx) */
  const int32_t *mask_ice_a, *mask_shelf_a, *mask_ocean_a, *mask_cf_a;
  double *dHi_dt_a, *Hi_tplusdt_a;                          /* (ny,nx) out */
  double *Qx_cx, *Qy_cy;                                    /* explicit scheme's fluxes (m^3 per step), out, may be NULL */
} imau_thickness_io;

void imau_thickness_default_config(imau_thickness_cfg *cfg);

/* Replaces calc_dHi_dt (src/ice_thickness_module.f90:25-58: the explicit
 * scheme :59-312 or the semi-implicit one :313-718, then
 * apply_ice_thickness_BC :719-909) over one time step dt [yr]. */
int  imau_thickness_update(imau_ctx *ctx, const imau_grid *grid, const imau_thickness_cfg *cfg,
                           imau_thickness_io *io, double dt);

/* The same with the velocities of a device-resident problem's member g (the
 * result of its last solve) instead of io->u_vav_cx / v_vav_cy, so they need
 * not leave the device; io holds the other fields of member g. */
int  imau_vel_problem_thickness(imau_ctx *ctx, imau_problem *prob, int32_t g, const imau_thickness_cfg *cfg,
                                imau_thickness_io *io, double dt);

/* Replaces calc_critical_timestep_adv (src/ice_dynamics_module.f90:726-762):
 * min(dt_max, 0.9 min over interior points of dx / (|u| + |v|)). */
int  imau_critical_timestep_adv(imau_ctx *ctx, const imau_grid *grid, const double *u_vav_cx,
                                const double *v_vav_cy, double dt_max, double *dt_crit_adv);

/* ---- input producers: the fields the velocity solve reads, from the
 *      geometry and the temperature ----
 * choice_ice_rheology (src/configuration_module.f90:425) */
enum { IMAU_RHEO_UNIFORM = 0, IMAU_RHEO_HUYBRECHTS1992 = 1, IMAU_RHEO_MISMIP_MOD = 2 };
/* mask_a values (C%type_*, src/configuration_module.f90:3171-3179) */
enum { IMAU_TYPE_LAND = 0, IMAU_TYPE_OCEAN = 1, IMAU_TYPE_LAKE = 2, IMAU_TYPE_SHEET = 3, IMAU_TYPE_SHELF = 4,
       IMAU_TYPE_COAST = 5, IMAU_TYPE_MARGIN = 6, IMAU_TYPE_GROUNDINGLINE = 7, IMAU_TYPE_CALVINGFRONT = 8 };

/* The C% fields the producers read (src/configuration_module.f90:168-171,
 * 273-274, 425-426, 702). */
typedef struct imau_producer_cfg {
  int32_t nz;
  double  zeta[IMAU_NZ_MAX];
  int32_t choice_ice_margin;        /* the flow factor of ice-free cells under Huybrechts1992 */
  int32_t do_ocean_floodfill;       /* 1: ocean = floating cells connected to the domain edge */
  int32_t choice_ice_rheology;
  double  uniform_flow_factor;      /* [Pa^-3 yr^-1] */
  double  m_enh_sheet, m_enh_shelf;
} imau_producer_cfg;

/* The fields update_general_ice_model_data and calc_ice_rheology read and
 * write (type_ice_model, src/data_types_module.f90:56-325), host arrays in
 * the layouts above (2D (ny,nx) a-grid unless noted, 3D (nz,ny,nx)).  An
 * output pointer may be NULL (not written). */
typedef struct imau_producer_io {
  const double *Hi_a, *Hb_a, *SL_a;
  const double *dHi_dt_a, *dHb_dt_a;        /* NULL: dHs_dt_a is not computed; it reads
                                               mask_land_a as it is on entry (the previous
                                               call's), as the reference does */
  const double *Ti_a;                       /* 3D; Huybrechts1992 only */
  double  *Hs_a, *TAF_a, *dHs_dt_a;
  int32_t *mask_land_a, *mask_ocean_a, *mask_ice_a, *mask_sheet_a, *mask_shelf_a;
  int32_t *mask_coast_a, *mask_margin_a, *mask_gl_a, *mask_cf_a, *mask_a;
  double  *Hi_cx, *Hi_cy, *Hi_b;            /* (ny,nx-1), (ny-1,nx), (ny-1,nx-1) */
  double  *A_flow_3D_a, *A_flow_vav_a;      /* in for update_general_ice_model_data, out for calc_ice_rheology */
  double  *A_flow_3D_cx, *A_flow_3D_cy;     /* (nz,ny,nx-1), (nz,ny-1,nx) */
  double  *A_flow_vav_cx, *A_flow_vav_cy;
  double  *dHi_dx_a, *dHi_dy_a, *dHs_dx_a, *dHs_dy_a;
  double  *dHs_dx_cx, *dHs_dy_cx, *dHs_dx_cy, *dHs_dy_cy;
  double  *float_margin_frac_a, *Hi_eff_cf_a;  /* time loop only: ice%float_margin_frac_a /
                                                  Hi_eff_cf_a as the last update_ice_thickness
                                                  left them (rows 1 .. ny-2) */
} imau_producer_io;

void imau_producer_default_config(imau_producer_cfg *cfg);

/* Replaces update_general_ice_model_data (src/general_ice_model_data_module.f90:30-84):
 * Hs, TAF and dHs/dt, the masks (determine_masks :87-346, with the ocean
 * flood fill :868-986 when do_ocean_floodfill), H on cx/cy/b, the flow factor
 * in io->A_flow_3D_a / A_flow_vav_a mapped to cx/cy, and the surface and
 * thickness slopes.  The masks read by calc_ice_rheology are written first. */
int  imau_update_general_ice_model_data(imau_ctx *ctx, const imau_grid *grid, const imau_producer_cfg *cfg,
                                        imau_producer_io *io);

/* Replaces calc_ice_rheology (src/thermodynamics_module.f90:641-753) at model
 * time `time` [yr]: the 3D flow factor (uniform, Huybrechts (1992) Arrhenius
 * law of Ti_a, or MISMIP_mod's step function of time) times the sheet/shelf
 * enhancement factors (io->mask_ice_a, mask_sheet_a, mask_shelf_a are read),
 * its vertical average, and both mapped to cx/cy. */
int  imau_calc_ice_rheology(imau_ctx *ctx, const imau_grid *grid, const imau_producer_cfg *cfg,
                            imau_producer_io *io, double time);

/* ---- time stepping: choice_timestepping = 'direct' ----
 * C%choice_calving_law (src/configuration_module.f90:412, calving_module.f90:30-91) */
#define IMAU_CALVING_NONE                0
#define IMAU_CALVING_THRESHOLD_THICKNESS 1

/* The C% fields of the time loop (src/configuration_module.f90:50-55) and of
 * update_ice_thickness (:335-337, :412-417); flags are 0 / 1. */
typedef struct imau_timestep_cfg {
  double start_time_of_run, end_time_of_run;   /* [yr] */
  double dt_min, dt_max, dt_startup_phase;     /* [yr] */
  int32_t choice_calving_law;                  /* IMAU_CALVING_* */
  int32_t fixed_shelf_geometry, fixed_sheet_geometry, fixed_grounding_line;
  int32_t do_remove_shelves, remove_shelves_larger_than_PD, continental_shelf_calving;
  int32_t pad0_;
  double calving_threshold_thickness;          /* [m] */
  double continental_shelf_min_height;         /* [m] */
  /* choice_timestepping = 'pc' (src/configuration_module.f90:317-320; Robinson
   * et al. 2020, Eqs. 32-33): target truncation error [m/yr], the two
   * exponents of the step controller, the floor of the error estimate */
  double pc_epsilon, pc_k_I, pc_k_p, pc_eta_min;
} imau_timestep_cfg;

/* The fixed inputs update_ice_thickness reads beside the ice model
 * (src/ice_dynamics_module.f90:435-445), all (ny,nx): mask_noice (1 = no ice
 * allowed; NULL = none), and the reference geometries that
 * remove_shelves_larger_than_PD (refgeo_PD) and continental_shelf_calving
 * (refgeo_GIAeq) read; NULL where the option is off. */
typedef struct imau_refgeo_io {
  const int32_t *mask_noice_a;
  const double *Hi_PD, *Hb_PD;
  const double *Hi_GIAeq, *Hb_GIAeq;
} imau_refgeo_io;

/* The timers of type_model_region the loop reads and writes
 * (src/data_types_module.f90: region%time, t_next_SIA, t_next_SSA, dt_crit_*, dt). */
typedef struct imau_timestep_state {
  double time, t_next_SIA, t_next_SSA;          /* in / out */
  double dt_crit_SIA, dt_crit_SSA, dt;          /* out: the last values */
  int32_t nsteps;                               /* out: time steps taken by the call */
} imau_timestep_state;

/* Replaces the ice part of the main time loop (src/IMAU_ICE_main_model.f90:90-214)
 * with choice_timestepping = 'direct' and no other model component:
 * while time < t_end, run_ice_dynamics_direct (src/ice_dynamics_module.f90:63-242:
 * the SIA and SSA solves when their timers are due, each with its critical
 * time step calc_critical_timestep_SIA :651-725 / _adv :726-762, the time
 * step of determine_timesteps_and_actions :817-941 limited to the dynamics,
 * then calc_dHi_dt), update_ice_thickness (:435-648: the fixed-geometry
 * switches, H = max(0, H_t+dt), the masks, determine_floating_margin_fraction
 * src/general_ice_model_data_module.f90:784-865, twice the calving law,
 * remove_unconnected_shelves src/ice_thickness_module.f90:932-1044, the
 * mask_noice / shelf / PD / continental-shelf removals, then
 * update_general_ice_model_data) and time += dt.  Everything stays on the
 * device between steps.  do_apply_prescribed_retreat_mask (a time series
 * read from NetCDF) is not an option here.  ref may be NULL when no option
 * that reads it is on.
 * cfg->choice_ice_dynamics: 'none', 'SIA', 'SSA' or 'SIA/SSA' (DIVA needs
 * the predictor-corrector loop, as in the reference).  io holds the state at
 * st->time (geometry and its derived fields consistent, as after
 * update_general_ice_model_data); its tolerances are the C%DIVA_* values the
 * start-up phase scales.  SMB_year and BMB (ny,nx) are held fixed.  On
 * return: io's velocity outputs, the final thickness in Hi_end (ny,nx) and
 * the fields of geo that are not NULL; geo's input pointers are ignored. */
int  imau_run_ice_dynamics_direct(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg,
                                  const imau_thickness_cfg *tcfg, const imau_producer_cfg *pcfg,
                                  const imau_timestep_cfg *scfg, imau_ice_io *io, const imau_refgeo_io *ref,
                                  const double *SMB_year, const double *BMB, double *Hi_end,
                                  imau_producer_io *geo, imau_timestep_state *st, double t_end);

/* The predictor-corrector state of type_model_region / type_ice_model that
 * run_ice_dynamics_pc carries from one time step to the next
 * (src/ice_dynamics_module.f90:243-432).  At the start of a run
 * (IMAU_ICE_main_model.f90:478-541, ice_dynamics_module.f90:993-997):
 * pc_zeta = 1, pc_eta = pc_eta_prev = pc_epsilon, dt_crit_ice =
 * dt_crit_ice_prev = dt = dt_min, t_next = start_time_of_run, and the two
 * fields zero. */
typedef struct imau_pc_state {
  double pc_zeta, pc_eta, pc_eta_prev;     /* in / out */
  double dt_crit_ice, dt_crit_ice_prev;    /* in / out [yr] */
  double dt;                               /* in / out: region%dt, the last time step [yr] */
  double t_next;                           /* in / out: t_next_SIA / _SSA / _DIVA, the next velocity update */
  double *dHidt_Hn_un;                     /* (ny,nx) in / out: dHi/dt of the last predictor [m/yr] */
  double *dHi_dt_a;                        /* (ny,nx) in / out: the corrected rate the steps apply [m/yr] */
} imau_pc_state;

/* The same loop with choice_timestepping = 'pc' (run_ice_dynamics_pc,
 * src/ice_dynamics_module.f90:243-432), for every choice_ice_dynamics
 * including DIVA: when time reaches pc->t_next, the step dt_crit_ice from
 * calc_critical_timestep_adv (:726-762) and the truncation-error controller,
 * the predictor dHi/dt at (H_n, u_n) and H* = max(0, H + dt ((1 + zeta/2)
 * dHi/dt_n - zeta/2 dHi/dt_n-1)), the velocity solve at H*, dHi/dt at
 * (H*, u_n+1), the corrector H = max(0, H_n + dt/2 (both rates)) and
 * calc_pc_truncation_error (:763-816); every step then applies dHi_dt_a
 * through update_ice_thickness as the direct loop does.  st->t_next_SIA /
 * _SSA are not used; st->dt is the last step. */
int  imau_run_ice_dynamics_pc(imau_ctx *ctx, const imau_grid *grid, const imau_cfg *cfg,
                              const imau_thickness_cfg *tcfg, const imau_producer_cfg *pcfg,
                              const imau_timestep_cfg *scfg, imau_ice_io *io, const imau_refgeo_io *ref,
                              const double *SMB_year, const double *BMB, double *Hi_end, imau_producer_io *geo,
                              imau_timestep_state *st, imau_pc_state *pc, double t_end);

/* ---- ensembles (no reference counterpart: the reference solves one region
 *      grid per call, src/IMAU_ICE_program.f90:247-250) ----
 * Solves ngrids independent DIVA problems of possibly different sizes in one
 * batched device solve.  grids[g] / ios[g] describe member g; the
 * configuration is shared. */
int  imau_vel_solve_diva_batch(imau_ctx *ctx, int32_t ngrids, const imau_grid *grids,
                               const imau_cfg *cfg, imau_ice_io *ios);

#ifdef __cplusplus
}
#endif
#endif /* IMAU_VEL_H */
