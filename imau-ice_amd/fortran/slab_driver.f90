! slab_driver.f90 -- Fortran host exercising libimauvel.so through
! imau_vel_bindings exactly as a Fortran ice_velocity_module would:
! Fortran-owned column-major arrays, C_LOC pointers, BIND(C) structs.
!
! Uniform grounded slab (8x8, 10 km, H = 2000 m, surface slope 5e-3,
! A = 1e-16, Weertman beta_sq = 1e4): the DIVA basal stress must equal the
! driving stress rho g H |s'| and the velocity must be uniform.
PROGRAM slab_driver
  USE ISO_C_BINDING
  USE imau_vel_bindings
  IMPLICIT NONE
  INTEGER, PARAMETER :: nx = 8, ny = 8
  REAL(C_DOUBLE), PARAMETER :: dx = 10000._C_DOUBLE, H = 2000._C_DOUBLE, slope = 5e-3_C_DOUBLE, A = 1e-16_C_DOUBLE
  REAL(C_DOUBLE), TARGET :: x(nx), y(ny)
  REAL(C_DOUBLE), TARGET :: Hi(ny,nx), Hb(ny,nx), SL(ny,nx), Hs(ny,nx), TAF(ny,nx), zero_a(ny,nx), Avav(ny,nx)
  REAL(C_DOUBLE), TARGET :: dHs_dx_a(ny,nx), beta_sq(ny,nx), alpha_sq(ny,nx), phi(ny,nx), tauc(ny,nx)
  INTEGER(C_INT32_T), TARGET :: m_land(ny,nx), m_ocean(ny,nx), m_ice(ny,nx), m_sheet(ny,nx), m_shelf(ny,nx)
  REAL(C_DOUBLE), TARGET :: Hi_cx(ny,nx-1), sx_cx(ny,nx-1), sy_cx(ny,nx-1), u(ny,nx-1), tbx(ny,nx-1)
  REAL(C_DOUBLE), TARGET :: Hi_cy(ny-1,nx), sx_cy(ny-1,nx), sy_cy(ny-1,nx), v(ny-1,nx), tby(ny-1,nx)
  REAL(C_DOUBLE), TARGET :: Hi_b(ny-1,nx-1)
  REAL(C_DOUBLE), TARGET, ALLOCATABLE :: A3(:,:,:), visc(:,:,:)
  TYPE(imau_grid) :: grid
  TYPE(imau_cfg) :: cfg
  TYPE(imau_ice_io) :: io
  TYPE(C_PTR) :: ctx
  INTEGER :: i, j, rc
  REAL(C_DOUBLE) :: tau

  CALL imau_vel_default_config(cfg)
  cfg%choice_ice_dynamics = IMAU_DYN_DIVA
  cfg%choice_sliding_law = IMAU_SLID_WEERTMAN
  cfg%choice_basal_roughness = IMAU_ROUGH_UNIFORM
  cfg%choice_basal_hydrology = IMAU_HYDRO_DRY
  cfg%include_basal_freezing = 0
  cfg%stop_rule = IMAU_STOP_RESIDUAL
  cfg%nonlinear_rtol = 1e-10_C_DOUBLE
  cfg%DIVA_visc_it_nit = 500

  DO i = 1, nx
    x(i) = (i-1) * dx
  END DO
  DO j = 1, ny
    y(j) = (j-1) * dx
  END DO
  DO i = 1, nx
    DO j = 1, ny
      Hb(j,i) = -slope * x(i)
    END DO
  END DO
  Hi = H; SL = -1e4_C_DOUBLE; Hs = Hi + Hb; TAF = Hi; zero_a = 0._C_DOUBLE; Avav = A
  dHs_dx_a = -slope; beta_sq = 1e4_C_DOUBLE; alpha_sq = 0.5_C_DOUBLE; phi = 15._C_DOUBLE; tauc = 0._C_DOUBLE
  m_land = 1; m_ocean = 0; m_ice = 1; m_sheet = 1; m_shelf = 0
  Hi_cx = H; sx_cx = -slope; sy_cx = 0._C_DOUBLE; u = 0._C_DOUBLE; tbx = 0._C_DOUBLE
  Hi_cy = H; sx_cy = -slope; sy_cy = 0._C_DOUBLE; v = 0._C_DOUBLE; tby = 0._C_DOUBLE
  Hi_b = H
  ALLOCATE(A3(cfg%nz,ny,nx), visc(cfg%nz,ny,nx))
  A3 = A; visc = 0._C_DOUBLE

  grid%nx = nx; grid%ny = ny; grid%dx = dx; grid%x = C_LOC(x); grid%y = C_LOC(y)
  io%Hi_a = C_LOC(Hi); io%Hb_a = C_LOC(Hb); io%SL_a = C_LOC(SL); io%Hs_a = C_LOC(Hs); io%TAF_a = C_LOC(TAF)
  io%dHs_dx_a = C_LOC(dHs_dx_a); io%dHs_dy_a = C_LOC(zero_a); io%dHi_dx_a = C_LOC(zero_a); io%dHi_dy_a = C_LOC(zero_a)
  io%A_flow_vav_a = C_LOC(Avav)
  io%mask_land_a = C_LOC(m_land); io%mask_ocean_a = C_LOC(m_ocean); io%mask_ice_a = C_LOC(m_ice)
  io%mask_sheet_a = C_LOC(m_sheet); io%mask_shelf_a = C_LOC(m_shelf)
  io%Hi_cx = C_LOC(Hi_cx); io%dHs_dx_cx = C_LOC(sx_cx); io%dHs_dy_cx = C_LOC(sy_cx)
  io%Hi_cy = C_LOC(Hi_cy); io%dHs_dx_cy = C_LOC(sx_cy); io%dHs_dy_cy = C_LOC(sy_cy)
  io%Hi_b = C_LOC(Hi_b); io%A_flow_3D_a = C_LOC(A3)
  io%beta_sq_a = C_LOC(beta_sq); io%alpha_sq_a = C_LOC(alpha_sq); io%phi_fric_a = C_LOC(phi); io%tauc_a = C_LOC(tauc)
  io%u_vav_cx = C_LOC(u); io%v_vav_cy = C_LOC(v); io%taub_cx = C_LOC(tbx); io%taub_cy = C_LOC(tby)
  io%visc_eff_3D_a = C_LOC(visc)
  io%DIVA_SOR_nit = 100000; io%DIVA_SOR_tol = 1e-9_C_DOUBLE; io%DIVA_SOR_omega = 1.3_C_DOUBLE
  io%DIVA_PETSc_rtol = 0.01_C_DOUBLE; io%DIVA_PETSc_abstol = 0._C_DOUBLE

  rc = imau_vel_init(grid, cfg, ctx)
  IF (rc /= IMAU_OK) THEN
    WRITE(*,*) 'imau_vel_init failed: ', TRIM(imau_vel_error_text(C_NULL_PTR))
    STOP 2
  END IF
  rc = imau_vel_solve_diva(ctx, grid, cfg, io)
  IF (rc /= IMAU_OK) THEN
    WRITE(*,*) 'imau_vel_solve_diva failed: ', TRIM(imau_vel_error_text(ctx))
    STOP 3
  END IF
  tau = 910._C_DOUBLE * 9.81_C_DOUBLE * H * slope
  WRITE(*,'(A,ES24.16)') 'u_vav = ', u(4,4)
  WRITE(*,'(A,ES24.16)') 'taub  = ', tbx(4,4)
  WRITE(*,'(A,ES24.16)') 'tau_d = ', tau
  WRITE(*,'(A,I6)')      'picard = ', io%stat_picard_iterations
  CALL imau_vel_finalize(ctx)
  IF (ABS(tbx(4,4) - tau) > 1e-6_C_DOUBLE * tau) STOP 4
  IF (MAXVAL(u) - MINVAL(u) > 1e-8_C_DOUBLE * MAXVAL(ABS(u))) STOP 5
  IF (MAXVAL(ABS(v)) > 1e-8_C_DOUBLE * MAXVAL(ABS(u))) STOP 6
  WRITE(*,*) 'slab_driver OK'
END PROGRAM slab_driver
