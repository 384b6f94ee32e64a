!------------------------------------------------------------------------------
! MODULE solver_2d for the MI355X stage: the drop-in for the reference's
! src/solver_2d.f90.  It keeps every name the reference's other modules take
! from solver_2d -- the procedures allocate_solver_variables,
! deallocate_solver_variables, check_solve, timestep, imex_RK_solver and
! update_erosion_deposition_cell (src/IMEX_SfloW2D.f90:66-71), and the module
! data q, qp, t, dt, hmax, pdynmax, mod_vel_max, thck_table, pdyn_table,
! vuln_table, solve_mask, solve_cells, j_cent, k_cent (:98-105), hpos,
! hpos_old, solve_mask_time (src/inpout_2d.f90:5265, :7174), Z and
! conv_kernel (src/stochastic_module.f90:19-21), and verbose_level,
! comp_cells_x, comp_cells_y passed through -- so that init_2d.f90,
! stochastic_module.f90, inpout_2d.f90 and IMEX_SfloW2D.f90 compile against
! it unchanged; the time step itself runs on the GPU through
! imex_hip_binding (libimexhip.so).
!
! The state lives on the device.  The module arrays are host mirrors:
! check_solve(.TRUE.) of the driver's set-up (src/IMEX_SfloW2D.f90:265)
! uploads the static fields and q that init_2d filled, and
! solver_sync_host refreshes the mirrors from the device (before the
! outputs, src/IMEX_SfloW2D.f90:546-589).  The driver's epilogue loop
! (:449-505) becomes CALL solver_post_step, the one edit of the driver.
!
! Physics errors: the device returns the conditions on which the reference
! prints and pauses (src/solver_2d.f90:1382-1560) as bits; check() prints
! them and pauses with READ(*,*) as the reference does.  An API or HIP error
! prints imex_last_error and stops.
!------------------------------------------------------------------------------
MODULE solver_2d

  USE, INTRINSIC :: iso_c_binding
  USE imex_hip_binding

  USE parameters_2d, ONLY : wp, n_vars, n_solid, verbose_level
  USE geometry_2d, ONLY : comp_cells_x, comp_cells_y

  IMPLICIT NONE

  REAL(wp) :: t                                !< time
  REAL(wp) :: dt                               !< time step
  REAL(wp), ALLOCATABLE, TARGET :: q(:,:,:)    !< conservative variables (host mirror)
  REAL(wp), ALLOCATABLE :: qp(:,:,:)           !< primitive variables (host mirror)
  LOGICAL, ALLOCATABLE :: hpos(:,:), hpos_old(:,:)
  REAL(wp), ALLOCATABLE, TARGET :: hmax(:,:), pdynmax(:,:), mod_vel_max(:,:)
  LOGICAL, ALLOCATABLE :: vuln_table(:,:,:)
  LOGICAL, ALLOCATABLE :: thck_table(:,:), pdyn_table(:,:)
  REAL(wp), ALLOCATABLE :: solve_mask_time(:,:)
  LOGICAL, ALLOCATABLE :: solve_mask(:,:)
  INTEGER :: solve_cells
  INTEGER, ALLOCATABLE :: j_cent(:), k_cent(:)
  REAL(wp), ALLOCATABLE :: Z(:,:)
  REAL(wp), ALLOCATABLE :: conv_kernel(:,:)

  !> the device run of this module
  TYPE(c_ptr), SAVE :: imex_h = c_null_ptr
  LOGICAL, SAVE :: uploaded = .FALSE.

  PRIVATE :: check, fill_config, upload

CONTAINS

  !----------------------------------------------------------------------------
  !> allocate_solver_variables (src/solver_2d.f90:264-541): the host mirrors
  !> as the reference allocates them, and the device run from the namelist
  !> values read_param left in the modules.
  !----------------------------------------------------------------------------
  SUBROUTINE allocate_solver_variables

    USE parameters_2d, ONLY : n_thickness_levels, n_dyn_pres_levels
    USE constitutive_2d, ONLY : T_ambient

    TYPE(imex_config) :: cfg
    INTEGER(c_int) :: rc

    ALLOCATE( q(n_vars, comp_cells_x, comp_cells_y) )
    ALLOCATE( qp(n_vars+2, comp_cells_x, comp_cells_y) )
    ALLOCATE( hpos(comp_cells_x, comp_cells_y), hpos_old(comp_cells_x, comp_cells_y) )
    q = 0.0_wp
    qp = 0.0_wp
    qp(4,:,:) = T_ambient
    ALLOCATE( hmax(comp_cells_x, comp_cells_y), pdynmax(comp_cells_x, comp_cells_y) )
    ALLOCATE( mod_vel_max(comp_cells_x, comp_cells_y) )
    ALLOCATE( vuln_table(n_thickness_levels * n_dyn_pres_levels, comp_cells_x, comp_cells_y) )
    ALLOCATE( thck_table(comp_cells_x, comp_cells_y), pdyn_table(comp_cells_x, comp_cells_y) )
    ALLOCATE( solve_mask_time(comp_cells_x, comp_cells_y) )
    ALLOCATE( solve_mask(comp_cells_x, comp_cells_y) )
    ALLOCATE( j_cent(comp_cells_x * comp_cells_y), k_cent(comp_cells_x * comp_cells_y) )
    ALLOCATE( Z(comp_cells_x, comp_cells_y) )
    hmax = 0.0_wp
    pdynmax = 0.0_wp
    mod_vel_max = 0.0_wp
    vuln_table = .FALSE.
    solve_mask_time = 0.0_wp
    solve_mask = .FALSE.
    solve_cells = 0
    Z = 0.0_wp

    CALL fill_config(cfg)
    rc = imex_create(cfg, imex_h)
    CALL check(rc, 'imex_create')

  END SUBROUTINE allocate_solver_variables

  !----------------------------------------------------------------------------
  !> deallocate_solver_variables (src/solver_2d.f90:548-650)
  !----------------------------------------------------------------------------
  SUBROUTINE deallocate_solver_variables

    IF ( C_ASSOCIATED(imex_h) ) CALL imex_destroy(imex_h)
    imex_h = c_null_ptr
    uploaded = .FALSE.
    DEALLOCATE( q, qp, hpos, hpos_old, hmax, pdynmax, mod_vel_max, vuln_table )
    DEALLOCATE( thck_table, pdyn_table, solve_mask_time, solve_mask, j_cent, k_cent, Z )
    IF ( ALLOCATED(conv_kernel) ) DEALLOCATE( conv_kernel )

  END SUBROUTINE deallocate_solver_variables

  !----------------------------------------------------------------------------
  !> check_solve (src/solver_2d.f90:667-858).  The first call, the driver's
  !> check_solve(.TRUE.) after the initial condition and the source set-up
  !> (src/IMEX_SfloW2D.f90:265), hands the static fields, q, t and dt0 to the
  !> device; the active set itself is built there.
  !----------------------------------------------------------------------------
  SUBROUTINE check_solve(solve_all)

    LOGICAL, INTENT(IN) :: solve_all
    INTEGER(c_int) :: rc

    IF ( .NOT. uploaded ) THEN
       CALL upload
       RETURN        ! the upload runs check_solve(.TRUE.) and the first epilogue
    END IF
    rc = imex_check_solve(imex_h, MERGE(1_c_int, 0_c_int, solve_all))
    CALL check(rc, 'imex_check_solve')

  END SUBROUTINE check_solve

  !----------------------------------------------------------------------------
  !> timestep (src/solver_2d.f90:873-1006): the CFL dt into the module's dt
  !----------------------------------------------------------------------------
  SUBROUTINE timestep

    INTEGER(c_int) :: rc
    REAL(c_double) :: dt_c

    rc = imex_timestep(imex_h, REAL(t, c_double), dt_c)
    CALL check(rc, 'imex_timestep')
    dt = dt_c

  END SUBROUTINE timestep

  !----------------------------------------------------------------------------
  !> imex_RK_solver (src/solver_2d.f90:1022-1634) with the driver's dt
  !----------------------------------------------------------------------------
  SUBROUTINE imex_RK_solver

    INTEGER(c_int) :: rc

    rc = imex_rk_step(imex_h, REAL(dt, c_double))
    CALL check(rc, 'imex_rk_step')

  END SUBROUTINE imex_RK_solver

  !----------------------------------------------------------------------------
  !> update_erosion_deposition_cell (src/solver_2d.f90:2382-2597)
  !----------------------------------------------------------------------------
  SUBROUTINE update_erosion_deposition_cell(dt)

    REAL(wp), INTENT(IN) :: dt
    INTEGER(c_int) :: rc

    rc = imex_erosion_deposition(imex_h, REAL(dt, c_double))
    CALL check(rc, 'imex_erosion_deposition')

  END SUBROUTINE update_erosion_deposition_cell

  !----------------------------------------------------------------------------
  !> The driver's epilogue (src/IMEX_SfloW2D.f90:449-505: qp of the active
  !> cells, the running maxima and the vulnerability table) on the device,
  !> with the stdout diagnostics of :507-544 in stats.
  !----------------------------------------------------------------------------
  SUBROUTINE solver_post_step(stats)

    TYPE(imex_step_stats), INTENT(OUT), TARGET :: stats
    INTEGER(c_int) :: rc

    rc = imex_post_step(imex_h, REAL(t, c_double), C_LOC(stats))
    CALL check(rc, 'imex_post_step')

  END SUBROUTINE solver_post_step

  !----------------------------------------------------------------------------
  !> The host mirrors from the device: q, hmax, pdynmax, mod_vel_max,
  !> vuln_table, the deposit of geometry_2d, solve_mask with its cell lists
  !> (k outer, j inner, src/solver_2d.f90:788-854) and qp as the driver
  !> leaves it (qc_to_qp where q(1) > 0, else zero with T_ambient).
  !----------------------------------------------------------------------------
  SUBROUTINE solver_sync_host

    USE geometry_2d, ONLY : deposit
    USE constitutive_2d, ONLY : qc_to_qp, T_ambient

    INTEGER(c_int32_t), ALLOCATABLE, TARGET :: vuln_i(:,:,:), mask_i(:,:)
    REAL(wp), ALLOCATABLE, TARGET :: dep(:,:,:)
    INTEGER(c_int) :: rc, what
    INTEGER :: j, k
    REAL(wp) :: p_dyn

    ALLOCATE( vuln_i(MAX(1, SIZE(vuln_table, 1)), comp_cells_x, comp_cells_y) )
    ALLOCATE( mask_i(comp_cells_x, comp_cells_y) )
    ALLOCATE( dep(comp_cells_x, comp_cells_y, n_solid) )
    what = IOR(IOR(IMEX_DL_Q, IMEX_DL_STATS), IOR(IMEX_DL_DEPOSIT, IMEX_DL_MASK))
    IF ( SIZE(vuln_table, 1) > 0 ) what = IOR(what, IMEX_DL_VULN)
    rc = imex_download(imex_h, what, C_LOC(q), C_LOC(hmax), C_LOC(pdynmax),             &
         C_LOC(mod_vel_max), C_LOC(dep), C_LOC(vuln_i), C_LOC(mask_i))
    CALL check(rc, 'imex_download')
    IF ( SIZE(vuln_table, 1) > 0 ) vuln_table = ( vuln_i /= 0 )
    IF ( ALLOCATED(deposit) ) deposit = dep
    solve_mask = ( mask_i /= 0 )
    solve_cells = 0
    DO k = 1, comp_cells_y
       DO j = 1, comp_cells_x
          IF ( solve_mask(j,k) ) THEN
             solve_cells = solve_cells + 1
             j_cent(solve_cells) = j
             k_cent(solve_cells) = k
          END IF
          IF ( q(1,j,k) .GT. 0.0_wp ) THEN
             CALL qc_to_qp( q(1:n_vars,j,k), qp(1:n_vars+2,j,k), p_dyn )
          ELSE
             qp(1:n_vars+2,j,k) = 0.0_wp
             qp(4,j,k) = T_ambient
          END IF
       END DO
    END DO

  END SUBROUTINE solver_sync_host

  !----------------------------------------------------------------------------
  ! the namelist values of the configured path, from the modules read_param
  ! filled (src/inpout_2d.f90:986-5134)
  !----------------------------------------------------------------------------
  SUBROUTINE fill_config(c)

    USE parameters_2d, ONLY : energy_flag, radial_source_flag, n_thickness_levels,       &
         n_dyn_pres_levels, thickness_levels, dyn_pres_levels, cfl, max_dt, vel_source,   &
         T_source, h_source, time_param, alphas_source, erodible_deposit_flag,            &
         topo_change_flag, rheology_model, stochastic_flag, pore_pressure_flag,           &
         stoch_transport_flag, gas_loss_flag, bcW, bcE, bcS, bcN
    USE constitutive_2d, ONLY : settling_flag, entrainment_flag, grav, mu, xi, pres,       &
         T_ambient, sp_heat_a, sp_gas_const_a, kin_visc_a, rho_s, diam_s, sp_heat_s,      &
         erosion_coeff, erodible_porosity, T_erodible, erodible_fract, mu_s, mu_2,        &
         muI_inf, I_0, f_inhibit_mode, N_inh, alpha_trans, hydraulic_permeability,        &
         dynamic_permeability_flag, maximum_solid_packing, sphericity_s
    USE geometry_2d, ONLY : dx, x0, y0

    TYPE(imex_config), INTENT(OUT) :: c
    INTEGER :: i

    c%nx = comp_cells_x
    c%ny = comp_cells_y
    c%n_solid = n_solid
    c%energy_flag = MERGE(1, 0, energy_flag)
    c%settling_flag = MERGE(1, 0, settling_flag)
    c%entrainment_flag = MERGE(1, 0, entrainment_flag)
    c%radial_source_flag = MERGE(1, 0, radial_source_flag)
    c%n_thickness_levels = n_thickness_levels
    c%n_dyn_pres_levels = n_dyn_pres_levels
    c%precision = 0
    c%dx = dx
    c%x0 = x0
    c%y0 = y0
    c%cfl = cfl
    c%max_dt = max_dt
    c%grav = grav
    c%mu = mu
    c%xi = xi
    c%pres = pres
    c%T_ambient = T_ambient
    c%sp_heat_a = sp_heat_a
    c%sp_gas_const_a = sp_gas_const_a
    c%kin_visc_a = kin_visc_a
    c%rho_s = 0.0_c_double
    c%diam_s = 0.0_c_double
    c%sp_heat_s = 0.0_c_double
    c%alphas_source = 0.0_c_double
    c%erodible_fract = 0.0_c_double
    DO i = 1, MIN(n_solid, 3)
       c%rho_s(i) = rho_s(i)
       c%diam_s(i) = diam_s(i)
       c%sp_heat_s(i) = sp_heat_s(i)
       c%alphas_source(i) = alphas_source(i)
       IF ( ALLOCATED(erodible_fract) ) c%erodible_fract(i) = erodible_fract(i)
    END DO
    c%vel_source = vel_source
    c%T_source = T_source
    c%h_source = h_source
    c%time_param = time_param
    c%thickness_levels = 0.0_c_double
    c%dyn_pres_levels = 0.0_c_double
    DO i = 1, MIN(n_thickness_levels, 8)
       c%thickness_levels(i) = thickness_levels(i)
    END DO
    DO i = 1, MIN(n_dyn_pres_levels, 8)
       c%dyn_pres_levels(i) = dyn_pres_levels(i)
    END DO
    c%erosion_coeff = 0.0_c_double
    IF ( ALLOCATED(erosion_coeff) ) c%erosion_coeff = erosion_coeff
    c%erodible_porosity = erodible_porosity
    c%T_erodible = T_erodible
    c%erodible_deposit_flag = MERGE(1, 0, erodible_deposit_flag)
    c%topo_change_flag = MERGE(1, 0, topo_change_flag)
    c%rheology_model = rheology_model
    c%mu_s = mu_s
    c%mu_2 = mu_2
    c%muI_inf = muI_inf
    c%I_0 = I_0
    c%stochastic_flag = MERGE(1, 0, stochastic_flag)
    ! pore pressure and the transported noise (the general path)
    c%pore_pressure_flag = MERGE(1, 0, pore_pressure_flag)
    c%stoch_transport_flag = MERGE(1, 0, stoch_transport_flag)
    c%gas_loss_flag = 0
    c%f_inhibit_mode = 0
    c%N_inh = 0
    c%dynamic_permeability_flag = 0
    c%hydraulic_permeability = 0.0_c_double
    c%alpha_trans = 0.0_c_double
    c%maximum_solid_packing = 0.0_c_double
    c%sphericity_s = 1.0_c_double
    IF ( pore_pressure_flag ) THEN
       c%gas_loss_flag = MERGE(1, 0, gas_loss_flag)
       SELECT CASE ( TRIM(f_inhibit_mode) )
       CASE ('STEP')
          c%f_inhibit_mode = 1
       CASE ('STATIC')
          c%f_inhibit_mode = 2
       CASE ('DYNAMIC')
          c%f_inhibit_mode = 3
       CASE DEFAULT
          c%f_inhibit_mode = 0
       END SELECT
       c%N_inh = N_inh
       c%dynamic_permeability_flag = MERGE(1, 0, dynamic_permeability_flag)
       c%hydraulic_permeability = hydraulic_permeability
       c%alpha_trans = alpha_trans
       c%maximum_solid_packing = maximum_solid_packing
       IF ( ALLOCATED(sphericity_s) ) THEN
          DO i = 1, MIN(n_solid, 3)
             c%sphericity_s(i) = sphericity_s(i)
          END DO
       END IF
    END IF
    ! the per-variable boundary conditions the namelists set
    ! (src/inpout_2d.f90:2495-3123); the general path takes them
    c%bc_custom = 0
    c%bc_flag = 1
    c%bc_value = 0.0_c_double
    IF ( ( pore_pressure_flag .OR. stoch_transport_flag ) .AND. ALLOCATED(bcW) ) THEN
       c%bc_custom = 1
       DO i = 1, MIN(n_vars, 8)
          c%bc_flag(i) = bcW(i)%flag
          c%bc_flag(8 + i) = bcE(i)%flag
          c%bc_flag(16 + i) = bcS(i)%flag
          c%bc_flag(24 + i) = bcN(i)%flag
          c%bc_value(i) = bcW(i)%value
          c%bc_value(8 + i) = bcE(i)%value
          c%bc_value(16 + i) = bcS(i)%value
          c%bc_value(24 + i) = bcN(i)%value
       END DO
    END IF

  END SUBROUTINE fill_config

  !----------------------------------------------------------------------------
  ! the static fields of geometry_2d, the erodible layer, the noise, B_cent
  ! and q, t, dt0 to the device (the upload runs check_solve(.TRUE.) and the
  ! initial epilogue, src/IMEX_SfloW2D.f90:265, :306-353)
  !----------------------------------------------------------------------------
  SUBROUTINE upload

    USE parameters_2d, ONLY : dt0, radial_source_flag, stochastic_flag
    USE geometry_2d, ONLY : B_cent, B_prime_x_geom, B_prime_y_geom, source_cell,          &
         sourceE, sourceW, sourceN, sourceS, cell_arc_perim, cell_arc_n_x, cell_arc_n_y, &
         erodible
    USE constitutive_2d, ONLY : erosion_coeff

    TYPE(imex_static) :: s
    REAL(c_double), ALLOCATABLE, TARGET :: bx(:,:), by(:,:), ap(:,:), ax(:,:), ay(:,:)
    INTEGER(c_int32_t), ALLOCATABLE, TARGET :: sc(:,:), sd(:,:)
    INTEGER(c_int) :: rc

    ALLOCATE( bx(comp_cells_x, comp_cells_y), by(comp_cells_x, comp_cells_y) )
    bx = B_prime_x_geom
    by = B_prime_y_geom
    s%B_prime_x = C_LOC(bx)
    s%B_prime_y = C_LOC(by)
    s%source_cell = c_null_ptr
    s%source_sides = c_null_ptr
    s%arc_perim = c_null_ptr
    s%arc_n_x = c_null_ptr
    s%arc_n_y = c_null_ptr
    IF ( radial_source_flag ) THEN
       ALLOCATE( sc(comp_cells_x, comp_cells_y), sd(comp_cells_x, comp_cells_y) )
       ALLOCATE( ap(comp_cells_x, comp_cells_y), ax(comp_cells_x, comp_cells_y) )
       ALLOCATE( ay(comp_cells_x, comp_cells_y) )
       sc = source_cell
       sd = MERGE(1, 0, sourceE) + MERGE(2, 0, sourceW) + MERGE(4, 0, sourceN)            &
            + MERGE(8, 0, sourceS)
       ap = cell_arc_perim
       ax = cell_arc_n_x
       ay = cell_arc_n_y
       s%source_cell = C_LOC(sc)
       s%source_sides = C_LOC(sd)
       s%arc_perim = C_LOC(ap)
       s%arc_n_x = C_LOC(ax)
       s%arc_n_y = C_LOC(ay)
    END IF
    rc = imex_set_static(imex_h, s)
    CALL check(rc, 'imex_set_static')
    rc = imex_set_topography(imex_h, B_cent)
    CALL check(rc, 'imex_set_topography')
    IF ( ALLOCATED(erosion_coeff) ) THEN
       ! erodible(n_solid, nx, ny) (src/inpout_2d.f90:1616): the order the call takes
       IF ( erosion_coeff > 0.0_wp .AND. ALLOCATED(erodible) ) THEN
          rc = imex_set_erodible(imex_h, erodible)
          CALL check(rc, 'imex_set_erodible')
       END IF
    END IF
    IF ( stochastic_flag ) THEN
       rc = imex_set_noise(imex_h, Z)
       CALL check(rc, 'imex_set_noise')
    END IF
    rc = imex_upload_state(imex_h, q, REAL(t, c_double), REAL(dt0, c_double))
    CALL check(rc, 'imex_upload_state')
    uploaded = .TRUE.

  END SUBROUTINE upload

  !----------------------------------------------------------------------------
  ! rc > 0: the physics error bits, printed with the pause of the reference;
  ! rc < 0: imex_last_error, then STOP
  !----------------------------------------------------------------------------
  SUBROUTINE check(rc, what)

    INTEGER(c_int), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: what

    IF ( rc < 0 ) THEN
       WRITE(*,*) TRIM(what), ' failed (', rc, '): ', TRIM(imex_error_text(imex_h))
       STOP 1
    ELSE IF ( rc > 0 ) THEN
       IF ( IAND(rc, IMEX_ERR_NAN) /= 0 ) WRITE(*,*) TRIM(what), ': NaN in the state'
       IF ( IAND(rc, IMEX_ERR_NEG_H) /= 0 ) WRITE(*,*) TRIM(what), ': negative thickness'
       IF ( IAND(rc, IMEX_ERR_NEG_SOLID) /= 0 ) WRITE(*,*) TRIM(what), ': negative solid mass'
       IF ( IAND(rc, IMEX_ERR_COLD) /= 0 ) WRITE(*,*) TRIM(what), ': T < 273 K'
       IF ( IAND(rc, IMEX_ERR_NEG_DEP) /= 0 ) WRITE(*,*) TRIM(what), ': negative deposition'
       IF ( IAND(rc, IMEX_ERR_SOLID_GT_TOT) /= 0 ) WRITE(*,*) TRIM(what),                 &
            ': solid mass above total mass'
       READ(*,*)
    END IF

  END SUBROUTINE check

END MODULE solver_2d
