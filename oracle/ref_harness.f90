!------------------------------------------------------------------------------
! Test harness around the reference IMEX_SfloW2D solver modules.
!
! TEST INFRASTRUCTURE ONLY.  This file is compiled together with the unmodified
! reference sources (parameters_2d, complexify, geometry_2d, constitutive_2d,
! solver_2d, init_2d from src/) into oracle/_ref/libimexref.so by
! oracle/Makefile.  Only tests/ and bench.py's cpu_baseline leg load it, as the
! checker; no product code path calls it.
!
! It replaces what the reference driver does for the configured path:
!   - read_param (src/inpout_2d.f90:986-5134): the harness sets the module
!     variables a namelist read would set for the gas-carrier / Voellmy / KT /
!     minmod / IMEX-RK2 configuration (derived constants as at
!     src/inpout_2d.f90:1356, :1379, :1455-2049, :2070-2080, :2271-2272);
!   - the DEM reader: the DEM grid is laid on the cell centres, so the bilinear
!     interpolation of init_grid (src/geometry_2d.f90:336-362) returns the
!     pixel values exactly;
!   - the time loop of PROGRAM IMEX_SfloW2D (src/IMEX_SfloW2D.f90:395-505)
!     with no output inside the loop and a fixed step count.
!------------------------------------------------------------------------------
MODULE imexref_harness

  USE, INTRINSIC :: iso_c_binding
  USE parameters_2d, ONLY : wp, n_solid, n_add_gas, n_stoch_vars, n_pore_vars,  &
       n_vars, n_eqns, n_RK, idx_h, idx_hu, idx_hv, idx_T, idx_alfas_first,      &
       idx_alfas_last, idx_addGas_first, idx_addGas_last, idx_stoch, idx_pore,   &
       idx_u, idx_v, idx_totMassEqn, idx_uEqn, idx_vEqn, idx_engyEqn,            &
       idx_solidEqn_first, idx_solidEqn_last, idx_addGasEqn_first,               &
       idx_addGasEqn_last, idx_stochEqn, idx_poreEqn, verbose_level,             &
       energy_flag, rheology_flag, rheology_model, liquid_flag, gas_flag,        &
       alpha_flag, slope_correction_flag, curvature_term_flag,                   &
       vertical_profiles_flag, stochastic_flag, stoch_transport_flag,            &
       length_spatial_corr,                                                      &
       pore_pressure_flag, sutherland_flag, topo_change_flag,                    &
       radial_source_flag, lateral_source_flag, bottom_radial_source_flag,       &
       erodible_deposit_flag, liquid_vaporization_flag, mean_field_flag,         &
       gas_loss_flag, collapsing_volume_flag, limiter, solver_scheme, cfl, dt0,  &
       max_dt, reconstr_coeff, theta, interfaces_relaxation, bcW, bcE, bcS, bcN, &
       x_collapse, y_collapse, r_collapse, h_collapse, T_collapse,               &
       alphas_collapse, alphag_collapse, x_source, y_source, r_source,           &
       r2_source, angle_source, vel_source, T_source, h_source, alphas_source,   &
       alphag_source, alphal_source, time_param, azimuth_source,                 &
       arc_width_source, mfr_source, Ri_source, thickness_levels,                &
       dyn_pres_levels, n_thickness_levels, n_dyn_pres_levels, t_start, t_end,   &
       n_intervals, bottom_conc_flag, eps_sing
  USE constitutive_2d, ONLY : grav, inv_grav, mu, xi, pres, inv_pres, T_ambient, &
       sp_heat_a, sp_gas_const_a, kin_visc_a, kin_visc_c, rho_a_amb, rho_s,      &
       inv_rho_s, c_inv_rho_s, diam_s, sp_heat_s, settling_flag,                 &
       entrainment_flag, erosion_coeff, erodible_porosity, erodible_fract,       &
       T_erodible, rho_c_sub, coeff_porosity, alphastot_min,                     &
       maximum_solid_packing, collective_settling_flag, sp_heat_g,               &
       sp_gas_const_g, loss_rate, init_problem_param, qc_to_qp, qp_to_qc,        &
       mixt_var, eval_fluxes, eval_local_speeds_x, eval_local_speeds_y,          &
       eval_expl_terms, eval_nh_semi_impl_terms, eval_implicit_terms,            &
       settling_velocity, eval_mass_exchange_terms, sphericity_s, mu_s, mu_2,   &
       muI_inf, I_0, hydraulic_permeability, dynamic_permeability_flag,          &
       alpha_trans, N_inh, f_inhibit_mode, pascal_coeff_precomputed,             &
       precompute_pascal_coefficient
  USE geometry_2d, ONLY : comp_cells_x, comp_cells_y, x0, y0, cell_size,        &
       topography_profile, nodata_topo, n_topography_profile_x,                  &
       n_topography_profile_y, init_grid, init_source, deposit, erosion,         &
       erodible, B_cent, B_prime_x_geom, B_prime_y_geom, source_cell, sourceE,   &
       sourceW, sourceN, sourceS, cell_arc_perim, cell_arc_n_x, cell_arc_n_y,    &
       x_comp, x_stag, y_comp, y_stag, dx, dy, pi_g, topography_reconstruction
  USE solver_2d, ONLY : allocate_solver_variables, check_solve, timestep,       &
       imex_RK_solver, update_erosion_deposition_cell, solve_rk_step, q, qp, t,  &
       dt, hmax, q_rk, divFlux, NH, SI_NH, expl_terms,                           &
       pdynmax, mod_vel_max, vuln_table, thck_table, pdyn_table, solve_mask,     &
       solve_cells, j_cent, k_cent, Z
  USE init_2d, ONLY : collapsing_volume

  IMPLICIT NONE

  REAL(wp) :: dt_old, dt_old_old

CONTAINS

  !----------------------------------------------------------------------------
  ! Set up the reference modules for one run.
  !   ip(0:54)  integer parameters, rp(0:110) real parameters (layout in
  !   oracle/refpy.py), bdem(nx*ny) bed elevation at the cell centres
  !   (x index fastest).
  !----------------------------------------------------------------------------
  INTEGER(c_int) FUNCTION imexref_setup(ip, rp, bdem) BIND(C, name='imexref_setup')
    INTEGER(c_int), INTENT(IN) :: ip(0:54)
    REAL(c_double), INTENT(IN) :: rp(0:110)
    REAL(c_double), INTENT(IN) :: bdem(*)

    INTEGER :: j, k, i, nx, ny
    REAL(wp) :: h_ell
    REAL(wp) :: source_length, Ri, rho_m, rho_c, red_grav, sp_heat_c,           &
         sp_heat_mix
    REAL(wp), ALLOCATABLE :: qp_src(:)
    REAL(wp) :: p_dyn, r_u, r_v, mod_vel2, mod_vel
    INTEGER :: l, i_table, i_thk_lev, i_pdyn_lev

    nx = ip(0)
    ny = ip(1)

    verbose_level = -1

    ! ---- model switches for the configured path -------------------------
    n_solid = ip(2)
    n_add_gas = 0
    ! the transported noise (src/inpout_2d.f90:1466-1475): ip(54)
    n_stoch_vars = 0
    IF ( ip(54) .NE. 0 ) n_stoch_vars = 1
    ! pore pressure (PORE_PRESSURE_PARAMETERS): ip(16:20), rp(73:78)
    pore_pressure_flag = ( ip(16) .NE. 0 )
    n_pore_vars = 0
    IF ( pore_pressure_flag ) n_pore_vars = 1
    gas_flag = .TRUE.
    liquid_flag = .FALSE.
    rheology_flag = .TRUE.
    ! 1 = Voellmy-Salm, 11 = mu(I), 12 = mu(I) + Voellmy turbulent term
    rheology_model = ip(14)
    IF ( rheology_model .EQ. 0 ) rheology_model = 1
    energy_flag = ( ip(3) .NE. 0 )
    alpha_flag = .FALSE.
    slope_correction_flag = .FALSE.
    curvature_term_flag = .FALSE.
    vertical_profiles_flag = .FALSE.
    bottom_conc_flag = .FALSE.
    ! stochastic friction: Z is set by imexref_set_noise (0 after
    ! allocate_solver_variables); with transport the host gives q1 * Z too
    stochastic_flag = ( ip(15) .NE. 0 )
    length_spatial_corr = -1.0_wp
    stoch_transport_flag = ( ip(54) .NE. 0 )
    mean_field_flag = .FALSE.
    gas_loss_flag = ( ip(17) .NE. 0 )
    sutherland_flag = .FALSE.
    topo_change_flag = ( ip(13) .NE. 0 )
    lateral_source_flag = .FALSE.
    bottom_radial_source_flag = .FALSE.
    erodible_deposit_flag = ( ip(12) .NE. 0 )
    liquid_vaporization_flag = .FALSE.
    radial_source_flag = ( ip(6) .NE. 0 )
    collapsing_volume_flag = ( ip(7) .NE. 0 )
    n_intervals = 0

    ! ---- variable / equation index map (src/inpout_2d.f90:584-655,1455-2049)
    n_vars = 4 + n_solid + n_stoch_vars + n_pore_vars
    n_eqns = n_vars
    idx_h = 1
    idx_hu = 2
    idx_hv = 3
    idx_T = 4
    idx_alfas_first = 5
    idx_alfas_last = 4 + n_solid
    idx_addGas_first = 5 + n_solid
    idx_addGas_last = 4 + n_solid
    idx_stoch = 5 + n_solid
    idx_pore = 5 + n_solid + n_stoch_vars    ! src/inpout_2d.f90:654 (no added gas)
    idx_u = n_vars + 1
    idx_v = n_vars + 2
    idx_totMassEqn = 1
    idx_uEqn = 2
    idx_vEqn = 3
    idx_engyEqn = 4
    idx_solidEqn_first = 5
    idx_solidEqn_last = 4 + n_solid
    idx_addGasEqn_first = 5 + n_solid
    idx_addGasEqn_last = 4 + n_solid
    idx_stochEqn = idx_stoch
    idx_poreEqn = idx_pore
    n_RK = 2

    ! ---- numerics --------------------------------------------------------
    solver_scheme = 'KT'
    limiter = -1
    limiter(1:n_vars) = ip(10)
    limiter(n_vars+1) = limiter(2)
    limiter(n_vars+2) = limiter(3)
    reconstr_coeff = 1.0_wp
    theta = 1.0_wp
    interfaces_relaxation = .FALSE.
    cell_size = rp(0)
    x0 = rp(1)
    y0 = rp(2)
    cfl = rp(3)
    dt0 = rp(4)
    max_dt = rp(5)
    t_start = rp(6)
    t_end = rp(7)
    comp_cells_x = nx
    comp_cells_y = ny

    IF ( ALLOCATED(bcW) ) DEALLOCATE(bcW, bcE, bcS, bcN)
    ALLOCATE( bcW(n_vars), bcE(n_vars), bcS(n_vars), bcN(n_vars) )
    bcW(1:n_vars)%flag = ip(11)
    bcE(1:n_vars)%flag = ip(11)
    bcS(1:n_vars)%flag = ip(11)
    bcN(1:n_vars)%flag = ip(11)
    bcW(1:n_vars)%value = 0.0_wp
    bcE(1:n_vars)%value = 0.0_wp
    bcS(1:n_vars)%value = 0.0_wp
    bcN(1:n_vars)%value = 0.0_wp
    ! per-variable boundary conditions (the *_BOUNDARY_CONDITIONS namelists):
    ! ip(21) switches them on, flags ip(22 + 8 side + i - 1) and values
    ! rp(79 + 8 side + i - 1), side 0 W, 1 E, 2 S, 3 N
    IF ( ip(21) .NE. 0 ) THEN
       DO i = 1, n_vars
          bcW(i)%flag = ip(22 + i - 1)
          bcE(i)%flag = ip(30 + i - 1)
          bcS(i)%flag = ip(38 + i - 1)
          bcN(i)%flag = ip(46 + i - 1)
          bcW(i)%value = rp(79 + i - 1)
          bcE(i)%value = rp(87 + i - 1)
          bcS(i)%value = rp(95 + i - 1)
          bcN(i)%value = rp(103 + i - 1)
       END DO
    END IF

    ! ---- physics constants (src/inpout_2d.f90:485,1313,1356,1379) -------
    grav = rp(8)
    inv_grav = 1.0_wp / grav
    mu = rp(9)
    xi = rp(10)
    mu_s = rp(69)
    mu_2 = rp(70)
    muI_inf = rp(71)
    I_0 = rp(72)
    pres = rp(11)
    inv_pres = 1.0_wp / pres
    T_ambient = rp(12)
    sp_heat_a = rp(13)
    sp_gas_const_a = rp(14)
    kin_visc_a = rp(15)
    kin_visc_c = kin_visc_a
    rho_a_amb = pres / ( sp_gas_const_a * T_ambient )

    IF ( ALLOCATED(rho_s) ) DEALLOCATE( rho_s, inv_rho_s, c_inv_rho_s, diam_s,   &
         sp_heat_s, erodible_fract, sphericity_s )
    ALLOCATE( rho_s(n_solid), inv_rho_s(n_solid), c_inv_rho_s(n_solid),          &
         diam_s(n_solid), sp_heat_s(n_solid), erodible_fract(n_solid),           &
         sphericity_s(n_solid) )
    IF ( .NOT. ALLOCATED(sp_heat_g) ) ALLOCATE( sp_heat_g(0), sp_gas_const_g(0) )
    DO i = 1, n_solid
       rho_s(i) = rp(30+i)
       diam_s(i) = rp(33+i)
       sp_heat_s(i) = rp(36+i)
       alphas_collapse(i) = rp(39+i)
       alphas_source(i) = rp(42+i)
    END DO
    sphericity_s = 1.0_wp
    DO i = 1, n_solid
       IF ( rp(75+i) .GT. 0.0_wp ) sphericity_s(i) = rp(75+i)
    END DO
    inv_rho_s(1:n_solid) = 1.0_wp / rho_s(1:n_solid)
    DO i = 1, n_solid
       c_inv_rho_s(i) = CMPLX( inv_rho_s(i), 0.0_wp, wp )
    END DO
    alphag_collapse = 0.0_wp
    alphag_source = 0.0_wp
    alphal_source = 0.0_wp

    settling_flag = ( ip(4) .NE. 0 )
    entrainment_flag = ( ip(5) .NE. 0 )
    collective_settling_flag = .FALSE.
    IF ( .NOT. ALLOCATED(erosion_coeff) ) ALLOCATE( erosion_coeff )
    IF ( .NOT. ALLOCATED(loss_rate) ) ALLOCATE( loss_rate )
    erosion_coeff = rp(63)
    loss_rate = 0.0_wp
    IF ( erosion_coeff .EQ. 0.0_wp ) THEN
       ! src/inpout_2d.f90:1827-1832
       erodible_porosity = 0.0_wp
       erodible_fract(1:n_solid) = 1.0_wp / n_solid
       T_erodible = 300.0_wp
       coeff_porosity = 0.0_wp
    ELSE
       ! src/inpout_2d.f90:1835-1890 (fractions given summing to one)
       erodible_porosity = rp(64)
       coeff_porosity = erodible_porosity / ( 1.0_wp - erodible_porosity )
       IF ( n_solid .EQ. 1 ) THEN
          erodible_fract(1) = 1.0_wp
       ELSE
          erodible_fract(1:n_solid) = rp(66:65+n_solid)
       END IF
       T_erodible = rp(65)
    END IF
    rho_c_sub = pres / ( sp_gas_const_a * T_erodible )
    alphastot_min = 0.0_wp
    maximum_solid_packing = 0.64_wp
    IF ( rp(75) .GT. 0.0_wp ) maximum_solid_packing = rp(75)
    hydraulic_permeability = rp(73)
    alpha_trans = rp(74)
    dynamic_permeability_flag = ( ip(20) .NE. 0 )
    N_inh = ip(19)
    SELECT CASE ( ip(18) )
    CASE ( 1 )
       f_inhibit_mode = 'STEP'
    CASE ( 2 )
       f_inhibit_mode = 'STATIC'
    CASE ( 3 )
       f_inhibit_mode = 'DYNAMIC'
    CASE DEFAULT
       f_inhibit_mode = 'OFF'
    END SELECT
    ! src/inpout_2d.f90:688-696
    pascal_coeff_precomputed = .FALSE.
    IF ( pore_pressure_flag .AND. ( ip(18) .GE. 2 ) )                              &
         CALL precompute_pascal_coefficient(N_inh)

    ! ---- initial volume and source ---------------------------------------
    x_collapse = rp(16)
    y_collapse = rp(17)
    r_collapse = rp(18)
    h_collapse = rp(19)
    T_collapse = rp(20)
    x_source = rp(21)
    y_source = rp(22)
    r_source = rp(23)
    r2_source = r_source
    angle_source = 0.0_wp
    azimuth_source = 0.0_wp
    arc_width_source = 360.0_wp
    vel_source = rp(24)
    T_source = rp(25)
    h_source = rp(26)
    time_param(1:4) = rp(27:30)
    mfr_source = rp(62)
    Ri_source = -1.0_wp

    ! ---- vulnerability levels -------------------------------------------
    n_thickness_levels = ip(8)
    n_dyn_pres_levels = ip(9)
    IF ( ALLOCATED(thickness_levels) ) DEALLOCATE( thickness_levels )
    IF ( ALLOCATED(dyn_pres_levels) ) DEALLOCATE( dyn_pres_levels )
    ALLOCATE( thickness_levels(MAX(n_thickness_levels,0)) )
    ALLOCATE( dyn_pres_levels(MAX(n_dyn_pres_levels,0)) )
    DO i = 1, n_thickness_levels
       thickness_levels(i) = rp(45+i)
    END DO
    DO i = 1, n_dyn_pres_levels
       dyn_pres_levels(i) = rp(53+i)
    END DO

    ! ---- DEM on the cell centres ----------------------------------------
    n_topography_profile_x = nx
    n_topography_profile_y = ny
    nodata_topo = -9999.0_wp
    IF ( ALLOCATED(topography_profile) ) DEALLOCATE( topography_profile )
    ALLOCATE( topography_profile(3,nx,ny) )
    DO k = 1, ny
       DO j = 1, nx
          topography_profile(1,j,k) = x0 + ( DBLE(j) - 0.5_wp ) * cell_size
          topography_profile(2,j,k) = y0 + ( DBLE(k) - 0.5_wp ) * cell_size
          topography_profile(3,j,k) = bdem( (k-1)*nx + j )
       END DO
    END DO

    CALL init_grid

    ! erodible/deposit/erosion arrays (src/inpout_2d.f90:1616-1619, :2083-2093)
    ALLOCATE( erodible(n_solid, nx, ny) )
    erodible = 0.0_wp
    ALLOCATE( deposit(nx, ny, n_solid) )
    deposit = 0.0_wp
    ALLOCATE( erosion(nx, ny, n_solid) )
    erosion = 0.0_wp

    ! ---- radial source thickness from the mass flow rate
    !      (MFR and VEL given, RI not: src/inpout_2d.f90:4108-4125, :4435-4437,
    !      :4522-4531, :4620-4632, :4732-4747)
    IF ( radial_source_flag .AND. ( mfr_source .GT. 0.0_wp ) ) THEN
       h_ell = ( r_source - r2_source )**2 / ( r_source + r2_source )**2
       source_length = pi_g * ( r_source + r2_source ) * ( 1.0_wp + 0.25_wp *    &
            h_ell + h_ell**2 / 64.0_wp + h_ell**3 / 256.0_wp +                   &
            h_ell**4 * 25.0_wp / 16384.0_wp + h_ell**5 * 49.0_wp / 65536.0_wp )
       ALLOCATE( qp_src(n_vars+2) )
       qp_src(1) = 1.0_wp
       qp_src(2) = 1.0_wp * vel_source
       qp_src(3) = 0.0_wp
       qp_src(4) = T_source
       qp_src(5:4+n_solid) = alphas_source(1:n_solid) * 1.0_wp
       qp_src(n_vars+1) = vel_source
       qp_src(n_vars+2) = 0.0_wp
       CALL mixt_var( qp_src, Ri, rho_m, rho_c, red_grav, .FALSE., sp_heat_c,    &
            sp_heat_mix )
       h_source = mfr_source / ( source_length * rho_m * vel_source * 1.0_wp )
       DEALLOCATE( qp_src )
    END IF

    CALL init_problem_param
    CALL allocate_solver_variables
    CALL zero_fill_unset_arrays

    t = t_start
    IF ( collapsing_volume_flag ) THEN
       CALL collapsing_volume
    ELSE
       q = 0.0_wp
    END IF

    IF ( radial_source_flag ) CALL init_source

    CALL check_solve(.TRUE.)

    dt = dt0
    dt_old = dt0
    dt_old_old = dt0

    ! initial epilogue (src/IMEX_SfloW2D.f90:306-353); the statistics arrays
    ! are left unset by allocate_solver_variables and are zeroed here
    hmax = 0.0_wp
    pdynmax = 0.0_wp
    mod_vel_max = 0.0_wp
    vuln_table = .FALSE.
    DO l = 1, solve_cells
       j = j_cent(l)
       k = k_cent(l)
       IF ( q(1,j,k) .GT. 0.0_wp ) THEN
          CALL qc_to_qp( q(1:n_vars,j,k), qp(1:n_vars+2,j,k), p_dyn )
          hmax(j,k) = qp(1,j,k)
          r_u = qp(n_vars+1,j,k)
          r_v = qp(n_vars+2,j,k)
          mod_vel2 = r_u**2 + r_v**2
          mod_vel = SQRT( mod_vel2 )
          IF ( qp(1,j,k) .GT. 0.001_wp ) THEN
             pdynmax(j,k) = p_dyn
             mod_vel_max(j,k) = mod_vel
          END IF
          i_table = 0
          DO i_thk_lev = 1, n_thickness_levels
             thck_table(j,k) = ( qp(1,j,k) .GE. thickness_levels(i_thk_lev) )
             DO i_pdyn_lev = 1, n_dyn_pres_levels
                pdyn_table(j,k) = ( p_dyn .GE. dyn_pres_levels(i_pdyn_lev) )
                i_table = i_table + 1
                vuln_table(i_table,j,k) = ( thck_table(j,k) .AND. pdyn_table(j,k) )
             END DO
          END DO
       ELSE
          qp(1:n_vars,j,k) = 0.0_wp
          qp(4,j,k) = T_ambient
          hmax(j,k) = 0.0_wp
          pdynmax(j,k) = 0.0_wp
       END IF
    END DO

    imexref_setup = 0

  END FUNCTION imexref_setup

  !----------------------------------------------------------------------------
  ! allocate_solver_variables (src/solver_2d.f90:264-541) leaves most work
  ! arrays unset, and cells outside the active set read face states and stage
  ! states that were never written (and d_grav_coeff_dy is never assigned when
  ! the slope correction is off, src/geometry_2d.f90:885-886).  A fresh
  ! process sees those allocations as zero pages; this makes that explicit so
  ! that several runs in one process see the same memory a first run does.
  !----------------------------------------------------------------------------
  SUBROUTINE zero_fill_unset_arrays
    USE solver_2d, ONLY : q0, q_fv, q_interfaceL, q_interfaceR, q_interfaceB,   &
         q_interfaceT, qp_interfaceL, qp_interfaceR, qp_interfaceB,              &
         qp_interfaceT, H_interface_x, H_interface_y, a_interface_x_max,         &
         a_interface_y_max, q_rk, qp_rk, divFlux, NH, SI_NH, expl_terms,         &
         residual_term, solve_mask_x, solve_mask_y, source_xy, hpos, hpos_old,   &
         diverg_interfaceL, diverg_interfaceR, diverg_interfaceB,                &
         diverg_interfaceT, thck_table, pdyn_table, solve_mask, solve_mask_temp
    USE geometry_2d, ONLY : d_grav_coeff_dy
    q0 = 0.0_wp
    q_fv = 0.0_wp
    q_interfaceL = 0.0_wp
    q_interfaceR = 0.0_wp
    q_interfaceB = 0.0_wp
    q_interfaceT = 0.0_wp
    qp_interfaceL = 0.0_wp
    qp_interfaceR = 0.0_wp
    qp_interfaceB = 0.0_wp
    qp_interfaceT = 0.0_wp
    H_interface_x = 0.0_wp
    H_interface_y = 0.0_wp
    a_interface_x_max = 0.0_wp
    a_interface_y_max = 0.0_wp
    q_rk = 0.0_wp
    qp_rk = 0.0_wp
    divFlux = 0.0_wp
    NH = 0.0_wp
    SI_NH = 0.0_wp
    expl_terms = 0.0_wp
    residual_term = 0.0_wp
    source_xy = 0.0_wp
    solve_mask_x = .FALSE.
    solve_mask_y = .FALSE.
    hpos = .FALSE.
    hpos_old = .FALSE.
    diverg_interfaceL = .FALSE.
    diverg_interfaceR = .FALSE.
    diverg_interfaceB = .FALSE.
    diverg_interfaceT = .FALSE.
    thck_table = .FALSE.
    pdyn_table = .FALSE.
    ! only the boundary ring is set by allocate_solver_variables
    solve_mask(2:comp_cells_x-1,2:comp_cells_y-1) = .FALSE.
    solve_mask_temp = .FALSE.
    d_grav_coeff_dy = 0.0_wp
  END SUBROUTINE zero_fill_unset_arrays

  !----------------------------------------------------------------------------
  ! Advance nsteps time steps exactly as the driver loop does
  ! (src/IMEX_SfloW2D.f90:395-505), without output.  dts(1:nsteps) receives
  ! the dt of each step and ncells(1:nsteps) the active-cell count.
  !----------------------------------------------------------------------------
  INTEGER(c_int) FUNCTION imexref_run(nsteps, dts, ncells) BIND(C, name='imexref_run')
    INTEGER(c_int), VALUE :: nsteps
    REAL(c_double), INTENT(OUT) :: dts(*)
    INTEGER(c_int), INTENT(OUT) :: ncells(*)

    INTEGER :: istep, l, j, k, i_table, i_thk_lev, i_pdyn_lev
    REAL(wp) :: p_dyn, r_u, r_v, mod_vel2, mod_vel

    DO istep = 1, nsteps

       CALL check_solve(.FALSE.)
       ncells(istep) = solve_cells

       CALL timestep

       IF ( t+dt .GT. t_end ) dt = t_end - t
       dt = MIN( dt, 1.1_wp * 0.5_wp * ( dt_old + dt_old_old ) )
       dt_old_old = dt_old
       dt_old = dt
       dts(istep) = dt

       CALL imex_RK_solver
       CALL update_erosion_deposition_cell(dt)
       IF ( topo_change_flag ) CALL topography_reconstruction

       t = t + dt

       DO l = 1, solve_cells
          j = j_cent(l)
          k = k_cent(l)
          IF ( q(1,j,k) .GT. 0.0_wp ) THEN
             CALL qc_to_qp( q(1:n_vars,j,k), qp(1:n_vars+2,j,k), p_dyn )
             hmax(j,k) = MAX( hmax(j,k), qp(1,j,k) )
             r_u = qp(n_vars+1,j,k)
             r_v = qp(n_vars+2,j,k)
             mod_vel2 = r_u**2 + r_v**2
             mod_vel = SQRT( mod_vel2 )
             IF ( qp(1,j,k) .GT. 0.001_wp ) THEN
                pdynmax(j,k) = MAX( pdynmax(j,k), p_dyn )
                mod_vel_max(j,k) = MAX( mod_vel_max(j,k), mod_vel )
             END IF
             i_table = 0
             DO i_thk_lev = 1, n_thickness_levels
                thck_table(j,k) = ( qp(1,j,k) .GE. thickness_levels(i_thk_lev) )
                DO i_pdyn_lev = 1, n_dyn_pres_levels
                   pdyn_table(j,k) = ( p_dyn .GE. dyn_pres_levels(i_pdyn_lev) )
                   i_table = i_table + 1
                   vuln_table(i_table,j,k) = vuln_table(i_table,j,k) .OR.       &
                        ( thck_table(j,k) .AND. pdyn_table(j,k) )
                END DO
             END DO
          ELSE
             qp(1:n_vars+2,j,k) = 0.0_wp
             qp(4,j,k) = T_ambient
          END IF
       END DO

    END DO

    imexref_run = 0

  END FUNCTION imexref_run

  !----------------------------------------------------------------------------
  ! Accessors (all arrays in Fortran order: variable fastest, then x, then y).
  !----------------------------------------------------------------------------
  SUBROUTINE imexref_get_q(qout) BIND(C, name='imexref_get_q')
    REAL(c_double), INTENT(OUT) :: qout(n_vars, comp_cells_x, comp_cells_y)
    qout = q
  END SUBROUTINE imexref_get_q

  ! The erodible layer erodible(n_solid, nx, ny) (read_erodible's result).
  SUBROUTINE imexref_set_erodible(e) BIND(C, name='imexref_set_erodible')
    REAL(c_double), INTENT(IN) :: e(n_solid, comp_cells_x, comp_cells_y)
    erodible = e
  END SUBROUTINE imexref_set_erodible

  ! solver_2d's stochastic values Z(nx, ny)
  SUBROUTINE imexref_set_noise(zin) BIND(C, name='imexref_set_noise')
    REAL(c_double), INTENT(IN) :: zin(comp_cells_x, comp_cells_y)
    Z = zin
  END SUBROUTINE imexref_set_noise

  ! B_cent(nx, ny)
  SUBROUTINE imexref_get_bcent(b) BIND(C, name='imexref_get_bcent')
    REAL(c_double), INTENT(OUT) :: b(comp_cells_x, comp_cells_y)
    b = B_cent
  END SUBROUTINE imexref_get_bcent

  ! erosion(nx, ny, n_solid) and erodible(n_solid, nx, ny)
  SUBROUTINE imexref_get_erosion(ers, erd) BIND(C, name='imexref_get_erosion')
    REAL(c_double), INTENT(OUT) :: ers(comp_cells_x, comp_cells_y, n_solid)
    REAL(c_double), INTENT(OUT) :: erd(n_solid, comp_cells_x, comp_cells_y)
    ers = erosion
    erd = erodible
  END SUBROUTINE imexref_get_erosion

  ! The driver's module array qp (primitive variables kept by the epilogue).
  SUBROUTINE imexref_get_qp(qpo) BIND(C, name='imexref_get_qp')
    REAL(c_double), INTENT(OUT) :: qpo(n_vars+2, comp_cells_x, comp_cells_y)
    qpo = qp
  END SUBROUTINE imexref_get_qp

  ! interp_2d_scalarB (src/geometry_2d.f90:1535-1599) of the field f on the
  ! cell centres x_comp, y_comp at n points xy(1:2, p) (output_probes).
  SUBROUTINE imexref_interp(n, xy, f, out) BIND(C, name='imexref_interp')
    USE geometry_2d, ONLY : x_comp, y_comp, interp_2d_scalarB
    INTEGER(c_int), VALUE :: n
    REAL(c_double), INTENT(IN) :: xy(2, n)
    REAL(c_double), INTENT(IN) :: f(comp_cells_x, comp_cells_y)
    REAL(c_double), INTENT(OUT) :: out(n)
    INTEGER :: i
    DO i = 1, n
       CALL interp_2d_scalarB( x_comp, y_comp, f, xy(1,i), xy(2,i), out(i) )
    END DO
  END SUBROUTINE imexref_interp

  SUBROUTINE imexref_set_q(qin) BIND(C, name='imexref_set_q')
    REAL(c_double), INTENT(IN) :: qin(n_vars, comp_cells_x, comp_cells_y)
    q = qin
  END SUBROUTINE imexref_set_q

  SUBROUTINE imexref_set_time(tt, dtt, d1, d2) BIND(C, name='imexref_set_time')
    REAL(c_double), VALUE :: tt, dtt, d1, d2
    t = tt
    dt = dtt
    dt_old = d1
    dt_old_old = d2
  END SUBROUTINE imexref_set_time

  SUBROUTINE imexref_get_fields(hm, pdm, mvm, dep, tt, dtt) BIND(C, name='imexref_get_fields')
    REAL(c_double), INTENT(OUT) :: hm(comp_cells_x, comp_cells_y)
    REAL(c_double), INTENT(OUT) :: pdm(comp_cells_x, comp_cells_y)
    REAL(c_double), INTENT(OUT) :: mvm(comp_cells_x, comp_cells_y)
    REAL(c_double), INTENT(OUT) :: dep(comp_cells_x, comp_cells_y, n_solid)
    REAL(c_double), INTENT(OUT) :: tt, dtt
    hm = hmax
    pdm = pdynmax
    mvm = mod_vel_max
    dep = deposit
    tt = t
    dtt = dt
  END SUBROUTINE imexref_get_fields

  SUBROUTINE imexref_get_vuln(v) BIND(C, name='imexref_get_vuln')
    INTEGER(c_int), INTENT(OUT) :: v(MAX(1,n_thickness_levels*n_dyn_pres_levels), &
         comp_cells_x, comp_cells_y)
    INTEGER :: i, j, k
    v = 0
    DO k = 1, comp_cells_y
       DO j = 1, comp_cells_x
          DO i = 1, n_thickness_levels*n_dyn_pres_levels
             IF ( vuln_table(i,j,k) ) v(i,j,k) = 1
          END DO
       END DO
    END DO
  END SUBROUTINE imexref_get_vuln

  SUBROUTINE imexref_get_mask(m) BIND(C, name='imexref_get_mask')
    INTEGER(c_int), INTENT(OUT) :: m(comp_cells_x, comp_cells_y)
    INTEGER :: j, k
    DO k = 1, comp_cells_y
       DO j = 1, comp_cells_x
          m(j,k) = MERGE( 1, 0, solve_mask(j,k) )
       END DO
    END DO
  END SUBROUTINE imexref_get_mask

  SUBROUTINE imexref_get_static(bx, by, bc, scell, sew, arcp, arcx, arcy, hsrc) &
       BIND(C, name='imexref_get_static')
    REAL(c_double), INTENT(OUT) :: bx(comp_cells_x, comp_cells_y)
    REAL(c_double), INTENT(OUT) :: by(comp_cells_x, comp_cells_y)
    REAL(c_double), INTENT(OUT) :: bc(comp_cells_x, comp_cells_y)
    INTEGER(c_int), INTENT(OUT) :: scell(comp_cells_x, comp_cells_y)
    INTEGER(c_int), INTENT(OUT) :: sew(comp_cells_x, comp_cells_y)
    REAL(c_double), INTENT(OUT) :: arcp(comp_cells_x, comp_cells_y)
    REAL(c_double), INTENT(OUT) :: arcx(comp_cells_x, comp_cells_y)
    REAL(c_double), INTENT(OUT) :: arcy(comp_cells_x, comp_cells_y)
    REAL(c_double), INTENT(OUT) :: hsrc
    INTEGER :: j, k
    bx = B_prime_x_geom
    by = B_prime_y_geom
    bc = B_cent
    scell = source_cell
    sew = 0
    IF ( radial_source_flag ) THEN
       DO k = 1, comp_cells_y
          DO j = 1, comp_cells_x
             IF ( sourceE(j,k) ) sew(j,k) = IOR( sew(j,k), 1 )
             IF ( sourceW(j,k) ) sew(j,k) = IOR( sew(j,k), 2 )
             IF ( sourceN(j,k) ) sew(j,k) = IOR( sew(j,k), 4 )
             IF ( sourceS(j,k) ) sew(j,k) = IOR( sew(j,k), 8 )
          END DO
       END DO
    END IF
    arcp = cell_arc_perim
    arcx = cell_arc_n_x
    arcy = cell_arc_n_y
    hsrc = h_source
  END SUBROUTINE imexref_get_static

  !----------------------------------------------------------------------------
  ! Single-function entry points for the per-function known-answer tests.
  ! They run the reference routines on caller-provided states (after setup).
  !----------------------------------------------------------------------------
  SUBROUTINE imexref_qc_to_qp(n, qc, qpo, pdyn) BIND(C, name='imexref_qc_to_qp')
    INTEGER(c_int), VALUE :: n
    REAL(c_double), INTENT(IN) :: qc(n_vars, n)
    REAL(c_double), INTENT(OUT) :: qpo(n_vars+2, n)
    REAL(c_double), INTENT(OUT) :: pdyn(n)
    INTEGER :: i
    DO i = 1, n
       CALL qc_to_qp( qc(:,i), qpo(:,i), pdyn(i) )
    END DO
  END SUBROUTINE imexref_qc_to_qp

  SUBROUTINE imexref_qp_to_qc(n, qpi, qco) BIND(C, name='imexref_qp_to_qc')
    INTEGER(c_int), VALUE :: n
    REAL(c_double), INTENT(IN) :: qpi(n_vars+2, n)
    REAL(c_double), INTENT(OUT) :: qco(n_vars, n)
    INTEGER :: i
    DO i = 1, n
       CALL qp_to_qc( qpi(:,i), qco(:,i) )
    END DO
  END SUBROUTINE imexref_qp_to_qc

  SUBROUTINE imexref_mixt_var(n, qpi, out) BIND(C, name='imexref_mixt_var')
    INTEGER(c_int), VALUE :: n
    REAL(c_double), INTENT(IN) :: qpi(n_vars+2, n)
    REAL(c_double), INTENT(OUT) :: out(6, n)
    INTEGER :: i
    DO i = 1, n
       CALL mixt_var( qpi(:,i), out(1,i), out(2,i), out(3,i), out(4,i), .TRUE.,  &
            out(5,i), out(6,i) )
    END DO
  END SUBROUTINE imexref_mixt_var

  SUBROUTINE imexref_fluxes(n, dir, qci, qpi, out) BIND(C, name='imexref_fluxes')
    INTEGER(c_int), VALUE :: n, dir
    REAL(c_double), INTENT(IN) :: qci(n_vars, n)
    REAL(c_double), INTENT(IN) :: qpi(n_vars+2, n)
    REAL(c_double), INTENT(OUT) :: out(n_eqns, n)
    INTEGER :: i
    DO i = 1, n
       CALL eval_fluxes( qci(:,i), qpi(:,i), 0.0_wp, 0.0_wp, 1.0_wp, dir, out(:,i) )
    END DO
  END SUBROUTINE imexref_fluxes

  SUBROUTINE imexref_speeds(n, dir, qpi, out) BIND(C, name='imexref_speeds')
    INTEGER(c_int), VALUE :: n, dir
    REAL(c_double), INTENT(IN) :: qpi(n_vars+2, n)
    REAL(c_double), INTENT(OUT) :: out(2, n)
    REAL(wp) :: vmin(n_vars), vmax(n_vars)
    INTEGER :: i
    DO i = 1, n
       IF ( dir .EQ. 1 ) THEN
          CALL eval_local_speeds_x( qpi(:,i), 1.0_wp, vmin, vmax )
       ELSE
          CALL eval_local_speeds_y( qpi(:,i), 1.0_wp, vmin, vmax )
       END IF
       out(1,i) = vmin(1)
       out(2,i) = vmax(1)
    END DO
  END SUBROUTINE imexref_speeds

  SUBROUTINE imexref_implicit(n, qci, out) BIND(C, name='imexref_implicit')
    INTEGER(c_int), VALUE :: n
    REAL(c_double), INTENT(IN) :: qci(n_vars, n)
    REAL(c_double), INTENT(OUT) :: out(n_eqns, n)
    REAL(wp) :: fv
    INTEGER :: i
    DO i = 1, n
       CALL eval_implicit_terms( 0.0_wp, 0.0_wp, 0.0_wp, fv, r_qj = qci(:,i),    &
            r_nh_term_impl = out(:,i) )
    END DO
  END SUBROUTINE imexref_implicit

  SUBROUTINE imexref_implicit_cplx(n, qre, qim, ore, oim) BIND(C, name='imexref_implicit_cplx')
    INTEGER(c_int), VALUE :: n
    REAL(c_double), INTENT(IN) :: qre(n_vars, n), qim(n_vars, n)
    REAL(c_double), INTENT(OUT) :: ore(n_eqns, n), oim(n_eqns, n)
    COMPLEX(wp) :: cq(n_vars), cnh(n_eqns)
    REAL(wp) :: fv
    INTEGER :: i, m
    DO i = 1, n
       DO m = 1, n_vars
          cq(m) = CMPLX( qre(m,i), qim(m,i), wp )
       END DO
       CALL eval_implicit_terms( 0.0_wp, 0.0_wp, 0.0_wp, fv, c_qj = cq,          &
            c_nh_term_impl = cnh )
       ore(:,i) = REAL( cnh )
       oim(:,i) = AIMAG( cnh )
    END DO
  END SUBROUTINE imexref_implicit_cplx

  ! solve_rk_step (src/solver_2d.f90:1662-1996) at the stage-2 coefficients
  ! (a_diag = 1) for n cells: q_si the first guess, q0 the old state, R the
  ! right-hand side without the implicit terms; out the solution
  SUBROUTINE imexref_newton(n, dtt, qsi, q0i, ri, out) BIND(C, name='imexref_newton')
    INTEGER(c_int), VALUE :: n
    REAL(c_double), VALUE :: dtt
    REAL(c_double), INTENT(IN) :: qsi(n_vars, n), q0i(n_vars, n), ri(n_eqns, n)
    REAL(c_double), INTENT(OUT) :: out(n_vars, n)
    REAL(wp) :: a_t(n_RK), a_d(n_RK), dF(n_eqns, n_RK), ex(n_eqns, n_RK), nh(n_eqns, n_RK)
    REAL(wp) :: qj(n_vars), fv, dt_keep
    INTEGER :: i
    dt_keep = dt
    dt = dtt
    a_t = 0.0_wp
    a_d = 0.0_wp
    dF = 0.0_wp
    ex = 0.0_wp
    nh = 0.0_wp
    DO i = 1, n
       qj = qsi(:,i)
       CALL solve_rk_step( qj, q0i(:,i), a_t, a_d, 1.0_wp, ri(:,i), dF, ex, nh, &
            0.0_wp, 0.0_wp, 0.0_wp, fv )
       out(:,i) = qj
    END DO
    dt = dt_keep
  END SUBROUTINE imexref_newton

  ! OpenMP threads of the reference's cell loops.  With pore pressure and
  ! several threads the loops race on the module variable
  ! hydraulic_permeability (src/constitutive_2d.f90:266), which every
  ! thread writes and then reads (:3870-3888, :4998-5020) with dynamic
  ! permeability: the parity runs take one thread.
  SUBROUTINE imexref_set_threads(n) BIND(C, name='imexref_set_threads')
    USE omp_lib, ONLY : omp_set_num_threads
    INTEGER(c_int), VALUE :: n
    CALL omp_set_num_threads(n)
  END SUBROUTINE imexref_set_threads

  ! the stage arrays of cell (j, k) (0-based) after the last step: q_rk,
  ! divFlux, NH, SI_NH, expl_terms, each (n_eqns, n_RK)
  SUBROUTINE imexref_get_stage(j, k, out) BIND(C, name='imexref_get_stage')
    INTEGER(c_int), VALUE :: j, k
    REAL(c_double), INTENT(OUT) :: out(n_eqns, n_RK, 5)
    out(:,:,1) = q_rk(1:n_eqns, j+1, k+1, 1:n_RK)
    out(:,:,2) = divFlux(1:n_eqns, j+1, k+1, 1:n_RK)
    out(:,:,3) = NH(1:n_eqns, j+1, k+1, 1:n_RK)
    out(:,:,4) = SI_NH(1:n_eqns, j+1, k+1, 1:n_RK)
    out(:,:,5) = expl_terms(1:n_eqns, j+1, k+1, 1:n_RK)
  END SUBROUTINE imexref_get_stage

  SUBROUTINE imexref_semi_impl(n, qci, qpi, out) BIND(C, name='imexref_semi_impl')
    INTEGER(c_int), VALUE :: n
    REAL(c_double), INTENT(IN) :: qci(n_vars, n)
    REAL(c_double), INTENT(IN) :: qpi(n_vars+2, n)
    REAL(c_double), INTENT(OUT) :: out(n_eqns, n)
    REAL(wp) :: fv
    INTEGER :: i
    DO i = 1, n
       CALL eval_nh_semi_impl_terms( 0.0_wp, 0.0_wp, 0.0_wp, 0.0_wp, 0.0_wp,     &
            1.0_wp, qci(:,i), qpi(:,i), out(:,i), 0.0_wp, fv )
    END DO
  END SUBROUTINE imexref_semi_impl

  SUBROUTINE imexref_expl(n, bx, by, tt, qpi, out) BIND(C, name='imexref_expl')
    INTEGER(c_int), VALUE :: n
    REAL(c_double), INTENT(IN) :: bx(n), by(n)
    REAL(c_double), VALUE :: tt
    REAL(c_double), INTENT(IN) :: qpi(n_vars+2, n)
    REAL(c_double), INTENT(OUT) :: out(n_eqns, n)
    INTEGER :: i
    DO i = 1, n
       CALL eval_expl_terms( bx(i), by(i), 0.0_wp, 0.0_wp, 0.0_wp, 1.0_wp,       &
            0.0_wp, 0.0_wp, 0.0_wp, qpi(:,i), out(:,i), tt, 0.0_wp, 0.0_wp,     &
            0.0_wp, 0.0_wp, dx*dy )
    END DO
  END SUBROUTINE imexref_expl

  SUBROUTINE imexref_settling(n, diam, rhos, rhoc, ikv, out) BIND(C, name='imexref_settling')
    INTEGER(c_int), VALUE :: n
    REAL(c_double), INTENT(IN) :: diam(n), rhos(n), rhoc(n), ikv(n)
    REAL(c_double), INTENT(OUT) :: out(n)
    INTEGER :: i
    DO i = 1, n
       out(i) = settling_velocity( diam(i), rhos(i), rhoc(i), ikv(i) )
    END DO
  END SUBROUTINE imexref_settling

  SUBROUTINE imexref_mass_exchange(n, qpi, ddt, dep, eq) BIND(C, name='imexref_mass_exchange')
    INTEGER(c_int), VALUE :: n
    REAL(c_double), INTENT(IN) :: qpi(n_vars+2, n)
    REAL(c_double), VALUE :: ddt
    REAL(c_double), INTENT(OUT) :: dep(n_solid, n)
    REAL(c_double), INTENT(OUT) :: eq(n_eqns, n)
    REAL(wp) :: ers(n_solid), cpe, cpl, topo, erod(n_solid)
    INTEGER :: i
    erod = 0.0_wp
    DO i = 1, n
       CALL eval_mass_exchange_terms( qpi(:,i), 0, 0.0_wp, 0.0_wp, erod, ddt,    &
            ers, dep(:,i), cpe, cpl, eq(:,i), topo )
    END DO
  END SUBROUTINE imexref_mass_exchange

  INTEGER(c_int) FUNCTION imexref_n_vars() BIND(C, name='imexref_n_vars')
    imexref_n_vars = n_vars
  END FUNCTION imexref_n_vars

  REAL(c_double) FUNCTION imexref_eps_sing() BIND(C, name='imexref_eps_sing')
    imexref_eps_sing = eps_sing
  END FUNCTION imexref_eps_sing

  ! Release every allocatable so that setup can run again in the same process.
  SUBROUTINE imexref_teardown() BIND(C, name='imexref_teardown')
    USE geometry_2d
    USE solver_2d, ONLY : deallocate_solver_variables
    USE constitutive_2d, ONLY : implicit_map
    CALL deallocate_solver_variables
    IF ( ALLOCATED(implicit_map) ) DEALLOCATE( implicit_map )
    DEALLOCATE( x_comp, x_stag, y_comp, y_stag, source_cell, sourceE, sourceW,  &
         sourceN, sourceS, dist_sourceE, dist_sourceW, dist_sourceN,             &
         dist_sourceS, sourceE_vect_x, sourceE_vect_y, sourceW_vect_x,           &
         sourceW_vect_y, sourceS_vect_x, sourceS_vect_y, sourceN_vect_x,         &
         sourceN_vect_y, B_cent, B_cent_extended, B_nodata, B_prime_x,           &
         B_prime_y, B_second_xx, B_second_yy, B_second_xy, B_prime_x_geom,       &
         B_prime_y_geom, B_second_xx_geom, B_second_yy_geom, B_second_xy_geom,   &
         grid_output, grid_output_int, grav_coeff, d_grav_coeff_dx,              &
         d_grav_coeff_dy, grav_coeff_stag_x, grav_coeff_stag_y,                  &
         cell_source_fractions, cell_arc_perim, cell_arc_n_x, cell_arc_n_y,      &
         B_zone, erodible, deposit, erosion )
  END SUBROUTINE imexref_teardown

END MODULE imexref_harness
