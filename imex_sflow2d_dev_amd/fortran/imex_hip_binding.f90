!------------------------------------------------------------------------------
! imex_hip_binding: ISO_C_BINDING interface of libimexhip.so, the MI355X
! IMEX-RK stage of IMEX_SfloW2D (C declarations in include/imex_hip.h).
!
! The derived types are BIND(C) copies of the C structs (same members, same
! order, so the same offsets and padding); the interfaces declare every
! entry point of the header except the per-function test entry points
! (imex_test_*), which only the known-answer tests call.  Fortran has no
! unsigned integers: the unsigned bit-mask arguments of the C functions
! (download, snapshot and runout selections) are INTEGER(c_int) here, which
! has the size and calling convention of C's unsigned int.  A C pointer that
! may be NULL is TYPE(c_ptr), VALUE (pass C_LOC(array) or c_null_ptr).
!
! The reference's module procedures it replaces are named beside each group
! (src/solver_2d.f90, src/IMEX_SfloW2D.f90, src/inpout_2d.f90).
!------------------------------------------------------------------------------
MODULE imex_hip_binding

  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE

  ! physics error bits (the conditions on which the reference prints and
  ! pauses, src/solver_2d.f90:1382-1560, src/constitutive_2d.f90:4860-4866)
  INTEGER(c_int), PARAMETER :: IMEX_ERR_NAN = 1, IMEX_ERR_NEG_H = 2
  INTEGER(c_int), PARAMETER :: IMEX_ERR_NEG_SOLID = 4, IMEX_ERR_COLD = 8
  INTEGER(c_int), PARAMETER :: IMEX_ERR_NEG_DEP = 16, IMEX_ERR_SOLID_GT_TOT = 32
  INTEGER(c_int), PARAMETER :: IMEX_WARN_RESCALE = 64

  ! imex_download selections
  INTEGER(c_int), PARAMETER :: IMEX_DL_Q = 1, IMEX_DL_STATS = 2, IMEX_DL_DEPOSIT = 4
  INTEGER(c_int), PARAMETER :: IMEX_DL_VULN = 8, IMEX_DL_MASK = 16

  ! imex_runout_reduce flags
  INTEGER(c_int), PARAMETER :: IMEX_RUNOUT_STORE = 1, IMEX_RUNOUT_B = 2

  ! imex_snapshot_begin selections
  INTEGER(c_int), PARAMETER :: IMEX_SNAP_Q = 1, IMEX_SNAP_Q2D = 2, IMEX_SNAP_DEP2D = 4
  INTEGER(c_int), PARAMETER :: IMEX_SNAP_P2D = 8, IMEX_SNAP_FIELDS = 16
  INTEGER(c_int), PARAMETER :: IMEX_SNAP_ERO2D = 32, IMEX_SNAP_ERD2D = 64

  ! imex_write_records / imex_format_value formats
  INTEGER(c_int), PARAMETER :: IMEX_FMT_E20_12 = 0, IMEX_FMT_ES15_7 = 1
  INTEGER(c_int), PARAMETER :: IMEX_FMT_ES14_7 = 2

  ! typedef struct imex_config: the namelist values of the configured path
  ! (the module variables read_param sets, src/inpout_2d.f90:986-5134)
  TYPE, BIND(C) :: imex_config
     INTEGER(c_int32_t) :: nx, ny, n_solid, energy_flag, settling_flag
     INTEGER(c_int32_t) :: entrainment_flag, radial_source_flag
     INTEGER(c_int32_t) :: n_thickness_levels, n_dyn_pres_levels, precision
     REAL(c_double) :: dx, x0, y0, cfl, max_dt
     REAL(c_double) :: grav, mu, xi, pres, T_ambient, sp_heat_a, sp_gas_const_a
     REAL(c_double) :: kin_visc_a
     REAL(c_double) :: rho_s(3), diam_s(3), sp_heat_s(3)
     REAL(c_double) :: vel_source, T_source, h_source, time_param(4), alphas_source(3)
     REAL(c_double) :: thickness_levels(8), dyn_pres_levels(8)
     REAL(c_double) :: erosion_coeff, erodible_porosity, T_erodible, erodible_fract(3)
     INTEGER(c_int32_t) :: erodible_deposit_flag, topo_change_flag
     INTEGER(c_int32_t) :: rheology_model
     REAL(c_double) :: mu_s, mu_2, muI_inf, I_0
     INTEGER(c_int32_t) :: stochastic_flag
     ! pore pressure (PORE_PRESSURE_PARAMETERS): f_inhibit_mode 0 OFF,
     ! 1 STEP, 2 STATIC, 3 DYNAMIC; device-driven steps only
     INTEGER(c_int32_t) :: pore_pressure_flag, gas_loss_flag, f_inhibit_mode, N_inh
     INTEGER(c_int32_t) :: dynamic_permeability_flag
     REAL(c_double) :: hydraulic_permeability, alpha_trans, maximum_solid_packing
     REAL(c_double) :: sphericity_s(3)
     ! per-variable boundary conditions (pore-pressure runs): with bc_custom,
     ! bc_flag(8 side + var) and bc_value for the sides W, E, S, N
     INTEGER(c_int32_t) :: bc_custom
     INTEGER(c_int32_t) :: bc_flag(32)
     REAL(c_double) :: bc_value(32)
     ! the transported noise q1 * Z (stoch_transport_flag; needs stochastic_flag)
     INTEGER(c_int32_t) :: stoch_transport_flag
  END TYPE imex_config

  ! typedef struct imex_static: borrowed pointers to nx*ny arrays (x
  ! fastest), copied by imex_set_static; the source arrays may be null
  TYPE, BIND(C) :: imex_static
     TYPE(c_ptr) :: B_prime_x, B_prime_y, source_cell, source_sides
     TYPE(c_ptr) :: arc_perim, arc_n_x, arc_n_y
  END TYPE imex_static

  ! typedef struct imex_step_stats: the driver's stdout diagnostics
  ! (src/IMEX_SfloW2D.f90:507-544)
  TYPE, BIND(C) :: imex_step_stats
     REAL(c_double) :: t, dt, mass, volume, area, solid_mass
     INTEGER(c_int64_t) :: active_cells, dried_cells
     INTEGER(c_int32_t) :: err
  END TYPE imex_step_stats

  ! typedef struct imex_runout_red: the whole-array expressions of one
  ! output_runout call (src/inpout_2d.f90:7170-7427); indices 0-based linear
  TYPE, BIND(C) :: imex_runout_red
     REAL(c_double) :: max_h, sum_q1, sum_xq1, sum_yq1
     INTEGER(c_int64_t) :: n_hpos, n_hpos_old, n_new, imax, imax_x, imax_y
     REAL(c_double) :: dist_max, dist_x_max, dist_y_max, dist_at_imax_x, dist_at_imax_y
     INTEGER(c_int64_t) :: ib
     REAL(c_double) :: b_max
  END TYPE imex_runout_red

  INTERFACE

     ! ---- allocate_solver_variables / deallocate_solver_variables
     INTEGER(c_int) FUNCTION imex_create(cfg, h) BIND(C, name='imex_create')
       IMPORT :: c_int, c_ptr, imex_config
       TYPE(imex_config), INTENT(IN) :: cfg
       TYPE(c_ptr), INTENT(OUT) :: h
     END FUNCTION imex_create
     INTEGER(c_int) FUNCTION imex_set_static(h, s) BIND(C, name='imex_set_static')
       IMPORT :: c_int, c_ptr, imex_static
       TYPE(c_ptr), VALUE :: h
       TYPE(imex_static), INTENT(IN) :: s
     END FUNCTION imex_set_static
     INTEGER(c_int) FUNCTION imex_upload_state(h, q, t, dt0) BIND(C, name='imex_upload_state')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), INTENT(IN) :: q(*)
       REAL(c_double), VALUE :: t, dt0
     END FUNCTION imex_upload_state
     SUBROUTINE imex_destroy(h) BIND(C, name='imex_destroy')
       IMPORT :: c_ptr
       TYPE(c_ptr), VALUE :: h
     END SUBROUTINE imex_destroy
     INTEGER(c_int) FUNCTION imex_last_error(h, buf, n) BIND(C, name='imex_last_error')
       IMPORT :: c_int, c_ptr, c_char, c_size_t
       TYPE(c_ptr), VALUE :: h
       CHARACTER(kind=c_char), INTENT(OUT) :: buf(*)
       INTEGER(c_size_t), VALUE :: n
     END FUNCTION imex_last_error

     ! ---- check_solve, timestep, imex_RK_solver, update_erosion_deposition_cell
     INTEGER(c_int) FUNCTION imex_check_solve(h, solve_all) BIND(C, name='imex_check_solve')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int), VALUE :: solve_all
     END FUNCTION imex_check_solve
     INTEGER(c_int) FUNCTION imex_timestep(h, t, dt) BIND(C, name='imex_timestep')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), VALUE :: t
       REAL(c_double), INTENT(OUT) :: dt
     END FUNCTION imex_timestep
     INTEGER(c_int) FUNCTION imex_rk_step(h, dt) BIND(C, name='imex_rk_step')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), VALUE :: dt
     END FUNCTION imex_rk_step
     INTEGER(c_int) FUNCTION imex_erosion_deposition(h, dt) &
          BIND(C, name='imex_erosion_deposition')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), VALUE :: dt
     END FUNCTION imex_erosion_deposition

     ! ---- the driver's epilogue and loop (src/IMEX_SfloW2D.f90:395-544)
     INTEGER(c_int) FUNCTION imex_post_step(h, t_new, stats) BIND(C, name='imex_post_step')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), VALUE :: t_new
       TYPE(c_ptr), VALUE :: stats              ! C_LOC(imex_step_stats) or c_null_ptr
     END FUNCTION imex_post_step
     ! imex_post_step without allocations (no counterpart in the reference)
     INTEGER(c_int) FUNCTION imex_post_step_noalloc(h, t_new, stats) &
          BIND(C, name='imex_post_step_noalloc')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), VALUE :: t_new
       TYPE(c_ptr), VALUE :: stats
     END FUNCTION imex_post_step_noalloc
     INTEGER(c_int) FUNCTION imex_run_steps(h, nsteps, t_end, dts) BIND(C, name='imex_run_steps')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int), VALUE :: nsteps
       REAL(c_double), VALUE :: t_end
       TYPE(c_ptr), VALUE :: dts                ! nsteps doubles or c_null_ptr
     END FUNCTION imex_run_steps
     INTEGER(c_int) FUNCTION imex_run_steps_timed(h, nsteps, t_end, dts, sample_every, &
          kernel_ms, n_sampled, cells) BIND(C, name='imex_run_steps_timed')
       IMPORT :: c_int, c_ptr, c_double, c_int64_t
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int), VALUE :: nsteps, sample_every
       REAL(c_double), VALUE :: t_end
       TYPE(c_ptr), VALUE :: dts
       REAL(c_double), INTENT(OUT) :: kernel_ms(3)
       INTEGER(c_int), INTENT(OUT) :: n_sampled
       INTEGER(c_int64_t), INTENT(OUT) :: cells(*)
     END FUNCTION imex_run_steps_timed
     INTEGER(c_int) FUNCTION imex_calib_copy(h, n_doubles, reps) BIND(C, name='imex_calib_copy')
       IMPORT :: c_int, c_ptr, c_int64_t
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int64_t), VALUE :: n_doubles
       INTEGER(c_int), VALUE :: reps
     END FUNCTION imex_calib_copy
     INTEGER(c_int) FUNCTION imex_get_time(h, t, dt) BIND(C, name='imex_get_time')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), INTENT(OUT) :: t, dt
     END FUNCTION imex_get_time
     INTEGER(c_int) FUNCTION imex_set_time(h, t, dt_old, dt_old_old) BIND(C, name='imex_set_time')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), VALUE :: t, dt_old, dt_old_old
     END FUNCTION imex_set_time
     INTEGER(c_int) FUNCTION imex_get_time_history(h, t, dt_old, dt_old_old) &
          BIND(C, name='imex_get_time_history')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), INTENT(OUT) :: t, dt_old, dt_old_old
     END FUNCTION imex_get_time_history
     INTEGER(c_int) FUNCTION imex_get_counters(h, active_cells, steps) &
          BIND(C, name='imex_get_counters')
       IMPORT :: c_int, c_ptr, c_int64_t
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int64_t), INTENT(OUT) :: active_cells, steps
     END FUNCTION imex_get_counters
     INTEGER(c_int) FUNCTION imex_get_footprint(h, bytes) BIND(C, name='imex_get_footprint')
       IMPORT :: c_int, c_ptr, c_size_t
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_size_t), INTENT(OUT) :: bytes
     END FUNCTION imex_get_footprint
     INTEGER(c_int) FUNCTION imex_get_work_counters(h, work_cells, frame_cells) &
          BIND(C, name='imex_get_work_counters')
       IMPORT :: c_int, c_ptr, c_int64_t
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int64_t), INTENT(OUT) :: work_cells, frame_cells
     END FUNCTION imex_get_work_counters

     ! ---- erosion, topography change, stochastic friction
     INTEGER(c_int) FUNCTION imex_topography_reconstruction(h) &
          BIND(C, name='imex_topography_reconstruction')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h
     END FUNCTION imex_topography_reconstruction
     INTEGER(c_int) FUNCTION imex_set_erodible(h, erodible) BIND(C, name='imex_set_erodible')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), INTENT(IN) :: erodible(*)   ! erodible(n_solid, nx, ny)
     END FUNCTION imex_set_erodible
     INTEGER(c_int) FUNCTION imex_set_noise(h, Z) BIND(C, name='imex_set_noise')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), INTENT(IN) :: Z(*)           ! Z(nx, ny)
     END FUNCTION imex_set_noise
     INTEGER(c_int) FUNCTION imex_get_erosion(h, erosion, erodible) BIND(C, name='imex_get_erosion')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h, erosion, erodible
     END FUNCTION imex_get_erosion

     ! ---- host mirrors (before output_solution, output_runout, write_restart_file)
     INTEGER(c_int) FUNCTION imex_download(h, what, q, hmax, pdynmax, modvelmax, deposit, &
          vuln, mask) BIND(C, name='imex_download')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int), VALUE :: what
       TYPE(c_ptr), VALUE :: q, hmax, pdynmax, modvelmax, deposit, vuln, mask
     END FUNCTION imex_download

     ! ---- device-side outputs (output_runout, output_probes, output_solution,
     !      write_restart_file)
     INTEGER(c_int) FUNCTION imex_set_topography(h, B_cent) BIND(C, name='imex_set_topography')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), INTENT(IN) :: B_cent(*)
     END FUNCTION imex_set_topography
     INTEGER(c_int) FUNCTION imex_get_topography(h, B_cent) BIND(C, name='imex_get_topography')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), INTENT(OUT) :: B_cent(*)
     END FUNCTION imex_get_topography
     INTEGER(c_int) FUNCTION imex_runout_reduce(h, x0, y0, thr, flags, red) &
          BIND(C, name='imex_runout_reduce')
       IMPORT :: c_int, c_ptr, c_double, imex_runout_red
       TYPE(c_ptr), VALUE :: h
       REAL(c_double), VALUE :: x0, y0, thr
       INTEGER(c_int), VALUE :: flags
       TYPE(imex_runout_red), INTENT(OUT) :: red
     END FUNCTION imex_runout_reduce
     INTEGER(c_int) FUNCTION imex_probes(h, n, xy, vals) BIND(C, name='imex_probes')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int), VALUE :: n
       REAL(c_double), INTENT(IN) :: xy(2, *)
       REAL(c_double), INTENT(OUT) :: vals(*)
     END FUNCTION imex_probes
     INTEGER(c_int) FUNCTION imex_snapshot_cols(h, kind) BIND(C, name='imex_snapshot_cols')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h
       INTEGER(c_int), VALUE :: kind
     END FUNCTION imex_snapshot_cols
     INTEGER(c_int) FUNCTION imex_snapshot_begin(h, what, q, q2d, dep2d, p2d, fields, &
          ero2d, erd2d) BIND(C, name='imex_snapshot_begin')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h, q, q2d, dep2d, p2d, fields, ero2d, erd2d
       INTEGER(c_int), VALUE :: what
     END FUNCTION imex_snapshot_begin
     INTEGER(c_int) FUNCTION imex_snapshot_wait(h) BIND(C, name='imex_snapshot_wait')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: h
     END FUNCTION imex_snapshot_wait
     INTEGER(c_int) FUNCTION imex_host_alloc(bytes, p) BIND(C, name='imex_host_alloc')
       IMPORT :: c_int, c_ptr, c_size_t
       INTEGER(c_size_t), VALUE :: bytes
       TYPE(c_ptr), INTENT(OUT) :: p
     END FUNCTION imex_host_alloc
     INTEGER(c_int) FUNCTION imex_host_free(p) BIND(C, name='imex_host_free')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: p
     END FUNCTION imex_host_free
     INTEGER(c_int) FUNCTION imex_write_records(path, rec, n_rows, n_col, nx, fmt, n_threads) &
          BIND(C, name='imex_write_records')
       IMPORT :: c_int, c_char, c_double, c_int64_t
       CHARACTER(kind=c_char), INTENT(IN) :: path(*)   ! NUL-terminated
       REAL(c_double), INTENT(IN) :: rec(*)
       INTEGER(c_int64_t), VALUE :: n_rows
       INTEGER(c_int), VALUE :: n_col, nx, fmt, n_threads
     END FUNCTION imex_write_records
     INTEGER(c_int) FUNCTION imex_write_restart(path, nx, ny, n_vars, n_solid, t, dt, q, &
          B_cent, erodible, deposit, erosion, hmax, pdynmax, mod_vel_max, Z) &
          BIND(C, name='imex_write_restart')
       IMPORT :: c_int, c_char, c_double, c_ptr
       CHARACTER(kind=c_char), INTENT(IN) :: path(*)
       INTEGER(c_int), VALUE :: nx, ny, n_vars, n_solid
       REAL(c_double), VALUE :: t, dt
       TYPE(c_ptr), VALUE :: q, B_cent, erodible, deposit, erosion, hmax, pdynmax
       TYPE(c_ptr), VALUE :: mod_vel_max, Z
     END FUNCTION imex_write_restart
     INTEGER(c_int) FUNCTION imex_format_value(fmt, x, out) BIND(C, name='imex_format_value')
       IMPORT :: c_int, c_double, c_char
       INTEGER(c_int), VALUE :: fmt
       REAL(c_double), VALUE :: x
       CHARACTER(kind=c_char), INTENT(OUT) :: out(*)    ! at least 21 characters
     END FUNCTION imex_format_value

     ! ---- batches of independent runs (the ensemble of
     !      EXAMPLES/EXAMPLE_ENSEMBLE_ETNA/launch_jobs.py in one process)
     INTEGER(c_int) FUNCTION imex_batch_create(n_solid, b) BIND(C, name='imex_batch_create')
       IMPORT :: c_int, c_ptr
       INTEGER(c_int), VALUE :: n_solid
       TYPE(c_ptr), INTENT(OUT) :: b
     END FUNCTION imex_batch_create
     INTEGER(c_int) FUNCTION imex_batch_create_form(n_solid, precision, b)                   &
          BIND(C, name='imex_batch_create_form')
       IMPORT :: c_int, c_ptr
       INTEGER(c_int), VALUE :: n_solid, precision
       TYPE(c_ptr), INTENT(OUT) :: b
     END FUNCTION imex_batch_create_form
     INTEGER(c_int) FUNCTION imex_batch_set_blocks(b, blocks) BIND(C, name='imex_batch_set_blocks')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: b
       INTEGER(c_int), VALUE :: blocks
     END FUNCTION imex_batch_set_blocks
     INTEGER(c_int) FUNCTION imex_batch_set_launch(b, flux_threads, graphs)                   &
          BIND(C, name='imex_batch_set_launch')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: b
       INTEGER(c_int), VALUE :: flux_threads, graphs
     END FUNCTION imex_batch_set_launch
     INTEGER(c_int) FUNCTION imex_batch_add(b, h, t_end) BIND(C, name='imex_batch_add')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: b, h
       REAL(c_double), VALUE :: t_end
     END FUNCTION imex_batch_add
     INTEGER(c_int) FUNCTION imex_batch_remove(b, h) BIND(C, name='imex_batch_remove')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: b, h
     END FUNCTION imex_batch_remove
     INTEGER(c_int) FUNCTION imex_batch_run(b, nsteps) BIND(C, name='imex_batch_run')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: b
       INTEGER(c_int), VALUE :: nsteps
     END FUNCTION imex_batch_run
     INTEGER(c_int) FUNCTION imex_batch_status(b, n, done, steps, t) BIND(C, name='imex_batch_status')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: b
       INTEGER(c_int), VALUE :: n
       TYPE(c_ptr), VALUE :: done, steps, t
     END FUNCTION imex_batch_status
     INTEGER(c_int) FUNCTION imex_batch_last_error(b, buf, n) BIND(C, name='imex_batch_last_error')
       IMPORT :: c_int, c_ptr, c_char, c_size_t
       TYPE(c_ptr), VALUE :: b
       CHARACTER(kind=c_char), INTENT(OUT) :: buf(*)
       INTEGER(c_size_t), VALUE :: n
     END FUNCTION imex_batch_last_error
     SUBROUTINE imex_batch_destroy(b) BIND(C, name='imex_batch_destroy')
       IMPORT :: c_ptr
       TYPE(c_ptr), VALUE :: b
     END SUBROUTINE imex_batch_destroy

  END INTERFACE

CONTAINS

  !> The text of the handle's last error as a Fortran string.
  FUNCTION imex_error_text(h) RESULT(msg)
    TYPE(c_ptr), INTENT(IN) :: h
    CHARACTER(LEN=512) :: msg
    CHARACTER(kind=c_char) :: buf(512)
    INTEGER :: i, rc
    buf = c_null_char
    rc = imex_last_error(h, buf, INT(512, c_size_t))
    msg = ' '
    DO i = 1, 512
       IF (buf(i) == c_null_char) EXIT
       msg(i:i) = buf(i)
    END DO
  END FUNCTION imex_error_text

END MODULE imex_hip_binding
