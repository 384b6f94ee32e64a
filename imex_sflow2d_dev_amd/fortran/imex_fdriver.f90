!------------------------------------------------------------------------------
! imex_fdriver: a Fortran host program that runs the stage through
! imex_hip_binding, the way the reference's driver loop calls MODULE
! solver_2d (src/IMEX_SfloW2D.f90:395-505): per step check_solve, timestep,
! the dt clip to t_end and the growth limit (:417-439), imex_RK_solver,
! update_erosion_deposition_cell, t = t + dt and the epilogue (here the
! device's imex_post_step).  Mode 1 runs the same loop as one imex_run_steps
! call instead.
!
! Usage: imex_fdriver INPUT OUTPUT
!   INPUT  (stream, native byte order): int32 nsteps, int32 mode,
!          int32 cfg_bytes, the bytes of the C struct imex_config,
!          f64 t0, dt0, t_end, then nx*ny values of B_prime_x, B_prime_y
!          (f64), source_cell, source_sides (int32), arc_perim, arc_n_x,
!          arc_n_y (f64), then q0(n_vars, nx, ny) (f64)
!   OUTPUT (stream): int32 rc, int32 error bits, f64 t, dts(nsteps),
!          q(n_vars, nx, ny), hmax(nx, ny), pdynmax, mod_vel_max,
!          deposit(nx, ny, n_solid), then the step statistics of the last
!          step: mass, volume, area, solid_mass
!------------------------------------------------------------------------------
PROGRAM imex_fdriver

  USE, INTRINSIC :: iso_c_binding
  USE imex_hip_binding
  IMPLICIT NONE

  CHARACTER(LEN=1024) :: fin, fout
  INTEGER :: u, nsteps, mode, cfg_bytes, n, nv, ns, istep
  INTEGER(c_int) :: rc, err
  INTEGER(c_int8_t), ALLOCATABLE :: raw(:)
  TYPE(imex_config) :: cfg
  TYPE(imex_static), TARGET :: st
  TYPE(imex_step_stats), TARGET :: stats
  TYPE(c_ptr) :: h
  REAL(c_double) :: t, dt, dt0, t_end, dt_old, dt_old_old
  REAL(c_double), ALLOCATABLE, TARGET :: bx(:), by(:), ap(:), ax(:), ay(:)
  INTEGER(c_int32_t), ALLOCATABLE, TARGET :: sc(:), sd(:)
  REAL(c_double), ALLOCATABLE, TARGET :: q(:), dts(:), hm(:), pd(:), mv(:), dep(:)

  CALL get_command_argument(1, fin)
  CALL get_command_argument(2, fout)

  OPEN(newunit=u, file=TRIM(fin), access='stream', form='unformatted', status='old')
  READ(u) nsteps, mode, cfg_bytes
  IF (cfg_bytes /= INT(C_SIZEOF(cfg))) THEN
     WRITE(*,*) 'imex_fdriver: imex_config is', cfg_bytes, 'bytes in C and', &
          C_SIZEOF(cfg), 'in Fortran'
     STOP 2
  END IF
  ALLOCATE(raw(cfg_bytes))
  READ(u) raw
  ! the memory image of the C struct, padding included
  cfg = TRANSFER(raw, cfg)
  READ(u) t, dt0, t_end
  n = cfg%nx * cfg%ny
  ns = cfg%n_solid
  ! the transported noise and the pore pressure after the solid classes
  nv = 4 + ns + MERGE(1, 0, cfg%stoch_transport_flag /= 0) + &
       MERGE(1, 0, cfg%pore_pressure_flag /= 0)
  ALLOCATE(bx(n), by(n), ap(n), ax(n), ay(n), sc(n), sd(n))
  ALLOCATE(q(nv * n), dts(MAX(nsteps, 1)), hm(n), pd(n), mv(n), dep(n * ns))
  READ(u) bx, by, sc, sd, ap, ax, ay
  READ(u) q
  CLOSE(u)

  rc = imex_create(cfg, h)
  CALL check('imex_create')
  st%B_prime_x = C_LOC(bx)
  st%B_prime_y = C_LOC(by)
  st%source_cell = C_LOC(sc)
  st%source_sides = C_LOC(sd)
  st%arc_perim = C_LOC(ap)
  st%arc_n_x = C_LOC(ax)
  st%arc_n_y = C_LOC(ay)
  rc = imex_set_static(h, st)
  CALL check('imex_set_static')
  rc = imex_upload_state(h, q, t, dt0)
  CALL check('imex_upload_state')

  err = 0
  dts = 0.0_c_double
  IF (mode == 1) THEN
     rc = imex_run_steps(h, INT(nsteps, c_int), t_end, C_LOC(dts))
     IF (rc > 0) err = IOR(err, rc)
     CALL check('imex_run_steps')
     rc = imex_get_time(h, t, dt)
     CALL check('imex_get_time')
  ELSE
     dt_old = dt0
     dt_old_old = dt0
     DO istep = 1, nsteps
        rc = imex_check_solve(h, 0_c_int)
        CALL check('imex_check_solve')
        rc = imex_timestep(h, t, dt)
        CALL check('imex_timestep')
        IF (t + dt > t_end) dt = t_end - t
        dt = MIN(dt, 1.1_c_double * 0.5_c_double * (dt_old + dt_old_old))
        dt_old_old = dt_old
        dt_old = dt
        dts(istep) = dt
        rc = imex_rk_step(h, dt)
        CALL check('imex_rk_step')
        rc = imex_erosion_deposition(h, dt)
        CALL check('imex_erosion_deposition')
        t = t + dt
        rc = imex_post_step(h, t, c_null_ptr)
        IF (rc > 0) err = IOR(err, rc)
        CALL check('imex_post_step')
     END DO
  END IF

  rc = imex_download(h, IOR(IOR(IMEX_DL_Q, IMEX_DL_STATS), IMEX_DL_DEPOSIT), C_LOC(q), &
       C_LOC(hm), C_LOC(pd), C_LOC(mv), C_LOC(dep), c_null_ptr, c_null_ptr)
  CALL check('imex_download')
  rc = imex_post_step(h, t, C_LOC(stats))
  IF (rc > 0) err = IOR(err, rc)
  CALL check('imex_post_step')
  CALL imex_destroy(h)

  OPEN(newunit=u, file=TRIM(fout), access='stream', form='unformatted', status='replace')
  WRITE(u) 0_c_int, err, t, dts(1:nsteps), q, hm, pd, mv, dep
  WRITE(u) stats%mass, stats%volume, stats%area, stats%solid_mass
  CLOSE(u)

CONTAINS

  ! rc < 0: an API or HIP error, reported with the handle's text; rc > 0
  ! (physics error bits) is the caller's to record
  SUBROUTINE check(what)
    CHARACTER(LEN=*), INTENT(IN) :: what
    IF (rc < 0) THEN
       WRITE(*,*) 'imex_fdriver: ', what, ' failed (', rc, '): ', TRIM(imex_error_text(h))
       OPEN(newunit=u, file=TRIM(fout), access='stream', form='unformatted', status='replace')
       WRITE(u) rc, err
       CLOSE(u)
       STOP 1
    END IF
  END SUBROUTINE check

END PROGRAM imex_fdriver
