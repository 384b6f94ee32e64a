! TEST INFRASTRUCTURE ONLY (the checker of the host writers, and the timed
! CPU writer of the outputs measurement): writes records with the Fortran
! runtime the way the reference's output routines do --
!   q2d      '(2e20.12,100(e20.12))' per cell, a list-directed blank record
!            after each row and two at the end (output_solution,
!            src/inpout_2d.f90:6056-6058, :6084-6109, :6111-6112)
!   p2d      '(100ES15.7E2)' per cell, the same blank records (:6390-6406)
!   restart  the unformatted sequential records of write_restart_file
!            (:8059-8081); restart_z adds the stochastic_flag record Z
! Usage: records_writer q2d|p2d <in.raw> <out> <nx> <ny> <ncol>
!        records_writer restart|restart_z <in.raw> <out> <nx> <ny> <n_vars> <n_solid>
! in.raw holds the values as raw doubles: the records (ncol per cell, cells
! x fastest), or for restart t, dt, q, B_cent, erodible, deposit, erosion,
! hmax, pdynmax, mod_vel_max (and Z for restart_z) in the reference's array
! orders.
PROGRAM records_writer
  IMPLICIT NONE
  CHARACTER(LEN=512) :: mode, fin, fout, arg
  INTEGER :: nx, ny, ncol, nv, ns, j, k, i
  REAL(8), ALLOCATABLE :: rec(:, :, :), q(:, :, :), b(:, :), er(:, :, :), dep(:, :, :), &
       ero(:, :, :), hm(:, :), pd(:, :), mv(:, :), z(:, :)
  REAL(8) :: t, dt
  CALL GET_COMMAND_ARGUMENT(1, mode)
  CALL GET_COMMAND_ARGUMENT(2, fin)
  CALL GET_COMMAND_ARGUMENT(3, fout)
  CALL GET_COMMAND_ARGUMENT(4, arg)
  READ(arg, *) nx
  CALL GET_COMMAND_ARGUMENT(5, arg)
  READ(arg, *) ny
  OPEN(10, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  IF (TRIM(mode) == 'restart' .OR. TRIM(mode) == 'restart_z') THEN
     CALL GET_COMMAND_ARGUMENT(6, arg)
     READ(arg, *) nv
     CALL GET_COMMAND_ARGUMENT(7, arg)
     READ(arg, *) ns
     ALLOCATE(q(nv, nx, ny), b(nx, ny), er(ns, nx, ny), dep(nx, ny, ns), ero(nx, ny, ns), &
          hm(nx, ny), pd(nx, ny), mv(nx, ny), z(nx, ny))
     READ(10) t, dt, q, b, er, dep, ero, hm, pd, mv
     IF (TRIM(mode) == 'restart_z') READ(10) z
     CLOSE(10)
     OPEN(11, FILE=TRIM(fout), FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
     WRITE(11) nx, ny, nv, ns
     WRITE(11) t, dt
     WRITE(11) q
     WRITE(11) b
     WRITE(11) er
     WRITE(11) dep
     WRITE(11) ero
     WRITE(11) hm
     WRITE(11) pd
     WRITE(11) mv
     IF (TRIM(mode) == 'restart_z') WRITE(11) z
     CLOSE(11)
  ELSE
     CALL GET_COMMAND_ARGUMENT(6, arg)
     READ(arg, *) ncol
     ALLOCATE(rec(ncol, nx, ny))
     READ(10) rec
     CLOSE(10)
     OPEN(11, FILE=TRIM(fout), STATUS='REPLACE', FORM='FORMATTED')
     DO k = 1, ny
        DO j = 1, nx
           IF (TRIM(mode) == 'q2d') THEN
              WRITE(11, '(2e20.12,100(e20.12))') rec(1, j, k), rec(2, j, k), &
                   (rec(i, j, k), i = 3, ncol)
           ELSE
              WRITE(11, '(100ES15.7E2)') (rec(i, j, k), i = 1, ncol)
           END IF
        END DO
        WRITE(11, *) ' '
     END DO
     WRITE(11, *) ' '
     WRITE(11, *) ' '
     CLOSE(11)
  END IF
END PROGRAM records_writer
