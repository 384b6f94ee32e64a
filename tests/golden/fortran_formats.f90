! Writes each value of fortran_formats_in.txt (one 64-bit pattern per line,
! as a signed decimal integer) with the edit descriptors of the reference's
! output files, one line per value and descriptor.  Built and run by
! make_fortran_formats.py; its output is the fixture fortran_formats.json.
PROGRAM fortran_formats
  IMPLICIT NONE
  INTEGER(8) :: bits
  REAL(8) :: x
  INTEGER :: ios
  OPEN(10, FILE='fortran_formats_in.txt', STATUS='old')
  DO
     READ(10, *, IOSTAT=ios) bits
     IF (ios /= 0) EXIT
     x = TRANSFER(bits, x)
     WRITE(*, '(A,F12.3,A)') '[', x, ']'
     WRITE(*, '(A,F14.3,A)') '[', x, ']'
     WRITE(*, '(A,F17.3,A)') '[', x, ']'
     WRITE(*, '(A,F10.2,A)') '[', x, ']'
     WRITE(*, '(A,ES14.7E2,A)') '[', x, ']'
     WRITE(*, '(A,ES15.7E2,A)') '[', x, ']'
     WRITE(*, '(A,e20.12,A)') '[', x, ']'
  END DO
  CLOSE(10)
  WRITE(*, '(A,A13,A)') '[', 'Time (s),', ']'
  WRITE(*, '(A,A17,A)') '[', 'X_runout (m),', ']'
  WRITE(*, '(A,A18,A)') '[', 'maximum runout =', ']'
  WRITE(*, '(A,A15,A)') '[', 'alfa_s(01),', ']'
  WRITE(*, '(A,A4,A)') '[', 'abcdefgh', ']'
END PROGRAM fortran_formats
