! Test driver of mod_brams_chem_history: writes the records of two grids to
! the file named by the first argument, reads them back and compares.  The
! grids are 16x16x10 (40 blocks of 64) and 40x30x12 (225 blocks).
PROGRAM history_driver
  USE mod_brams_chem_history
  IMPLICIT NONE
  CHARACTER(LEN=512) :: path
  DOUBLE PRECISION :: dt1(40), dt2(225), r1(40), r2(225)
  INTEGER :: u, i

  CALL get_command_argument(1, path)
  DO i = 1, 40
     dt1(i) = 10.0d0 + 0.5d0 * i
  END DO
  DO i = 1, 225
     dt2(i) = 120.0d0 / i
  END DO
  OPEN(NEWUNIT=u, FILE=TRIM(path), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  CALL chem_dt_history_write(u, 12, 18, 18, 2, 17, 2, 17, 64, 40, dt1)
  CALL chem_dt_history_write(u, 14, 42, 32, 2, 41, 2, 31, 64, 225, dt2)
  CLOSE(u)
  OPEN(NEWUNIT=u, FILE=TRIM(path), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  CALL chem_dt_history_read(u, 12, 18, 18, 2, 17, 2, 17, 64, 40, r1)
  CALL chem_dt_history_read(u, 14, 42, 32, 2, 41, 2, 31, 64, 225, r2)
  CLOSE(u)
  IF (ALL(r1 == dt1) .AND. ALL(r2 == dt2)) THEN
     PRINT '(A)', 'OK'
  ELSE
     STOP 1
  END IF
END PROGRAM history_driver
