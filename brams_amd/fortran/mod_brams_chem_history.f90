! Chemistry step-size history in the model's history (restart) file.
! AllocIndex keeps, per grid, the last accepted step of every block
! (index_g(ngr)%last_accepted_dt, chem_spack_utils.f90:11,137) and starts it at
! dtlt*N_DYN_CHEM (chem_spack_utils.f90:170-172); its comment there plans to
! keep it in the history file, so that a run restarted from history takes the
! same step sizes as the uninterrupted one.  A history writer calls
! chem_dt_history_write for each grid; a restart calls chem_dt_history_read
! after AllocIndex, in place of that initialisation.  The library packs and
! checks the record (brams_chem_dt_history_pack/unpack, include/brams_chem.h):
! the grid, block size and mechanism it belongs to and a checksum travel with
! the step sizes, and a record of another grid or mechanism, or a corrupt
! one, stops the run with the library's message.
MODULE mod_brams_chem_history

  USE iso_c_binding

  IMPLICIT NONE

  PRIVATE

  PUBLIC chem_dt_history_write, chem_dt_history_read

  ! mirror of brams_dt_history_grid (include/brams_chem.h)
  TYPE, BIND(C) :: brams_dt_history_grid
     INTEGER(C_INT32_T) :: m1, m2, m3, ia, iz, ja, jz, block_size
  END TYPE brams_dt_history_grid

  INTERFACE
     FUNCTION brams_chem_dt_history_bytes(nob) BIND(C, name='brams_chem_dt_history_bytes') RESULT(n)
       IMPORT :: C_INT32_T, C_INT64_T
       INTEGER(C_INT32_T), VALUE :: nob
       INTEGER(C_INT64_T) :: n
     END FUNCTION brams_chem_dt_history_bytes
     FUNCTION brams_chem_dt_history_pack(grid, dt, nob, buf, cap) &
          BIND(C, name='brams_chem_dt_history_pack') RESULT(n)
       IMPORT :: brams_dt_history_grid, C_DOUBLE, C_INT32_T, C_INT64_T, C_INT8_T
       TYPE(brams_dt_history_grid), INTENT(IN) :: grid
       REAL(C_DOUBLE), INTENT(IN) :: dt(*)
       INTEGER(C_INT32_T), VALUE :: nob
       INTEGER(C_INT8_T), INTENT(OUT) :: buf(*)
       INTEGER(C_INT64_T), VALUE :: cap
       INTEGER(C_INT64_T) :: n
     END FUNCTION brams_chem_dt_history_pack
     FUNCTION brams_chem_dt_history_unpack(buf, nbytes, grid, dt, max_nob) &
          BIND(C, name='brams_chem_dt_history_unpack') RESULT(rc)
       IMPORT :: brams_dt_history_grid, C_DOUBLE, C_INT, C_INT32_T, C_INT64_T, C_INT8_T
       INTEGER(C_INT8_T), INTENT(IN) :: buf(*)
       INTEGER(C_INT64_T), VALUE :: nbytes
       TYPE(brams_dt_history_grid), INTENT(IN) :: grid
       REAL(C_DOUBLE), INTENT(OUT) :: dt(*)
       INTEGER(C_INT32_T), VALUE :: max_nob
       INTEGER(C_INT) :: rc
     END FUNCTION brams_chem_dt_history_unpack
     FUNCTION brams_chem_last_error() BIND(C, name='brams_chem_last_error') RESULT(p)
       IMPORT :: C_PTR
       TYPE(C_PTR) :: p
     END FUNCTION brams_chem_last_error
     FUNCTION c_strlen(s) BIND(C, name='strlen') RESULT(n)
       IMPORT :: C_PTR, C_SIZE_T
       TYPE(C_PTR), VALUE :: s
       INTEGER(C_SIZE_T) :: n
     END FUNCTION c_strlen
  END INTERFACE

CONTAINS

  ! Appends one grid's record to iunit, an unformatted stream unit open for
  ! writing: the record's length in bytes (8-byte integer), then the record.
  SUBROUTINE chem_dt_history_write(iunit, m1, m2, m3, ia, iz, ja, jz, block_size, nob, last_accepted_dt)
    INTEGER, INTENT(IN) :: iunit, m1, m2, m3, ia, iz, ja, jz, block_size, nob
    DOUBLE PRECISION, INTENT(IN) :: last_accepted_dt(nob)
    TYPE(brams_dt_history_grid) :: grid
    INTEGER(C_INT8_T), ALLOCATABLE :: buf(:)
    INTEGER(C_INT64_T) :: n, got

    grid = brams_dt_history_grid(m1, m2, m3, ia, iz, ja, jz, block_size)
    n = brams_chem_dt_history_bytes(INT(nob, C_INT32_T))
    IF (n < 0) CALL fatal('chem_dt_history_write', INT(n))
    ALLOCATE(buf(n))
    got = brams_chem_dt_history_pack(grid, last_accepted_dt, INT(nob, C_INT32_T), buf, n)
    IF (got /= n) CALL fatal('chem_dt_history_write', INT(got))
    WRITE(iunit) n
    WRITE(iunit) buf
    DEALLOCATE(buf)
  END SUBROUTINE chem_dt_history_write

  ! Reads the next record from iunit (written by chem_dt_history_write) into
  ! last_accepted_dt, after the library has checked it against this grid.
  SUBROUTINE chem_dt_history_read(iunit, m1, m2, m3, ia, iz, ja, jz, block_size, nob, last_accepted_dt)
    INTEGER, INTENT(IN) :: iunit, m1, m2, m3, ia, iz, ja, jz, block_size, nob
    DOUBLE PRECISION, INTENT(OUT) :: last_accepted_dt(nob)
    TYPE(brams_dt_history_grid) :: grid
    INTEGER(C_INT8_T), ALLOCATABLE :: buf(:)
    INTEGER(C_INT64_T) :: n
    INTEGER(C_INT) :: rc

    grid = brams_dt_history_grid(m1, m2, m3, ia, iz, ja, jz, block_size)
    READ(iunit) n
    IF (n < 16 .OR. n > 80_C_INT64_T + 8_C_INT64_T * INT(nob, C_INT64_T)) THEN
       PRINT *, 'chem_dt_history_read: a record of ', n, ' bytes cannot hold ', nob, ' blocks'
       STOP 'chem_dt_history_read'
    END IF
    ALLOCATE(buf(n))
    READ(iunit) buf
    rc = brams_chem_dt_history_unpack(buf, n, grid, last_accepted_dt, INT(nob, C_INT32_T))
    IF (rc /= nob) CALL fatal('chem_dt_history_read', INT(rc))
    DEALLOCATE(buf)
  END SUBROUTINE chem_dt_history_read

  SUBROUTINE fatal(who, rc)
    CHARACTER(LEN=*), INTENT(IN) :: who
    INTEGER, INTENT(IN) :: rc
    TYPE(C_PTR) :: p
    CHARACTER(KIND=C_CHAR), POINTER :: s(:)
    CHARACTER(LEN=512) :: msg
    INTEGER :: i, n

    p = brams_chem_last_error()
    n = MIN(INT(c_strlen(p)), LEN(msg))
    CALL C_F_POINTER(p, s, [n])
    msg = ' '
    DO i = 1, n
       msg(i:i) = s(i)
    END DO
    PRINT *, who, ' (', rc, '): ', TRIM(msg)
    STOP 'chem dt history'
  END SUBROUTINE fatal

END MODULE mod_brams_chem_history
