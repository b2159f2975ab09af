! Host-array call of the chemistry library, shared by the drop-in drivers
! mod_chem_spack_rodas3_dyndt (CHEMISTRY=4), mod_chem_spack_ros_dyndt
! (CHEMISTRY=3) and mod_chem_spack_qssa (CHEMISTRY=1): fills brams_chem_args
! (include/brams_chem.h) with C_LOC pointers to the caller's arrays, calls
! the solver's C-ABI entry (brams_chem_rodas3_dyndt, brams_chem_ros_dyndt or
! brams_chem_qssa) and handles its status.
! The library returns a status instead of aborting: fatal errors STOP here
! with the library's message; blocks that hit a zero pivot or a non-finite
! state (where the reference's Sparse1.3 would abort the run) are reported
! and left at their last accepted state.
MODULE mod_brams_chem_host

  USE mem_chem1, ONLY: &
       chem1_vars        ! Type

  USE iso_c_binding

  IMPLICIT NONE

  PRIVATE

  PUBLIC brams_chem_host_step, brams_chem_host_qssa

  INTEGER, PARAMETER, PUBLIC :: brams_solver_rodas3 = 0, brams_solver_ros2 = 1

  ! mirror of brams_chem_args (include/brams_chem.h)
  TYPE, BIND(C) :: brams_chem_args
     INTEGER(C_INT32_T) :: m1, m2, m3, ia, iz, ja, jz
     REAL(C_FLOAT)      :: dtlt
     INTEGER(C_INT32_T) :: n_dyn_chem, split_method, photoj_method
     REAL(C_FLOAT)      :: cp, cpor, p00
     TYPE(C_PTR)        :: pp, pi0, theta, rv
     TYPE(C_PTR)        :: jphoto
     INTEGER(C_INT32_T) :: nspecies, nr, nr_photo
     TYPE(C_PTR)        :: weight
     INTEGER(C_INT32_T) :: n_transported, n_not_transported
     TYPE(C_PTR)        :: transp_chem_index, no_transp_chem_index
     TYPE(C_PTR)        :: sc_p, sc_t, sc_t_dyn
     INTEGER(C_INT32_T) :: nob, maxblock_size
     TYPE(C_PTR)        :: block_end, last_accepted_dt
     TYPE(C_PTR)        :: atol, rtol
  END TYPE brams_chem_args

  INTERFACE
     FUNCTION brams_chem_rodas3_dyndt(args) BIND(C, name='brams_chem_rodas3_dyndt') RESULT(rc)
       IMPORT :: brams_chem_args, C_INT
       TYPE(brams_chem_args), INTENT(IN) :: args
       INTEGER(C_INT) :: rc
     END FUNCTION brams_chem_rodas3_dyndt
     FUNCTION brams_chem_ros_dyndt(args) BIND(C, name='brams_chem_ros_dyndt') RESULT(rc)
       IMPORT :: brams_chem_args, C_INT
       TYPE(brams_chem_args), INTENT(IN) :: args
       INTEGER(C_INT) :: rc
     END FUNCTION brams_chem_ros_dyndt
     FUNCTION brams_chem_qssa(args, itchim) BIND(C, name='brams_chem_qssa') RESULT(rc)
       IMPORT :: brams_chem_args, C_INT, C_INT32_T
       TYPE(brams_chem_args), INTENT(IN) :: args
       INTEGER(C_INT32_T), VALUE :: itchim
       INTEGER(C_INT) :: rc
     END FUNCTION brams_chem_qssa
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

  !========================================================================================
  ! One chemistry call on the caller's host arrays.  who names the calling
  ! routine in messages; solver is brams_solver_rodas3 or brams_solver_ros2.
  ! The arguments are those of chem_rodas3_dyndt / chem_ros_dyndt that the
  ! RELACS_TUV path reads, with the reference's names and meaning.
  SUBROUTINE brams_chem_host_step(who, solver, m1, m2, m3, dtlt, pp, pi0, theta, rv, &
                                  nspecies, nr, nr_photo, weight, PhotojMethod, &
                                  nspecies_chem_transported, nspecies_chem_no_transported, &
                                  transp_chem_index, no_transp_chem_index, chem1_g, nob, maxblock_size, &
                                  block_end, indexk, indexi, indexj, last_accepted_dt, atol, rtol, &
                                  N_DYN_CHEM, split_method, cp, cpor, p00, jphoto)
    CHARACTER(LEN=*) , INTENT(IN) :: who
    INTEGER          , INTENT(IN) :: solver
    INTEGER          , INTENT(IN) :: m1, m2, m3
    REAL             , INTENT(IN) :: dtlt
    REAL             , INTENT(IN), TARGET :: pp(m1,m2,m3), pi0(m1,m2,m3), theta(m1,m2,m3), rv(m1,m2,m3)
    INTEGER          , INTENT(IN) :: nspecies, nr, nr_photo
    REAL             , INTENT(IN) :: weight(nspecies)
    CHARACTER(LEN=*) , INTENT(IN) :: PhotojMethod
    INTEGER          , INTENT(IN) :: nspecies_chem_transported, nspecies_chem_no_transported
    INTEGER          , INTENT(IN) :: transp_chem_index(:), no_transp_chem_index(:)
    TYPE (chem1_vars), INTENT(INOUT) :: chem1_g(nspecies)
    INTEGER          , INTENT(IN) :: nob, maxblock_size
    INTEGER          , INTENT(IN) :: block_end(:), indexk(:,:), indexi(:,:), indexj(:,:)
    DOUBLE PRECISION , INTENT(INOUT), TARGET :: last_accepted_dt(nob)
    DOUBLE PRECISION , INTENT(IN) :: atol(nspecies), rtol(nspecies)
    INTEGER          , INTENT(IN) :: N_DYN_CHEM
    CHARACTER(LEN=*) , INTENT(IN) :: split_method
    REAL             , INTENT(IN) :: cp, cpor, p00
    DOUBLE PRECISION , INTENT(IN), TARGET :: jphoto(nr_photo,m1,m2,m3)

    TYPE(brams_chem_args) :: a
    TYPE(C_PTR), ALLOCATABLE, TARGET :: tab_p(:), tab_t(:), tab_td(:)
    REAL(C_FLOAT), ALLOCATABLE, TARGET :: w32(:)
    INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: tidx(:), nidx(:), bend(:)
    DOUBLE PRECISION, ALLOCATABLE, TARGET :: atol64(:), rtol64(:)
    INTEGER :: n, rc
    LOGICAL :: par, par_n

    CALL fill_common(who, a, w32, m1, m2, m3, dtlt, pp, pi0, theta, rv, nspecies, nr, nr_photo, &
                     weight, PhotojMethod, nob, block_end, indexk, indexi, indexj, cp, cpor, p00, jphoto)
    a%n_dyn_chem = N_DYN_CHEM
    SELECT CASE (TRIM(split_method))
    CASE ('SYMMETRIC');  a%split_method = 0
    CASE ('SEQUENTIAL'); a%split_method = 1
    CASE ('PARALLEL');   a%split_method = 2
    CASE DEFAULT
       PRINT *, who, ': unknown split_method ', split_method
       STOP
    END SELECT
    ALLOCATE(tidx(max(1,nspecies_chem_transported)), nidx(max(1,nspecies_chem_no_transported)))
    tidx(1:nspecies_chem_transported) = transp_chem_index(1:nspecies_chem_transported)
    nidx(1:nspecies_chem_no_transported) = no_transp_chem_index(1:nspecies_chem_no_transported)
    a%n_transported = nspecies_chem_transported
    a%n_not_transported = nspecies_chem_no_transported
    a%transp_chem_index = C_LOC(tidx)
    a%no_transp_chem_index = C_LOC(nidx)
    ! species state: chem1_vars holds pointer components, so pass a table of
    ! their addresses (the type itself is not interoperable)
    par   = TRIM(split_method) == 'PARALLEL'
    par_n = par .AND. N_DYN_CHEM > 1
    ALLOCATE(tab_p(nspecies), tab_t(nspecies), tab_td(nspecies))
    DO n = 1, nspecies
       tab_p(n) = C_LOC(chem1_g(n)%sc_p(1,1,1))
       tab_t(n) = C_NULL_PTR
       tab_td(n) = C_NULL_PTR
       IF (par .AND. .NOT. par_n) tab_t(n) = C_LOC(chem1_g(n)%sc_t(1))
       IF (par_n) tab_td(n) = C_LOC(chem1_g(n)%sc_t_dyn(1))
    END DO
    a%sc_p = C_LOC(tab_p)
    a%sc_t = C_NULL_PTR
    a%sc_t_dyn = C_NULL_PTR
    IF (par .AND. .NOT. par_n) a%sc_t = C_LOC(tab_t)
    IF (par_n) a%sc_t_dyn = C_LOC(tab_td)
    a%nob = nob
    a%maxblock_size = maxblock_size
    ALLOCATE(bend(nob)); bend = block_end(1:nob)
    a%block_end = C_LOC(bend)
    a%last_accepted_dt = C_LOC(last_accepted_dt)
    ALLOCATE(atol64(nspecies), rtol64(nspecies))
    atol64 = atol; rtol64 = rtol
    a%atol = C_LOC(atol64); a%rtol = C_LOC(rtol64)

    IF (solver == brams_solver_ros2) THEN
       rc = brams_chem_ros_dyndt(a)
    ELSE
       rc = brams_chem_rodas3_dyndt(a)
    END IF
    IF (rc < 0) THEN
       PRINT *, who, ': ', TRIM(last_error())
       STOP
    ELSE IF (rc > 0) THEN
       PRINT *, who, ': ', rc, ' block(s) hit a zero pivot or a non-finite state'
    END IF
  END SUBROUTINE brams_chem_host_step

  !========================================================================================
  ! The fields every solver reads: grid and interior bounds (from the block
  ! index of AllocIndex, j -> i -> k order), time step, photolysis method,
  ! constants, meteorology, J and the molar weights (w32 keeps the REAL
  ! copy alive for the call).
  SUBROUTINE fill_common(who, a, w32, m1, m2, m3, dtlt, pp, pi0, theta, rv, nspecies, nr, nr_photo, &
                         weight, PhotojMethod, nob, block_end, indexk, indexi, indexj, cp, cpor, p00, jphoto)
    CHARACTER(LEN=*) , INTENT(IN) :: who
    TYPE(brams_chem_args), INTENT(INOUT) :: a
    REAL(C_FLOAT), ALLOCATABLE, TARGET, INTENT(INOUT) :: w32(:)
    INTEGER          , INTENT(IN) :: m1, m2, m3
    REAL             , INTENT(IN) :: dtlt
    REAL             , INTENT(IN), TARGET :: pp(m1,m2,m3), pi0(m1,m2,m3), theta(m1,m2,m3), rv(m1,m2,m3)
    INTEGER          , INTENT(IN) :: nspecies, nr, nr_photo
    REAL             , INTENT(IN) :: weight(nspecies)
    CHARACTER(LEN=*) , INTENT(IN) :: PhotojMethod
    INTEGER          , INTENT(IN) :: nob
    INTEGER          , INTENT(IN) :: block_end(:), indexk(:,:), indexi(:,:), indexj(:,:)
    REAL             , INTENT(IN) :: cp, cpor, p00
    DOUBLE PRECISION , INTENT(IN), TARGET :: jphoto(nr_photo,m1,m2,m3)
    INTEGER :: lastb, lastc

    lastb = nob
    lastc = block_end(nob)
    a%m1 = m1; a%m2 = m2; a%m3 = m3
    a%ia = indexi(1,1);         a%ja = indexj(1,1)
    a%iz = indexi(lastc,lastb); a%jz = indexj(lastc,lastb)
    IF (indexk(1,1) /= 2 .OR. indexk(lastc,lastb) /= m1-1) THEN
       PRINT *, who, ': block index does not span k=2..m1-1'
       STOP
    END IF
    a%dtlt = dtlt
    SELECT CASE (TRIM(PhotojMethod))
    CASE ('FAST-TUV'); a%photoj_method = 0
    CASE ('FAST-JX');  a%photoj_method = 1
    CASE ('LUT');      a%photoj_method = 2
    CASE DEFAULT;      a%photoj_method = 0
    END SELECT
    a%cp = cp; a%cpor = cpor; a%p00 = p00
    a%pp = C_LOC(pp); a%pi0 = C_LOC(pi0); a%theta = C_LOC(theta); a%rv = C_LOC(rv)
    a%jphoto = C_LOC(jphoto)
    a%nspecies = nspecies; a%nr = nr; a%nr_photo = nr_photo
    ALLOCATE(w32(nspecies)); w32 = weight
    a%weight = C_LOC(w32)
  END SUBROUTINE fill_common

  !========================================================================================
  ! One QSSA call (chem_qssa) on the caller's host arrays: every species'
  ! sc_p is read (ppbm) and sc_t receives the tendency; itchim sub-steps.
  SUBROUTINE brams_chem_host_qssa(who, m1, m2, m3, itchim, dtlt, pp, pi0, theta, rv, jphoto, &
                                  nspecies, nr, nr_photo, weight, PhotojMethod, chem1_g, nob, &
                                  block_end, indexk, indexi, indexj, cp, cpor, p00)
    CHARACTER(LEN=*) , INTENT(IN) :: who
    INTEGER          , INTENT(IN) :: m1, m2, m3, itchim
    REAL             , INTENT(IN) :: dtlt
    REAL             , INTENT(IN), TARGET :: pp(m1,m2,m3), pi0(m1,m2,m3), theta(m1,m2,m3), rv(m1,m2,m3)
    INTEGER          , INTENT(IN) :: nspecies, nr, nr_photo
    DOUBLE PRECISION , INTENT(IN), TARGET :: jphoto(nr_photo,m1,m2,m3)
    REAL             , INTENT(IN) :: weight(nspecies)
    CHARACTER(LEN=*) , INTENT(IN) :: PhotojMethod
    TYPE (chem1_vars), INTENT(INOUT) :: chem1_g(nspecies)
    INTEGER          , INTENT(IN) :: nob
    INTEGER          , INTENT(IN) :: block_end(:), indexk(:,:), indexi(:,:), indexj(:,:)
    REAL             , INTENT(IN) :: cp, cpor, p00

    TYPE(brams_chem_args) :: a
    TYPE(C_PTR), ALLOCATABLE, TARGET :: tab_p(:), tab_t(:)
    REAL(C_FLOAT), ALLOCATABLE, TARGET :: w32(:)
    INTEGER :: n, rc

    CALL fill_common(who, a, w32, m1, m2, m3, dtlt, pp, pi0, theta, rv, nspecies, nr, nr_photo, &
                     weight, PhotojMethod, nob, block_end, indexk, indexi, indexj, cp, cpor, p00, jphoto)
    a%n_dyn_chem = 1
    a%split_method = 0
    ALLOCATE(tab_p(nspecies), tab_t(nspecies))
    DO n = 1, nspecies
       tab_p(n) = C_LOC(chem1_g(n)%sc_p(1,1,1))
       tab_t(n) = C_LOC(chem1_g(n)%sc_t(1))
    END DO
    a%sc_p = C_LOC(tab_p)
    a%sc_t = C_LOC(tab_t)
    a%sc_t_dyn = C_NULL_PTR
    a%nob = nob
    rc = brams_chem_qssa(a, INT(itchim, C_INT32_T))
    IF (rc < 0) THEN
       PRINT *, who, ': ', TRIM(last_error())
       STOP
    ELSE IF (rc > 0) THEN
       PRINT *, who, ': ', rc, ' cell(s) reached 1e20, tendencies left unchanged'
    END IF
  END SUBROUTINE brams_chem_host_qssa

  FUNCTION last_error() RESULT(msg)
    CHARACTER(LEN=512) :: msg
    TYPE(C_PTR) :: p
    CHARACTER(KIND=C_CHAR), POINTER :: s(:)
    INTEGER :: i, n
    msg = ' '
    p = brams_chem_last_error()
    n = INT(c_strlen(p))
    CALL C_F_POINTER(p, s, [n])
    DO i = 1, MIN(n, 512)
       msg(i:i) = s(i)
    END DO
  END FUNCTION last_error

END MODULE mod_brams_chem_host
