! Drop-in replacement of BRAMS's RODAS3 chemistry driver for MI355X.
!
! Same module name, same PUBLIC routine and the same argument list, types,
! species order and units as the reference
!   src/brams/ccatt/chem_spack_rodas3_dyndt.f90:1-131 (MODULE mod_chem_spack_rodas3_dyndt,
!   SUBROUTINE chem_rodas3_dyndt), called from src/brams/ccatt/chemistry.f90:443-453.
! The body hands the caller's arrays to the C-ABI entry point
! brams_chem_rodas3_dyndt (include/brams_chem.h) through ISO_C_BINDING; the
! integration itself runs in the library's HIP kernels.  The library returns
! a status instead of aborting: fatal errors STOP here with the library's
! message, blocks that hit a zero pivot (where the reference's Sparse1.3
! would abort the run) are reported and left unchanged for the caller.
!
! Arguments the RELACS_TUV path never reads (dn0, rcp, cosz, att, spack,
! spack_2d, extra3d, get_non_zeros, maxnspecies) are kept in the signature
! and ignored, as in the reference.
MODULE mod_chem_spack_rodas3_dyndt

  USE mem_chem1, ONLY: &
       chem1_vars        ! Type

  USE mem_spack, ONLY: &
       spack_type,     & ! Type
       spack_type_2d     ! Type

  USE Extras, ONLY: &
       ext3d             ! Type

  USE iso_c_binding

  IMPLICIT NONE

  PRIVATE

  PUBLIC chem_rodas3_dyndt

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
     FUNCTION brams_chem_init(device) BIND(C, name='brams_chem_init') RESULT(rc)
       IMPORT :: C_INT
       INTEGER(C_INT), VALUE :: device
       INTEGER(C_INT) :: rc
     END FUNCTION brams_chem_init
     FUNCTION brams_chem_rodas3_dyndt(args) BIND(C, name='brams_chem_rodas3_dyndt') RESULT(rc)
       IMPORT :: brams_chem_args, C_INT
       TYPE(brams_chem_args), INTENT(IN) :: args
       INTEGER(C_INT) :: rc
     END FUNCTION brams_chem_rodas3_dyndt
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
  SUBROUTINE chem_rodas3_dyndt(m1,m2,m3,dtlt,pp,pi0,theta,rv,dn0,cosz,rcp, &
                               nspecies,nr,nr_photo,weight,PhotojMethod,maxnspecies, &
                               nspecies_chem_transported,nspecies_chem_no_transported, &
                               transp_chem_index,no_transp_chem_index,chem1_g,nob,maxblock_size, &
                               block_end,indexk,indexi,indexj,kij_index,last_accepted_dt, &
                               spack,spack_2d,atol,rtol,get_non_zeros,&
                               N_DYN_CHEM,split_method,cp,cpor,p00,na_extra3d,extra3d,&
                               jphoto,att)

    INTEGER , INTENT(IN) :: m1
    INTEGER , INTENT(IN) :: m2
    INTEGER , INTENT(IN) :: m3
    REAL             , INTENT(IN) :: dtlt
    REAL             , INTENT(IN), TARGET :: pp(m1,m2,m3)
    REAL             , INTENT(IN), TARGET :: pi0(m1,m2,m3)
    REAL             , INTENT(IN), TARGET :: theta(m1,m2,m3)
    REAL             , INTENT(IN), TARGET :: rv(m1,m2,m3)
    REAL             , INTENT(IN) :: dn0(m1,m2,m3)
    REAL             , INTENT(IN) :: cosz(m2,m3)
    REAL             , INTENT(IN) :: rcp(m1,m2,m3)
    INTEGER          , INTENT(IN) :: nspecies
    INTEGER          , INTENT(IN) :: nr_photo
    INTEGER          , INTENT(IN) :: nr
    REAL             , INTENT(IN) :: weight(nspecies)
    CHARACTER(LEN=10), INTENT(IN) :: PhotojMethod
    INTEGER          , INTENT(IN) :: maxnspecies
    INTEGER          , INTENT(IN) :: nspecies_chem_transported
    INTEGER          , INTENT(IN) :: nspecies_chem_no_transported
    INTEGER          , INTENT(IN) :: transp_chem_index(maxnspecies)
    INTEGER          , INTENT(IN) :: no_transp_chem_index(maxnspecies)
    CHARACTER(LEN=20), INTENT(IN) :: split_method
    INTEGER          , INTENT(IN) :: N_DYN_CHEM
    TYPE (chem1_vars), INTENT(INOUT) :: chem1_g(nspecies)
    INTEGER         ,   INTENT(IN)    :: nob
    INTEGER         ,   INTENT(IN)    :: maxblock_size
    INTEGER         ,   INTENT(IN)    :: block_end(:)
    INTEGER         ,   INTENT(IN)    :: indexk(:,:)
    INTEGER         ,   INTENT(IN)    :: indexi(:,:)
    INTEGER         ,   INTENT(IN)    :: indexj(:,:)
    INTEGER         ,   INTENT(IN)    :: kij_index(:,:)
    DOUBLE PRECISION,   INTENT(INOUT), TARGET :: last_accepted_dt(:)
    TYPE(spack_type)   , INTENT(INOUT) :: spack(1)
    TYPE(spack_type_2d), INTENT(INOUT) :: spack_2d(:,:)
    DOUBLE PRECISION   , INTENT(IN)    :: atol(nspecies)
    DOUBLE PRECISION   , INTENT(IN)    :: rtol(nspecies)
    LOGICAL , INTENT(INOUT) :: get_non_zeros
    REAL, INTENT(IN) :: cp
    REAL, INTENT(IN) :: cpor
    REAL, INTENT(IN) :: p00
    INTEGER,     INTENT(IN)    :: na_extra3d
    TYPE(ext3d), INTENT(INOUT) :: extra3d(na_extra3d)
    DOUBLE PRECISION, INTENT(IN), TARGET :: jphoto(nr_photo,m1,m2,m3)
    REAL, INTENT(IN) :: att(m2,m3)

    TYPE(brams_chem_args) :: a
    TYPE(C_PTR), ALLOCATABLE, TARGET :: tab_p(:), tab_t(:), tab_td(:)
    REAL(C_FLOAT), ALLOCATABLE, TARGET :: w32(:)
    INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: tidx(:), nidx(:), bend(:)
    DOUBLE PRECISION, ALLOCATABLE, TARGET :: atol64(:), rtol64(:), dt64(:)
    INTEGER :: n, rc, lastb, lastc
    LOGICAL :: par, par_n

    ! interior columns from the block index of AllocIndex (j -> i -> k order)
    lastb = nob
    lastc = block_end(nob)
    a%m1 = m1; a%m2 = m2; a%m3 = m3
    a%ia = indexi(1,1);         a%ja = indexj(1,1)
    a%iz = indexi(lastc,lastb); a%jz = indexj(lastc,lastb)
    IF (indexk(1,1) /= 2 .OR. indexk(lastc,lastb) /= m1-1) THEN
       PRINT *, 'chem_rodas3_dyndt: block index does not span k=2..m1-1'
       STOP
    END IF
    a%dtlt = dtlt
    a%n_dyn_chem = N_DYN_CHEM
    SELECT CASE (TRIM(split_method))
    CASE ('SYMMETRIC');  a%split_method = 0
    CASE ('SEQUENTIAL'); a%split_method = 1
    CASE ('PARALLEL');   a%split_method = 2
    CASE DEFAULT
       PRINT *, 'chem_rodas3_dyndt: unknown split_method ', split_method
       STOP
    END SELECT
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

    rc = brams_chem_rodas3_dyndt(a)
    IF (rc < 0) THEN
       PRINT *, 'chem_rodas3_dyndt: ', last_error()
       STOP
    ELSE IF (rc > 0) THEN
       PRINT *, 'chem_rodas3_dyndt: ', rc, ' block(s) hit a zero pivot or a non-finite state'
    END IF
  END SUBROUTINE chem_rodas3_dyndt

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

END MODULE mod_chem_spack_rodas3_dyndt
