! Device-resident coupling for a BRAMS build that keeps its tracers in HBM.
!
! chemistry_driver (src/brams/ccatt/chemistry.f90:383-460) calls, per grid
! and chemistry step, chem_accum (ntask 1) every dynamics step when the split
! is PARALLEL with N_DYN_CHEM > 1, chem_rodas3_dyndt every N_DYN_CHEM-th
! step, and chem_accum (ntask 2) after it.  This module gives the same three
! calls on device arrays, so chem1_g%sc_p, sc_t, sc_t_dyn and J stay in device
! memory across timesteps; each hands its arguments to the C-ABI
! (include/brams_chem.h) through ISO_C_BINDING:
!   chem_rodas3_dyndt_device -> brams_chem_rodas3_dyndt_device
!   chem_accum_device        -> brams_chem_accum_device (chemistry.f90:528-604)
!   chem_tuv_jphoto_device   -> brams_chem_tuv_jphoto_device (the FAST-TUV output
!                               mapping into jphoto, ModTuvDriver2.7.f90:1302-1368)
! Device arrays are passed as C pointers: the species as one
! [nspecies][m1*m2*m3] REAL buffer each for sc_p, sc_t and sc_t_dyn.
MODULE mod_brams_chem_device

  USE iso_c_binding

  IMPLICIT NONE

  PRIVATE

  PUBLIC brams_chem_dev_args, chem_rodas3_dyndt_device, chem_accum_device, chem_tuv_jphoto_device

  INTEGER, PARAMETER, PUBLIC :: brams_split_symmetric = 0, brams_split_sequential = 1, &
                                brams_split_parallel = 2

  ! mirror of brams_chem_dev_args (include/brams_chem.h)
  TYPE, BIND(C) :: brams_chem_dev_args
     INTEGER(C_INT32_T) :: m1, m2, m3, ia, iz, ja, jz
     REAL(C_FLOAT)      :: dtlt
     INTEGER(C_INT32_T) :: n_dyn_chem, split_method, photoj_method
     REAL(C_FLOAT)      :: cp, cpor, p00
     TYPE(C_PTR)        :: pp, pi0, theta, rv
     TYPE(C_PTR)        :: jphoto
     TYPE(C_PTR)        :: weight, transported
     TYPE(C_PTR)        :: sc_p, sc_t, sc_t_dyn
     INTEGER(C_INT32_T) :: nob
     TYPE(C_PTR)        :: last_accepted_dt
     TYPE(C_PTR)        :: atol, rtol
     TYPE(C_PTR)        :: stream
  END TYPE brams_chem_dev_args

  ! mirror of brams_tuv_map_args (include/brams_chem.h)
  TYPE, BIND(C) :: brams_tuv_map_args
     INTEGER(C_INT32_T) :: m1, m2, m3, ia, iz, ja, jz
     INTEGER(C_INT32_T) :: kj, kz
     TYPE(C_PTR)        :: valj
     TYPE(C_PTR)        :: xref
     INTEGER(C_INT32_T) :: nr_photo
     TYPE(C_PTR)        :: jphoto
     TYPE(C_PTR)        :: stream
  END TYPE brams_tuv_map_args

  INTERFACE
     FUNCTION brams_chem_tuv_jphoto_device(args) BIND(C, name='brams_chem_tuv_jphoto_device') RESULT(rc)
       IMPORT :: brams_tuv_map_args, C_INT
       TYPE(brams_tuv_map_args), INTENT(IN) :: args
       INTEGER(C_INT) :: rc
     END FUNCTION brams_chem_tuv_jphoto_device
     FUNCTION brams_chem_rodas3_dyndt_device(args) BIND(C, name='brams_chem_rodas3_dyndt_device') RESULT(rc)
       IMPORT :: brams_chem_dev_args, C_INT
       TYPE(brams_chem_dev_args), INTENT(IN) :: args
       INTEGER(C_INT) :: rc
     END FUNCTION brams_chem_rodas3_dyndt_device
     FUNCTION brams_chem_accum_device(m1, m2, m3, ntask, n_dyn_chem, n_transported, transp_chem_index, &
                                      sc_t_dyn, sc_t, stream) BIND(C, name='brams_chem_accum_device') RESULT(rc)
       IMPORT :: C_INT, C_INT32_T, C_PTR
       INTEGER(C_INT32_T), VALUE :: m1, m2, m3, ntask, n_dyn_chem, n_transported
       INTEGER(C_INT32_T), INTENT(IN) :: transp_chem_index(*)
       TYPE(C_PTR), VALUE :: sc_t_dyn, sc_t, stream
       INTEGER(C_INT) :: rc
     END FUNCTION brams_chem_accum_device
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
  ! One chemistry step on device arrays (the call of chemistry.f90:443-453).
  ! nfailed: blocks that hit a zero pivot or a non-finite state, where the
  ! reference's Sparse1.3 would stop the run; they keep their last accepted
  ! state.  A fatal error (invalid arguments, a HIP failure) STOPs.
  SUBROUTINE chem_rodas3_dyndt_device(args, nfailed)
    TYPE(brams_chem_dev_args), INTENT(IN) :: args
    INTEGER, INTENT(OUT) :: nfailed
    INTEGER(C_INT) :: rc
    rc = brams_chem_rodas3_dyndt_device(args)
    IF (rc < 0) CALL fatal('chem_rodas3_dyndt_device', rc)
    nfailed = rc
  END SUBROUTINE chem_rodas3_dyndt_device

  !========================================================================================
  ! chem_accum (chemistry.f90:528-604) with the same arguments except that the
  ! species are device buffers: ntask 1 adds REAL(1./N_DYN_CHEM)*sc_t to
  ! sc_t_dyn for the transported species, ntask 2 sets sc_t_dyn to zero.
  SUBROUTINE chem_accum_device(m1, m2, m3, ntask, nspecies_chem_transported, transp_chem_index, &
                               N_DYN_CHEM, sc_t_dyn, sc_t, stream)
    INTEGER, INTENT(IN) :: m1, m2, m3, ntask, nspecies_chem_transported, N_DYN_CHEM
    INTEGER, INTENT(IN) :: transp_chem_index(nspecies_chem_transported)
    TYPE(C_PTR), INTENT(IN) :: sc_t_dyn, sc_t, stream
    INTEGER(C_INT32_T) :: idx(nspecies_chem_transported)
    INTEGER(C_INT) :: rc
    idx = INT(transp_chem_index, C_INT32_T)
    rc = brams_chem_accum_device(INT(m1, C_INT32_T), INT(m2, C_INT32_T), INT(m3, C_INT32_T), &
                                 INT(ntask, C_INT32_T), INT(N_DYN_CHEM, C_INT32_T), &
                                 INT(nspecies_chem_transported, C_INT32_T), idx, sc_t_dyn, sc_t, stream)
    IF (rc < 0) CALL fatal('chem_accum_device', rc)
  END SUBROUTINE chem_accum_device

  !========================================================================================
  ! TUV's j-values to jphoto on the device (ModTuvDriver2.7.f90:1302-1332) and
  ! RELACS's J14/J16 relations (:1363-1368): valjloc is the device copy of
  ! valJLoc(ia:iz,ja:jz,kj,kz) (double), xref TUV's cross-reference (host;
  ! 1-based jphoto index of reaction k, 0 = unused), jphoto the device
  ! (nr_photo,m1,m2,m3) array, updated in place.
  SUBROUTINE chem_tuv_jphoto_device(m1, m2, m3, ia, iz, ja, jz, kj, kz, valjloc, xref, nr_photo, &
                                    jphoto, stream)
    INTEGER, INTENT(IN) :: m1, m2, m3, ia, iz, ja, jz, kj, kz, nr_photo
    INTEGER, INTENT(IN) :: xref(kj)
    TYPE(C_PTR), INTENT(IN) :: valjloc, jphoto, stream
    TYPE(brams_tuv_map_args) :: a
    INTEGER(C_INT32_T), TARGET :: x(kj)
    INTEGER(C_INT) :: rc
    x = INT(xref, C_INT32_T)
    a%m1 = m1; a%m2 = m2; a%m3 = m3; a%ia = ia; a%iz = iz; a%ja = ja; a%jz = jz
    a%kj = kj; a%kz = kz
    a%valj = valjloc
    a%xref = C_LOC(x)
    a%nr_photo = nr_photo
    a%jphoto = jphoto
    a%stream = stream
    rc = brams_chem_tuv_jphoto_device(a)
    IF (rc < 0) CALL fatal('chem_tuv_jphoto_device', rc)
  END SUBROUTINE chem_tuv_jphoto_device

  !========================================================================================
  SUBROUTINE fatal(who, rc)
    CHARACTER(LEN=*), INTENT(IN) :: who
    INTEGER(C_INT), INTENT(IN) :: rc
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
    PRINT *, who, ': error ', rc, ': ', TRIM(msg)
    STOP
  END SUBROUTINE fatal

END MODULE mod_brams_chem_device
