! Drop-in replacement of BRAMS's RODAS3 chemistry driver for MI355X.
!
! Same module name, same PUBLIC routine and the same argument list, types,
! species order and units as the reference
!   src/brams/ccatt/chem_spack_rodas3_dyndt.f90:1-131 (MODULE mod_chem_spack_rodas3_dyndt,
!   SUBROUTINE chem_rodas3_dyndt), called from src/brams/ccatt/chemistry.f90:443-453.
! The body hands the caller's arrays to the C-ABI entry point
! brams_chem_rodas3_dyndt (include/brams_chem.h) through mod_brams_chem_host;
! the integration itself runs in the library's HIP kernels.  Fatal errors
! STOP with the library's message; blocks that hit a zero pivot (where the
! reference's Sparse1.3 would abort the run) are reported and left at their
! last accepted state.
!
! Arguments the RELACS_TUV path never reads (dn0, rcp, cosz, att, spack,
! spack_2d, extra3d, get_non_zeros, maxnspecies, kij_index) are kept in the
! signature and ignored, as in the reference.
MODULE mod_chem_spack_rodas3_dyndt

  USE mem_chem1, ONLY: &
       chem1_vars        ! Type

  USE mem_spack, ONLY: &
       spack_type,     & ! Type
       spack_type_2d     ! Type

  USE Extras, ONLY: &
       ext3d             ! Type

  USE mod_brams_chem_host, ONLY: &
       brams_chem_host_step, brams_solver_rodas3

  IMPLICIT NONE

  PRIVATE

  PUBLIC chem_rodas3_dyndt

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

    CALL brams_chem_host_step('chem_rodas3_dyndt', brams_solver_rodas3, m1, m2, m3, dtlt, &
                              pp, pi0, theta, rv, nspecies, nr, nr_photo, weight, PhotojMethod, &
                              nspecies_chem_transported, nspecies_chem_no_transported, &
                              transp_chem_index, no_transp_chem_index, chem1_g, nob, maxblock_size, &
                              block_end, indexk, indexi, indexj, last_accepted_dt, atol, rtol, &
                              N_DYN_CHEM, split_method, cp, cpor, p00, jphoto)
  END SUBROUTINE chem_rodas3_dyndt


END MODULE mod_chem_spack_rodas3_dyndt
