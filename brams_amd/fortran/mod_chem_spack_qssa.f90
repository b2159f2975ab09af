! Drop-in replacement of BRAMS's QSSA chemistry driver (CHEMISTRY=1) for
! MI355X.
!
! Same module name, the same PUBLIC chem_qssa with the same argument list,
! types, species order and units as the reference
!   src/brams/ccatt/chem_spack_qssa.f90:37-43 (MODULE mod_chem_spack_qssa,
!   SUBROUTINE chem_qssa), called from src/brams/ccatt/chemistry.f90:399-407.
! The body hands the caller's arrays to the C-ABI entry point brams_chem_qssa
! (include/brams_chem.h) through mod_brams_chem_host; the integration runs in
! the library's HIP kernel.  chem1_g(n)%sc_t receives every species'
! tendency; sc_p is not changed.  Fatal errors STOP with the library's
! message; cells where a species reaches 1e20 (where the reference's solveur
! STOPs) are reported and keep their tendency.
!
! Arguments the RELACS_TUV path never reads (nrmax, mynum, cosz, maxblock_size,
! kij_index, att) are kept in the signature and ignored.  Of the reference
! module's other PUBLIC names (solveur, cvmgm_dp, cvmgp_dp, cvmgz_dp) nothing
! outside it makes use (chemistry.f90:198-199 imports chem_qssa only), so
! they are not provided.
MODULE mod_chem_spack_qssa

  USE mem_chem1, ONLY: &
       chem1_vars ! Type

  USE mod_brams_chem_host, ONLY: &
       brams_chem_host_qssa

  IMPLICIT NONE

  PRIVATE

  PUBLIC :: chem_qssa

CONTAINS

  !========================================================================================
  SUBROUTINE chem_qssa(m1,m2,m3,itchim,nrmax,mynum,dtlt,pp,pi0,theta,rv, &
                       cosz,jphoto,nspecies,nr_photo,weight,PhotoJMethod,chem1_g, &
                       nob,maxblock_size, &
                       block_end,indexk,indexi,indexj,kij_index, &
                       cp,cpor,p00,att)

    REAL             , INTENT(IN)    :: dtlt
    INTEGER          , INTENT(IN)    :: mynum
    INTEGER          , INTENT(IN)    :: m1
    INTEGER          , INTENT(IN)    :: m2
    INTEGER          , INTENT(IN)    :: m3
    REAL             , INTENT(IN)    :: pp(m1,m2,m3)
    REAL             , INTENT(IN)    :: pi0(m1,m2,m3)
    REAL             , INTENT(IN)    :: theta(m1,m2,m3)
    REAL             , INTENT(IN)    :: rv(m1,m2,m3)
    INTEGER          , INTENT(IN)    :: itchim
    REAL             , INTENT(IN)    :: cosz(m2,m3)
    INTEGER          , INTENT(IN)    :: nspecies
    INTEGER          , INTENT(IN)    :: nr_photo
    INTEGER          , INTENT(IN)    :: nrmax
    REAL             , INTENT(IN)    :: weight(nspecies)
    CHARACTER(LEN=10), INTENT(IN)    :: PhotojMethod
    DOUBLE PRECISION , INTENT(IN)    :: jphoto(nr_photo,m1,m2,m3)
    TYPE (chem1_vars), INTENT(INOUT) :: chem1_g(nspecies)
    INTEGER          , INTENT(IN)    :: nob
    INTEGER          , INTENT(IN)    :: maxblock_size
    INTEGER          , INTENT(IN)    :: block_end(:)
    INTEGER          , INTENT(IN)    :: indexk(:,:)
    INTEGER          , INTENT(IN)    :: indexi(:,:)
    INTEGER          , INTENT(IN)    :: indexj(:,:)
    INTEGER          , INTENT(IN)    :: kij_index(:,:)
    REAL             , INTENT(IN)    :: cp
    REAL             , INTENT(IN)    :: cpor
    REAL             , INTENT(IN)    :: p00
    REAL             , INTENT(IN)    :: att(m2,m3)

    CALL brams_chem_host_qssa('chem_qssa', m1, m2, m3, itchim, dtlt, pp, pi0, theta, rv, jphoto, &
                              nspecies, nrmax, nr_photo, weight, PhotojMethod, chem1_g, nob, &
                              block_end, indexk, indexi, indexj, cp, cpor, p00)
  END SUBROUTINE chem_qssa

END MODULE mod_chem_spack_qssa
