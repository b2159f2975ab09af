! Minimal copies of the three BRAMS derived types that appear in the
! interface of chem_rodas3_dyndt, so the shim can be compiled and exercised
! outside a BRAMS build (test harness only).  Inside BRAMS the real modules
! are used:
!   mem_chem1 :: chem1_vars       src/brams/ccatt/mem_chem1.f90:17-23
!   mem_spack :: spack_type(_2d)  src/brams/ccatt/mem_spack.f90:16-66 (unused by the shim)
!   Extras    :: ext3d            src/brams/memory/extra.f90:16-18   (unused by the shim)
MODULE mem_chem1
  IMPLICIT NONE
  TYPE chem1_vars
     REAL, POINTER, DIMENSION(:,:,:)  :: sc_p, sc_pp, sc_pf
     REAL, POINTER, DIMENSION(:,:  )  :: sc_dd, sc_wd
     REAL, POINTER, DIMENSION(:    )  :: sc_t, sc_t_dyn
  END TYPE chem1_vars
END MODULE mem_chem1

MODULE mem_spack
  IMPLICIT NONE
  TYPE, PUBLIC :: spack_type
     DOUBLE PRECISION, POINTER, DIMENSION(:,:) :: sc_p
  END TYPE spack_type
  TYPE, PUBLIC :: spack_type_2d
     DOUBLE PRECISION, POINTER, DIMENSION(:,:) :: DLmat
  END TYPE spack_type_2d
END MODULE mem_spack

MODULE Extras
  IMPLICIT NONE
  TYPE ext3d
     REAL, POINTER :: d3(:,:,:)
  END TYPE ext3d
END MODULE Extras
