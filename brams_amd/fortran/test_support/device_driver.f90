! Test driver of the device-resident coupling (mod_brams_chem_device): reads a
! grid, its state and the dynamic tendencies of every dynamics step from a
! stream file written by tests/test_gpu_fortran_device.py, keeps everything in
! device memory, maps TUV j-values into jphoto when the file holds them
! (kj > 0, chem_tuv_jphoto_device), and runs the PARALLEL-split sequence of
! chemistry.f90:383-460 for nchem chemistry steps of N_DYN_CHEM dynamics
! steps each:
!   every dynamics step:      sc_t <- that step's tendency; chem_accum ntask 1
!   every N_DYN_CHEM-th step: chem_rodas3_dyndt; chem_accum ntask 2
! then writes sc_p and last_accepted_dt.  Device memory comes from the HIP
! runtime through ISO_C_BINDING.
PROGRAM device_driver
  USE iso_c_binding
  USE mod_brams_chem_device
  IMPLICIT NONE
  INTERFACE
     FUNCTION hipMalloc(p, n) BIND(C, name='hipMalloc') RESULT(rc)
       IMPORT :: C_PTR, C_SIZE_T, C_INT
       TYPE(C_PTR) :: p
       INTEGER(C_SIZE_T), VALUE :: n
       INTEGER(C_INT) :: rc
     END FUNCTION hipMalloc
     FUNCTION hipMemcpy(dst, src, n, kind) BIND(C, name='hipMemcpy') RESULT(rc)
       IMPORT :: C_PTR, C_SIZE_T, C_INT
       TYPE(C_PTR), VALUE :: dst, src
       INTEGER(C_SIZE_T), VALUE :: n
       INTEGER(C_INT), VALUE :: kind
       INTEGER(C_INT) :: rc
     END FUNCTION hipMemcpy
     FUNCTION hipFree(p) BIND(C, name='hipFree') RESULT(rc)
       IMPORT :: C_PTR, C_INT
       TYPE(C_PTR), VALUE :: p
       INTEGER(C_INT) :: rc
     END FUNCTION hipFree
     FUNCTION hipDeviceSynchronize() BIND(C, name='hipDeviceSynchronize') RESULT(rc)
       IMPORT :: C_INT
       INTEGER(C_INT) :: rc
     END FUNCTION hipDeviceSynchronize
  END INTERFACE
  INTEGER(C_INT), PARAMETER :: h2d = 1, d2h = 2
  INTEGER :: m1, m2, m3, ia, iz, ja, jz, nob, nsp, ntr, ndyn, nchem, c, d, u, nf, n3, kj, kz
  REAL :: dtlt, cp, cpor, p00
  REAL(C_FLOAT), ALLOCATABLE, TARGET :: pp(:), pi0(:), theta(:), rv(:), weight(:), sc_p(:), tend(:), zero(:)
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: jphoto(:), last_dt(:), atol(:), rtol(:), valj(:)
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: transported(:)
  INTEGER, ALLOCATABLE :: tidx(:), xref(:)
  TYPE(C_PTR) :: d_pp, d_pi0, d_theta, d_rv, d_jph, d_w, d_tr, d_scp, d_sct, d_sctd, d_dt, d_atol, d_rtol
  TYPE(C_PTR) :: d_valj
  TYPE(brams_chem_dev_args) :: a
  CHARACTER(LEN=256) :: fin, fout

  CALL get_command_argument(1, fin)
  CALL get_command_argument(2, fout)
  OPEN(newunit=u, file=TRIM(fin), access='stream', form='unformatted', status='old')
  READ(u) m1, m2, m3, ia, iz, ja, jz, nob, nsp, ntr, ndyn, nchem, kj, kz
  READ(u) dtlt, cp, cpor, p00
  n3 = m1 * m2 * m3
  ALLOCATE(pp(n3), pi0(n3), theta(n3), rv(n3), weight(nsp), transported(nsp), jphoto(17 * n3))
  ALLOCATE(last_dt(nob), atol(nsp), rtol(nsp), tidx(ntr), sc_p(nsp * n3), tend(nsp * n3), zero(nsp * n3))
  READ(u) pp, pi0, theta, rv, weight, transported, jphoto, last_dt, atol, rtol, tidx, sc_p
  zero = 0.0

  d_pp = upload(C_LOC(pp), 4_C_SIZE_T * n3)
  d_pi0 = upload(C_LOC(pi0), 4_C_SIZE_T * n3)
  d_theta = upload(C_LOC(theta), 4_C_SIZE_T * n3)
  d_rv = upload(C_LOC(rv), 4_C_SIZE_T * n3)
  d_jph = upload(C_LOC(jphoto), 8_C_SIZE_T * 17 * n3)
  d_w = upload(C_LOC(weight), 4_C_SIZE_T * nsp)
  d_tr = upload(C_LOC(transported), 4_C_SIZE_T * nsp)
  d_scp = upload(C_LOC(sc_p), 4_C_SIZE_T * nsp * n3)
  d_sct = upload(C_LOC(zero), 4_C_SIZE_T * nsp * n3)
  d_sctd = upload(C_LOC(zero), 4_C_SIZE_T * nsp * n3)
  d_dt = upload(C_LOC(last_dt), 8_C_SIZE_T * nob)
  d_atol = upload(C_LOC(atol), 8_C_SIZE_T * nsp)
  d_rtol = upload(C_LOC(rtol), 8_C_SIZE_T * nsp)

  a%m1 = m1; a%m2 = m2; a%m3 = m3; a%ia = ia; a%iz = iz; a%ja = ja; a%jz = jz
  a%dtlt = dtlt; a%n_dyn_chem = ndyn; a%split_method = brams_split_parallel; a%photoj_method = 0
  a%cp = cp; a%cpor = cpor; a%p00 = p00
  a%pp = d_pp; a%pi0 = d_pi0; a%theta = d_theta; a%rv = d_rv; a%jphoto = d_jph
  a%weight = d_w; a%transported = d_tr
  a%sc_p = d_scp; a%sc_t = d_sct; a%sc_t_dyn = d_sctd
  a%nob = nob; a%last_accepted_dt = d_dt; a%atol = d_atol; a%rtol = d_rtol
  a%stream = C_NULL_PTR

  IF (kj > 0) THEN                 ! TUV j-values -> jphoto on the device
     ALLOCATE(valj((iz - ia + 1) * (jz - ja + 1) * kj * kz), xref(kj))
     READ(u) valj, xref
     d_valj = upload(C_LOC(valj), 8_C_SIZE_T * SIZE(valj))
     CALL chem_tuv_jphoto_device(m1, m2, m3, ia, iz, ja, jz, kj, kz, d_valj, xref, 17, d_jph, C_NULL_PTR)
     CALL check(hipFree(d_valj), 'hipFree')
  END IF

  DO c = 1, nchem
     DO d = 1, ndyn
        READ(u) tend
        CALL check(hipMemcpy(d_sct, C_LOC(tend), 4_C_SIZE_T * nsp * n3, h2d), 'hipMemcpy tendency')
        CALL chem_accum_device(m1, m2, m3, 1, ntr, tidx, ndyn, d_sctd, d_sct, C_NULL_PTR)
     END DO
     CALL chem_rodas3_dyndt_device(a, nf)
     IF (nf /= 0) PRINT *, 'device_driver: ', nf, ' block(s) failed in chemistry step ', c
     CALL chem_accum_device(m1, m2, m3, 2, ntr, tidx, ndyn, d_sctd, C_NULL_PTR, C_NULL_PTR)
  END DO
  CLOSE(u)
  CALL check(hipDeviceSynchronize(), 'hipDeviceSynchronize')
  CALL check(hipMemcpy(C_LOC(sc_p), d_scp, 4_C_SIZE_T * nsp * n3, d2h), 'hipMemcpy sc_p')
  CALL check(hipMemcpy(C_LOC(last_dt), d_dt, 8_C_SIZE_T * nob, d2h), 'hipMemcpy last_accepted_dt')
  CALL check(hipMemcpy(C_LOC(tend), d_sctd, 4_C_SIZE_T * nsp * n3, d2h), 'hipMemcpy sc_t_dyn')
  OPEN(newunit=u, file=TRIM(fout), access='stream', form='unformatted', status='replace')
  CALL check(hipMemcpy(C_LOC(jphoto), d_jph, 8_C_SIZE_T * 17 * n3, d2h), 'hipMemcpy jphoto')
  WRITE(u) sc_p, last_dt, tend, jphoto
  CLOSE(u)
  CALL check(hipFree(d_pp), 'hipFree')
  CALL check(hipFree(d_pi0), 'hipFree')
  CALL check(hipFree(d_theta), 'hipFree')
  CALL check(hipFree(d_rv), 'hipFree')
  CALL check(hipFree(d_jph), 'hipFree')
  CALL check(hipFree(d_w), 'hipFree')
  CALL check(hipFree(d_tr), 'hipFree')
  CALL check(hipFree(d_scp), 'hipFree')
  CALL check(hipFree(d_sct), 'hipFree')
  CALL check(hipFree(d_sctd), 'hipFree')
  CALL check(hipFree(d_dt), 'hipFree')
  CALL check(hipFree(d_atol), 'hipFree')
  CALL check(hipFree(d_rtol), 'hipFree')

CONTAINS

  FUNCTION upload(src, n) RESULT(p)
    TYPE(C_PTR), INTENT(IN) :: src
    INTEGER(C_SIZE_T), INTENT(IN) :: n
    TYPE(C_PTR) :: p
    p = C_NULL_PTR
    CALL check(hipMalloc(p, n), 'hipMalloc')
    CALL check(hipMemcpy(p, src, n, h2d), 'hipMemcpy')
  END FUNCTION upload

  SUBROUTINE check(rc, what)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: what
    IF (rc /= 0) THEN
       PRINT *, 'device_driver: ', what, ' returned ', rc
       STOP 1
    END IF
  END SUBROUTINE check

END PROGRAM device_driver
