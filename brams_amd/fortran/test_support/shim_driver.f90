! Test driver of the Fortran drop-ins: reads one chemistry call from a stream
! file written by tests/test_gpu_fortran_shim.py, calls chem_rodas3_dyndt
! exactly as chemistry.f90:443-453 does (with a third argument "ros2",
! chem_ros_dyndt as chemistry.f90:430-440 does; with "qssa", chem_qssa as
! chemistry.f90:399-407 does, itchim 75), and writes the updated state (sc_p,
! or for qssa the tendencies sc_t) and last_accepted_dt.
PROGRAM shim_driver
  USE mem_chem1, ONLY: chem1_vars
  USE mem_spack, ONLY: spack_type, spack_type_2d
  USE Extras, ONLY: ext3d
  USE mod_chem_spack_rodas3_dyndt, ONLY: chem_rodas3_dyndt
  USE mod_chem_spack_ros_dyndt, ONLY: chem_ros_dyndt
  USE mod_chem_spack_qssa, ONLY: chem_qssa
  IMPLICIT NONE
  INTEGER :: m1, m2, m3, nob, mbs, nsp, nr, nrp, ntr, nnt, ndyn, n, u
  REAL :: dtlt, cp, cpor, p00
  REAL, ALLOCATABLE :: pp(:,:,:), pi0(:,:,:), theta(:,:,:), rv(:,:,:), dn0(:,:,:), rcp(:,:,:)
  REAL, ALLOCATABLE :: cosz(:,:), att(:,:), weight(:)
  DOUBLE PRECISION, ALLOCATABLE :: jphoto(:,:,:,:), last_dt(:), atol(:), rtol(:)
  INTEGER, ALLOCATABLE :: tidx(:), nidx(:), block_end(:), ik(:,:), ii(:,:), ij(:,:), kij(:,:)
  TYPE(chem1_vars), ALLOCATABLE :: chem1_g(:)
  TYPE(spack_type) :: spack(1)
  TYPE(spack_type_2d) :: spack_2d(1,1)
  TYPE(ext3d) :: extra3d(1)
  LOGICAL :: gnz
  CHARACTER(LEN=256) :: fin, fout, solver
  CHARACTER(LEN=10) :: photoj
  CHARACTER(LEN=20) :: split

  CALL get_command_argument(1, fin)
  CALL get_command_argument(2, fout)
  solver = 'rodas3'
  IF (command_argument_count() >= 3) CALL get_command_argument(3, solver)
  OPEN(newunit=u, file=TRIM(fin), access='stream', form='unformatted', status='old')
  READ(u) m1, m2, m3, nob, mbs, nsp, nr, nrp, ntr, nnt, ndyn
  READ(u) dtlt, cp, cpor, p00
  ALLOCATE(pp(m1,m2,m3), pi0(m1,m2,m3), theta(m1,m2,m3), rv(m1,m2,m3))
  ALLOCATE(dn0(m1,m2,m3), rcp(m1,m2,m3), cosz(m2,m3), att(m2,m3))
  ALLOCATE(weight(nsp), jphoto(nrp,m1,m2,m3), last_dt(nob), atol(nsp), rtol(nsp))
  ALLOCATE(tidx(200), nidx(200), block_end(nob), ik(mbs,nob), ii(mbs,nob), ij(mbs,nob), kij(mbs,nob))
  tidx = 0; nidx = 0; dn0 = 1.0; rcp = 0.0; cosz = 0.0; att = 1.0
  READ(u) pp, pi0, theta, rv, weight, jphoto, last_dt, atol, rtol
  READ(u) tidx(1:ntr), nidx(1:nnt), block_end, ik, ii, ij
  kij = 0
  ALLOCATE(chem1_g(nsp))
  DO n = 1, nsp
     ALLOCATE(chem1_g(n)%sc_p(m1,m2,m3))
     READ(u) chem1_g(n)%sc_p
  END DO
  CLOSE(u)
  photoj = 'FAST-TUV'
  split = 'SYMMETRIC'
  gnz = .FALSE.
  IF (TRIM(solver) == 'qssa') THEN
     DO n = 1, nsp
        ALLOCATE(chem1_g(n)%sc_t(m1*m2*m3))
        chem1_g(n)%sc_t = 0.0
     END DO
     CALL chem_qssa(m1,m2,m3,75,nr,0,dtlt,pp,pi0,theta,rv, &
                    cosz,jphoto,nsp,nrp,weight,photoj,chem1_g, &
                    nob,mbs,block_end,ik,ii,ij,kij,cp,cpor,p00,att)
  ELSE IF (TRIM(solver) == 'ros2') THEN
     CALL chem_ros_dyndt(m1,m2,m3,dtlt,pp,pi0,theta,rv,dn0,cosz,rcp, &
                         nsp,nr,nrp,weight,photoj,200,ntr,nnt,tidx,nidx,chem1_g,nob,mbs, &
                         block_end,ik,ii,ij,kij,last_dt,spack,spack_2d,atol,rtol,gnz, &
                         ndyn,split,cp,cpor,p00,jphoto,att)
  ELSE
     CALL chem_rodas3_dyndt(m1,m2,m3,dtlt,pp,pi0,theta,rv,dn0,cosz,rcp, &
                            nsp,nr,nrp,weight,photoj,200,ntr,nnt,tidx,nidx,chem1_g,nob,mbs, &
                            block_end,ik,ii,ij,kij,last_dt,spack,spack_2d,atol,rtol,gnz, &
                            ndyn,split,cp,cpor,p00,1,extra3d,jphoto,att)
  END IF
  OPEN(newunit=u, file=TRIM(fout), access='stream', form='unformatted', status='replace')
  DO n = 1, nsp
     IF (TRIM(solver) == 'qssa') THEN
        WRITE(u) chem1_g(n)%sc_t
     ELSE
        WRITE(u) chem1_g(n)%sc_p
     END IF
  END DO
  WRITE(u) last_dt
  CLOSE(u)
END PROGRAM shim_driver
