! C-callable harness around the reference's own RELACS_TUV kernels and block
! partition, compiled from the reference source tree by oracle/Makefile into
! oracle/_ref/libbrams_ref.so.  Test infrastructure only: it lets the tests
! compare the oracle restatement with the reference's compiled arithmetic.
!
! Wrapped routines (reference paths):
!   kinetic     src/brams/ccatt/RELACS_TUV/chem_spack_kinetic.f90:7
!   fexchem     src/brams/ccatt/RELACS_TUV/chem_spack_fexchem.f90:10
!   jacdchemdc  src/brams/ccatt/RELACS_TUV/chem_spack_jacdchemdc.f90:10
!   rates, dratedc, fexprod, fexloss (the P - L c split of QSSA)
!               src/brams/ccatt/RELACS_TUV/chem_spack_{rates,dratedc,fexprod,fexloss}.f90:9
!   AllocIndex  src/brams/ccatt/chem_spack_utils.f90:39
module brams_ref_harness
  use iso_c_binding
  use mod_chem_spack_kinetic,    only: kinetic
  use mod_chem_spack_fexchem,    only: fexchem
  use mod_chem_spack_jacdchemdc, only: jacdchemdc
  use mod_chem_spack_rates,      only: rates
  use mod_chem_spack_dratedc,    only: dratedc
  use mod_chem_spack_fexprod,    only: fexprod
  use mod_chem_spack_fexloss,    only: fexloss
  use Spack_utils,               only: AllocIndex, index_g, nob, maxblock_size
  implicit none
contains

  ! rk(n,nr) from jphoto(n,nphot), temp(n), vapp(n), press(n); column-major
  subroutine ref_kinetic(n, nphot, nr, jphoto, rk, temp, vapp, press) bind(C, name="ref_kinetic")
    integer(c_int), value :: n, nphot, nr
    real(c_double), intent(in)  :: jphoto(n, nphot), temp(n), vapp(n), press(n)
    real(c_double), intent(out) :: rk(n, nr)
    real(c_double) :: cosz(n), att(n)
    cosz = 0.d0
    att = 0.d0
    rk = 0.d0        ! the reference zero-initialises rk once (slot 97 is never set)
    call kinetic(nphot, jphoto, rk, temp, vapp, press, cosz, att, 1, n, n, nr)
  end subroutine

  subroutine ref_fexchem(n, ns, nr, y, rk, chem) bind(C, name="ref_fexchem")
    integer(c_int), value :: n, ns, nr
    real(c_double), intent(in)  :: y(n, ns), rk(n, nr)
    real(c_double), intent(out) :: chem(n, ns)
    chem = 0.d0      ! rows never written by fexchem stay at their allocation value 0
    call fexchem(y, rk, chem, ns, 1, n, n, nr)
  end subroutine

  subroutine ref_jacdchemdc(n, ns, nr, y, rk, jac) bind(C, name="ref_jacdchemdc")
    integer(c_int), value :: n, ns, nr
    real(c_double), intent(in)  :: y(n, ns), rk(n, nr)
    real(c_double), intent(out) :: jac(n, ns, ns)
    jac = 0.d0       ! unassigned entries stay at their allocation value 0
    call jacdchemdc(y, rk, jac, ns, 1, n, n, nr)
  end subroutine

  ! production and loss terms of chem_spack_qssa.f90:251-263 (rates, dratedc,
  ! fexprod, fexloss in that order)
  subroutine ref_prod_loss(n, ns, nr, y, rk, prod, loss) bind(C, name="ref_prod_loss")
    integer(c_int), value :: n, ns, nr
    real(c_double), intent(in)  :: y(n, ns), rk(n, nr)
    real(c_double), intent(out) :: prod(n, ns), loss(n, ns)
    real(c_double), allocatable :: w(:,:), dw(:,:,:)
    allocate(w(n, nr), dw(n, nr, ns))
    w = 0.d0; dw = 0.d0
    prod = 0.d0      ! rows fexprod / fexloss never write stay at 0
    loss = 0.d0
    call rates(rk, y, w, ns, 1, n, n, nr)
    call dratedc(rk, y, dw, ns, 1, n, n, nr)
    call fexprod(w, prod, ns, 1, n, n, nr)
    call fexloss(dw, loss, ns, 1, n, n, nr)
    deallocate(w, dw)
  end subroutine

  ! Block partition of one grid.  Outputs are sized by the caller:
  ! block_end(maxnob), indexk/i/j(maxbs, maxnob) (column-major), dt0(maxnob).
  subroutine ref_allocindex(block_size, m1, m2, m3, ia, iz, ja, jz, dtlong, ndyn, &
                            maxnob, maxbs, out_nob, out_maxbs, block_end, indexk, &
                            indexi, indexj, dt0) bind(C, name="ref_allocindex")
    integer(c_int), value :: block_size, m1, m2, m3, ia, iz, ja, jz, ndyn, maxnob, maxbs
    real(c_float),  value :: dtlong
    integer(c_int), intent(out) :: out_nob, out_maxbs
    integer(c_int), intent(out) :: block_end(maxnob), indexk(maxbs, maxnob)
    integer(c_int), intent(out) :: indexi(maxbs, maxnob), indexj(maxbs, maxnob)
    real(c_double), intent(out) :: dt0(maxnob)
    integer :: mmxp(1), mmyp(1), mmzp(1), mia(1), miz(1), mja(1), mjz(1), b, c
    real :: dtl(1)
    if (allocated(index_g)) then
      do b = 1, size(index_g)
        deallocate(index_g(b)%block_end, index_g(b)%indexi, index_g(b)%indexj, &
                   index_g(b)%indexk, index_g(b)%ijkindex, index_g(b)%bindex, &
                   index_g(b)%kij_index)
        deallocate(index_g(b)%last_accepted_dt)
      end do
      deallocate(index_g, nob)
    end if
    mmxp(1) = m2; mmyp(1) = m3; mmzp(1) = m1
    mia(1) = ia; miz(1) = iz; mja(1) = ja; mjz(1) = jz
    dtl(1) = dtlong
    call AllocIndex(block_size, mmxp, mmyp, mmzp, mia, miz, mja, mjz, 1, dtl, ndyn)
    out_nob = nob(1)
    out_maxbs = maxblock_size
    block_end = 0; indexk = 0; indexi = 0; indexj = 0; dt0 = 0.d0
    if (nob(1) > maxnob .or. maxblock_size > maxbs) return
    do b = 1, nob(1)
      block_end(b) = index_g(1)%block_end(b)
      dt0(b) = index_g(1)%last_accepted_dt(b)
      do c = 1, index_g(1)%block_end(b)
        indexk(c, b) = index_g(1)%indexk(c, b)
        indexi(c, b) = index_g(1)%indexi(c, b)
        indexj(c, b) = index_g(1)%indexj(c, b)
      end do
    end do
  end subroutine
end module brams_ref_harness
