! Test-only C bindings around the reference's own weight tables and banded
! LU routines, so the C oracle can be checked bit-for-bit against them.
! Links against pyranda/parcop/stencils.f90 and pentadiagonal.f90 (plus their
! module dependencies nrtype/blockmath/nrutil) compiled in place; see
! oracle/Makefile. The modules used here need no MPI.
module pyr_ref_driver
  use iso_c_binding
  use LES_stencils
  use LES_pentadiagonal
  implicit none
contains

  ! which: 0=c10d1, 1=c10d2, 2=c10d8, 3=c8ff8 (ff(2)), 4=cgfs4 (ff(6))
  ! Arrays are returned zero-padded to the widest stencil, variants -1..1.
  subroutine ref_get_weight(which, ali, ari, alb1, alb2, arb1, arb2, dims) &
      bind(C, name="ref_get_weight")
    integer(c_int), value :: which
    real(c_double), intent(out) :: ali(5), ari(9)
    real(c_double), intent(out) :: alb1(5,4,3), alb2(5,4,3), arb1(9,4,3), arb2(9,4,3)
    integer(c_int), intent(out) :: dims(6)
    type(compact_weight), pointer :: w
    if (.not. stencils_set) call setup_stencils(.false.)
    select case (which)
    case (0)
      w => compact_weight_d1(1)
    case (1)
      w => compact_weight_d2(1)
    case (2)
      w => compact_weight_d8(1)
    case (3)
      w => compact_weight_ff(2)
    case default
      w => compact_weight_ff(6)
    end select
    ali = 0.0d0; ari = 0.0d0
    alb1 = 0.0d0; alb2 = 0.0d0; arb1 = 0.0d0; arb2 = 0.0d0
    ali(1:w%ncl) = w%ali
    ari(1:w%ncr) = w%ari
    alb1(1:w%ncl,:,:) = w%alb1(:,:,-1:1)
    alb2(1:w%ncl,:,:) = w%alb2(:,:,-1:1)
    arb1(1:w%ncr,:,:) = w%arb1(:,:,-1:1)
    arb2(1:w%ncr,:,:) = w%arb2(:,:,-1:1)
    dims = [w%nol, w%nor, w%ncl, w%ncr, w%null_option, merge(1, 0, logical(w%implicit_op))]
  end subroutine ref_get_weight

  subroutine ref_bpentLUD1(c, n) bind(C, name="ref_bpentLUD1")
    integer(c_int), value :: n
    real(c_double), intent(inout) :: c(n,5)
    call bpentLUD1(c, n)
  end subroutine ref_bpentLUD1

  subroutine ref_ppentLUD1(c, n) bind(C, name="ref_ppentLUD1")
    integer(c_int), value :: n
    real(c_double), intent(inout) :: c(n,9)
    call ppentLUD1(c, n)
  end subroutine ref_ppentLUD1

  ! dir: 1=x, 2=y, 3=z sweep over an (n1,n2,n3) array
  subroutine ref_bpentLUS3(dir, c, r, n, n1, n2, n3) bind(C, name="ref_bpentLUS3")
    integer(c_int), value :: dir, n, n1, n2, n3
    real(c_double), intent(in) :: c(n,5)
    real(c_double), intent(inout) :: r(n1,n2,n3)
    select case (dir)
    case (1)
      call bpentLUS3x(c, r, n, n1, n2, n3)
    case (2)
      call bpentLUS3y(c, r, n, n1, n2, n3)
    case default
      call bpentLUS3z(c, r, n, n1, n2, n3)
    end select
  end subroutine ref_bpentLUS3

  subroutine ref_ppentLUS3(dir, c, r, n, n1, n2, n3) bind(C, name="ref_ppentLUS3")
    integer(c_int), value :: dir, n, n1, n2, n3
    real(c_double), intent(in) :: c(n,9)
    real(c_double), intent(inout) :: r(n1,n2,n3)
    select case (dir)
    case (1)
      call ppentLUS3x(c, r, n, n1, n2, n3)
    case (2)
      call ppentLUS3y(c, r, n, n1, n2, n3)
    case default
      call ppentLUS3z(c, r, n, n1, n2, n3)
    end select
  end subroutine ref_ppentLUS3

end module pyr_ref_driver
