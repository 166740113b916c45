!
! solucionador: tridiagonal solves (upstream module source/mod_solucionador.f90
! and the uncompiled alternate source/tri.f90).  Same arguments and in-place
! semantics as upstream: r returns the solution, b is overwritten; the solve
! runs on the GPU (Thomas order, so results equal upstream's).  A line goes
! through ixh_tridiagonal_noalloc, which stages it in the library's device
! scratch instead of allocating device memory per call.
!
module solucionador
  use iso_c_binding
  use malla, only : DBL
  use ixchel_hip
  implicit none
contains

  ! tridiagonal(a,b,c,r,n), a(1) and c(n) unused (mod_solucionador.f90:25-47)
  subroutine tridiagonal(a, b, c, r, n)
    integer, intent(in) :: n
    real(DBL), intent(in) :: a(n), c(n)
    real(DBL), intent(inout) :: b(n), r(n)
    call ixh_check(ixh_tridiagonal_noalloc(a, b, c, r, int(n, c_int32_t), 1_c_int32_t, IXH_SOLVER_THOMAS), &
         'ixh_tridiagonal_noalloc')
  end subroutine tridiagonal

  ! tri(a,b,c,r,n) with the sub-diagonal stored as a(1..n-1) (tri.f90:8-25)
  subroutine tri(a, b, c, r, n)
    integer, intent(in) :: n
    real(DBL), intent(in) :: a(n-1), c(n-1)
    real(DBL), intent(inout) :: b(n), r(n)
    real(DBL) :: aa(n), cc(n)
    aa(1) = 0.0_DBL
    aa(2:n) = a(1:n-1)
    cc(1:n-1) = c(1:n-1)
    cc(n) = 0.0_DBL
    call tridiagonal(aa, b, cc, r, n)
  end subroutine tri

end module solucionador
