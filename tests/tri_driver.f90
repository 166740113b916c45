!
! Test driver of the Fortran host's line solvers (module solucionador):
! reads nsys systems of n rows from sistemas.raw (stream: n, nsys as 32-bit
! integers, then a, b, c, r, each n x nsys, column-major FP64), solves each
! with tridiagonal(a,b,c,r,n) (mod_solucionador.f90:25-47 convention: a(1)
! and c(n) unused) and with tri(a,b,c,r,n) (tri.f90:8-25 convention: the
! sub-diagonal in a(1..n-1)), both on the GPU, and writes soluciones.raw:
! r and b after tridiagonal, then r and b after tri.
!
program tri_driver
  use malla, only : DBL
  use solucionador
  implicit none
  integer(4) :: n, nsys
  real(DBL), allocatable :: a(:, :), b(:, :), c(:, :), r(:, :)
  real(DBL), allocatable :: b1(:, :), r1(:, :), b2(:, :), r2(:, :)
  integer :: k, un
  open(newunit=un, file='sistemas.raw', access='stream', status='old', action='read')
  read(un) n, nsys
  allocate(a(n, nsys), b(n, nsys), c(n, nsys), r(n, nsys))
  read(un) a, b, c, r
  close(un)
  b1 = b
  r1 = r
  b2 = b
  r2 = r
  do k = 1, nsys
     call tridiagonal(a(:, k), b1(:, k), c(:, k), r1(:, k), int(n))
     call tri(a(2:n, k), b2(:, k), c(1:n-1, k), r2(:, k), int(n))
  end do
  open(newunit=un, file='soluciones.raw', access='stream', status='replace', action='write')
  write(un) r1, b1, r2, b2
  close(un)
end program tri_driver
