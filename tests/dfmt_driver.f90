!
! Test driver of the host's D23.15 record formatter (ixchel2d_amd/fortran/
! d_format.c): reads n values from valores.raw (stream: n as a 32-bit
! integer, then n FP64 values) and writes them as records of 3 and of 4
! values twice, through Fortran's formatted write ("(3D23.15)",
! "(4D23.15)": f3.txt, f4.txt) and through d23_records (c3.txt, c4.txt),
! which must give the same bytes.
!
program dfmt_driver
  use iso_c_binding
  implicit none
  interface
     integer(c_int) function d23_records(vals, nrec, ncol, out) bind(C, name='d23_records')
       import :: c_double, c_int64_t, c_int32_t, c_int, c_char
       real(c_double), intent(in) :: vals(*)
       integer(c_int64_t), value :: nrec
       integer(c_int32_t), value :: ncol
       character(kind=c_char), intent(out) :: out(*)
     end function d23_records
  end interface
  integer(4) :: n
  real(c_double), allocatable :: x(:)
  integer :: un
  open(newunit=un, file='valores.raw', access='stream', status='old', action='read')
  read(un) n
  allocate(x(n))
  read(un) x
  close(un)
  call both(3, '(3D23.15)', 'f3.txt', 'c3.txt')
  call both(4, '(4D23.15)', 'f4.txt', 'c4.txt')
contains
  subroutine both(ncol, form, ffile, cfile)
    integer, intent(in) :: ncol
    character(len=*), intent(in) :: form, ffile, cfile
    character(kind=c_char), allocatable :: buf(:)
    integer :: nrec, r, u
    nrec = n / ncol
    open(newunit=u, file=ffile, status='replace')
    do r = 1, nrec
       write(u, form) x((r - 1) * ncol + 1:r * ncol)
    end do
    close(u)
    allocate(buf(nrec * (23 * ncol + 1)))
    if (d23_records(x, int(nrec, c_int64_t), int(ncol, c_int32_t), buf) /= 0) stop 3
    open(newunit=u, file=cfile, access='stream', form='unformatted', status='replace')
    write(u) buf
    close(u)
  end subroutine both
end program dfmt_driver
