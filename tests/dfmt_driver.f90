!
! Test driver of the host's D23.15 record formatter and reader
! (ixchel2d_amd/fortran/d_format.c).  Without arguments: reads n values from
! valores.raw (stream: n as a 32-bit integer, then n FP64 values) and writes
! them as records of 3 and of 4 values twice, through Fortran's formatted
! write ("(3D23.15)", "(4D23.15)": f3.txt, f4.txt) and through d23_records
! (c3.txt, c4.txt), which must give the same bytes.  With the arguments
! "leer <file> <nrec> <ncol>": reads the nrec records of ncol values after
! the file's header line twice, through Fortran's formatted read
! (fleer.raw) and through d23_read (cleer.raw, and its status on stdout).
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
     integer(c_int) function d23_read(path, pathlen, nrec, ncol, vals) bind(C, name='d23_read')
       import :: c_char, c_int32_t, c_int64_t, c_double, c_int
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), value :: pathlen
       integer(c_int64_t), value :: nrec
       integer(c_int32_t), value :: ncol
       real(c_double), intent(out) :: vals(*)
     end function d23_read
  end interface
  integer(4) :: n
  real(c_double), allocatable :: x(:)
  integer :: un
  character(len=256) :: arg
  if (command_argument_count() >= 4) then
     call leer()
     stop
  end if
  open(newunit=un, file='valores.raw', access='stream', status='old', action='read')
  read(un) n
  allocate(x(n))
  read(un) x
  close(un)
  call both(3, '(3D23.15)', 'f3.txt', 'c3.txt')
  call both(4, '(4D23.15)', 'f4.txt', 'c4.txt')
contains
  subroutine leer()
    character(len=256) :: archivo
    integer :: nrec, ncol, r, u, st
    real(c_double), allocatable :: a(:, :), b(:, :)
    character(len=9) :: form
    call get_command_argument(2, archivo)
    call get_command_argument(3, arg)
    read(arg, *) nrec
    call get_command_argument(4, arg)
    read(arg, *) ncol
    write(form, '(A,I1,A)') '(', ncol, 'D23.15)'
    allocate(a(ncol, nrec), b(ncol, nrec))
    a = 0
    b = 0
    open(newunit=u, file=trim(archivo), status='old', action='read')
    read(u, *)
    do r = 1, nrec
       read(u, form) a(:, r)
    end do
    close(u)
    st = d23_read(trim(archivo), int(len_trim(archivo), c_int32_t), int(nrec, c_int64_t), int(ncol, c_int32_t), b)
    print '(A,I0)', 'status ', st
    open(newunit=u, file='fleer.raw', access='stream', status='replace')
    write(u) a
    close(u)
    open(newunit=u, file='cleer.raw', access='stream', status='replace')
    write(u) b
    close(u)
  end subroutine leer

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
