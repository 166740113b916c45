!
! Test driver of the Fortran host's writers (module postproceso): reads the
! mesh from the three mesh files named on the command line (u, v, p files of
! a case directory) and the fields from the column-major FP64 files u.raw,
! v.raw, t.raw, p.raw of the working directory, and writes out.vtk with
! postproceso_vtk.  With a fourth argument, the restart-file header values
! "placa_min placa_max itera_total ao", it also writes outu.dat, outv.dat,
! outp.dat with postproceso_dat and outu.bin, outv.bin, outp.bin with
! postproceso_bin.  No GPU call is made.
!
program vtk_driver
  use malla
  use postproceso
  implicit none
  character(len=256) :: fu, fv, ftp, cab
  real(DBL), allocatable :: u(:, :), v(:, :), t(:, :), p(:, :)
  real(DBL) :: ao
  integer :: placa_min, placa_max, itera_inicial
  call get_command_argument(1, fu)
  call get_command_argument(2, fv)
  call get_command_argument(3, ftp)
  call dimensiones_malla(trim(fu), trim(ftp))
  allocate(u(mi, nj+1), v(mi+1, nj), t(mi+1, nj+1), p(mi+1, nj+1))
  call lectura_mallas_escalonadas(trim(fu), trim(fv), trim(ftp), u, v, p, t, ao, placa_min, placa_max, &
       itera_inicial)
  call lee('u.raw', u)
  call lee('v.raw', v)
  call lee('t.raw', t)
  call lee('p.raw', p)
  call postproceso_vtk('out.vtk', u, v, t, p)
  if (command_argument_count() >= 4) then
     call get_command_argument(4, cab)
     read(cab, *) placa_min, placa_max, itera_inicial, ao
     call postproceso_dat('out', u, v, t, p, placa_min, placa_max, itera_inicial, ao)
     call postproceso_bin('out', u, v, t, p)
  end if
contains
  subroutine lee(archivo, a)
    character(len=*), intent(in) :: archivo
    real(DBL), intent(out) :: a(:, :)
    integer :: un
    open(newunit=un, file=archivo, access='stream', status='old', action='read')
    read(un) a
    close(un)
  end subroutine lee
end program vtk_driver
