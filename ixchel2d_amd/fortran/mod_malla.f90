!
! malla: mesh and geometry of the staggered grid (upstream module
! source/mod_malla.f90).  The extents mi, nj are run-time values here (upstream
! fixes them at compile time, mod_malla.f90:15): they are taken from the line
! counts of the mesh files, which hold mi*(nj+1) u-nodes and (mi+1)*(nj+1)
! p-nodes.  File formats are upstream's (3D23.15 / 4D23.15 after a header line
! "placa_min placa_max iteration ao", mod_malla.f90:137-162).
!
module malla
  use iso_c_binding
  implicit none
  integer :: mi = 0, nj = 0
  integer, parameter :: DBL = selected_real_kind(p=15, r=307)
  character(len=9), parameter :: form24 = "(3D23.15)", form25 = "(4D23.15)"
  character(len=9), parameter :: form26 = "(6D23.15)"
  character(len=20) :: mic, njc
  real(DBL), allocatable, target :: xu(:), yv(:), xp(:), yp(:)

  interface
     ! d_format.c: the records after the header line read as (ncolD23.15);
     ! 1 when the file needs Fortran's own formatted read
     integer(c_int) function d23_read(path, pathlen, nrec, ncol, vals) bind(C, name='d23_read')
       import :: c_char, c_int32_t, c_int64_t, c_double, c_int
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), value :: pathlen
       integer(c_int64_t), value :: nrec
       integer(c_int32_t), value :: ncol
       real(c_double), intent(out) :: vals(*)
     end function d23_read
     ! d_format.c: the non-blank lines after the first
     integer(c_int) function d23_count_lines(path, pathlen, n) bind(C, name='d23_count_lines')
       import :: c_char, c_int32_t, c_int64_t, c_int
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), value :: pathlen
       integer(c_int64_t), intent(out) :: n
     end function d23_count_lines
  end interface

contains

  integer function cuenta_lineas(archivo) result(n)
    character(len=*), intent(in) :: archivo
    integer :: ios, u
    integer(c_int64_t) :: nc
    character(len=256) :: linea
    n = 0
    if (d23_count_lines(trim(archivo), int(len_trim(archivo), c_int32_t), nc) == 0) then
       n = int(nc)
       return
    end if
    open(newunit=u, file=trim(archivo), status='old', action='read')
    read(u, '(A)', iostat=ios) linea          ! header
    do
       read(u, '(A)', iostat=ios) linea
       if (ios /= 0) exit
       if (len_trim(linea) > 0) n = n + 1
    end do
    close(u)
  end function cuenta_lineas

  ! mi, nj from the node counts of the u and p files
  subroutine dimensiones_malla(entrada_u, entrada_tp)
    character(len=*), intent(in) :: entrada_u, entrada_tp
    integer :: nu, np
    nu = cuenta_lineas(entrada_u)
    np = cuenta_lineas(entrada_tp)
    nj = np - nu - 1
    if (nj < 4) error stop 'malla: cannot infer nj from the mesh files'
    mi = nu / (nj + 1)
    if (mi * (nj + 1) /= nu .or. (mi + 1) * (nj + 1) /= np) error stop 'malla: inconsistent mesh files'
  end subroutine dimensiones_malla

  ! lectura_mallas_escalonadas (source/mod_malla.f90:62-162): coordinates and
  ! the initial u, v, T, p; the geometry factors are built by the library.
  ! Each file's header is read here; its records by d23_read (d_format.c),
  ! or by Fortran's formatted read when d23_read declines the file.
  subroutine lectura_mallas_escalonadas(entrada_u, entrada_v, entrada_tp, u_ant, v_ant, pres, temp_ant, &
       ao, placa_min, placa_max, itera_inicial)
    character(len=*), intent(in) :: entrada_u, entrada_v, entrada_tp
    real(DBL), intent(out) :: u_ant(:,:), v_ant(:,:), pres(:,:), temp_ant(:,:)
    real(DBL), intent(out) :: ao
    integer, intent(out) :: placa_min, placa_max, itera_inicial
    real(DBL), allocatable :: rec(:, :)
    integer :: ii, jj, k
    if (.not. allocated(xu)) allocate(xu(mi), yv(nj), xp(mi+1), yp(nj+1))
    call registros(entrada_u, 3, mi * (nj + 1))
    k = 0
    do jj = 1, nj + 1
       do ii = 1, mi
          k = k + 1
          xu(ii) = rec(1, k)
          yp(jj) = rec(2, k)
          u_ant(ii, jj) = rec(3, k)
       end do
    end do
    call registros(entrada_v, 3, (mi + 1) * nj)
    k = 0
    do jj = 1, nj
       do ii = 1, mi + 1
          k = k + 1
          xp(ii) = rec(1, k)
          yv(jj) = rec(2, k)
          v_ant(ii, jj) = rec(3, k)
       end do
    end do
    call registros(entrada_tp, 4, (mi + 1) * (nj + 1))
    k = 0
    do jj = 1, nj + 1
       do ii = 1, mi + 1
          k = k + 1
          xp(ii) = rec(1, k)
          yp(jj) = rec(2, k)
          temp_ant(ii, jj) = rec(3, k)
          pres(ii, jj) = rec(4, k)
       end do
    end do
  contains
    ! the header and the nrec records of ncol values of one file into rec
    subroutine registros(archivo, ncol, nrec)
      character(len=*), intent(in) :: archivo
      integer, intent(in) :: ncol, nrec
      integer :: un, r
      if (allocated(rec)) deallocate(rec)
      allocate(rec(ncol, nrec))
      open(newunit=un, file=trim(archivo), status='old', action='read')
      read(un, *) placa_min, placa_max, itera_inicial, ao
      if (d23_read(trim(archivo), int(len_trim(archivo), c_int32_t), int(nrec, c_int64_t), &
           int(ncol, c_int32_t), rec) /= 0) then
         do r = 1, nrec
            if (ncol == 3) then
               read(un, form24) rec(:, r)
            else
               read(un, form25) rec(:, r)
            end if
         end do
      end if
      close(un)
    end subroutine registros
  end subroutine lectura_mallas_escalonadas

end module malla
