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

contains

  integer function cuenta_lineas(archivo) result(n)
    character(len=*), intent(in) :: archivo
    integer :: ios, u
    character(len=256) :: linea
    n = 0
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
  ! the initial u, v, T, p; the geometry factors are built by the library
  subroutine lectura_mallas_escalonadas(entrada_u, entrada_v, entrada_tp, u_ant, v_ant, pres, temp_ant, &
       ao, placa_min, placa_max, itera_inicial)
    character(len=*), intent(in) :: entrada_u, entrada_v, entrada_tp
    real(DBL), intent(out) :: u_ant(:,:), v_ant(:,:), pres(:,:), temp_ant(:,:)
    real(DBL), intent(out) :: ao
    integer, intent(out) :: placa_min, placa_max, itera_inicial
    integer :: ii, jj, un
    if (.not. allocated(xu)) allocate(xu(mi), yv(nj), xp(mi+1), yp(nj+1))
    open(newunit=un, file=trim(entrada_u), status='old', action='read')
    read(un, *) placa_min, placa_max, itera_inicial, ao
    do jj = 1, nj + 1
       do ii = 1, mi
          read(un, form24) xu(ii), yp(jj), u_ant(ii, jj)
       end do
    end do
    close(un)
    open(newunit=un, file=trim(entrada_v), status='old', action='read')
    read(un, *) placa_min, placa_max, itera_inicial, ao
    do jj = 1, nj
       do ii = 1, mi + 1
          read(un, form24) xp(ii), yv(jj), v_ant(ii, jj)
       end do
    end do
    close(un)
    open(newunit=un, file=trim(entrada_tp), status='old', action='read')
    read(un, *) placa_min, placa_max, itera_inicial, ao
    do jj = 1, nj + 1
       do ii = 1, mi + 1
          read(un, form25) xp(ii), yp(jj), temp_ant(ii, jj), pres(ii, jj)
       end do
    end do
    close(un)
  end subroutine lectura_mallas_escalonadas

end module malla
