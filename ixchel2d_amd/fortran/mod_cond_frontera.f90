!
! cond_frontera: boundary conditions from files (upstream module
! source/mod_cond_frontera.f90).  The file format and the segment-to-node
! index rule are upstream's (lectura_cond_frontera, :70-245); imposing them
! (impone_cond_frontera, :254-420) is done on the device, so here the segments
! are flattened to one (type, value) pair per node of each side and handed to
! the library with ixh_set_bc.
!
module cond_frontera
  use iso_c_binding
  use malla, only : DBL
  use ixchel_hip
  implicit none

  type tipo_cond_front
     character(len=1) :: lado_front = 'z'
     integer :: ndivis = 0
     character(len=4), allocatable :: tipo_condi(:)
     real(DBL), allocatable :: valor_cond(:)
     integer, allocatable :: indice_div(:)
  end type tipo_cond_front

contains

  subroutine inicializa_cond_front(cf)
    type(tipo_cond_front), intent(inout) :: cf
    cf % lado_front = 'z'
    cf % ndivis = 0
    if (allocated(cf % tipo_condi)) deallocate(cf % tipo_condi, cf % valor_cond, cf % indice_div)
  end subroutine inicializa_cond_front

  ! four blocks "variable side ndiv" + ndiv lines "x0 x1 type value"; segment
  ! jj ends at the first node whose coordinate exceeds its x1, the first
  ! starts at node 2 and the last ends at node nn-1
  subroutine lectura_cond_frontera(archivo, cfa, cfb, cfc, cfd, xx, yy, mm, nn)
    character(len=*), intent(in) :: archivo
    type(tipo_cond_front), intent(inout) :: cfa, cfb, cfc, cfd
    integer, intent(in) :: mm, nn
    real(DBL), intent(in) :: xx(mm), yy(nn)
    character(len=4) :: variable, tipo
    character(len=1) :: lado
    real(DBL) :: x0, x1, valor
    integer :: ii, ndiv, un
    open(newunit=un, file=trim(archivo), status='old', action='read')
    do ii = 1, 4
       read(un, *) variable, lado, ndiv
       select case (lado)
       case ('a'); call lee_lado(cfa, yy, nn)
       case ('b'); call lee_lado(cfb, xx, mm)
       case ('c'); call lee_lado(cfc, yy, nn)
       case ('d'); call lee_lado(cfd, xx, mm)
       case default; error stop 'cond_frontera: side must be a, b, c or d'
       end select
    end do
    close(un)
  contains
    subroutine lee_lado(cf, coord, n)
      type(tipo_cond_front), intent(inout) :: cf
      integer, intent(in) :: n
      real(DBL), intent(in) :: coord(n)
      integer :: jj, kk, m
      call inicializa_cond_front(cf)
      allocate(cf % tipo_condi(ndiv), cf % valor_cond(ndiv), cf % indice_div(ndiv + 1))
      cf % lado_front = lado
      cf % ndivis = ndiv
      cf % indice_div(1) = 2
      do jj = 1, ndiv
         read(un, *) x0, x1, tipo, valor
         cf % tipo_condi(jj) = tipo
         cf % valor_cond(jj) = valor
         kk = n + 1
         do m = 1, n
            if (coord(m) > x1) then
               kk = m
               exit
            end if
         end do
         cf % indice_div(jj + 1) = kk
      end do
      cf % indice_div(ndiv + 1) = n - 1
    end subroutine lee_lado
  end subroutine lectura_cond_frontera

  ! flatten one side to per-node (type, value) and send it to the device
  ! (node k of the side at index k; later segments overwrite shared nodes)
  subroutine impone_cond_frontera(cf, var, lado, n)
    type(tipo_cond_front), intent(in) :: cf
    integer(c_int32_t), intent(in) :: var, lado
    integer, intent(in) :: n
    integer(c_int32_t) :: tipo(n)
    real(c_double) :: valor(n)
    integer :: ld, k
    tipo = 0
    valor = 0.0_DBL
    do ld = 1, cf % ndivis
       do k = cf % indice_div(ld), cf % indice_div(ld + 1)
          if (k < 1 .or. k > n) error stop 'cond_frontera: segment outside the side'
          select case (cf % tipo_condi(ld))
          case ('diri'); tipo(k) = 1
          case ('neum'); tipo(k) = 2
          case default; error stop 'cond_frontera: type must be diri or neum'
          end select
          valor(k) = cf % valor_cond(ld)
       end do
    end do
    call ixh_check(ixh_set_bc(estado, var, lado, tipo, valor), 'ixh_set_bc')
  end subroutine impone_cond_frontera

end module cond_frontera
