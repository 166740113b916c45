!
! frontera_inmersa: immersed solids by large-source penalty (upstream module
! source/mod_front_inmersa.f90).  Inside a body Sp = -1e51 and Sc = 1e51*1e-11
! for u and v (so u, v ~ 1e-11) and Sc = 1e51*T_b for T; the mask is tested at
! the pressure-node coordinates (xp(ii), yp(jj)) for every array, as upstream
! does.  'sincu' (what upstream's driver passes) sets nothing.  Besides
! upstream's shapes this adds 'cilin': a circular cylinder held at T = 1.
!
module frontera_inmersa
  use iso_c_binding
  use malla, only : mi, nj, DBL, xp, yp
  use ixchel_hip
  implicit none
  real(DBL), parameter :: penal = 10.0e50_DBL
contains

  logical function dentro(opcion, x, y, xc, yc, rc)
    character(len=*), intent(in) :: opcion
    real(DBL), intent(in) :: x, y, xc, yc, rc
    real(DBL) :: yv, hh, xv, m
    dentro = .false.
    select case (opcion)
    case ('trian')
       ! triangle 0 <= y <= 0.1 between the lines x = 2y + 5.8 and
       ! x = 6.2 - 2y (upstream's recta_mxb1 / recta_mxb2)
       yv = 0.1_DBL; hh = 0.1_DBL; m = 0.5_DBL
       if (yv - hh <= y .and. y <= yv) then
          dentro = (y/m - (0.1_DBL - m*6.0_DBL)/m <= x) .and. (x <= y/(-m) + (0.1_DBL + m*6.0_DBL)/m)
       end if
    case ('cuadr')
       yv = 0.05_DBL; xv = 6.0_DBL; hh = 0.1_DBL
       dentro = (yv - hh/2 <= y .and. y <= yv + hh/2) .and. (xv - hh/2 <= x .and. x <= xv + hh/2)
    case ('chime')
       yv = 0.5_DBL; xv = 8.5_DBL; hh = 0.3_DBL
       dentro = (yv - hh/2 <= y .and. y <= yv + hh/2) .and. (xv - hh/2 <= x .and. x <= xv + hh/2)
    case ('cilin')
       dentro = (x - xc)**2 + (y - yc)**2 <= rc*rc
    end select
  end function dentro

  ! definir_cuerpo (mod_front_inmersa.f90:54-179); xc, yc, rc describe the
  ! 'cilin' body
  subroutine definir_cuerpo(opcion, xc, yc, rc)
    character(len=*), intent(in) :: opcion
    real(DBL), intent(in), optional :: xc, yc, rc
    real(c_double), allocatable, target :: scu(:,:), spu(:,:), scv(:,:), spv(:,:), sct(:,:), spt(:,:)
    real(DBL) :: x0, y0, r0
    integer :: ii, jj
    logical :: hay_t
    if (trim(opcion) == 'sincu') return
    x0 = 0.0_DBL; y0 = 0.0_DBL; r0 = 0.0_DBL
    if (present(xc)) x0 = xc
    if (present(yc)) y0 = yc
    if (present(rc)) r0 = rc
    allocate(scu(mi, nj+1), spu(mi, nj+1), scv(mi+1, nj), spv(mi+1, nj), sct(mi+1, nj+1), spt(mi+1, nj+1))
    scu = 0; spu = 0; scv = 0; spv = 0; sct = 0; spt = 0
    hay_t = trim(opcion) == 'chime' .or. trim(opcion) == 'cilin'
    do jj = 1, nj + 1
       do ii = 1, mi + 1
          if (.not. dentro(trim(opcion), xp(ii), yp(jj), x0, y0, r0)) cycle
          if (ii <= mi) then
             spu(ii, jj) = -penal
             scu(ii, jj) = penal * 10.0e-12_DBL
          end if
          if (jj <= nj) then
             spv(ii, jj) = -penal
             scv(ii, jj) = penal * 10.0e-12_DBL
          end if
          if (hay_t) then
             spt(ii, jj) = -penal
             sct(ii, jj) = penal * 1.0_DBL
          end if
       end do
    end do
    call ixh_check(ixh_set_sources(estado, IXH_VAR_U, c_loc(scu), c_loc(spu)), 'ixh_set_sources')
    call ixh_check(ixh_set_sources(estado, IXH_VAR_V, c_loc(scv), c_loc(spv)), 'ixh_set_sources')
    if (hay_t) call ixh_check(ixh_set_sources(estado, IXH_VAR_T, c_loc(sct), c_loc(spt)), 'ixh_set_sources')
  end subroutine definir_cuerpo

end module frontera_inmersa
