!
! ec_energia: energy equation (upstream module source/mod_ec_energia.f90).
! ecuacion_energia replaces one loop nest of main.f90 (assembly by
! ensambla_energia_x/_y, boundary rows, line solves, copy-back).
!
module ec_energia
  use iso_c_binding
  use malla, only : mi, nj, DBL, xp, yp
  use cond_frontera
  use ixchel_hip
  implicit none

  real(DBL), allocatable, target :: temp(:,:), temp_ant(:,:)
  real(DBL) :: gamma_energ = 0.0_DBL, conv_t, rel_ener
  type(tipo_cond_front) :: cond_front_ta, cond_front_tb, cond_front_tc, cond_front_td

contains

  subroutine reserva_energia()
    allocate(temp(mi+1, nj+1), temp_ant(mi+1, nj+1))
    temp = 0.0_DBL
    temp_ant = 0.0_DBL
  end subroutine reserva_energia

  ! condicion_inicial_tempe (mod_ec_energia.f90:46-80)
  subroutine condicion_inicial_tempe(cond_inicial)
    character(len=*), intent(in) :: cond_inicial
    select case (trim(cond_inicial))
    case ('baja'); temp_ant = 0.0_DBL
    case ('alta'); temp_ant = 1.0_DBL
    case ('archivo')
    end select
  end subroutine condicion_inicial_tempe

  ! ini_frontera_t (mod_ec_energia.f90:91-109)
  subroutine ini_frontera_t()
    call lectura_cond_frontera('cond_fronterat.dat', cond_front_ta, cond_front_tb, cond_front_tc, &
         cond_front_td, xp, yp, mi+1, nj+1)
    call impone_cond_frontera(cond_front_ta, IXH_VAR_T, 0, nj+1)
    call impone_cond_frontera(cond_front_tb, IXH_VAR_T, 1, mi+1)
    call impone_cond_frontera(cond_front_tc, IXH_VAR_T, 2, nj+1)
    call impone_cond_frontera(cond_front_td, IXH_VAR_T, 3, mi+1)
  end subroutine ini_frontera_t

  ! one T sweep (main.f90:974-1049 / 1056-1121); the x sweep returns the
  ! energy error (main.f90:1128-1136)
  subroutine ecuacion_energia(dir, error)
    integer(c_int32_t), intent(in) :: dir
    real(c_double), intent(out), target :: error
    error = 0.0_DBL
    if (dir == IXH_DIR_X) then
       call ixh_check(ixh_energy(estado, dir, c_loc(error)), 'ixh_energy')
    else
       call ixh_check(ixh_energy(estado, dir, c_null_ptr), 'ixh_energy')
    end if
  end subroutine ecuacion_energia

end module ec_energia
