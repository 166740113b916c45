!
! ec_momento: momentum and pressure-correction equations (upstream module
! source/mod_ec_momento.f90).  Upstream's per-cell assembly routines
! (ensambla_velu_x/_y, ensambla_velv_x/_y, ensambla_corr_pres_x/_y) are called
! from loop nests in main.f90 and followed there by the boundary rows, the
! line solves and the copy-back; each routine below replaces one such loop
! nest and runs it on the GPU through the C-ABI.
!
module ec_momento
  use iso_c_binding
  use malla, only : mi, nj, DBL, xu, yv, xp, yp
  use cond_frontera
  use ixchel_hip
  implicit none

  real(DBL), allocatable, target :: u(:,:), u_ant(:,:), v(:,:), v_ant(:,:)
  real(DBL), allocatable, target :: pres(:,:), b_o(:,:)
  real(DBL) :: gamma_momen = 0.0_DBL, Ri = 0.0_DBL, Riy = 0.0_DBL
  real(DBL) :: maxbo = 0.0_DBL, conv_u, conv_p, conv_paso, rel_pres, rel_vel
  type(tipo_cond_front) :: cond_front_ua, cond_front_ub, cond_front_uc, cond_front_ud
  type(tipo_cond_front) :: cond_front_va, cond_front_vb, cond_front_vc, cond_front_vd

contains

  subroutine reserva_momento()
    allocate(u(mi, nj+1), u_ant(mi, nj+1), v(mi+1, nj), v_ant(mi+1, nj))
    allocate(pres(mi+1, nj+1), b_o(mi+1, nj+1))
    u = 0.0_DBL; u_ant = 0.0_DBL; v = 0.0_DBL; v_ant = 0.0_DBL
    pres = 0.0_DBL; b_o = 0.0_DBL
  end subroutine reserva_momento

  ! condicion_inicial_flujo (mod_ec_momento.f90:59-89)
  subroutine condicion_inicial_flujo(cond_inicial)
    character(len=*), intent(in) :: cond_inicial
    select case (trim(cond_inicial))
    case ('nulo')
       u_ant = 0.0_DBL
       v_ant = 0.0_DBL
    case ('archivo')
    end select
  end subroutine condicion_inicial_flujo

  ! ini_frontera_uv (mod_ec_momento.f90:100-135): read and impose the u and v
  ! boundary conditions
  subroutine ini_frontera_uv()
    call lectura_cond_frontera('cond_fronterau.dat', cond_front_ua, cond_front_ub, cond_front_uc, &
         cond_front_ud, xu, yp, mi, nj+1)
    call lectura_cond_frontera('cond_fronterav.dat', cond_front_va, cond_front_vb, cond_front_vc, &
         cond_front_vd, xp, yv, mi+1, nj)
    call impone_cond_frontera(cond_front_ua, IXH_VAR_U, 0, nj+1)
    call impone_cond_frontera(cond_front_ub, IXH_VAR_U, 1, mi)
    call impone_cond_frontera(cond_front_uc, IXH_VAR_U, 2, nj+1)
    call impone_cond_frontera(cond_front_ud, IXH_VAR_U, 3, mi)
    call impone_cond_frontera(cond_front_va, IXH_VAR_V, 0, nj)
    call impone_cond_frontera(cond_front_vb, IXH_VAR_V, 1, mi+1)
    call impone_cond_frontera(cond_front_vc, IXH_VAR_V, 2, nj)
    call impone_cond_frontera(cond_front_vd, IXH_VAR_V, 3, mi+1)
  end subroutine ini_frontera_uv

  ! one u sweep (main.f90:420-484 for dir = IXH_DIR_Y, 492-556 for IXH_DIR_X)
  subroutine ecuacion_velu(dir)
    integer(c_int32_t), intent(in) :: dir
    call ixh_check(ixh_momentum_u(estado, dir), 'ixh_momentum_u')
  end subroutine ecuacion_velu

  ! one v sweep (main.f90:570-640 / 647-711); after the x sweep `error` is
  ! sqrt(sum (v - fv)^2) of the momentum pass (main.f90:716-725)
  subroutine ecuacion_velv(dir, error)
    integer(c_int32_t), intent(in) :: dir
    real(c_double), intent(out), target :: error
    error = 0.0_DBL
    if (dir == IXH_DIR_X) then
       call ixh_check(ixh_momentum_v(estado, dir, c_loc(error)), 'ixh_momentum_v')
    else
       call ixh_check(ixh_momentum_v(estado, dir, c_null_ptr), 'ixh_momentum_v')
    end if
  end subroutine ecuacion_velv

  ! p' = 0 (main.f90:752-758)
  subroutine inicializa_corrector_presion()
    call ixh_check(ixh_pcorr_zero(estado), 'ixh_pcorr_zero')
  end subroutine inicializa_corrector_presion

  ! one p' sweep (main.f90:762-879); the x sweep returns the p' error and
  ! maxbo (main.f90:884-906)
  subroutine ecuacion_corr_pres(dir, error, maxbo_o)
    integer(c_int32_t), intent(in) :: dir
    real(c_double), intent(out), target :: error, maxbo_o
    error = 0.0_DBL
    maxbo_o = 0.0_DBL
    if (dir == IXH_DIR_X) then
       call ixh_check(ixh_pcorr(estado, dir, c_loc(error), c_loc(maxbo_o)), 'ixh_pcorr')
    else
       call ixh_check(ixh_pcorr(estado, dir, c_null_ptr, c_null_ptr), 'ixh_pcorr')
    end if
  end subroutine ecuacion_corr_pres

  ! p += 0.6 p', u and v corrections (main.f90:931-962)
  subroutine corrige_presion_velocidad()
    call ixh_check(ixh_correct(estado), 'ixh_correct')
  end subroutine corrige_presion_velocidad

end module ec_momento
