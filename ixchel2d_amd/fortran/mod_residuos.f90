!
! residuos: residual norms (upstream module source/mod_residuos.f90, which is
! empty; upstream computes the u residual with ec_momento::residuo_u and
! reduces it in main.f90:1169-1199).
!
module residuos
  use iso_c_binding
  use malla, only : DBL
  use ixchel_hip
  implicit none
contains

  ! residuo = sqrt(sum Resu^2) over ii 2..mi-1, jj 2..nj of the unrelaxed
  ! u-momentum residual (mod_ec_momento.f90:1107-1266)
  subroutine residuo_u(residuo)
    real(c_double), intent(out), target :: residuo
    call ixh_check(ixh_residual_u(estado, c_loc(residuo)), 'ixh_residual_u')
  end subroutine residuo_u

end module residuos
