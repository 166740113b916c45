!
! postproceso: Nusselt numbers and output files (upstream module
! source/mod_postproceso.f90).  The Nusselt integrals run on the GPU; the
! writers produce upstream's text (.dat, which are also restart files) and
! stream (.bin) formats.
!
module postproceso
  use iso_c_binding
  use malla, only : mi, nj, DBL, xu, yv, xp, yp, form24, form25
  use ixchel_hip
  implicit none
contains

  ! nusselt_promedio_y (mod_postproceso.f90:100-215) on the device's T
  subroutine nusselt_promedio_y(nusselt0, nusselt1, placa_min, placa_max)
    real(DBL), intent(out) :: nusselt0, nusselt1
    integer, intent(in) :: placa_min, placa_max
    call ixh_check(ixh_nusselt_y(estado, int(placa_min, c_int32_t), int(placa_max, c_int32_t), &
         nusselt0, nusselt1), 'ixh_nusselt_y')
  end subroutine nusselt_promedio_y

  ! mean Nusselt numbers of the vertical walls (hot x = 0, cold x = W)
  subroutine nusselt_promedio_x(nusselt_w, nusselt_e)
    real(DBL), intent(out) :: nusselt_w, nusselt_e
    call ixh_check(ixh_nusselt_x(estado, nusselt_w, nusselt_e), 'ixh_nusselt_x')
  end subroutine nusselt_promedio_x

  ! entero_caracter (mod_postproceso.f90:375-399): 5-digit zero-padded
  function entero_caracter(entero) result(s)
    integer, intent(in) :: entero
    character(len=6) :: s
    write(s, '(I5.5,1X)') mod(entero, 100000)
  end function entero_caracter

  ! out_n<nj>m<mi>_R<Re>{u,v,p}.dat, readable by lectura_mallas_escalonadas
  subroutine postproceso_dat(prefijo, u, v, temp, pres, placa_min, placa_max, itera_total, ao)
    character(len=*), intent(in) :: prefijo
    real(DBL), intent(in) :: u(mi, nj+1), v(mi+1, nj), temp(mi+1, nj+1), pres(mi+1, nj+1), ao
    integer, intent(in) :: placa_min, placa_max, itera_total
    integer :: ii, jj, un
    open(newunit=un, file=prefijo//'u.dat')
    write(un, *) placa_min, placa_max, itera_total, ao
    do jj = 1, nj + 1
       do ii = 1, mi
          write(un, form24) xu(ii), yp(jj), u(ii, jj)
       end do
    end do
    close(un)
    open(newunit=un, file=prefijo//'v.dat')
    write(un, *) placa_min, placa_max, itera_total, ao
    do jj = 1, nj
       do ii = 1, mi + 1
          write(un, form24) xp(ii), yv(jj), v(ii, jj)
       end do
    end do
    close(un)
    open(newunit=un, file=prefijo//'p.dat')
    write(un, *) placa_min, placa_max, itera_total, ao
    do jj = 1, nj + 1
       do ii = 1, mi + 1
          write(un, form25) xp(ii), yp(jj), temp(ii, jj), pres(ii, jj)
       end do
    end do
    close(un)
  end subroutine postproceso_dat

  ! postproceso_bin (mod_postproceso.f90:310-365): exact FP64 dumps
  subroutine postproceso_bin(prefijo, u, v, temp, pres)
    character(len=*), intent(in) :: prefijo
    real(DBL), intent(in) :: u(mi, nj+1), v(mi+1, nj), temp(mi+1, nj+1), pres(mi+1, nj+1)
    integer :: ii, jj, un
    open(newunit=un, file=prefijo//'u.bin', access='stream', status='replace')
    do jj = 1, nj + 1
       do ii = 1, mi
          write(un) xu(ii), yp(jj), u(ii, jj)
       end do
    end do
    close(un)
    open(newunit=un, file=prefijo//'v.bin', access='stream', status='replace')
    do jj = 1, nj
       do ii = 1, mi + 1
          write(un) xp(ii), yv(jj), v(ii, jj)
       end do
    end do
    close(un)
    open(newunit=un, file=prefijo//'p.bin', access='stream', status='replace')
    do jj = 1, nj + 1
       do ii = 1, mi + 1
          write(un) temp(ii, jj), pres(ii, jj)
       end do
    end do
    close(un)
  end subroutine postproceso_bin

end module postproceso
