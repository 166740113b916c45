!
! postproceso: Nusselt numbers and output files (upstream module
! source/mod_postproceso.f90).  The Nusselt integrals run on the GPU; the
! writers produce upstream's text (.dat, which are also restart files),
! stream (.bin) and big-endian VTK formats.
!
module postproceso
  use iso_c_binding
  use malla, only : mi, nj, DBL, xu, yv, xp, yp, form24, form25
  use ixchel_hip
  implicit none

  interface
     ! d_format.c: nrec records of ncol values appended as (ncolD23.15) lines
     integer(c_int) function d23_append(path, pathlen, vals, nrec, ncol) bind(C, name='d23_append')
       import :: c_double, c_int64_t, c_int32_t, c_int, c_char
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), value :: pathlen
       real(c_double), intent(in) :: vals(*)
       integer(c_int64_t), value :: nrec
       integer(c_int32_t), value :: ncol
     end function d23_append
  end interface

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
  ! (upstream's postproceso_dat: a list-directed header, then one (3D23.15)
  ! or (4D23.15) record per node).  The records are formatted and appended
  ! by d23_append (d_format.c); a field with a non-finite value goes through
  ! Fortran's formatted write instead.
  subroutine postproceso_dat(prefijo, u, v, temp, pres, placa_min, placa_max, itera_total, ao)
    character(len=*), intent(in) :: prefijo
    real(DBL), intent(in) :: u(mi, nj+1), v(mi+1, nj), temp(mi+1, nj+1), pres(mi+1, nj+1), ao
    integer, intent(in) :: placa_min, placa_max, itera_total
    real(DBL), allocatable :: rec(:, :)
    integer :: ii, jj, k
    allocate(rec(3, mi * (nj + 1)))
    k = 0
    do jj = 1, nj + 1
       do ii = 1, mi
          k = k + 1
          rec(1, k) = xu(ii)
          rec(2, k) = yp(jj)
          rec(3, k) = u(ii, jj)
       end do
    end do
    call write_dat(prefijo//'u.dat', rec)
    deallocate(rec)
    allocate(rec(3, (mi + 1) * nj))
    k = 0
    do jj = 1, nj
       do ii = 1, mi + 1
          k = k + 1
          rec(1, k) = xp(ii)
          rec(2, k) = yv(jj)
          rec(3, k) = v(ii, jj)
       end do
    end do
    call write_dat(prefijo//'v.dat', rec)
    deallocate(rec)
    allocate(rec(4, (mi + 1) * (nj + 1)))
    k = 0
    do jj = 1, nj + 1
       do ii = 1, mi + 1
          k = k + 1
          rec(1, k) = xp(ii)
          rec(2, k) = yp(jj)
          rec(3, k) = temp(ii, jj)
          rec(4, k) = pres(ii, jj)
       end do
    end do
    call write_dat(prefijo//'p.dat', rec)
  contains
    subroutine write_dat(archivo, rec)
      character(len=*), intent(in) :: archivo
      real(DBL), intent(in) :: rec(:, :)
      integer :: un, r, st
      open(newunit=un, file=archivo, status='replace')
      write(un, *) placa_min, placa_max, itera_total, ao
      close(un)
      st = d23_append(archivo, int(len(archivo), c_int32_t), rec, int(size(rec, 2), c_int64_t), &
           int(size(rec, 1), c_int32_t))
      if (st == 2) error stop 'postproceso_dat: cannot write the records'
      if (st == 0) return
      open(newunit=un, file=archivo, position='append', status='old')
      do r = 1, size(rec, 2)
         if (size(rec, 1) == 3) then
            write(un, form24) rec(:, r)
         else
            write(un, form25) rec(:, r)
         end if
      end do
      close(un)
    end subroutine write_dat
  end subroutine postproceso_dat

  ! postproceso_bin (mod_postproceso.f90:310-365): exact FP64 dumps
  subroutine postproceso_bin(prefijo, u, v, temp, pres)
    character(len=*), intent(in) :: prefijo
    real(DBL), intent(in) :: u(mi, nj+1), v(mi+1, nj), temp(mi+1, nj+1), pres(mi+1, nj+1)
    real(DBL), allocatable :: rec(:, :)
    integer :: ii, jj, k, un
    ! each file's records gathered first and written in one statement: the
    ! same bytes as one write per record, without the runtime's per-record cost
    allocate(rec(3, mi * (nj + 1)))
    k = 0
    do jj = 1, nj + 1
       do ii = 1, mi
          k = k + 1
          rec(1, k) = xu(ii)
          rec(2, k) = yp(jj)
          rec(3, k) = u(ii, jj)
       end do
    end do
    open(newunit=un, file=prefijo//'u.bin', access='stream', status='replace')
    write(un) rec
    close(un)
    deallocate(rec)
    allocate(rec(3, (mi + 1) * nj))
    k = 0
    do jj = 1, nj
       do ii = 1, mi + 1
          k = k + 1
          rec(1, k) = xp(ii)
          rec(2, k) = yv(jj)
          rec(3, k) = v(ii, jj)
       end do
    end do
    open(newunit=un, file=prefijo//'v.bin', access='stream', status='replace')
    write(un) rec
    close(un)
    deallocate(rec)
    allocate(rec(2, (mi + 1) * (nj + 1)))
    k = 0
    do jj = 1, nj + 1
       do ii = 1, mi + 1
          k = k + 1
          rec(1, k) = temp(ii, jj)
          rec(2, k) = pres(ii, jj)
       end do
    end do
    open(newunit=un, file=prefijo//'p.bin', access='stream', status='replace')
    write(un) rec
    close(un)
  end subroutine postproceso_bin

  ! Node velocities of the VTK output (main.f90:1295-1313).  Interior p nodes
  ! take the mean of their two faces; the side columns take the boundary
  ! face of u and the ghost column of v; the bottom and top rows take the
  ! wall faces of v, while u there stays 0 apart from the corners, as
  ! upstream leaves it.
  subroutine velocidad_nodal(u, v, uf, vf)
    real(DBL), intent(in) :: u(mi, nj+1), v(mi+1, nj)
    real(DBL), intent(out) :: uf(mi+1, nj+1), vf(mi+1, nj+1)
    integer :: ii, jj
    uf = 0.0_DBL
    vf = 0.0_DBL
    do jj = 2, nj
       do ii = 2, mi
          uf(ii, jj) = (u(ii, jj) + u(ii-1, jj)) / 2._DBL
          vf(ii, jj) = (v(ii, jj) + v(ii, jj-1)) / 2._DBL
       end do
       vf(1, jj) = v(1, jj)
       vf(mi+1, jj) = v(mi+1, jj)
    end do
    uf(1, :) = u(1, :)
    uf(mi+1, :) = u(mi, :)
    vf(:, 1) = v(:, 1)
    vf(:, nj+1) = v(:, nj)
  end subroutine velocidad_nodal

  ! postproceso_vtk (mod_postproceso.f90:225-300): legacy VTK, binary and
  ! big-endian, a structured grid on the p nodes (ghosts included) carrying
  ! the pressure, the temperature and the node velocities in FP32.  The
  ! header bytes are upstream's: its integers in list-directed form, and no
  ! line break between the POINT_DATA count and the next keyword (VTK
  ! readers take the count and the keyword apart).
  subroutine postproceso_vtk(archivo, u, v, temp, pres)
    character(len=*), intent(in) :: archivo
    real(DBL), intent(in) :: u(mi, nj+1), v(mi+1, nj), temp(mi+1, nj+1), pres(mi+1, nj+1)
    real(DBL), allocatable :: uf(:, :), vf(:, :)
    real, allocatable :: trip(:, :)  ! point coordinates, then node velocities, as written
    character(len=32) :: cx, cy, cz, cn
    character(len=1) :: nl
    integer :: ii, jj, k, un
    allocate(uf(mi+1, nj+1), vf(mi+1, nj+1), trip(3, (mi+1) * (nj+1)))
    call velocidad_nodal(u, v, uf, vf)
    nl = new_line('a')
    write(cx, *) mi + 1
    write(cy, *) nj + 1
    write(cz, *) 1
    write(cn, *) (mi + 1) * (nj + 1)
    ! native stream; the FP32 arrays go out big-endian through write_be
    open(newunit=un, file=archivo, access='stream', status='replace')
    write(un) '# vtk DataFile Version 2.3'//nl, '3D Mesh'//nl, 'BINARY'//nl, 'DATASET STRUCTURED_GRID'//nl, &
         'DIMENSIONS '//trim(cx)//trim(cy)//trim(cz)//nl, 'POINTS '//trim(cn)//' float'//nl
    k = 0
    do jj = 1, nj + 1
       do ii = 1, mi + 1
          k = k + 1
          trip(1, k) = real(xp(ii))
          trip(2, k) = real(yp(jj))
          trip(3, k) = 0.0
       end do
    end do
    call write_be(un, trip, size(trip))
    write(un) nl//'POINT_DATA '//trim(cn), 'SCALARS PRESS float'//nl, 'LOOKUP_TABLE default'//nl
    call write_be(un, real(pres), size(pres))
    write(un) nl//'SCALARS TEMPER float'//nl, 'LOOKUP_TABLE default'//nl
    call write_be(un, real(temp), size(temp))
    write(un) nl//'VECTORS VELOCITY float'//nl
    k = 0
    do jj = 1, nj + 1
       do ii = 1, mi + 1
          k = k + 1
          trip(1, k) = real(uf(ii, jj))
          trip(2, k) = real(vf(ii, jj))
          trip(3, k) = 0.0
       end do
    end do
    call write_be(un, trip, size(trip))
    close(un)
  end subroutine postproceso_vtk

  ! n FP32 values to a native stream unit as big-endian bytes (the bytes of
  ! convert='big_endian', which swaps one value at a time and costs about
  ! 90 ns a value in flang's runtime)
  subroutine write_be(un, x, n)
    integer, intent(in) :: un, n
    real, intent(in) :: x(n)
    integer(4), allocatable :: w(:)
    w = transfer(x, 0_4, n)
    if (transfer(1_4, 'a') == achar(1)) &  ! little-endian host: swap the bytes
         w = ior(ior(ishft(w, 24), iand(ishft(w, 8), int(z'00FF0000', 4))), &
                 ior(iand(ishft(w, -8), int(z'0000FF00', 4)), ishft(w, -24)))
    write(un) w
  end subroutine write_be

  ! Sample tag of the VTK file after itera_total iterations in packages of
  ! paq (main.f90:1282-1294): the thousands, hundreds, tens and units digits
  ! of the package count, then the tenth of a package, one I1 field each.
  function muestra_vtk(itera_total, paq) result(s)
    integer, intent(in) :: itera_total, paq
    character(len=5) :: s
    integer :: k, r
    k = itera_total / paq
    r = mod(itera_total, paq)
    write(s, '(5I1)') k / 1000, mod(k, 1000) / 100, mod(k, 100) / 10, mod(k, 10), r / (paq / 10)
  end function muestra_vtk

end module postproceso
