!
! IXCHEL2D on one MI355X: the driver of upstream's source/main.f90 with the
! time-step loop body on the GPU.
!
! Reads the same case directory upstream reads (parametros.dat, the three
! mesh/initial-field files, cond_frontera{u,v,t}.dat) and writes the same
! outputs: fort.102 ('SIMPLE', iteration, maxbo, residuo), the nuss_sim file
! every 1000 iterations, and per package the out_*.dat restart files, the
! exact .bin dumps and the t_*.vtk sample when postprocesar is true.  The mesh extents come from the
! files (no recompilation per mesh).  Environment: IXCHEL2D_SOLVER = pcr
! (default) | thomas selects the line solver, IXCHEL2D_PRECISION = fp64
! (default) | fp32_mixed | fp32_refined the arithmetic of the sweeps (PCR
! only).  IXCHEL2D_LOOP = phases runs every time step through the
! phase-level procedures of ec_momento, ec_energia and residuos with
! upstream's own loops, shared pass counter and exit rules (as a main.f90
! that keeps its control flow would), instead of ixh_simple_iterations.
! IXCHEL2D_CUERPO = trian | cuadr | chime passes that immersed-body option
! to definir_cuerpo in place of upstream's 'sincu'.
!
program ixchel2d_gpu
  use iso_c_binding
  use malla
  use cond_frontera
  use ec_momento
  use ec_energia
  use residuos
  use frontera_inmersa
  use postproceso
  use ixchel_hip
  implicit none

  character(len=28) :: entrada_u, entrada_v, entrada_tp
  character(len=8) :: adimen, flujo_ini, tempe_ini
  character(len=16) :: solver_env, loop_env, cuerpo_env
  character(len=6) :: Rec
  character(len=64) :: prefijo, nussf
  logical :: postprocesar
  real(DBL) :: Ra, Pr, Ri_1, dt, conv_resi, ao, tiempo_inicial
  real(DBL) :: nusselt0, nusselt1, temp_med, residuo_ult, maxbo_ult
  integer :: paq_itera, itermax, iter_ecuaci_max, iter_simple_max
  integer :: placa_min, placa_max, itera_inicial, itera, itera_total, ll, resto, n, k, cap, un, ic
  integer :: iter_ecuaci, iter_simple
  logical :: por_fases
  type(ixh_grid_t) :: grid
  type(ixh_params_t) :: prm
  type(ixh_hist_t), allocatable :: hist(:)
  integer(c_int32_t) :: nrec

  write(*,*) "------------------------------------------"
  write(*,*) "          IXCHEL2D (MI355X, HIP)"
  write(*,*) "------------------------------------------"
  postprocesar = .false.
  open(newunit=un, file='parametros.dat', status='old', action='read')
  read(un, *) adimen
  read(un, *) Ra
  read(un, *) Pr
  read(un, *) Ri_1
  read(un, *) dt
  read(un, *) paq_itera
  read(un, *) itermax
  read(un, *) rel_pres
  read(un, *) rel_vel
  read(un, *) rel_ener
  read(un, *) conv_u
  read(un, *) conv_t
  read(un, *) conv_p
  read(un, *) conv_resi
  read(un, *) conv_paso
  read(un, *) iter_ecuaci_max
  read(un, *) iter_simple_max
  read(un, *) entrada_u
  read(un, *) entrada_v
  read(un, *) entrada_tp
  read(un, *) flujo_ini
  read(un, *) tempe_ini
  read(un, *) postprocesar
  close(un)

  ! non-dimensionalisation (main.f90:221-244)
  if (trim(adimen) == 'natural') then
     gamma_momen = sqrt(Pr/Ra)
     gamma_energ = sqrt(1._DBL/(Pr*Ra))
     Ri = 0.0_DBL
     Riy = -1.0_DBL
     Rec = entero_caracter(ceiling(sqrt(Ra)))
  else if (trim(adimen) == 'mixta') then
     gamma_momen = 1.0_DBL/Ra
     gamma_energ = 1.0_DBL/(Ra*Pr)
     Ri = 0.0_DBL
     Riy = Ri_1
     Rec = entero_caracter(ceiling(Ra))
  else
     error stop 'adimen must be natural or mixta'
  end if
  if (paq_itera < 1) error stop 'paq_itera must be positive'
  ! the VTK sample tag counts tenths of a package (upstream divides by
  ! paq_itera/10 there)
  if (postprocesar .and. paq_itera < 10) error stop 'paq_itera must be at least 10 to name the VTK samples'

  call dimensiones_malla(entrada_u, entrada_tp)
  mic = entero_caracter(mi)
  njc = entero_caracter(nj)
  call reserva_momento()
  call reserva_energia()
  call lectura_mallas_escalonadas(entrada_u, entrada_v, entrada_tp, u_ant, v_ant, pres, temp_ant, &
       ao, placa_min, placa_max, itera_inicial)
  tiempo_inicial = itera_inicial * dt
  call condicion_inicial_flujo(flujo_ini)
  call condicion_inicial_tempe(tempe_ini)
  u = u_ant
  v = v_ant
  temp = temp_ant

  ! device state (replaces the !$acc data region, main.f90:356-389)
  grid % mi = mi
  grid % nj = nj
  grid % xu = c_loc(xu)
  grid % yv = c_loc(yv)
  grid % xp = c_loc(xp)
  grid % yp = c_loc(yp)
  prm % dt = dt
  prm % rel_pres = rel_pres
  prm % rel_vel = rel_vel
  prm % rel_ener = rel_ener
  prm % gamma_momen = gamma_momen
  prm % gamma_energ = gamma_energ
  prm % ri_x = Ri
  prm % ri_y = Riy
  prm % conv_u = conv_u
  prm % conv_t = conv_t
  prm % conv_p = conv_p
  prm % conv_resi = conv_resi
  prm % conv_paso = conv_paso
  prm % iter_ecuaci_max = iter_ecuaci_max
  prm % iter_simple_max = iter_simple_max
  prm % solver = IXH_SOLVER_PCR
  prm % precision = IXH_FP64
  call get_environment_variable('IXCHEL2D_SOLVER', solver_env)
  if (trim(solver_env) == 'thomas') prm % solver = IXH_SOLVER_THOMAS
  call get_environment_variable('IXCHEL2D_PRECISION', solver_env)
  if (trim(solver_env) == 'fp32_mixed') prm % precision = IXH_FP32_MIXED
  if (trim(solver_env) == 'fp32_refined') prm % precision = IXH_FP32_REFINED
  if (trim(solver_env) == 'fp32_stochastic') prm % precision = IXH_FP32_STOCHASTIC
  call get_environment_variable('IXCHEL2D_LOOP', loop_env)
  por_fases = trim(loop_env) == 'phases'
  call ixh_check(ixh_create(grid, prm, estado), 'ixh_create')
  call ixh_check(ixh_upload_field(estado, IXH_F_U_ANT, u_ant, mi), 'upload u_ant')
  call ixh_check(ixh_upload_field(estado, IXH_F_V_ANT, v_ant, mi+1), 'upload v_ant')
  call ixh_check(ixh_upload_field(estado, IXH_F_T_ANT, temp_ant, mi+1), 'upload temp_ant')
  call ixh_check(ixh_upload_field(estado, IXH_F_U, u, mi), 'upload u')
  call ixh_check(ixh_upload_field(estado, IXH_F_V, v, mi+1), 'upload v')
  call ixh_check(ixh_upload_field(estado, IXH_F_T, temp, mi+1), 'upload temp')
  call ixh_check(ixh_upload_field(estado, IXH_F_P, pres, mi+1), 'upload pres')
  call ini_frontera_uv()
  call ini_frontera_t()
  ! the driver's body option (upstream passes 'sincu', main.f90:327)
  call get_environment_variable('IXCHEL2D_CUERPO', cuerpo_env)
  if (len_trim(cuerpo_env) == 0) cuerpo_env = 'sincu'
  call definir_cuerpo(trim(cuerpo_env))
  write(*,'(1X,A,I6,A,I6,A,A)') 'mi=', mi, ', nj=', nj, ', Re=', Rec

  ! SIMPLE records per time step are at most iter_simple_max + 2
  cap = 1000 * max(1, iter_simple_max + 2)
  allocate(hist(cap))
  open(unit=102, file='fort.102', status='replace')
  nussf = 'nuss_sim_n'//trim(njc)//'m'//trim(mic)//'_R'//trim(Rec)//'.dat'
  itera = 0
  residuo_ult = 0.0_DBL
  maxbo_ult = 0.0_DBL
  ! pass counters, shared by the inner loops (main.f90:313-314)
  iter_ecuaci = 0
  iter_simple = 0
  do ll = 1, itermax / paq_itera
     resto = paq_itera
     do while (resto > 0)
        ! stop at every multiple of 500 (progress line) and 1000 (Nusselt)
        if (por_fases) then
           n = 1
           call paso_por_fases(maxbo_ult, residuo_ult)
        else
           n = min(resto, 500 - mod(itera, 500), cap / max(1, iter_simple_max + 2))
           call ixh_check(ixh_simple_iterations(estado, int(n, c_int32_t), hist, int(cap, c_int32_t), nrec), &
                'ixh_simple_iterations')
           do k = 1, min(int(nrec), cap)
              if (hist(k) % logged /= 0) write(102, *) 'SIMPLE', hist(k) % simple_iter, hist(k) % maxbo, &
                   hist(k) % residuo
           end do
           if (nrec > 0) then
              residuo_ult = hist(min(int(nrec), cap)) % residuo
              maxbo_ult = hist(min(int(nrec), cap)) % maxbo
           end if
        end if
        itera = itera + n
        resto = resto - n
        if (mod(itera, 500) == 0) write(*,*) 'ITERACION: ', itera, residuo_ult, maxbo_ult
        if (mod(itera, 1000) == 0) then
           call nusselt_promedio_y(nusselt0, nusselt1, placa_min, placa_max)
           call ixh_check(ixh_download_field(estado, IXH_F_T, temp, mi+1), 'download temp')
           ic = (placa_min + placa_max) / 2
           temp_med = (temp(ic, nj/2+1) + temp(ic+1, nj/2+1)) / 2._DBL
           open(newunit=un, file=trim(nussf), position='append')
           write(un, form26) tiempo_inicial + itera*dt, nusselt0, -nusselt1, temp_med, 0.0_DBL, 0.0_DBL
           close(un)
        end if
     end do
     ! end of the package: fields back to the host (main.f90:1277-1354)
     call ixh_check(ixh_download_field(estado, IXH_F_U, u, mi), 'download u')
     call ixh_check(ixh_download_field(estado, IXH_F_V, v, mi+1), 'download v')
     call ixh_check(ixh_download_field(estado, IXH_F_T, temp, mi+1), 'download temp')
     call ixh_check(ixh_download_field(estado, IXH_F_P, pres, mi+1), 'download pres')
     itera_total = itera_inicial + itera
     write(*,*) 'itera_total=', itera_total
     write(*,'(1X,A,D23.15,A,D23.15)') 'b_o  =', maxbo_ult, ', Res_u=', residuo_ult
     if (postprocesar) then
        prefijo = 'out_n'//trim(njc)//'m'//trim(mic)//'_R'//trim(Rec)
        call postproceso_dat(trim(prefijo), u, v, temp, pres, placa_min, placa_max, itera_total, ao)
        call postproceso_bin(trim(prefijo), u, v, temp, pres)
        call postproceso_vtk('n'//trim(njc)//'m'//trim(mic)//'R'//trim(Rec)//'/t_'// &
             muestra_vtk(itera_total, paq_itera)//'.vtk', u, v, temp, pres)
     end if
  end do
  close(102)
  call ixh_check(ixh_destroy(estado), 'ixh_destroy')

contains

  ! One time step of main.f90:395-1270 through the phase procedures: the
  ! momentum loop (u then v sweeps, error of v), p' = 0 and the p' loop
  ! (error of p', maxbo), the correction, the energy loop, residuo_u, the
  ! SIMPLE exit test with its fort.102 line, and the lagged copies.  The
  ! inner loops exit on error < conv or after iter_ecuaci_max + 2 passes;
  ! iter_ecuaci is shared by all three and reset on every exit.
  subroutine paso_por_fases(maxbo_o, residuo_o)
    real(DBL), intent(out) :: maxbo_o, residuo_o
    real(c_double) :: error, mb, residuo
    simple: do
       momento: do
          call ecuacion_velu(IXH_DIR_Y)
          call ecuacion_velu(IXH_DIR_X)
          call ecuacion_velv(IXH_DIR_Y, error)
          call ecuacion_velv(IXH_DIR_X, error)
          if (salida(error, conv_u, 'ADVER. MOMEN: convergencia no alcanzada ')) exit
       end do momento
       call inicializa_corrector_presion()
       presion: do
          call ecuacion_corr_pres(IXH_DIR_Y, error, mb)
          call ecuacion_corr_pres(IXH_DIR_X, error, mb)
          maxbo = mb
          if (salida(error, conv_p, 'ADVER. PRES: convergencia no alcanzada ')) exit
       end do presion
       call corrige_presion_velocidad()
       energia: do
          call ecuacion_energia(IXH_DIR_Y, error)
          call ecuacion_energia(IXH_DIR_X, error)
          if (salida(error, conv_t, 'Adver. ENERG: convergencia no alcanzada ')) exit
       end do energia
       call residuo_u(residuo)
       if (maxbo < conv_paso .and. residuo < conv_resi) then
          iter_simple = 0
          write(102, *) 'SIMPLE', iter_simple, maxbo, residuo
          exit
       else if (iter_simple > iter_simple_max) then
          iter_simple = 0
          exit
       else
          iter_simple = iter_simple + 1
          write(102, *) 'SIMPLE', iter_simple, maxbo, residuo
       end if
    end do simple
    call ixh_check(ixh_lag(estado), 'ixh_lag')
    maxbo_o = maxbo
    residuo_o = residuo
  end subroutine paso_por_fases

  ! exit test of an inner loop (main.f90:729-740, 913-924, 1144-1156)
  logical function salida(error, conv, aviso)
    real(c_double), intent(in) :: error
    real(DBL), intent(in) :: conv
    character(len=*), intent(in) :: aviso
    salida = .true.
    if (error < conv) then
       iter_ecuaci = 0
    else if (iter_ecuaci > iter_ecuaci_max) then
       iter_ecuaci = 0
       write(*, *) aviso, error
    else
       iter_ecuaci = iter_ecuaci + 1
       salida = .false.
    end if
  end function salida

end program ixchel2d_gpu
