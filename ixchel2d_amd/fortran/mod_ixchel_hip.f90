!
! ixchel_hip: iso_c_binding interfaces to libixchel2d.so (include/ixchel2d.h).
!
! Every C entry point is bound here with its exact symbol name; the modules
! that keep Ixchel2D's names (malla, cond_frontera, ec_momento, ec_energia,
! solucionador, residuos, frontera_inmersa, postproceso) call through these.
! Status codes: 0 = ok, negative = error; ixh_check aborts with the library's
! message, matching upstream's fail-stop behaviour.
!
module ixchel_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: IXH_F_U = 0, IXH_F_V = 1, IXH_F_T = 2, IXH_F_P = 3
  integer(c_int), parameter :: IXH_F_U_ANT = 4, IXH_F_V_ANT = 5, IXH_F_T_ANT = 6
  integer(c_int), parameter :: IXH_F_PCORR = 7, IXH_F_BO = 8, IXH_F_AU = 9, IXH_F_AV = 10, IXH_F_RESU = 11
  integer(c_int), parameter :: IXH_VAR_U = 0, IXH_VAR_V = 1, IXH_VAR_T = 2
  integer(c_int), parameter :: IXH_DIR_Y = 0, IXH_DIR_X = 1
  integer(c_int), parameter :: IXH_SOLVER_PCR = 0, IXH_SOLVER_THOMAS = 1
  integer(c_int), parameter :: IXH_FP64 = 0, IXH_FP32_MIXED = 1, IXH_FP32_REFINED = 2, IXH_FP32_STOCHASTIC = 3

  type, bind(C) :: ixh_grid_t
     integer(c_int32_t) :: mi, nj
     type(c_ptr) :: xu, yv, xp, yp
  end type ixh_grid_t

  type, bind(C) :: ixh_params_t
     real(c_double) :: dt, rel_pres, rel_vel, rel_ener
     real(c_double) :: gamma_momen, gamma_energ
     real(c_double) :: ri_x, ri_y
     real(c_double) :: conv_u, conv_t, conv_p, conv_resi, conv_paso
     integer(c_int32_t) :: iter_ecuaci_max, iter_simple_max
     integer(c_int32_t) :: solver, precision
  end type ixh_params_t

  type, bind(C) :: ixh_hist_t
     real(c_double) :: err_v, err_p, maxbo, err_t, residuo
     integer(c_int32_t) :: passes_mom, passes_p, passes_t, simple_iter, logged, pad_
  end type ixh_hist_t

  interface
     function ixh_last_error() bind(C, name='ixh_last_error') result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function ixh_last_error
     function ixh_version() bind(C, name='ixh_version') result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function ixh_version
     function ixh_kernel_name(id) bind(C, name='ixh_kernel_name') result(p)
       import :: c_ptr, c_int32_t
       integer(c_int32_t), value :: id
       type(c_ptr) :: p
     end function ixh_kernel_name
     integer(c_int) function ixh_create(grid, params, state) bind(C, name='ixh_create')
       import :: c_int, c_ptr, ixh_grid_t, ixh_params_t
       type(ixh_grid_t), intent(in) :: grid
       type(ixh_params_t), intent(in) :: params
       type(c_ptr), intent(out) :: state
     end function ixh_create
     integer(c_int) function ixh_destroy(state) bind(C, name='ixh_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: state
     end function ixh_destroy
     integer(c_int) function ixh_set_params(state, params) bind(C, name='ixh_set_params')
       import :: c_int, c_ptr, ixh_params_t
       type(c_ptr), value :: state
       type(ixh_params_t), intent(in) :: params
     end function ixh_set_params
     integer(c_int) function ixh_set_stream(state, stream) bind(C, name='ixh_set_stream')
       import :: c_int, c_ptr
       type(c_ptr), value :: state, stream
     end function ixh_set_stream
     integer(c_int) function ixh_synchronize(state) bind(C, name='ixh_synchronize')
       import :: c_int, c_ptr
       type(c_ptr), value :: state
     end function ixh_synchronize
     integer(c_int) function ixh_upload_field(state, field, src, ld) bind(C, name='ixh_upload_field')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: state
       integer(c_int32_t), value :: field, ld
       real(c_double), intent(in) :: src(*)
     end function ixh_upload_field
     integer(c_int) function ixh_download_field(state, field, dst, ld) bind(C, name='ixh_download_field')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: state
       integer(c_int32_t), value :: field, ld
       real(c_double), intent(out) :: dst(*)
     end function ixh_download_field
     integer(c_int) function ixh_set_bc(state, var, side, tipo, valor) bind(C, name='ixh_set_bc')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: state
       integer(c_int32_t), value :: var, side
       integer(c_int32_t), intent(in) :: tipo(*)
       real(c_double), intent(in) :: valor(*)
     end function ixh_set_bc
     integer(c_int) function ixh_set_sources(state, var, sc, sp) bind(C, name='ixh_set_sources')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: state
       integer(c_int32_t), value :: var
       type(c_ptr), value :: sc, sp
     end function ixh_set_sources
     integer(c_int) function ixh_momentum_u(state, dir) bind(C, name='ixh_momentum_u')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: state
       integer(c_int32_t), value :: dir
     end function ixh_momentum_u
     integer(c_int) function ixh_momentum_v(state, dir, err_v) bind(C, name='ixh_momentum_v')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: state
       integer(c_int32_t), value :: dir
       type(c_ptr), value :: err_v
     end function ixh_momentum_v
     integer(c_int) function ixh_pcorr_zero(state) bind(C, name='ixh_pcorr_zero')
       import :: c_int, c_ptr
       type(c_ptr), value :: state
     end function ixh_pcorr_zero
     integer(c_int) function ixh_pcorr(state, dir, err_p, maxbo) bind(C, name='ixh_pcorr')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: state
       integer(c_int32_t), value :: dir
       type(c_ptr), value :: err_p, maxbo
     end function ixh_pcorr
     integer(c_int) function ixh_correct(state) bind(C, name='ixh_correct')
       import :: c_int, c_ptr
       type(c_ptr), value :: state
     end function ixh_correct
     integer(c_int) function ixh_energy(state, dir, err_t) bind(C, name='ixh_energy')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: state
       integer(c_int32_t), value :: dir
       type(c_ptr), value :: err_t
     end function ixh_energy
     integer(c_int) function ixh_residual_u(state, residuo) bind(C, name='ixh_residual_u')
       import :: c_int, c_ptr
       type(c_ptr), value :: state
       type(c_ptr), value :: residuo
     end function ixh_residual_u
     integer(c_int) function ixh_lag(state) bind(C, name='ixh_lag')
       import :: c_int, c_ptr
       type(c_ptr), value :: state
     end function ixh_lag
     integer(c_int) function ixh_simple_iterations(state, nsteps, hist, cap, nrec) &
          bind(C, name='ixh_simple_iterations')
       import :: c_int, c_ptr, c_int32_t, ixh_hist_t
       type(c_ptr), value :: state
       integer(c_int32_t), value :: nsteps, cap
       type(ixh_hist_t), intent(out) :: hist(*)
       integer(c_int32_t), intent(out) :: nrec
     end function ixh_simple_iterations
     integer(c_int) function ixh_nusselt_y(state, i0, i1, nu0, nu1) bind(C, name='ixh_nusselt_y')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: state
       integer(c_int32_t), value :: i0, i1
       real(c_double), intent(out) :: nu0, nu1
     end function ixh_nusselt_y
     integer(c_int) function ixh_nusselt_x(state, nu_w, nu_e) bind(C, name='ixh_nusselt_x')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: state
       real(c_double), intent(out) :: nu_w, nu_e
     end function ixh_nusselt_x
     integer(c_int) function ixh_tridiagonal(a, b, c, r, n, nsys, solver) bind(C, name='ixh_tridiagonal')
       import :: c_int, c_int32_t, c_double
       real(c_double), intent(in) :: a(*), c(*)
       real(c_double), intent(inout) :: b(*), r(*)
       integer(c_int32_t), value :: n, nsys, solver
     end function ixh_tridiagonal
     integer(c_int) function ixh_tridiagonal_noalloc(a, b, c, r, n, nsys, solver) &
          bind(C, name='ixh_tridiagonal_noalloc')
       import :: c_int, c_int32_t, c_double
       real(c_double), intent(in) :: a(*), c(*)
       real(c_double), intent(inout) :: b(*), r(*)
       integer(c_int32_t), value :: n, nsys, solver
     end function ixh_tridiagonal_noalloc
     integer(c_int) function ixh_set_profiling(state, on) bind(C, name='ixh_set_profiling')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: state
       integer(c_int32_t), value :: on
     end function ixh_set_profiling
     integer(c_int) function ixh_kernel_times(state, ms, launches, cap, nk) bind(C, name='ixh_kernel_times')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: state
       real(c_double), intent(out) :: ms(*)
       integer(c_int32_t), intent(out) :: launches(*)
       integer(c_int32_t), value :: cap
       integer(c_int32_t), intent(out) :: nk
     end function ixh_kernel_times
     ! batches of independent cavities (no upstream counterpart)
     integer(c_int) function ixh_batch_create(states, n, batch) bind(C, name='ixh_batch_create')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), intent(in) :: states(*)
       integer(c_int32_t), value :: n
       type(c_ptr), intent(out) :: batch
     end function ixh_batch_create
     integer(c_int) function ixh_batch_destroy(batch) bind(C, name='ixh_batch_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
     end function ixh_batch_destroy
     integer(c_int) function ixh_batch_set_schedule(batch, order, graph) bind(C, name='ixh_batch_set_schedule')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: batch
       integer(c_int32_t), value :: order, graph
     end function ixh_batch_set_schedule
     integer(c_int) function ixh_batch_iterations(batch, nsteps, hist, cap) bind(C, name='ixh_batch_iterations')
       import :: c_int, c_ptr, c_int32_t, ixh_hist_t
       type(c_ptr), value :: batch
       integer(c_int32_t), value :: nsteps, cap
       type(ixh_hist_t), intent(out) :: hist(*)
     end function ixh_batch_iterations
     integer(c_int) function ixh_batch_set_shape(batch, line_threads, lines_per_group, residual_with_t) &
          bind(C, name='ixh_batch_set_shape')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: batch
       integer(c_int32_t), value :: line_threads, lines_per_group, residual_with_t
     end function ixh_batch_set_shape
     integer(c_int) function ixh_batch_autotune(batch, trial_steps, order, graph, line_threads, lines_per_group, &
          residual_with_t, ms) bind(C, name='ixh_batch_autotune')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: batch
       integer(c_int32_t), value :: trial_steps
       integer(c_int32_t), intent(out) :: order, graph, line_threads, lines_per_group, residual_with_t
       real(c_double), intent(out) :: ms(*)
     end function ixh_batch_autotune
     ! discrete adjoint of the energy transport (no upstream counterpart)
     integer(c_int) function ixh_adjoint_record(state, nsteps, every) bind(C, name='ixh_adjoint_record')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: state
       integer(c_int32_t), value :: nsteps, every
     end function ixh_adjoint_record
     integer(c_int) function ixh_adjoint_energy(state, tbar, ld, grad_a, grad_b, grad_c, grad_d, grad_t0, ld0) &
          bind(C, name='ixh_adjoint_energy')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: state
       real(c_double), intent(in) :: tbar(*)
       integer(c_int32_t), value :: ld, ld0
       type(c_ptr), value :: grad_a, grad_b, grad_c, grad_d, grad_t0
     end function ixh_adjoint_energy
     integer(c_int) function ixh_adjoint_coupled(state, tbar, ld, grad_a, grad_b, grad_c, grad_d, grad_t0, ld0) &
          bind(C, name='ixh_adjoint_coupled')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: state
       real(c_double), intent(in) :: tbar(*)
       integer(c_int32_t), value :: ld, ld0
       type(c_ptr), value :: grad_a, grad_b, grad_c, grad_d, grad_t0
     end function ixh_adjoint_coupled
  end interface

  ! the device state of this program's one cavity
  type(c_ptr), save :: estado = c_null_ptr

contains

  function c_cadena(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: ch(:)
    integer :: n
    interface
       integer(c_size_t) function c_strlen(q) bind(C, name='strlen')
         import :: c_ptr, c_size_t
         type(c_ptr), value :: q
       end function c_strlen
    end interface
    if (.not. c_associated(p)) then
       s = ''
       return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, ch, [n])
    allocate(character(len=n) :: s)
    s = transfer(ch(1:n), s)
  end function c_cadena

  subroutine ixh_check(rc, donde)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: donde
    if (rc /= 0) then
       write(*,*) 'IXCHEL2D-HIP: error ', rc, ' en ', donde, ': ', c_cadena(ixh_last_error())
       error stop 1
    end if
  end subroutine ixh_check

end module ixchel_hip
