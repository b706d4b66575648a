! dgv_gpu.f90 -- iso_c_binding shim: the reference's collision / spatial-operator /
! stepper procedures, same names and argument lists, running on the GPU through
! the C ABI of include/dgv.h (libdgv_amd.so).
!
! Replaces, for a Fortran driver such as DGVblzmPL.f90's serial path:
!   EvalCollisionPeriodicAPlus_DGV(f,fc)  collision_mod.f90:143-229
!   EvalCollisionPeriodicA_DGV(f,fc)      collision_mod.f90:29-129
!   SpatialOperSH_DGV(f,fcol,dt)          spat_oper_mod.f90:28-68 (run_mode 0)
!   TimeIntegratorORK_SH_DGV(f,dt)        time_integr_mod.f90:26-152
!   TimeIntegratorMTS_SH_DGV(f,dt)        time_integr_mod.f90:314-364
!   PrepareMTS_SH_DGV(f,dt)               time_integr_mod.f90:443-569
!   MassCheckRec(f,n,ubar,vbar,wbar,temp) dgvtools_mod.f90:1437-1458
!   EvalCollisionPeriodicMixedTermsA_DGV(f,fm,fc)  collision_mod.f90:244-332
!   PrepareFMA_DGV                        collision_mod.f90:377-464 (module fm)
!   EvalCollisionLinear(f,fc)             collision_mod.f90:346-359
!   CheckSolutionMode_DGV(f,Df)           spat_oper_mod.f90:195-272
! run_mode, decomp_lev, linear_lev and fm are module variables here (commvar in
! the reference); SpatialOperSH_DGV and the steppers honour the regime switch
! when decomp_lev > 0, and leave run_mode and fm as the last check set them.
! The grid and A-array that the reference keeps in commvar are handed over once
! with dgv_gpu_init and dgv_gpu_set_aarray (or generated with
! dgv_gpu_generate_aarray).  A nonzero status from the library prints the
! library's message and stops, as the reference does on its errors.
module dgv_gpu
  use iso_c_binding
  implicit none
  private

  integer, parameter :: DP = kind(1.0d0)

  type, bind(C) :: dgv_grid_desc
    integer(c_int32_t) :: Mu, Mv, Mw, su, sv, sw
    real(c_double) :: u_L, u_R, v_L, v_R, w_L, w_R
  end type

  type, bind(C) :: dgv_agen_params
    real(c_double) :: Trad, min_sens
    integer(c_int32_t) :: nquad, moment_mode
  end type

  type(c_ptr), save :: ctx = c_null_ptr
  integer(c_int64_t), save :: nnodes = 0
  ! (mol_diam/L_inf)**2*N_inf of SpatialOperSH_DGV and the stepper order
  real(DP), save, public :: mol_diam = 1.0_DP, L_inf = 1.0_DP, N_inf = 1.0_DP
  integer, save, public :: rkmts = 2
  real(DP), save, public :: e_time = 0.0_DP
  real(DP), allocatable, save, public :: frhs1(:)
  real(DP), allocatable, save :: hist(:,:)      ! f1..f4, frhs1..frhs5
  ! decomposition / linearised regimes
  integer, save, public :: run_mode = 0
  real(DP), save, public :: decomp_lev = 0.0_DP, linear_lev = 0.0_DP
  real(DP), allocatable, save, public :: fm(:)
  integer, save :: pushed_mode = -1
  real(DP), save :: pushed_decomp = -1.0_DP, pushed_linear = -1.0_DP

  public :: dgv_gpu_init, dgv_gpu_set_aarray, dgv_gpu_generate_aarray, dgv_gpu_finalize
  public :: EvalCollisionPeriodicAPlus_DGV, EvalCollisionPeriodicA_DGV, SpatialOperSH_DGV
  public :: TimeIntegratorORK_SH_DGV, TimeIntegratorMTS_SH_DGV, PrepareMTS_SH_DGV, MassCheckRec
  public :: EvalCollisionPeriodicMixedTermsA_DGV, PrepareFMA_DGV, EvalCollisionLinear, CheckSolutionMode_DGV
  public :: MassCheckRecHighCMoments, dgv_gpu_get_history, dgv_gpu_set_history
  public :: IphiDGV, IphiDGV_decomp

  interface
    integer(c_int) function c_ctx_create(grid, device, out) bind(C, name="dgv_ctx_create")
      import :: c_int, c_ptr, dgv_grid_desc
      type(dgv_grid_desc), intent(in) :: grid
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: out
    end function
    subroutine c_ctx_destroy(c) bind(C, name="dgv_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: c
    end subroutine
    type(c_ptr) function c_last_error(c) bind(C, name="dgv_last_error")
      import :: c_ptr
      type(c_ptr), value :: c
    end function
    integer(c_int) function c_num_nodes(c, n) bind(C, name="dgv_ctx_num_nodes")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: c
      integer(c_int64_t), intent(out) :: n
    end function
    integer(c_int) function c_set_aarray(c, capphi, ncap, A, xi, xi1, nrec, base) bind(C, name="dgv_ctx_set_aarray")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: c
      integer(c_int32_t), intent(in) :: capphi(*), xi(*), xi1(*)
      integer(c_int64_t), value :: ncap, nrec
      real(c_double), intent(in) :: A(*)
      integer(c_int), value :: base
    end function
    integer(c_int) function c_generate(c, p) bind(C, name="dgv_ctx_generate_aarray")
      import :: c_int, c_ptr, dgv_agen_params
      type(c_ptr), value :: c
      type(dgv_agen_params), intent(in) :: p
    end function
    integer(c_int) function c_eval_q(c, f, fc, n) bind(C, name="dgv_eval_q")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: c
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: fc(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function c_eval_q_a(c, f, fc, n) bind(C, name="dgv_eval_q_a")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: c
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: fc(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function c_spatial_oper(c, f, fcol, n, dt, coef) bind(C, name="dgv_spatial_oper")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: c
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: fcol(*)
      integer(c_int64_t), value :: n
      real(c_double), value :: dt, coef
    end function
    integer(c_int) function c_rk_step(c, f, n, dt, coef, k, frhs) bind(C, name="dgv_rk_step")
      import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
      type(c_ptr), value :: c
      real(c_double), intent(inout) :: f(*)
      integer(c_int64_t), value :: n
      real(c_double), value :: dt, coef
      integer(c_int32_t), value :: k
      real(c_double), intent(out) :: frhs(*)
    end function
    integer(c_int) function c_ab_step(c, f, n, dt, coef, k, h) bind(C, name="dgv_ab_step")
      import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
      type(c_ptr), value :: c
      real(c_double), intent(inout) :: f(*)
      integer(c_int64_t), value :: n
      real(c_double), value :: dt, coef
      integer(c_int32_t), value :: k
      real(c_double), intent(inout) :: h(*)
    end function
    integer(c_int) function c_moments(c, f, n, out) bind(C, name="dgv_moments")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: c
      real(c_double), intent(in) :: f(*)
      integer(c_int64_t), value :: n
      real(c_double), intent(out) :: out(5)
    end function
    integer(c_int) function c_eval_mixed(c, f, g, fc, n) bind(C, name="dgv_eval_mixed")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: c
      real(c_double), intent(in) :: f(*), g(*)
      real(c_double), intent(out) :: fc(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function c_prepare_fma(c, g, n) bind(C, name="dgv_prepare_fma")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: c
      real(c_double), intent(in) :: g(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function c_eval_linear(c, f, fc, n) bind(C, name="dgv_eval_linear")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: c
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: fc(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function c_set_regime(c, mode, dlev, llev) bind(C, name="dgv_ctx_set_regime")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: c
      integer(c_int32_t), value :: mode
      real(c_double), value :: dlev, llev
    end function
    integer(c_int) function c_get_regime(c, mode, l1, l1fm, counts) bind(C, name="dgv_ctx_get_regime")
      import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: c
      integer(c_int32_t), intent(out) :: mode
      real(c_double), intent(out) :: l1, l1fm
      integer(c_int64_t), intent(out) :: counts(2)
    end function
    integer(c_int) function c_check_mode(c, f, df, n) bind(C, name="dgv_check_solution_mode")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: c
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: df(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function c_get_fm(c, g, n) bind(C, name="dgv_ctx_get_fm")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: c
      real(c_double), intent(out) :: g(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function c_moments_high(c, f, n, out) bind(C, name="dgv_moments_high")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: c
      real(c_double), intent(in) :: f(*)
      integer(c_int64_t), value :: n
      real(c_double), intent(out) :: out(20)
    end function
    integer(c_int) function c_iphi(c, f, out, n, coef) bind(C, name="dgv_iphi")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: c
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: out(*)
      integer(c_int64_t), value :: n
      real(c_double), value :: coef
    end function
    integer(c_int) function c_iphi_decomp(c, f, out, n, coef) bind(C, name="dgv_iphi_decomp")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: c
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: out(*)
      integer(c_int64_t), value :: n
      real(c_double), value :: coef
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == 0) return
    p = c_last_error(ctx)
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    print *, where, ": ", msg(1:n)
    stop
  end subroutine

  ! grid of SetDGVblzmmesh/Set3DCellsR_DGVblzm/SetNodesDGV; su,sv,sw = nodes per cell
  subroutine dgv_gpu_init(Mu, Mv, Mw, su, sv, sw, u_L, u_R, v_L, v_R, w_L, w_R, device)
    integer, intent(in) :: Mu, Mv, Mw, su, sv, sw
    real(DP), intent(in) :: u_L, u_R, v_L, v_R, w_L, w_R
    integer, intent(in), optional :: device
    type(dgv_grid_desc) :: g
    integer(c_int) :: dev
    dev = 0
    if (present(device)) dev = device
    g = dgv_grid_desc(Mu, Mv, Mw, su, sv, sw, u_L, u_R, v_L, v_R, w_L, w_R)
    call check(c_ctx_create(g, dev, ctx), "dgv_gpu_init")
    call check(c_num_nodes(ctx, nnodes), "dgv_gpu_init")
    if (allocated(frhs1)) deallocate(frhs1)
    if (allocated(hist)) deallocate(hist)
    allocate(frhs1(nnodes), hist(nnodes, 9))
    frhs1 = 0
    hist = 0
  end subroutine

  ! the commvar A-arrays (1-based node ids), as read by ReadAarraysChnksOfstNrecDGV
  subroutine dgv_gpu_set_aarray(A_capphi, A, A_xi, A_xi1)
    integer, intent(in) :: A_capphi(:), A_xi(:), A_xi1(:)
    real(DP), intent(in) :: A(:)
    integer(c_int32_t), allocatable :: cap(:), x(:), x1(:)
    cap = A_capphi
    x = A_xi
    x1 = A_xi1
    call check(c_set_aarray(ctx, cap, size(cap, kind=c_int64_t), A, x, x1, size(A, kind=c_int64_t), 1_c_int), &
               "dgv_gpu_set_aarray")
  end subroutine

  ! SetA_DGV for all canonical basis functions (Trad, min_sens as in parameters.dat)
  subroutine dgv_gpu_generate_aarray(Trad, min_sens)
    real(DP), intent(in) :: Trad, min_sens
    type(dgv_agen_params) :: p
    p = dgv_agen_params(Trad, min_sens, 0, 0)
    call check(c_generate(ctx, p), "dgv_gpu_generate_aarray")
  end subroutine

  subroutine dgv_gpu_finalize()
    if (c_associated(ctx)) call c_ctx_destroy(ctx)
    ctx = c_null_ptr
  end subroutine

  subroutine EvalCollisionPeriodicAPlus_DGV(f, fc)
    real(DP), dimension(:), intent(in) :: f
    real(DP), dimension(:), intent(out) :: fc
    call check(c_eval_q(ctx, f, fc, size(f, kind=c_int64_t)), "EvalCollisionPeriodicAPlus_DGV")
  end subroutine

  subroutine EvalCollisionPeriodicA_DGV(f, fc)
    real(DP), dimension(:), intent(in) :: f
    real(DP), dimension(:), intent(out) :: fc
    call check(c_eval_q_a(ctx, f, fc, size(f, kind=c_int64_t)), "EvalCollisionPeriodicA_DGV")
  end subroutine

  subroutine SpatialOperSH_DGV(f, fcol, dt)
    real(DP), dimension(:), intent(in) :: f
    real(DP), dimension(:), intent(out) :: fcol
    real(DP), intent(in) :: dt
    call regime_in("SpatialOperSH_DGV")
    call check(c_spatial_oper(ctx, f, fcol, size(f, kind=c_int64_t), dt, (mol_diam/L_inf)**2*N_inf), &
               "SpatialOperSH_DGV")
    call regime_out("SpatialOperSH_DGV")
  end subroutine

  ! module run_mode / levels -> the library, when the driver changed them
  subroutine regime_in(where)
    character(*), intent(in) :: where
    if (run_mode /= pushed_mode .or. decomp_lev /= pushed_decomp .or. linear_lev /= pushed_linear) then
      call check(c_set_regime(ctx, int(run_mode, c_int32_t), decomp_lev, linear_lev), where)
      pushed_mode = run_mode; pushed_decomp = decomp_lev; pushed_linear = linear_lev
    end if
  end subroutine

  ! the library's run_mode and Maxwellian -> the module, after a regime check
  subroutine regime_out(where)
    character(*), intent(in) :: where
    integer(c_int32_t) :: mode
    real(c_double) :: l1, l1fm
    integer(c_int64_t) :: counts(2)
    if (.not. (decomp_lev > 0.0_DP)) return
    call check(c_get_regime(ctx, mode, l1, l1fm, counts), where)
    if (mode /= 0 .or. counts(2) > 0) then
      if (.not. allocated(fm)) allocate(fm(nnodes))
      call check(c_get_fm(ctx, fm, nnodes), where)
    end if
    run_mode = mode
    pushed_mode = run_mode
  end subroutine

  subroutine EvalCollisionPeriodicMixedTermsA_DGV(f, fmx, fc)
    real(DP), dimension(:), intent(in) :: f, fmx
    real(DP), dimension(:), intent(out) :: fc
    call check(c_eval_mixed(ctx, f, fmx, fc, size(f, kind=c_int64_t)), "EvalCollisionPeriodicMixedTermsA_DGV")
  end subroutine

  ! the linearised operator at the module's fm
  subroutine PrepareFMA_DGV()
    if (.not. allocated(fm)) then
      print *, "dgv_gpu: PrepareFMA_DGV: fm is not set"
      stop
    end if
    call check(c_prepare_fma(ctx, fm, size(fm, kind=c_int64_t)), "PrepareFMA_DGV")
  end subroutine

  subroutine EvalCollisionLinear(f, fc)
    real(DP), dimension(:), intent(in) :: f
    real(DP), dimension(:), intent(out) :: fc
    call check(c_eval_linear(ctx, f, fc, size(f, kind=c_int64_t)), "EvalCollisionLinear")
  end subroutine

  subroutine CheckSolutionMode_DGV(f, Df)
    real(DP), dimension(:), intent(in) :: f
    real(DP), dimension(:), intent(out) :: Df
    call regime_in("CheckSolutionMode_DGV")
    call check(c_check_mode(ctx, f, Df, size(f, kind=c_int64_t)), "CheckSolutionMode_DGV")
    call regime_out("CheckSolutionMode_DGV")
  end subroutine

  ! f is intent(out) in the reference but read there; intent(inout) here
  subroutine TimeIntegratorORK_SH_DGV(f, dt)
    real(DP), dimension(:), intent(inout) :: f
    real(DP), intent(in) :: dt
    call regime_in("TimeIntegratorORK_SH_DGV")
    call check(c_rk_step(ctx, f, size(f, kind=c_int64_t), dt, (mol_diam/L_inf)**2*N_inf, &
                         int(rkmts, c_int32_t), frhs1), "TimeIntegratorORK_SH_DGV")
    call regime_out("TimeIntegratorORK_SH_DGV")
  end subroutine

  subroutine TimeIntegratorMTS_SH_DGV(f, dt)
    real(DP), dimension(:), intent(inout) :: f
    real(DP), intent(in) :: dt
    hist(:, 5) = frhs1
    call regime_in("TimeIntegratorMTS_SH_DGV")
    call check(c_ab_step(ctx, f, size(f, kind=c_int64_t), dt, (mol_diam/L_inf)**2*N_inf, &
                         int(rkmts, c_int32_t), hist), "TimeIntegratorMTS_SH_DGV")
    call regime_out("TimeIntegratorMTS_SH_DGV")
    frhs1 = hist(:, 5)
    e_time = e_time + dt
  end subroutine

  subroutine PrepareMTS_SH_DGV(f, dt)
    real(DP), dimension(:), intent(inout) :: f
    real(DP), intent(in) :: dt
    integer :: i
    do i = 2, rkmts
      hist(:, 5) = frhs1
      select case (rkmts)
      case (2)
        hist(:, 1) = f
      case (3)
        hist(:, 2) = hist(:, 1); hist(:, 1) = f; hist(:, 6) = hist(:, 5)
      case (4)
        hist(:, 3) = hist(:, 2); hist(:, 2) = hist(:, 1); hist(:, 1) = f
        hist(:, 7) = hist(:, 6); hist(:, 6) = hist(:, 5)
      case (5)
        hist(:, 4) = hist(:, 3); hist(:, 3) = hist(:, 2); hist(:, 2) = hist(:, 1); hist(:, 1) = f
        hist(:, 8) = hist(:, 7); hist(:, 7) = hist(:, 6); hist(:, 6) = hist(:, 5)
      end select
      call TimeIntegratorORK_SH_DGV(f, dt)
      e_time = e_time + dt
    end do
  end subroutine

  subroutine MassCheckRec(f, n, ubar, vbar, wbar, temp)
    real(DP), dimension(:), intent(in) :: f
    real(DP), intent(out) :: n, ubar, vbar, wbar, temp
    real(c_double) :: out(5)
    call check(c_moments(ctx, f, size(f, kind=c_int64_t), out), "MassCheckRec")
    n = out(1); ubar = out(2); vbar = out(3); wbar = out(4); temp = out(5)
  end subroutine

  subroutine MassCheckRecHighCMoments(f, n, ubar, vbar, wbar, temp, temp_u, temp_v, temp_w, mom3_u, mom3_v, &
                                      mom3_w, mom4_u, mom4_v, mom4_w, mom5_u, mom5_v, mom5_w, mom6_u, mom6_v, mom6_w)
    real(DP), dimension(:), intent(in) :: f
    real(DP), intent(out) :: n, ubar, vbar, wbar, temp, temp_u, temp_v, temp_w, mom3_u, mom3_v, mom3_w
    real(DP), intent(out) :: mom4_u, mom4_v, mom4_w, mom5_u, mom5_v, mom5_w, mom6_u, mom6_v, mom6_w
    real(c_double) :: o(20)
    call check(c_moments_high(ctx, f, size(f, kind=c_int64_t), o), "MassCheckRecHighCMoments")
    n = o(1); ubar = o(2); vbar = o(3); wbar = o(4); temp = o(5)
    temp_u = o(6); temp_v = o(7); temp_w = o(8)
    mom3_u = o(9); mom3_v = o(10); mom3_w = o(11); mom4_u = o(12); mom4_v = o(13); mom4_w = o(14)
    mom5_u = o(15); mom5_v = o(16); mom5_w = o(17); mom6_u = o(18); mom6_v = o(19); mom6_w = o(20)
  end subroutine

  ! IphiDGV (dgvtools_mod.f90:847) and IphiDGV_decomp (dgvtools_mod.f90:896): the
  ! moments of the collision operator for every node's basis function, scaled by
  ! (mol_diam/L_inf)**2*N_inf; the second splits f at its local Maxwellian.
  function IphiDGV(f) result(Iphi)
    real(DP), dimension(:), intent(in) :: f
    real(DP), dimension(size(f)) :: Iphi
    call check(c_iphi(ctx, f, Iphi, size(f, kind=c_int64_t), (mol_diam/L_inf)**2*N_inf), "IphiDGV")
  end function

  function IphiDGV_decomp(f) result(Iphi)
    real(DP), dimension(:), intent(in) :: f
    real(DP), dimension(size(f)) :: Iphi
    call check(c_iphi_decomp(ctx, f, Iphi, size(f, kind=c_int64_t), (mol_diam/L_inf)**2*N_inf), &
               "IphiDGV_decomp")
  end function

  ! The multistep history the steppers keep (commvar f1..f4, frhs1..frhs5 in the
  ! reference): copy it out before the reference's WriteSol_DGV and back in
  ! after its RestartSolutionDGV.
  subroutine dgv_gpu_get_history(f1, f2, f3, f4, rhs1, rhs2, rhs3, rhs4, rhs5)
    real(DP), dimension(:), intent(out), optional :: f1, f2, f3, f4, rhs1, rhs2, rhs3, rhs4, rhs5
    if (present(f1)) f1 = hist(:, 1)
    if (present(f2)) f2 = hist(:, 2)
    if (present(f3)) f3 = hist(:, 3)
    if (present(f4)) f4 = hist(:, 4)
    if (present(rhs1)) rhs1 = frhs1
    if (present(rhs2)) rhs2 = hist(:, 6)
    if (present(rhs3)) rhs3 = hist(:, 7)
    if (present(rhs4)) rhs4 = hist(:, 8)
    if (present(rhs5)) rhs5 = hist(:, 9)
  end subroutine

  subroutine dgv_gpu_set_history(f1, f2, f3, f4, rhs1, rhs2, rhs3, rhs4, rhs5)
    real(DP), dimension(:), intent(in), optional :: f1, f2, f3, f4, rhs1, rhs2, rhs3, rhs4, rhs5
    if (present(f1)) hist(:, 1) = f1
    if (present(f2)) hist(:, 2) = f2
    if (present(f3)) hist(:, 3) = f3
    if (present(f4)) hist(:, 4) = f4
    if (present(rhs1)) frhs1 = rhs1
    if (present(rhs2)) hist(:, 6) = rhs2
    if (present(rhs3)) hist(:, 7) = rhs3
    if (present(rhs4)) hist(:, 8) = rhs4
    if (present(rhs5)) hist(:, 9) = rhs5
  end subroutine

end module dgv_gpu
