! Test infrastructure: bind(C) entry points that fill the commvar state read by
! the reference's dgvtools_mod routines and call them.  The routines themselves
! are compiled unmodified from the reference's sources (the "ref" target of the
! Makefile beside this file, output in _ref/); this file only sets their inputs
! and copies their outputs.  Nothing in the library links or loads it.
module dgv_ref_driver
  use iso_c_binding
  use nrtype
  use commvar, only: nodes_u, nodes_v, nodes_w, nodes_gwts, A, A_capphi, A_xi, A_xi1, A_phi, &
                     mol_diam, L_inf, N_inf, k_c, su, sv, sw, Mu, Mv, Mw, N, current_solution_dir, &
                     current_solution_base_name, current_Aoperator_base_name, mesh_x_uniform, &
                     mesh_u_uniform, mesh_v_uniform, mesh_w_uniform, f, f1, f2, f3, f4, frhs1, frhs2, &
                     frhs3, frhs4, frhs5, e_time, dt, rkmts, L1_a, L1_t, t_L, t_R, num_save_solution, &
                     MomDerivRelaxT, MomDerivRelaxT_flag, num_eval_error
  use distributions_mod
  use dgvtools_mod
  use readwrite
  implicit none
contains
  ! a NUL-terminated C string as a Fortran string
  function cstr(c) result(s)
    character(kind=c_char), intent(in) :: c(*)
    character(len=256) :: s
    integer :: i
    s = ' '
    i = 1
    do while (c(i) /= c_null_char .and. i <= 256)
      s(i:i) = c(i)
      i = i + 1
    end do
  end function

  subroutine to_c(s, out, cap)
    character(len=*), intent(in) :: s
    integer(c_int32_t), intent(in) :: cap
    character(kind=c_char), intent(out) :: out(cap)
    integer :: i, l
    l = min(len_trim(s), int(cap) - 1)
    do i = 1, l
      out(i) = s(i:i)
    end do
    out(l + 1) = c_null_char
  end subroutine

  ! the commvar parameters of the reference's file names (uniform meshes)
  subroutine ref_set_names(dir, sbase, abase, kc, s_u, s_v, s_w, nx, m_u, m_v, m_w) bind(C, name="ref_set_names")
    character(kind=c_char), intent(in) :: dir(*), sbase(*), abase(*)
    integer(c_int32_t), value :: kc, s_u, s_v, s_w, nx, m_u, m_v, m_w
    current_solution_dir = cstr(dir)
    current_solution_base_name = cstr(sbase)
    current_Aoperator_base_name = cstr(abase)
    k_c = kc; su = s_u; sv = s_v; sw = s_w; N = nx; Mu = m_u; Mv = m_v; Mw = m_w
    mesh_x_uniform = .true.; mesh_u_uniform = .true.; mesh_v_uniform = .true.; mesh_w_uniform = .true.
  end subroutine

  ! kind 0: MakeBaseNameDGV, 1: MakeBaseNameAoperDGV, 2: MakeBaseNameAoperChnksDGV(i),
  ! 3: MakeBaseNameChnksDGV(i)
  subroutine ref_base_name(kind, i, out, cap) bind(C, name="ref_base_name")
    integer(c_int32_t), value :: kind, i, cap
    character(kind=c_char), intent(out) :: out(cap)
    character(len=132) :: fn
    integer(I4B) :: ii
    ii = i
    select case (kind)
    case (0)
      call MakeBaseNameDGV(fn)
    case (1)
      call MakeBaseNameAoperDGV(fn)
    case (2)
      call MakeBaseNameAoperChnksDGV(fn, ii)
    case default
      call MakeBaseNameChnksDGV(fn, ii)
    end select
    call to_c(trim(adjustl(fn)), out, cap)
  end subroutine

  subroutine free_aarray()
    if (allocated(A_capphi)) deallocate(A_capphi)
    if (allocated(A)) deallocate(A)
    if (allocated(A_xi)) deallocate(A_xi)
    if (allocated(A_xi1)) deallocate(A_xi1)
    if (allocated(A_phi)) deallocate(A_phi)
  end subroutine

  ! WriteAarraysDGV of the A-arrays set by ref_set_aarray / ref_set_phi
  subroutine ref_write_aarray() bind(C, name="ref_write_aarray")
    call WriteAarraysDGV
  end subroutine

  subroutine ref_set_phi(nrec, phi) bind(C, name="ref_set_phi")
    integer(c_int64_t), value :: nrec
    integer(c_int32_t), intent(in) :: phi(nrec)
    A_phi(1:nrec) = phi
  end subroutine

  subroutine ref_get_phi(nrec, phi) bind(C, name="ref_get_phi")
    integer(c_int64_t), value :: nrec
    integer(c_int32_t), intent(out) :: phi(nrec)
    phi = A_phi(1:nrec)
  end subroutine

  ! ReadAarraysDGV into freshly emptied A-arrays; mm = size(A_capphi)
  subroutine ref_read_aarray(mm) bind(C, name="ref_read_aarray")
    integer(c_int64_t), intent(out) :: mm
    call free_aarray()
    call ReadAarraysDGV
    mm = size(A_capphi, 1)
  end subroutine

  subroutine ref_scan_chunks(nch, act) bind(C, name="ref_scan_chunks")
    integer(c_int32_t), value :: nch
    integer(c_int32_t), intent(out) :: act(nch)
    call ScanAarrsChnks4Acapphi(act)
  end subroutine

  subroutine ref_read_chunks(fst, ofst, numrec, nch, mm) bind(C, name="ref_read_chunks")
    integer(c_int32_t), value :: fst, ofst, numrec, nch
    integer(c_int64_t), intent(out) :: mm
    integer(I4B) :: a1, a2, a3, a4
    a1 = fst; a2 = ofst; a3 = numrec; a4 = nch
    call free_aarray()
    call ReadAarraysChnksOfstNrecDGV(a1, a2, a3, a4)
    mm = size(A_capphi, 1)
  end subroutine

  ! WriteSol_DGV(suff) (init = 0) or WriteSol_init_DGV(suff) (init = 1) of f and
  ! the history hist(n, 9) = f1..f4, frhs1..frhs5
  subroutine ref_write_sol(suff, init, e, d, k, n, fv, hist) bind(C, name="ref_write_sol")
    character(kind=c_char), intent(in) :: suff(*)
    integer(c_int32_t), value :: init, k
    real(c_double), value :: e, d
    integer(c_int64_t), value :: n
    real(c_double), intent(in) :: fv(n), hist(n, 9)
    character(len=15) :: sf
    sf = cstr(suff)
    call free_sol()
    allocate(f(n), f1(n), f2(n), f3(n), f4(n), frhs1(n), frhs2(n), frhs3(n), frhs4(n), frhs5(n))
    f = fv; f1 = hist(:, 1); f2 = hist(:, 2); f3 = hist(:, 3); f4 = hist(:, 4)
    frhs1 = hist(:, 5); frhs2 = hist(:, 6); frhs3 = hist(:, 7); frhs4 = hist(:, 8); frhs5 = hist(:, 9)
    e_time = e; dt = d; rkmts = k
    if (init /= 0) then
      call WriteSol_init_DGV(sf)
    else
      call WriteSol_DGV(sf)
    end if
  end subroutine

  subroutine free_sol()
    if (allocated(f)) deallocate(f)
    if (allocated(f1)) deallocate(f1)
    if (allocated(f2)) deallocate(f2)
    if (allocated(f3)) deallocate(f3)
    if (allocated(f4)) deallocate(f4)
    if (allocated(frhs1)) deallocate(frhs1)
    if (allocated(frhs2)) deallocate(frhs2)
    if (allocated(frhs3)) deallocate(frhs3)
    if (allocated(frhs4)) deallocate(frhs4)
    if (allocated(frhs5)) deallocate(frhs5)
  end subroutine

  ! RestartSolutionDGV(txt) with rkmts = k set beforehand (it allocates the
  ! history for that order); returns e_time, dt, the file's rkmts, f and the
  ! history rows it allocated (the others zero)
  subroutine ref_restart(txt, k, n, e, d, kout, fv, hist) bind(C, name="ref_restart")
    character(kind=c_char), intent(in) :: txt(*)
    integer(c_int32_t), value :: k
    integer(c_int64_t), value :: n
    real(c_double), intent(out) :: e, d
    integer(c_int32_t), intent(out) :: kout
    real(c_double), intent(out) :: fv(n), hist(n, 9)
    character(len=20) :: t
    t = cstr(txt)
    call free_sol()
    allocate(f(n))
    rkmts = k
    call RestartSolutionDGV(t)
    e = e_time; d = dt; kout = rkmts; fv = f
    hist = 0.0_DP
    if (allocated(f1)) hist(:, 1) = f1
    if (allocated(f2)) hist(:, 2) = f2
    if (allocated(f3)) hist(:, 3) = f3
    if (allocated(f4)) hist(:, 4) = f4
    if (allocated(frhs1)) hist(:, 5) = frhs1
    if (allocated(frhs2)) hist(:, 6) = frhs2
    if (allocated(frhs3)) hist(:, 7) = frhs3
    if (allocated(frhs4)) hist(:, 8) = frhs4
    if (allocated(frhs5)) hist(:, 9) = frhs5
  end subroutine

  ! WriteErrHighCMoments_DGV of nrow rows of 21 columns (time, then the 20 moments)
  subroutine ref_write_macroerr(nrow, cols) bind(C, name="ref_write_macroerr")
    integer(c_int64_t), value :: nrow
    real(c_double), intent(in) :: cols(nrow, 21)
    call WriteErrHighCMoments_DGV(cols(:, 1), cols(:, 2), cols(:, 3), cols(:, 4), cols(:, 5), cols(:, 6), &
                                  cols(:, 7), cols(:, 8), cols(:, 9), cols(:, 10), cols(:, 11), cols(:, 12), &
                                  cols(:, 13), cols(:, 14), cols(:, 15), cols(:, 16), cols(:, 17), cols(:, 18), &
                                  cols(:, 19), cols(:, 20), cols(:, 21))
  end subroutine

  ! KeepTrackL1_err_DGV(err) at time e (t_L, t_R, num_save_solution as the
  ! driver's parameters); ref_l1_reset forgets the track
  subroutine ref_l1_track(e, tl, tr, nsave, err) bind(C, name="ref_l1_track")
    real(c_double), value :: e, tl, tr, err
    integer(c_int32_t), value :: nsave
    e_time = e; t_L = tl; t_R = tr; num_save_solution = nsave
    call KeepTrackL1_err_DGV(err)
  end subroutine

  ! RecordMomDerivRelaxTimes_DGV(t) from f1 and frhs1 (nodes set by ref_set_nodes)
  subroutine ref_momderiv_record(t, n, f1v, r1v, nee, tr, k) bind(C, name="ref_momderiv_record")
    real(c_double), value :: t, tr
    integer(c_int64_t), value :: n
    real(c_double), intent(in) :: f1v(n), r1v(n)
    integer(c_int32_t), value :: nee, k
    if (allocated(f1)) deallocate(f1)
    if (allocated(frhs1)) deallocate(frhs1)
    allocate(f1(n), frhs1(n))
    f1 = f1v; frhs1 = r1v
    num_eval_error = nee; t_R = tr; rkmts = k
    call RecordMomDerivRelaxTimes_DGV(t)
  end subroutine

  subroutine ref_momderiv_reset() bind(C, name="ref_momderiv_reset")
    if (allocated(MomDerivRelaxT)) deallocate(MomDerivRelaxT)
    MomDerivRelaxT_flag = 0
  end subroutine

  ! the store MomDerivRelaxT (count, columns, then the records); its length
  function ref_momderiv_get(cap, out) bind(C, name="ref_momderiv_get") result(m)
    integer(c_int64_t), value :: cap
    real(c_double), intent(out) :: out(cap)
    integer(c_int64_t) :: m
    m = size(MomDerivRelaxT, 1)
    out(1:min(cap, m)) = MomDerivRelaxT(1:min(cap, m))
  end function

  subroutine ref_momderiv_write() bind(C, name="ref_momderiv_write")
    call WriteMomDerivRelaxTimes_DGV
  end subroutine

  subroutine ref_l1_reset() bind(C, name="ref_l1_reset")
    if (allocated(L1_a)) deallocate(L1_a)
    if (allocated(L1_t)) deallocate(L1_t)
  end subroutine

  ! node coordinates and quadrature weights (commvar nodes_u/v/w, nodes_gwts)
  subroutine ref_set_nodes(n, u, v, w, gw) bind(C, name="ref_set_nodes")
    integer(c_int64_t), value :: n
    real(c_double), intent(in) :: u(n), v(n), w(n), gw(n)
    if (allocated(nodes_u)) deallocate(nodes_u)
    if (allocated(nodes_v)) deallocate(nodes_v)
    if (allocated(nodes_w)) deallocate(nodes_w)
    if (allocated(nodes_gwts)) deallocate(nodes_gwts)
    allocate(nodes_u(n), nodes_v(n), nodes_w(n), nodes_gwts(n))
    nodes_u = u; nodes_v = v; nodes_w = w; nodes_gwts = gw
  end subroutine

  ! (mol_diam, L_inf, N_inf): IphiDGV scales by (mol_diam/L_inf)**2*N_inf
  subroutine ref_set_coef(d, l, nn) bind(C, name="ref_set_coef")
    real(c_double), value :: d, l, nn
    mol_diam = d; L_inf = l; N_inf = nn
  end subroutine

  ! A-arrays in the reference layout (1-based node ids; records grouped by node);
  ! A_phi holds each record's owner node
  subroutine ref_set_aarray(n, capphi, nrec, av, xi, xi1) bind(C, name="ref_set_aarray")
    integer(c_int64_t), value :: n, nrec
    integer(c_int32_t), intent(in) :: capphi(n), xi(nrec), xi1(nrec)
    real(c_double), intent(in) :: av(nrec)
    integer(c_int64_t) :: i, j, k
    if (allocated(A_capphi)) deallocate(A_capphi)
    if (allocated(A)) deallocate(A)
    if (allocated(A_xi)) deallocate(A_xi)
    if (allocated(A_xi1)) deallocate(A_xi1)
    if (allocated(A_phi)) deallocate(A_phi)
    allocate(A_capphi(n), A(nrec), A_xi(nrec), A_xi1(nrec), A_phi(nrec))
    A_capphi = capphi; A = av; A_xi = xi; A_xi1 = xi1
    k = 0
    do i = 1, n
      do j = 1, capphi(i)
        k = k + 1
        A_phi(k) = int(i, c_int32_t)
      end do
    end do
  end subroutine

  function ref_nrec() bind(C, name="ref_nrec") result(r)
    integer(c_int64_t) :: r
    r = size(A, 1)
  end function

  subroutine ref_get_aarray(n, capphi, nrec, av, xi, xi1) bind(C, name="ref_get_aarray")
    integer(c_int64_t), value :: n, nrec
    integer(c_int32_t), intent(out) :: capphi(n), xi(nrec), xi1(nrec)
    real(c_double), intent(out) :: av(nrec)
    capphi = A_capphi(1:n); av = A(1:nrec); xi = A_xi(1:nrec); xi1 = A_xi1(1:nrec)
  end subroutine

  subroutine ref_iphi(n, f, out) bind(C, name="ref_iphi")
    integer(c_int64_t), value :: n
    real(c_double), intent(in) :: f(n)
    real(c_double), intent(out) :: out(n)
    out = IphiDGV(f)
  end subroutine

  subroutine ref_iphi_decomp(n, f, out) bind(C, name="ref_iphi_decomp")
    integer(c_int64_t), value :: n
    real(c_double), intent(in) :: f(n)
    real(c_double), intent(out) :: out(n)
    out = IphiDGV_decomp(f)
  end subroutine

  ! MassCheckRec: n, ubar, vbar, wbar, temp
  subroutine ref_masscheck(n, f, out) bind(C, name="ref_masscheck")
    integer(c_int64_t), value :: n
    real(c_double), intent(in) :: f(n)
    real(c_double), intent(out) :: out(5)
    call MassCheckRec(f, out(1), out(2), out(3), out(4), out(5))
  end subroutine

  ! MassCheckRecHighCMoments: the 20 outputs in argument order
  subroutine ref_masscheck_high(n, f, out) bind(C, name="ref_masscheck_high")
    integer(c_int64_t), value :: n
    real(c_double), intent(in) :: f(n)
    real(c_double), intent(out) :: out(20)
    call MassCheckRecHighCMoments(f, out(1), out(2), out(3), out(4), out(5), out(6), out(7), out(8), &
                                  out(9), out(10), out(11), out(12), out(13), out(14), out(15), out(16), &
                                  out(17), out(18), out(19), out(20))
  end subroutine

  function ref_maxwell(T, u0, v0, w0, nn, u, v, w) bind(C, name="ref_maxwell") result(y)
    real(c_double), value :: T, u0, v0, w0, nn, u, v, w
    real(c_double) :: y
    y = maxwelveldist(T, u0, v0, w0, nn, u, v, w)
  end function

  subroutine ref_trunc_threshold(t) bind(C, name="ref_trunc_threshold")
    real(c_double), value :: t
    call TruncAarrsTrhls_DGV(t)
  end subroutine

  subroutine ref_trunc_radius(r) bind(C, name="ref_trunc_radius")
    real(c_double), value :: r
    call TruncAarrsRadius_DGV(r)
  end subroutine
end module dgv_ref_driver
