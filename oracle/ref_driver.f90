! Test infrastructure: bind(C) entry points that fill the commvar state read by
! the reference's dgvtools_mod routines and call them.  The routines themselves
! are compiled unmodified from the reference's sources (the "ref" target of the
! Makefile beside this file, output in _ref/); this file only sets their inputs
! and copies their outputs.  Nothing in the library links or loads it.
module dgv_ref_driver
  use iso_c_binding
  use nrtype
  use commvar, only: nodes_u, nodes_v, nodes_w, nodes_gwts, A, A_capphi, A_xi, A_xi1, A_phi, &
                     mol_diam, L_inf, N_inf
  use distributions_mod
  use dgvtools_mod
  implicit none
contains
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
