! dgv_driver.f90 -- serial driver in the shape of DGVblzmPL.f90's time loop,
! on the GPU through the dgv_gpu shim: grid, A-array, one Q(f), nsteps RK steps.
! Usage: dgv_driver M box rkmts nsteps dt out.bin [decomp_lev linear_lev nsteps2]
! Writes f0, Q(f0), f after the steps, IphiDGV(f) and IphiDGV_decomp(f) (raw
! float64) to out.bin.  With the three
! optional arguments it then turns on the regime switch (CheckSolutionMode_DGV
! before every spatial operator), runs nsteps2 more steps and appends f,
! run_mode and EvalCollisionLinear(f) at the last Maxwellian.
program dgv_driver
  use dgv_gpu
  implicit none
  integer, parameter :: DP = kind(1.0d0)
  character(len=256) :: arg, outp
  integer :: M, nsteps, nsteps2, i, n, u
  real(DP) :: box, dt, dens, ub, vb, wb, T
  real(DP), allocatable :: f(:), fc(:), x(:)
  call get_command_argument(1, arg); read (arg, *) M
  call get_command_argument(2, arg); read (arg, *) box
  call get_command_argument(3, arg); read (arg, *) rkmts
  call get_command_argument(4, arg); read (arg, *) nsteps
  call get_command_argument(5, arg); read (arg, *) dt
  call get_command_argument(6, outp)
  call dgv_gpu_init(M, M, M, 3, 3, 3, -box, box, -box, box, -box, box)
  call dgv_gpu_generate_aarray(100.0_DP, 0.0_DP)
  n = (3*M)**3
  allocate(f(n), fc(n), x(n))
  do i = 1, n       ! a smooth positive f in reference node order
    x(i) = real(i, DP) / n
    f(i) = exp(-4.0_DP * (x(i) - 0.5_DP)**2) + 0.1_DP * sin(7.0_DP * x(i))**2
  end do
  call EvalCollisionPeriodicAPlus_DGV(f, fc)
  open (newunit=u, file=trim(outp), access="stream", form="unformatted", status="replace")
  write (u) f, fc
  call MassCheckRec(f, dens, ub, vb, wb, T)
  print *, "density", dens, "temperature", T
  do i = 1, nsteps
    call TimeIntegratorORK_SH_DGV(f, dt)
  end do
  write (u) f
  write (u) IphiDGV(f), IphiDGV_decomp(f)
  if (command_argument_count() >= 9) then
    call get_command_argument(7, arg); read (arg, *) decomp_lev
    call get_command_argument(8, arg); read (arg, *) linear_lev
    call get_command_argument(9, arg); read (arg, *) nsteps2
    do i = 1, nsteps2
      call TimeIntegratorORK_SH_DGV(f, dt)
    end do
    print *, "run_mode", run_mode
    write (u) f, real(run_mode, DP)
    call EvalCollisionLinear(f, fc)
    write (u) fc
  end if
  close (u)
  call dgv_gpu_finalize()
end program dgv_driver
