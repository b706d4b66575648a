"""The driver's files and moments, no GPU.

 * MassCheckRecHighCMoments (oracle): on a Maxwellian exp(-|u - ubar|^2 / T)
   sampled on a fine grid the closed forms hold -- T_d = T/3 (the reference's
   2/3 factor on a one-dimensional variance T/2), odd central moments 0,
   mom4 = 3 (T/2)^2, mom6 = 15 (T/2)^3 -- and T = T_u + T_v + T_w for any f;
 * the solution checkpoint of WriteSol_DGV / RestartSolutionDGV and the
   initial-data file of WriteSol_init_DGV: the library's files are byte for
   byte the files a Fortran program writes with the reference's write
   statements (built with amdflang here), read back for every rkmts, and a
   size mismatch is refused;
 * the file names of MakeBaseNameDGV, and the macro / L1 log lines against the
   same Fortran runtime's formatting of the reference's format strings.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from tests.helpers import bimodal

FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
needs_fc = pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")


def test_high_moments_closed_forms_on_a_maxwellian():
    G = oracle.Geometry(7, 6, (-6.0, 6.0))
    T, ub = 1.3, (0.2, -0.1, 0.05)
    f = np.array([oracle.maxwell(T, *ub, 1.0, u, v, w) for u, v, w in zip(G.nodes_u, G.nodes_v, G.nodes_w)])
    m = oracle.masscheck_high(G, f)
    s2 = T / 2
    assert abs(m[0] - 1.0) < 1e-6 and abs(m[4] / T - 1) < 1e-6
    assert np.allclose(m[1:4], ub, atol=1e-7)
    assert np.allclose(m[5:8], T / 3, rtol=1e-6)
    assert np.all(np.abs(m[8:11]) < 1e-6) and np.all(np.abs(m[14:17]) < 1e-6)
    assert np.allclose(m[11:14], 3 * s2 ** 2, rtol=1e-6)
    assert np.allclose(m[17:20], 15 * s2 ** 3, rtol=1e-6)
    # equals MassCheckRec on the common five
    assert np.array_equal(m[:5], oracle.masscheck(G, f))


def test_high_moments_temperature_splits_by_direction():
    G = oracle.Geometry(4, 3, (-3.0, 3.0))
    f = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, 2)
    m = oracle.masscheck_high(G, f)
    assert abs(m[4] - (m[5] + m[6] + m[7])) <= 1e-14 * m[4]


FORTRAN_SOL = """
program w
  implicit none
  integer, parameter :: DP = kind(1.0d0)
  integer :: rkmts, n, i
  real(DP) :: e_time, dt
  real(DP), allocatable :: f(:), f1(:), f2(:), frhs1(:), frhs2(:)
  n = 7; rkmts = 3; e_time = 0.125_DP; dt = 0.0025_DP
  allocate(f(n), f1(n), f2(n), frhs1(n), frhs2(n))
  do i = 1, n
    f(i) = 1.0_DP / i; f1(i) = 2.0_DP / i; f2(i) = 3.0_DP / i
    frhs1(i) = -1.0_DP * i; frhs2(i) = -2.0_DP * i
  end do
  open (15, file="ref_sol.dat", position="REWIND", action="WRITE", form="UNFORMATTED", access="SEQUENTIAL")
  write (15) e_time, dt
  write (15) rkmts
  write (15) size(f, 1)
  write (15) f, f1, f2, frhs1, frhs2
  close (15)
  open (17, file="ref_init.dat", position="REWIND", action="WRITE", form="UNFORMATTED", access="SEQUENTIAL")
  write (17) e_time, dt
  write (17) rkmts
  write (17) size(f, 1)
  write (17) f
  close (17)
  open (16, file="macro.txt", form="FORMATTED")
  write (16, "(F18.14, A2, F18.15, A2, F18.15, A2, F18.15, A2, F18.15, A2, F18.15, A2, F18.15, A2, F18.15,  A2, F18.15, " // &
            "A2, F18.14, A2, F18.15, A2, F18.15, A2, F18.15, A2, F18.15, A2, F18.15, A2," // &
            "F18.14, A2, F18.15, A2, F18.15, A2, F18.15, A2, F18.15, A2, F18.15 )") &
    0.5_DP, ",", 1.0_DP, ",", 0.1_DP, ",", -0.2_DP, ",", 0.3_DP, ",", 0.75_DP, ",", 0.25_DP, ",", 0.25_DP, ",", &
    0.25_DP, ",", -0.001_DP, ",", 0.002_DP, ",", 0.0_DP, ",", 0.1875_DP, ",", 0.19_DP, ",", 0.2_DP, ",", &
    -1.0e-5_DP, ",", 3.0e-4_DP, ",", 0.0_DP, ",", 0.25_DP, ",", 12.5_DP, ",", -12.5_DP
  write (16, "(F18.14, A2, F18.15)") 0.0125_DP, ",", 0.0421_DP
  close (16)
end program w
"""


@needs_fc
def test_solution_file_matches_fortran_and_reads_back(tmp_path):
    import dgv0d3v_direct_collision_amd as dgv
    (tmp_path / "w.f90").write_text(FORTRAN_SOL)
    subprocess.run([FC, "-O0", "w.f90", "-o", "w"], cwd=tmp_path, check=True)
    subprocess.run(["./w"], cwd=tmp_path, check=True)
    n = 7
    i = np.arange(1, n + 1, dtype=np.float64)
    f = 1.0 / i
    hist = np.zeros((9, n))
    hist[0], hist[1] = 2.0 / i, 3.0 / i
    hist[4], hist[5] = -1.0 * i, -2.0 * i
    ours = tmp_path / "ours.dat"
    dgv.write_sol_file(ours, 0.125, 0.0025, 3, f, hist)
    assert ours.read_bytes() == (tmp_path / "ref_sol.dat").read_bytes()
    e, dt, k, f2, h2 = dgv.read_sol_file(tmp_path / "ref_sol.dat", n)
    assert (e, dt, k) == (0.125, 0.0025, 3)
    assert np.array_equal(f2, f) and np.array_equal(h2[[0, 1, 4, 5]], hist[[0, 1, 4, 5]])
    with pytest.raises(dgv.DGVError):
        dgv.read_sol_file(tmp_path / "ref_sol.dat", n + 1)
    # WriteSol_init_DGV: f alone whatever rkmts
    dgv.write_sol_init_file(tmp_path / "ours_init.dat", 0.125, 0.0025, 3, f)
    assert (tmp_path / "ours_init.dat").read_bytes() == (tmp_path / "ref_init.dat").read_bytes()
    with pytest.raises(dgv.DGVError):
        dgv.write_sol_init_file(tmp_path / "bad.dat", 0.0, 0.1, 6, f)
    # the log lines against the Fortran runtime's formatting
    lines = (tmp_path / "macro.txt").read_text().splitlines()
    vals = [1.0, 0.1, -0.2, 0.3, 0.75, 0.25, 0.25, 0.25, -0.001, 0.002, 0.0, 0.1875, 0.19, 0.2, -1.0e-5, 3.0e-4,
            0.0, 0.25, 12.5, -12.5]
    from dgv0d3v_direct_collision_amd.driver_io import fortran_f, macro_line
    assert macro_line(0.5, vals) == lines[0]
    assert fortran_f(0.0125, 18, 14) + " ," + fortran_f(0.0421, 18, 15) == lines[1]


@pytest.mark.parametrize("rkmts", [1, 2, 3, 4, 5])
def test_solution_file_round_trip(tmp_path, rkmts):
    import dgv0d3v_direct_collision_amd as dgv
    rng = np.random.default_rng(rkmts)
    n = 31
    f = rng.standard_normal(n)
    hist = rng.standard_normal((9, n))
    p = tmp_path / "s.dat"
    dgv.write_sol_file(p, 1.5, 0.01, rkmts, f, hist)
    assert os.path.getsize(p) == (8 + 16) + (8 + 4) * 2 + 8 + 8 * n * (2 * rkmts - 1)
    e, dt, k, f2, h2 = dgv.read_sol_file(p, n)
    assert (e, dt, k) == (1.5, 0.01, rkmts) and np.array_equal(f2, f)
    rows = list(range(rkmts - 1)) + list(range(4, 4 + rkmts - 1))
    assert np.array_equal(h2[rows], hist[rows])


def test_file_names():
    import dgv0d3v_direct_collision_amd as dgv
    assert dgv.sol_file_name(9, 3, (-4.0, 4.0), " 0.0100000000", "sol/", "run", k_c=2, N_x=3) == \
        "sol/run2kc3su3sv3sw3NXU9MuUU9MvVU9MwWU_time0.0100000000_sol.dat"
    assert dgv.log_file_name(5, 2, (-1.0, 1.0), "_L1err.txt", "", "b", k_c=1, N_x=1) == \
        "b1kc2su2sv2sw1NXU5MuUU5MvVU5MwWU_L1err.txt"


# ---- RecordMomDerivRelaxTimes_DGV / WriteMomDerivRelaxTimes_DGV -----------------------------

def _grid_arrays(G):
    return G.nodes_u, G.nodes_v, G.nodes_w, G.nodes_gwts


def test_moment_derivatives_of_a_scaled_rhs():
    # frhs1 = c f1 scales every moment: dn = c n, the bulk velocity and the
    # temperature do not move, and each centred moment's derivative is c times
    # the moment itself (MassCheckRecHighCMoments of the oracle)
    from dgv0d3v_direct_collision_amd.driver_io import moment_derivatives
    G = oracle.Geometry(4, 3, (-3.0, 3.0))
    f1 = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, 4)
    c = -0.37
    d = moment_derivatives(f1, c * f1, *_grid_arrays(G))
    m = oracle.masscheck_high(G, f1)
    assert abs(d[0] - c * m[0]) <= 1e-14 * abs(c * m[0])
    assert np.all(np.abs(d[1:5]) <= 1e-14)
    assert np.allclose(d[5:20], c * np.asarray(m[5:20]), rtol=1e-12, atol=1e-15)


def test_moment_derivatives_against_finite_differences():
    from dgv0d3v_direct_collision_amd.driver_io import moment_derivatives
    G = oracle.Geometry(4, 3, (-3.0, 3.0))
    u, v, w, gw = _grid_arrays(G)
    f1 = bimodal(u, v, w, 5)
    r = bimodal(u, v, w, 6) - 0.8 * f1
    d = moment_derivatives(f1, r, u, v, w, gw)
    h = 1e-5
    n = lambda f: np.sum(f * gw)
    ub = lambda f: np.array([np.sum(f * gw * x) for x in (u, v, w)]) / n(f)
    assert abs(d[0] - (n(f1 + h * r) - n(f1 - h * r)) / (2 * h)) <= 1e-10 * abs(d[0])
    assert np.allclose(d[1:4], (ub(f1 + h * r) - ub(f1 - h * r)) / (2 * h), rtol=1e-7, atol=1e-12)
    # symmetric f1 and an even rhs: the bulk velocity stays 0 and dT is the
    # derivative of the temperature
    def temp(f):
        b = ub(f)
        return np.sum(f * gw * ((u - b[0]) ** 2 + (v - b[1]) ** 2 + (w - b[2]) ** 2)) / n(f) * 2.0 / 3.0
    fs = np.exp(-(u ** 2 + v ** 2 + w ** 2) / 0.7) + 0.3 * np.exp(-(u ** 2 + v ** 2 + w ** 2) / 2.1)
    rs = np.exp(-(u ** 2 + v ** 2 + w ** 2) / 1.4) - 0.5 * fs
    ds = moment_derivatives(fs, rs, u, v, w, gw)
    assert np.all(np.abs(ds[1:4]) <= 1e-13)
    assert abs(ds[4] - (temp(fs + h * rs) - temp(fs - h * rs)) / (2 * h)) <= 1e-7 * abs(ds[4])


def test_relaxation_times_of_a_bgk_rhs():
    # frhs1 = -(f1 - f_M)/tau with f_M the local Maxwellian of f1: every
    # relaxation time of the directional temperatures and the centred moments
    # is 1/tau (the moments of f_M are the Gaussian ones to quadrature
    # accuracy), and the density's is dn/mn = -1/tau
    from dgv0d3v_direct_collision_amd.driver_io import moment_derivatives
    G = oracle.Geometry(7, 6, (-6.0, 6.0))
    u, v, w, gw = _grid_arrays(G)
    f1 = bimodal(u, v, w, 7)
    n, ub, vb, wb, T = oracle.masscheck(G, f1)
    fm = n * np.exp(-((u - ub) ** 2 + (v - vb) ** 2 + (w - wb) ** 2) / T) / (np.pi * T) ** 1.5
    tau = 0.4
    d = moment_derivatives(f1, -(f1 - fm) / tau, u, v, w, gw)
    t = d[20:]
    # mn is itself a quadrature error (the Maxwellian's density defect), so
    # the density's time carries the rounding of the temperature sums
    assert abs(t[0] + 1 / tau) <= 1e-4 / tau
    assert np.allclose(t[5:8], 1 / tau, rtol=1e-4)                         # T_u, T_v, T_w
    # mom3 .. mom6: the discrete Maxwellian's centred moments are Gaussian to
    # about 1e-4 of these; a wrong constant or sign would be off by O(1)
    assert np.allclose(t[8:20], 1 / tau, rtol=1e-3)


def test_moment_derivative_store():
    from dgv0d3v_direct_collision_amd.driver_io import MomDerivLog
    G = oracle.Geometry(3, 3, (-1.5, 1.5))
    f1 = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, 1)
    log = MomDerivLog("unused", num_eval_error=1, rkmts=1, t_R=1.0)
    assert log.capacity == 6
    assert all(log.record(0.1 * k, f1, -f1, G) for k in range(6))
    assert not log.record(0.7, f1, -f1, G)                 # full: dropped
    assert len(log.rows) == 6 and log.rows[-1][0] == 0.5
    log.record(1.5, f1, -f1, G)                            # past t_R: released
    with pytest.raises(RuntimeError):
        log.write()


@needs_fc
def test_moment_derivative_files_match_fortran_formatting(tmp_path):
    from dgv0d3v_direct_collision_amd.driver_io import MomDerivLog
    G = oracle.Geometry(3, 3, (-1.5, 1.5))
    f1 = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, 2)
    r = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, 3) - f1
    log = MomDerivLog(str(tmp_path / "run"), num_eval_error=4, rkmts=2, t_R=10.0)
    log.record(0.25, f1, r, G)
    log.record(0.5, f1, 2 * r, G)
    log.write()
    rows = np.array(log.rows)
    vals = list(rows[0, :21]) + [0.0, -1.0e-120, 6.02e23, -9.9999996, 1e100] + [1.0] * 16
    fmt = "(" + "".join(["E18.6, A2, "] * 20) + "E18.6)"
    fmts = " &\n      ".join(f'"{fmt[i:i + 60]}"//' for i in range(0, len(fmt), 60))[:-2]
    items = lambda vs: ", &\n    ".join(f'{float(x)!r}_DP, ","' for x in vs[:-1]) + f", &\n    {float(vs[-1])!r}_DP"
    prog = ("program e\n  implicit none\n  integer, parameter :: DP = kind(1.0d0)\n"
            "  character(len=*), parameter :: fm = &\n      " + fmts + "\n"
            "  open (16, file=\"e.txt\", form=\"FORMATTED\")\n"
            f"  write (16, fm) &\n    {items(vals[:21])}\n"
            f"  write (16, fm) &\n    {items(vals[21:])}\n"
            "  close (16)\nend program e\n")
    (tmp_path / "e.f90").write_text(prog)
    subprocess.run([FC, "-O0", "e.f90", "-o", "e"], cwd=tmp_path, check=True)
    subprocess.run(["./e"], cwd=tmp_path, check=True)
    ref = (tmp_path / "e.txt").read_text().splitlines()
    from dgv0d3v_direct_collision_amd.driver_io import fortran_e
    assert (tmp_path / "run_dermom.txt").read_text().splitlines()[-2] == ref[0]   # the first of 2 rows
    assert " ,".join(fortran_e(x, 18, 6) for x in vals[21:]) == ref[1]
    rel = (tmp_path / "run_relaxtimes.txt").read_text().splitlines()
    assert len(rel) == 7 and rel[-2].split(",")[0].strip() == "0.250000E+00"   # 2 + 3 header records, 2 rows
