"""The iso_c_binding shim keeps the reference's procedure names and argument
lists (collision_mod, spat_oper_mod, time_integr_mod, dgvtools_mod) and reaches
the GPU library; a driver in the shape of DGVblzmPL.f90's loop runs on it."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.helpers import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dgv0d3v_direct_collision_amd")
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")
def test_shim_compiles_with_reference_signatures(tmp_path):
    src = open(os.path.join(PKG, "fortran", "dgv_gpu.f90")).read()
    for name in ("EvalCollisionPeriodicAPlus_DGV(f, fc)", "EvalCollisionPeriodicA_DGV(f, fc)",
                 "SpatialOperSH_DGV(f, fcol, dt)", "TimeIntegratorORK_SH_DGV(f, dt)",
                 "TimeIntegratorMTS_SH_DGV(f, dt)", "PrepareMTS_SH_DGV(f, dt)",
                 "MassCheckRec(f, n, ubar, vbar, wbar, temp)", "EvalCollisionLinear(f, fc)",
                 "PrepareFMA_DGV()", "CheckSolutionMode_DGV(f, Df)"):
        assert "subroutine " + name in src
    for name in ("IphiDGV(f) result(Iphi)", "IphiDGV_decomp(f) result(Iphi)"):
        assert "function " + name in src
    subprocess.run([FC, "-O1", f"-J{tmp_path}", "-c", os.path.join(PKG, "fortran", "dgv_gpu.f90"),
                    "-o", str(tmp_path / "s.o")], check=True)


@pytest.mark.gpu
def test_fortran_driver_matches_python_api(tmp_path):
    import dgv0d3v_direct_collision_amd as dgv
    exe = os.path.join(PKG, "dgv_driver")
    assert os.path.exists(exe), "build the driver with __graft_entry__.build()"
    out = tmp_path / "o.bin"
    print("\n[fortran] driver M=3 rk2 x2", flush=True)
    r = subprocess.run([exe, "3", "1.5", "2", "2", "0.01", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and out.exists(), r.stdout + r.stderr
    data = np.fromfile(out, dtype=np.float64)
    n = 729
    f0, q0, f2 = data[:n], data[n:2 * n], data[2 * n:3 * n]
    iphi, iphi_d = data[3 * n:4 * n], data[4 * n:5 * n]
    d = dgv.DGV(3, 3, (-1.5, 1.5))
    d.generate_aarray()
    assert rel_l2(q0, d.EvalCollisionPeriodicAPlus_DGV(f0)) <= 1e-13
    f = f0.copy()
    d.rkmts = 2
    for _ in range(2):
        d.TimeIntegratorORK_SH_DGV(f, 0.01)
    assert rel_l2(f2, f) <= 1e-13
    assert rel_l2(iphi, d.IphiDGV(f2)) <= 1e-13
    assert rel_l2(iphi_d, d.IphiDGV_decomp(f2)) <= 1e-13
    d.close()


@pytest.mark.gpu
def test_fortran_driver_regime_switch_matches_python_api(tmp_path):
    """decomp_lev 1e6, linear_lev 0: the first check enters run_mode 1 (fmA
    built at the local Maxwellian) and stays there; the driver's f, run_mode
    and EvalCollisionLinear equal the Python API's doing the same."""
    import dgv0d3v_direct_collision_amd as dgv
    exe = os.path.join(PKG, "dgv_driver")
    out = tmp_path / "o.bin"
    print("\n[fortran] driver M=3 rk2 x2 then run_mode switch x3", flush=True)
    r = subprocess.run([exe, "3", "1.5", "2", "2", "0.01", str(out), "1e6", "0.0", "3"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and out.exists(), r.stdout + r.stderr
    data = np.fromfile(out, dtype=np.float64)
    n = 729
    f0, f2, f5, mode, lin = data[:n], data[2 * n:3 * n], data[5 * n:6 * n], data[6 * n], data[6 * n + 1:]
    d = dgv.DGV(3, 3, (-1.5, 1.5))
    d.generate_aarray()
    f = f0.copy()
    d.rkmts = 2
    for _ in range(2):
        d.TimeIntegratorORK_SH_DGV(f, 0.01)
    assert rel_l2(f2, f) <= 1e-13
    d.set_regime(0, 1e6, 0.0)
    for _ in range(3):
        d.TimeIntegratorORK_SH_DGV(f, 0.01)
    assert d.run_mode == 1 and mode == 1.0
    assert rel_l2(f5, f) <= 1e-13
    assert rel_l2(lin, d.EvalCollisionLinear(f)) <= 1e-13
    d.close()
