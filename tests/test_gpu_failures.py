"""Error paths forced from the host: each returns its status through the C ABI
(and the Fortran shim prints the library's message and stops, as the reference
does on its errors), with no kernel launched on invalid input."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ERR_ARG, ERR_INDEX, ERR_NOMEM, ERR_STATE = -1, -3, -4, -6


def _grid():
    import dgv0d3v_direct_collision_amd as dgv
    g = dgv.DGV(3, 3, (-1.5, 1.5))
    g.generate_aarray(nquad=16)
    return g


def test_set_aarray_rejects_node_ids_outside_the_grid():
    """A_xi / A_xi1 must name nodes 1..N (the reference would index past its
    arrays); the refusal leaves the context's previous A-array in place."""
    import dgv0d3v_direct_collision_amd as dgv
    print("\n[fail] set_aarray with an out-of-range node id", flush=True)
    g = _grid()
    try:
        cap, A, xi, xi1 = g.get_aarray()
        f = np.exp(-(g.nodes_u ** 2 + g.nodes_v ** 2 + g.nodes_w ** 2))
        q0 = g.EvalCollisionPeriodicAPlus_DGV(f)
        for bad in (g.nnodes + 1, 0, -5):
            x = xi1.copy()
            x[len(x) // 2] = bad
            with pytest.raises(dgv.DGVError) as e:
                g.set_aarray(cap, A, xi, x)
            assert e.value.code == ERR_INDEX, str(e.value)
        assert np.array_equal(g.EvalCollisionPeriodicAPlus_DGV(f), q0)
        c2 = cap.copy()
        c2[np.flatnonzero(c2)[0]] += 1               # sum(A_capphi) != record count
        with pytest.raises(dgv.DGVError) as e:
            g.set_aarray(c2, A, xi, xi1)
        assert e.value.code == ERR_ARG
    finally:
        g.close()


def test_length_mismatch_is_an_argument_error():
    import dgv0d3v_direct_collision_amd as dgv
    from dgv0d3v_direct_collision_amd._lib import dptr
    print("\n[fail] f of the wrong length through the C ABI", flush=True)
    g = _grid()
    try:
        L = dgv.load()
        f = np.ones(g.nnodes)
        q = np.zeros(g.nnodes)
        for n in (g.nnodes - 1, g.nnodes + 1, 0):
            assert L.dgv_eval_q(g.ctx, dptr(f), dptr(q), n) == ERR_ARG
            assert L.dgv_spatial_oper(g.ctx, dptr(f), dptr(q), n, 0.01, 1.0) == ERR_ARG
        assert "length" in L.dgv_last_error(g.ctx).decode().lower() or L.dgv_last_error(g.ctx)
        assert np.all(q == 0)                   # nothing was written
    finally:
        g.close()


def test_regime_thresholds_must_be_ordered():
    """decomp_lev <= linear_lev stops the reference (spat_oper_mod.f90:331-334)."""
    import dgv0d3v_direct_collision_amd as dgv
    g = _grid()
    try:
        for dl, ll in ((0.1, 0.2), (0.1, 0.1)):
            with pytest.raises(dgv.DGVError) as e:
                g.set_regime(0, dl, ll)
            assert e.value.code == ERR_ARG and "treshold" in str(e.value)
        g.set_regime(0, 0.0, 0.0)               # switched off
        g.set_regime(0, 0.2, 0.1)
        g.set_regime(0, 0.0, 0.0)
    finally:
        g.close()


def test_failed_device_allocation_is_a_status():
    import dgv0d3v_direct_collision_amd as dgv
    print("\n[fail] a batch too large for the device", flush=True)
    g = _grid()
    try:
        L = dgv.load()
        n = 1 << 27                               # 729 nodes x 2^27 instances x 8 B: far beyond HBM
        which = np.zeros(n, np.int32)
        arr = (C.c_void_p * 1)(g.ctx)
        p = C.c_void_p()
        from dgv0d3v_direct_collision_amd._lib import iptr
        rc = L.dgv_batch_create(arr, 1, iptr(which), n, dgv.DGV_FP64, C.byref(p))
        assert rc == ERR_NOMEM and not p.value, rc
        # the device is still usable
        assert np.all(np.isfinite(g.EvalCollisionPeriodicAPlus_DGV(np.ones(g.nnodes))))
    finally:
        g.close()


FSRC = """
program bad
  use dgv_gpu
  implicit none
  integer, parameter :: DP = kind(1.0d0)
  real(DP), allocatable :: f(:), fc(:)
  call dgv_gpu_init(3, 3, 3, 3, 3, 3, -1.5_DP, 1.5_DP, -1.5_DP, 1.5_DP, -1.5_DP, 1.5_DP)
  call dgv_gpu_generate_aarray(100.0_DP, 0.0_DP)
  allocate(f(728), fc(728))
  f = 1.0_DP
  print *, "CALLING"
  call EvalCollisionPeriodicAPlus_DGV(f, fc)
  print *, "UNREACHED"
end program
"""


def test_fortran_shim_prints_and_stops(tmp_path):
    from dgv0d3v_direct_collision_amd import build
    fc = build.amdflang()
    if fc is None:
        pytest.skip("amdflang not found")
    pkg = os.path.join(ROOT, "dgv0d3v_direct_collision_amd")
    mod = str(tmp_path)                              # the shim module, compiled here
    subprocess.run([fc, "-O1", f"-J{mod}", "-c", os.path.join(pkg, "fortran", "dgv_gpu.f90"), "-o",
                    os.path.join(mod, "dgv_gpu.o")], check=True)
    src = tmp_path / "bad.f90"
    src.write_text(FSRC)
    exe = tmp_path / "bad"
    subprocess.run([fc, "-O1", f"-I{mod}", str(src), os.path.join(mod, "dgv_gpu.o"), f"-L{pkg}", "-ldgv_amd",
                    f"-Wl,-rpath,{pkg}", "-o", str(exe)], check=True)
    print("\n[fail] Fortran shim, EvalCollisionPeriodicAPlus_DGV with size(f) = N - 1", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    out = r.stdout + r.stderr
    assert "CALLING" in out and "UNREACHED" not in out, out
    assert "EvalCollisionPeriodicAPlus_DGV" in out, out
