"""Test infrastructure: the reference's own dgvtools_mod routines, compiled
unmodified from the reference's sources into _ref/libdgv_ref.so (the "ref" target
of the Makefile beside this file, run by __graft_entry__.build() where the
reference sources exist), driven through ref_driver.f90's bind(C) entry points.

Used only by tests/ (and tests/golden/'s fixture writer) to check the oracle's
restatements against the reference itself:
  IphiDGV          dgvtools_mod.f90:847-877
  IphiDGV_decomp   dgvtools_mod.f90:896-927
  MassCheckRec     dgvtools_mod.f90:1437-1458
  MassCheckRecHighCMoments dgvtools_mod.f90:1501-1549
  maxwelveldist    distributions_mod.f90:83-94
  TruncAarrsTrhls_DGV / TruncAarrsRadius_DGV dgvtools_mod.f90:1313-1433
and its readwrite routines:
  MakeBaseNameDGV / MakeBaseNameAoperDGV / MakeBaseNameChnksDGV /
  MakeBaseNameAoperChnksDGV          readwrite.f90:1505-1650, 2148-2290
  WriteAarraysDGV / ReadAarraysDGV   readwrite.f90:1459-1484, 1066-1118
  ScanAarrsChnks4Acapphi / ReadAarraysChnksOfstNrecDGV readwrite.f90:2292-2365, 1261-1449
  WriteSol_DGV / WriteSol_init_DGV / RestartSolutionDGV readwrite.f90:1756-1850, 2389-2513
  WriteErrHighCMoments_DGV / KeepTrackL1_err_DGV      readwrite.f90:2025-2147
  RecordMomDerivRelaxTimes_DGV (dgvtools_mod.f90:1589-2023) and
  WriteMomDerivRelaxTimes_DGV (readwrite.f90:2528-2609)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libdgv_ref.so")
_lib = None
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        L.ref_set_nodes.argtypes = [C.c_int64, _f64p, _f64p, _f64p, _f64p]
        L.ref_set_coef.argtypes = [C.c_double] * 3
        L.ref_set_aarray.argtypes = [C.c_int64, _i32p, C.c_int64, _f64p, _i32p, _i32p]
        L.ref_nrec.restype = C.c_int64
        L.ref_get_aarray.argtypes = [C.c_int64, _i32p, C.c_int64, _f64p, _i32p, _i32p]
        for name in ("ref_iphi", "ref_iphi_decomp", "ref_masscheck", "ref_masscheck_high"):
            getattr(L, name).argtypes = [C.c_int64, _f64p, _f64p]
        L.ref_maxwell.argtypes = [C.c_double] * 8
        L.ref_maxwell.restype = C.c_double
        L.ref_trunc_threshold.argtypes = [C.c_double]
        L.ref_trunc_radius.argtypes = [C.c_double]
        L.ref_set_names.argtypes = [C.c_char_p] * 3 + [C.c_int32] * 8
        L.ref_base_name.argtypes = [C.c_int32, C.c_int32, C.c_char_p, C.c_int32]
        L.ref_set_phi.argtypes = [C.c_int64, _i32p]
        L.ref_get_phi.argtypes = [C.c_int64, _i32p]
        L.ref_read_aarray.argtypes = [C.POINTER(C.c_int64)]
        L.ref_scan_chunks.argtypes = [C.c_int32, _i32p]
        L.ref_read_chunks.argtypes = [C.c_int32] * 4 + [C.POINTER(C.c_int64)]
        L.ref_write_sol.argtypes = [C.c_char_p, C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int64, _f64p,
                                    _f64p]
        L.ref_restart.argtypes = [C.c_char_p, C.c_int32, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_int32), _f64p, _f64p]
        L.ref_write_macroerr.argtypes = [C.c_int64, _f64p]
        L.ref_l1_track.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int32, C.c_double]
        L.ref_momderiv_record.argtypes = [C.c_double, C.c_int64, _f64p, _f64p, C.c_int32, C.c_double, C.c_int32]
        L.ref_momderiv_get.argtypes = [C.c_int64, _f64p]
        L.ref_momderiv_get.restype = C.c_int64
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def set_nodes(u, v, w, gwts):
    lib().ref_set_nodes(len(u), _c(u, np.float64), _c(v, np.float64), _c(w, np.float64), _c(gwts, np.float64))


def set_aarray(capphi, A, A_xi, A_xi1):
    lib().ref_set_aarray(len(capphi), _c(capphi, np.int32), len(A), _c(A, np.float64), _c(A_xi, np.int32),
                         _c(A_xi1, np.int32))


def get_aarray(n):
    nrec = lib().ref_nrec()
    cap, a = np.zeros(n, np.int32), np.zeros(nrec)
    x, y = np.zeros(nrec, np.int32), np.zeros(nrec, np.int32)
    lib().ref_get_aarray(n, cap, nrec, a, x, y)
    return cap, a, x, y


def iphi(u, v, w, gwts, capphi, A, A_xi, A_xi1, f, mol_diam=1.0, L_inf=1.0, N_inf=1.0, decomp=False):
    """IphiDGV(f) (or IphiDGV_decomp(f)) on a full A-array in the reference layout."""
    set_nodes(u, v, w, gwts)
    lib().ref_set_coef(mol_diam, L_inf, N_inf)
    set_aarray(capphi, A, A_xi, A_xi1)
    out = np.zeros(len(f))
    (lib().ref_iphi_decomp if decomp else lib().ref_iphi)(len(f), _c(f, np.float64), out)
    return out


def masscheck(u, v, w, gwts, f):
    set_nodes(u, v, w, gwts)
    out = np.zeros(5)
    lib().ref_masscheck(len(f), _c(f, np.float64), out)
    return out


def masscheck_high(u, v, w, gwts, f):
    set_nodes(u, v, w, gwts)
    out = np.zeros(20)
    lib().ref_masscheck_high(len(f), _c(f, np.float64), out)
    return out


def maxwell(T, u0, v0, w0, n, u, v, w):
    return lib().ref_maxwell(T, u0, v0, w0, n, u, v, w)


def trunc_threshold(capphi, A, A_xi, A_xi1, trhld):
    set_aarray(capphi, A, A_xi, A_xi1)
    lib().ref_trunc_threshold(trhld)
    return get_aarray(len(capphi))


def trunc_radius(u, v, w, gwts, capphi, A, A_xi, A_xi1, rad):
    set_nodes(u, v, w, gwts)
    set_aarray(capphi, A, A_xi, A_xi1)
    lib().ref_trunc_radius(rad)
    return get_aarray(len(capphi))


# ---- readwrite.f90 ------------------------------------------------------------------
def set_names(directory, sol_base, aoper_base, k_c, s, N_x, M):
    """commvar's name parameters (uniform meshes): current_solution_dir,
    current_solution_base_name, current_Aoperator_base_name, k_c, su/sv/sw, N,
    Mu/Mv/Mw."""
    lib().ref_set_names(directory.encode(), sol_base.encode(), aoper_base.encode(), k_c, s, s, s, N_x, M, M, M)


BASE, AOPER, AOPER_CHUNK, CHUNK = 0, 1, 2, 3


def base_name(kind=BASE, chunk=0):
    buf = C.create_string_buffer(512)
    lib().ref_base_name(kind, chunk, buf, 512)
    return buf.value.decode()


def write_aarray(capphi, A, A_xi, A_xi1, A_phi):
    """WriteAarraysDGV: the file base_name(BASE) + "_Aarrs.dat"."""
    set_aarray(capphi, A, A_xi, A_xi1)
    lib().ref_set_phi(len(A), _c(A_phi, np.int32))
    lib().ref_write_aarray()


def _arrays_out(mm):
    cap, a, x, y = get_aarray(mm)
    phi = np.zeros(len(a), np.int32)
    lib().ref_get_phi(len(a), phi)
    return cap, a, x, y, phi


def read_aarray():
    """ReadAarraysDGV: the file base_name(AOPER) + "_Aarrs.dat"."""
    mm = C.c_int64()
    lib().ref_read_aarray(C.byref(mm))
    return _arrays_out(mm.value)


def scan_chunks(nchnks):
    act = np.zeros(nchnks, np.int32)
    lib().ref_scan_chunks(nchnks, act)
    return act


def read_chunks(fstchnk, ofst, numrec, nchnks):
    mm = C.c_int64()
    lib().ref_read_chunks(fstchnk, ofst, numrec, nchnks, C.byref(mm))
    return _arrays_out(mm.value)


def write_sol(suff, e_time, dt, rkmts, f, hist, init=False):
    """WriteSol_DGV (or WriteSol_init_DGV) of f and hist (9, n) = f1..f4, frhs1..frhs5."""
    # the Fortran hist(n, 9) is column-major: its columns are hist's rows in order
    lib().ref_write_sol(suff.encode(), int(init), e_time, dt, rkmts, len(f), _c(f, np.float64),
                        np.ascontiguousarray(np.asarray(hist, np.float64)).reshape(-1))


def restart(txt, rkmts, n):
    """RestartSolutionDGV(txt) with commvar rkmts set to rkmts first: (e_time, dt,
    rkmts read, f, hist (9, n))."""
    e, d, k = C.c_double(), C.c_double(), C.c_int32()
    f = np.zeros(n)
    hist = np.zeros(9 * n)
    lib().ref_restart(txt.encode(), rkmts, n, C.byref(e), C.byref(d), C.byref(k), f, hist)
    return e.value, d.value, k.value, f, hist.reshape(9, n)


def write_macroerr(rows):
    """WriteErrHighCMoments_DGV of rows (nrow, 21): time, then the 20 moments."""
    rows = np.asarray(rows, np.float64)
    lib().ref_write_macroerr(len(rows), np.ascontiguousarray(rows.T).reshape(-1))


def l1_track(e_time, t_L, t_R, num_save_solution, l1_err):
    lib().ref_l1_track(e_time, t_L, t_R, num_save_solution, l1_err)


def l1_reset():
    lib().ref_l1_reset()


def momderiv_reset():
    lib().ref_momderiv_reset()


def momderiv_record(curr_time, f1, frhs1, num_eval_error, t_R, rkmts):
    """RecordMomDerivRelaxTimes_DGV(curr_time) (nodes from set_nodes)."""
    lib().ref_momderiv_record(curr_time, len(f1), _c(f1, np.float64), _c(frhs1, np.float64), num_eval_error, t_R,
                              rkmts)


def momderiv_rows():
    """The stored records (time + 40 values each)."""
    buf = np.zeros(1 << 16)
    m = lib().ref_momderiv_get(len(buf), buf)
    count, k = int(buf[0]), int(buf[1])
    return buf[2:2 + count * k].reshape(count, k)


def momderiv_write():
    """WriteMomDerivRelaxTimes_DGV: base_name(BASE) + "_dermom.txt" / "_relaxtimes.txt"."""
    lib().ref_momderiv_write()
