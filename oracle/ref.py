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
