"""CPU oracle for the direct DGV collision path -- TEST INFRASTRUCTURE ONLY.

This package is the oracle.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
(``dgv0d3v_direct_collision_amd``) never does.  It wraps ``libdgv_oracle.so``, a
plain-C restatement of the reference's routines (alexmalekseenko/
DGV0D3V_direct_collision):

=====================  ==============================================================
function               reference routine restated
=====================  ==============================================================
``Geometry``           SetDGblzmGnodes/GauLeg, OldCreateUmesh, FillCellsArrsDGV,
                       SetNodesDGV, FindCellContainsPoint_DGV, SetCellsUGINdsSftArrs
                       (gaussian_mod.f90:4-53, miscset.f90:29-1311)
``eval_aplus``         EvalCollisionPeriodicAPlus_DGV (collision_mod.f90:143-229)
``eval_a``             EvalCollisionPeriodicA_DGV (collision_mod.f90:29-129)
``eval_aplus_list``    the same loop at a list of nodes (bench's CPU sample, 27^3 parity)
``eval_aplus_comp``    the same loop with Neumaier-compensated accumulation
``eval_aplus_dd_list`` the same loop at listed nodes, exact products and a
                       double-double sum (Q rounded once, whatever the conditioning)
``iphi``               IphiDGV (dgvtools_mod.f90:847-877)
``iphi_decomp``        IphiDGV_decomp (dgvtools_mod.f90:896-927), given fM (numpy)
``full_aarray``        the canonical records replicated to every node's basis function
                       (the full A-array IphiDGV walks; shift and bounds test of
                       dgvtools_mod.f90:1175-1247, collision_mod.f90:196-211)
``masscheck``          MassCheckRec (dgvtools_mod.f90:1437-1458)
``masscheck_high``     MassCheckRecHighCMoments (dgvtools_mod.f90:1501-1549)
``maxwell``            maxwelveldist (distributions_mod.f90:83-94)
``rk_step``            SpatialOperSH_DGV run_mode 0 + TimeIntegratorORK_SH_DGV
                       (spat_oper_mod.f90:28-68, time_integr_mod.f90:26-152)
``ab_step``            TimeIntegratorMTS_SH_DGV (time_integr_mod.f90:314-364)
``eval_mixed``          EvalCollisionPeriodicMixedTermsA_DGV (collision_mod.f90:244-332)
``prepare_fma``         PrepareFMA_DGV (collision_mod.f90:377-464)
``eval_linear``         EvalCollisionLinear (collision_mod.f90:346-359)
``Regime``              CheckSolutionMode_DGV and the run_mode switch of
                       SpatialOperSH_DGV (spat_oper_mod.f90:195-272, 28-68)
``trunc_threshold``     TruncAarrsTrhls_DGV (dgvtools_mod.f90:1313-1362)
``trunc_radius``        TruncAarrsRadius_DGV (dgvtools_mod.f90:1375-1429)
``read_chunks_ofst_nrec`` ReadAarraysChnksOfstNrecDGV (readwrite.f90:1261-1449) on
                       in-memory chunks
``seta`` / ``seta_raw``  SetA_DGV with the Lagrange basis restored, its adaptive
                       Simpson integration in chi and epsilon (dgvtools_mod.f90:377-681,
                       938-1122, 232-289)
=====================  ==============================================================

Index conventions are the reference's: node ids inside ``A_xi``/``A_xi1`` and cell
ids are 1-based.

Parity pins (what fixes each function, since the reference ships no tests or
golden vectors) live in ``tests/test_oracle_*.py``.  ``iphi``, ``iphi_decomp``,
``masscheck``, ``masscheck_high``, ``maxwell``, ``trunc_threshold`` and
``trunc_radius`` are also held bit for bit to the reference's own routines,
compiled unmodified into ``_ref/`` (``ref.py``, ``tests/test_ref_pins.py``), and
to their stored outputs (``tests/golden/``, ``tests/test_golden_ref.py``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DGV_ORACLE_LIB: a sanitizer-instrumented build of the same sources (tests only)
_LIB_PATH = os.environ.get("DGV_ORACLE_LIB") or os.path.join(_HERE, "libdgv_oracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class _Grid(C.Structure):
    _fields_ = [("Mu", C.c_int32), ("Mv", C.c_int32), ("Mw", C.c_int32),
                ("su", C.c_int32), ("sv", C.c_int32), ("sw", C.c_int32),
                ("u_L", C.c_double), ("u_R", C.c_double), ("v_L", C.c_double),
                ("v_R", C.c_double), ("w_L", C.c_double), ("w_R", C.c_double)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class _Geom(C.Structure):
    _fields_ = [("g", _Grid), ("nnodes", C.c_int64), ("ncells", C.c_int64),
                ("ccell", C.c_int32),
                ("cells_lu", _dp), ("cells_ru", _dp), ("cells_lv", _dp), ("cells_rv", _dp),
                ("cells_lw", _dp), ("cells_rw", _dp),
                ("cells_ugi", _ip), ("cells_vgi", _ip), ("cells_wgi", _ip),
                ("nodes_u", _dp), ("nodes_v", _dp), ("nodes_w", _dp), ("nodes_gwts", _dp),
                ("nodes_pcell", _ip), ("nodes_ui", _ip), ("nodes_vi", _ip), ("nodes_wi", _ip),
                ("nodes_dui", _ip), ("nodes_dvi", _ip), ("nodes_dwi", _ip),
                ("nodes_phican", _ip),
                ("gn_u", C.c_double * 16), ("gn_v", C.c_double * 16), ("gn_w", C.c_double * 16),
                ("gw_u", C.c_double * 16), ("gw_v", C.c_double * 16), ("gw_w", C.c_double * 16)]


_QFN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


def build(force: bool = False) -> str:
    """Compile libdgv_oracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        L.orc_gauleg.argtypes = [C.c_int, C.c_double, C.c_double, _f64p, _f64p]
        L.orc_geom_build.argtypes = [C.POINTER(_Grid), C.POINTER(C.POINTER(_Geom))]
        L.orc_geom_free.argtypes = [C.POINTER(_Geom)]
        L.orc_ashift.argtypes = [C.POINTER(_Geom), _i32p, _i64p]
        ev = [C.POINTER(_Geom), _i32p, _f64p, _i32p, _i32p, _f64p, _f64p, C.c_int64,
              C.c_int64, C.c_int]
        L.orc_eval_aplus.argtypes = ev
        L.orc_eval_a.argtypes = ev
        L.orc_eval_aplus_list.argtypes = [C.POINTER(_Geom), _i32p, _f64p, _i32p, _i32p, _f64p,
                                          _f64p, _i64p, C.c_int64, C.c_int]
        L.orc_eval_aplus_comp.argtypes = [C.POINTER(_Geom), _i32p, _f64p, _i32p, _i32p, _f64p,
                                          _f64p, C.c_void_p, C.c_int64, C.c_int64, C.c_int]
        L.orc_eval_aplus_dd_list.argtypes = [C.POINTER(_Geom), _i32p, _f64p, _i32p, _i32p, _f64p,
                                             _f64p, _f64p, _i64p, C.c_int64, C.c_int]
        L.orc_iphi.argtypes = [C.c_int64, _i32p, _f64p, _i32p, _i32p, _f64p, _f64p,
                               C.c_double, _f64p]
        L.orc_masscheck.argtypes = [C.POINTER(_Geom), _f64p, _f64p]
        L.orc_masscheck_high.argtypes = [C.POINTER(_Geom), _f64p, _f64p]
        L.orc_maxwell.argtypes = [C.c_double] * 8
        L.orc_maxwell.restype = C.c_double
        L.orc_spatial_oper.argtypes = [_QFN, C.c_void_p, C.c_int64, _f64p, _f64p, C.c_double,
                                       C.c_double]
        L.orc_rk_step.argtypes = [_QFN, C.c_void_p, C.c_int64, _f64p, C.c_double, C.c_double,
                                  C.c_int, _f64p]
        L.orc_ab_step.argtypes = [_QFN, C.c_void_p, C.c_int64, _f64p, C.c_double, C.c_double,
                                  C.c_int, _f64p]
        L.orc_basis.argtypes = [C.POINTER(_Geom), C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_int32]
        L.orc_basis.restype = C.c_double
        L.orc_prefilter.argtypes = [C.POINTER(_Geom), C.c_int32, C.c_int32, C.c_int32,
                                    C.c_double]
        L.orc_seta_raw.argtypes = [C.POINTER(_Geom), C.c_int, C.c_int32, C.c_int32, C.c_int32,
                                   C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.orc_seta.argtypes = [C.POINTER(_Geom), C.c_int, C.c_int32, C.c_double, C.c_double,
                               C.c_double, C.c_double, C.c_int64, _f64p, _i32p, _i32p, C.c_int]
        L.orc_seta.restype = C.c_int64
        L.orc_eval_mixed.argtypes = [C.POINTER(_Geom), _i32p, _f64p, _i32p, _i32p, _f64p, _f64p,
                                     _f64p, C.c_int]
        L.orc_prepare_fma.argtypes = [C.POINTER(_Geom), _i32p, _f64p, _i32p, _i32p, _f64p, _f64p]
        L.orc_eval_linear.argtypes = [C.c_int64, _f64p, _f64p, _f64p]
        L.orc_eval_linear.restype = None
        L.orc_trunc_threshold.argtypes = [C.c_int64, _i32p, _f64p, _i32p, _i32p, C.c_double]
        L.orc_trunc_threshold.restype = C.c_int64
        L.orc_trunc_radius.argtypes = [C.POINTER(_Geom), _i32p, _f64p, _i32p, _i32p, C.c_double]
        L.orc_trunc_radius.restype = C.c_int64
        _lib = L
    return _lib


def gauleg(n: int, x1: float = -1.0, x2: float = 1.0):
    x = np.zeros(n)
    w = np.zeros(n)
    if lib().orc_gauleg(n, x1, x2, x, w) != 0:
        raise RuntimeError("gauleg did not converge")
    return x, w


def _arr(ptr, n, dtype):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Geometry:
    """The commvar geometry arrays of one uniform grid (1-based ids kept)."""

    def __init__(self, M, s=3, box=(-1.0, 1.0)):
        if np.isscalar(M):
            M = (M, M, M)
        if np.isscalar(s):
            s = (s, s, s)
        if len(box) == 2:
            box = (box[0], box[1]) * 3
        self.grid = _Grid(int(M[0]), int(M[1]), int(M[2]), int(s[0]), int(s[1]), int(s[2]),
                          *[float(b) for b in box])
        self._p = C.POINTER(_Geom)()
        rc = lib().orc_geom_build(C.byref(self.grid), C.byref(self._p))
        if rc != 0:
            raise ValueError(f"orc_geom_build failed ({rc})")
        g = self._p.contents
        self.nnodes, self.ncells, self.ccell = g.nnodes, g.ncells, g.ccell
        nn, nc = self.nnodes, self.ncells
        for name in ("nodes_u", "nodes_v", "nodes_w", "nodes_gwts"):
            setattr(self, name, _arr(getattr(g, name), nn, np.float64))
        for name in ("nodes_pcell", "nodes_ui", "nodes_vi", "nodes_wi", "nodes_dui",
                     "nodes_dvi", "nodes_dwi", "nodes_phican"):
            setattr(self, name, _arr(getattr(g, name), nn, np.int32))
        for name in ("cells_lu", "cells_ru", "cells_lv", "cells_rv", "cells_lw", "cells_rw"):
            setattr(self, name, _arr(getattr(g, name), nc, np.float64))
        for name in ("cells_ugi", "cells_vgi", "cells_wgi"):
            setattr(self, name, _arr(getattr(g, name), nc, np.int32))
        self.M = tuple(M)
        self.s = tuple(s)

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.orc_geom_free(self._p)
            self._p = None

    @property
    def ptr(self):
        return self._p

    def ashift(self, capphi):
        out = np.zeros(self.nnodes, np.int64)
        lib().orc_ashift(self._p, np.ascontiguousarray(capphi, np.int32), out)
        return out


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def eval_aplus(G: Geometry, capphi, A, A_xi, A_xi1, f, nodes=None, nthreads=1):
    fc = np.zeros(G.nnodes)
    b, e = (0, G.nnodes) if nodes is None else nodes
    lib().orc_eval_aplus(G.ptr, _c(capphi, np.int32), _c(A, np.float64), _c(A_xi, np.int32),
                         _c(A_xi1, np.int32), _c(f, np.float64), fc, b, e, nthreads)
    return fc


def eval_aplus_list(G: Geometry, capphi, A, A_xi, A_xi1, f, nodes, nthreads=1):
    """a1 at the given 0-based nodes only; returns Q at those nodes."""
    nodes = _c(nodes, np.int64)
    fc = np.zeros(len(nodes))
    lib().orc_eval_aplus_list(G.ptr, _c(capphi, np.int32), _c(A, np.float64), _c(A_xi, np.int32),
                              _c(A_xi1, np.int32), _c(f, np.float64), fc, nodes, len(nodes),
                              nthreads)
    return fc


def eval_aplus_dd_list(G: Geometry, capphi, A, A_xi, A_xi1, f, nodes, nthreads=1):
    """a1 at the given 0-based nodes with exact products and a double-double
    sum (Q rounded once); returns (Q, sum|terms|/w) at those nodes."""
    nodes = _c(nodes, np.int64)
    fc = np.zeros(len(nodes))
    ab = np.zeros(len(nodes))
    lib().orc_eval_aplus_dd_list(G.ptr, _c(capphi, np.int32), _c(A, np.float64), _c(A_xi, np.int32),
                                 _c(A_xi1, np.int32), _c(f, np.float64), fc, ab, nodes, len(nodes),
                                 nthreads)
    return fc, ab


def eval_a(G: Geometry, capphi, A, A_xi, A_xi1, f, nodes=None, nthreads=1):
    fc = np.zeros(G.nnodes)
    b, e = (0, G.nnodes) if nodes is None else nodes
    rc = lib().orc_eval_a(G.ptr, _c(capphi, np.int32), _c(A, np.float64), _c(A_xi, np.int32),
                          _c(A_xi1, np.int32), _c(f, np.float64), fc, b, e, nthreads)
    if rc:
        raise MemoryError("orc_eval_a")
    return fc


def eval_aplus_comp(G: Geometry, capphi, A, A_xi, A_xi1, f, nodes=None, nthreads=1):
    """Returns (Q, sum|terms|/w) with a compensated accumulator."""
    fc = np.zeros(G.nnodes)
    ab = np.zeros(G.nnodes)
    b, e = (0, G.nnodes) if nodes is None else nodes
    lib().orc_eval_aplus_comp(G.ptr, _c(capphi, np.int32), _c(A, np.float64),
                              _c(A_xi, np.int32), _c(A_xi1, np.int32), _c(f, np.float64), fc,
                              ab.ctypes.data, b, e, nthreads)
    return fc, ab


def iphi(capphi, A, A_xi, A_xi1, gwts, f, coef=1.0):
    n = len(capphi)
    out = np.zeros(n)
    lib().orc_iphi(n, _c(capphi, np.int32), _c(A, np.float64), _c(A_xi, np.int32),
                   _c(A_xi1, np.int32), _c(gwts, np.float64), _c(f, np.float64), coef, out)
    return out


def full_aarray(G: Geometry, capphi, A, A_xi, A_xi1):
    """The canonical records replicated to every node's basis function (shift and
    bounds test applied once, dgvtools_mod.f90:1175-1247 and collision_mod.f90:
    196-211): (capphi, A, A_xi, A_xi1) of the full A-array that IphiDGV
    (dgvtools_mod.f90:847-877) walks, one basis function per node, 1-based ids."""
    D = int(np.prod(G.s))
    ash = G.ashift(capphi)
    M = G.M[0]
    node_cell = np.arange(G.nnodes) // D
    cc = np.stack([node_cell // (M * M), (node_cell // M) % M, node_cell % M], 1)
    fullcap, fa, fx, fx1 = np.zeros(G.nnodes, np.int32), [], [], []
    for i in range(G.nnodes):
        k = G.nodes_phican[i] - 1
        b, n = ash[i], capphi[k]
        d = np.array([G.nodes_dui[i], G.nodes_dvi[i], G.nodes_dwi[i]])
        X = cc[A_xi[b:b + n] - 1] + d
        Y = cc[A_xi1[b:b + n] - 1] + d
        ok = np.all((X >= 0) & (X < M) & (Y >= 0) & (Y < M), axis=1)
        nx = ((X[:, 0] * M + X[:, 1]) * M + X[:, 2]) * D + (A_xi[b:b + n] - 1) % D
        ny = ((Y[:, 0] * M + Y[:, 1]) * M + Y[:, 2]) * D + (A_xi1[b:b + n] - 1) % D
        fullcap[i] = ok.sum()
        fa.append(A[b:b + n][ok])
        fx.append(nx[ok] + 1)
        fx1.append(ny[ok] + 1)
    return (fullcap, np.concatenate(fa), np.concatenate(fx).astype(np.int32),
            np.concatenate(fx1).astype(np.int32))


def iphi_decomp(capphi, A, A_xi, A_xi1, gwts, f, fM, coef=1.0):
    """IphiDGV_decomp (dgvtools_mod.f90:896-927) on a full A-array, given the local
    Maxwellian fM: 2/w_i sum A (fM(xi) Df(xi1) + Df(xi) fM(xi1) + Df(xi) Df(xi1)),
    Df = f - fM, the records summed in order (numpy, for small grids)."""
    Df = np.asarray(f, np.float64) - np.asarray(fM, np.float64)
    x, y = np.asarray(A_xi) - 1, np.asarray(A_xi1) - 1
    a = np.asarray(A, np.float64)
    t1, t2, t3 = a * fM[x] * Df[y], a * Df[x] * fM[y], a * Df[x] * Df[y]
    out = np.zeros(len(capphi))
    ct = 0
    for i in range(len(capphi)):
        s = 0.0
        for j in range(ct, ct + capphi[i]):   # the statement's left-to-right additions (:919)
            s = ((s + t1[j]) + t2[j]) + t3[j]
        out[i] = 2 * s / gwts[i] * coef
        ct += capphi[i]
    return out


def masscheck(G: Geometry, f):
    out = np.zeros(5)
    lib().orc_masscheck(G.ptr, _c(f, np.float64), out)
    return out


def masscheck_high(G: Geometry, f):
    """MassCheckRecHighCMoments: n, ubar, vbar, wbar, T, Tu, Tv, Tw, mom3..6 (u, v, w)."""
    out = np.zeros(20)
    lib().orc_masscheck_high(G.ptr, _c(f, np.float64), out)
    return out


def maxwell(T, u0, v0, w0, n, u, v, w):
    return lib().orc_maxwell(T, u0, v0, w0, n, u, v, w)


def _wrap_q(q, n):
    def cb(_user, fp, fcp):
        f = np.ctypeslib.as_array(fp, shape=(n,))
        fc = np.ctypeslib.as_array(fcp, shape=(n,))
        fc[:] = q(f.copy())
        return 0
    return _QFN(cb)


def spatial_oper(q, f, dt, coef=1.0):
    n = len(f)
    out = np.zeros(n)
    cb = _wrap_q(q, n)
    rc = lib().orc_spatial_oper(cb, None, n, _c(f, np.float64), out, dt, coef)
    if rc:
        raise RuntimeError("spatial_oper")
    return out


def rk_step(q, f, dt, coef=1.0, rkmts=2):
    """One TimeIntegratorORK_SH_DGV step; q(f) -> Q(f).  Returns (f_new, frhs1)."""
    f = _c(f, np.float64).copy()
    n = len(f)
    frhs1 = np.zeros(n)
    cb = _wrap_q(q, n)
    rc = lib().orc_rk_step(cb, None, n, f, dt, coef, rkmts, frhs1)
    if rc:
        raise RuntimeError(f"rk_step rc={rc}")
    return f, frhs1


def ab_step(q, f, hist, dt, coef=1.0, rkmts=2):
    """One TimeIntegratorMTS_SH_DGV step; hist is (9, n) = f1..f4, frhs1..frhs5."""
    f = _c(f, np.float64).copy()
    n = len(f)
    h = _c(hist, np.float64).reshape(-1).copy()
    cb = _wrap_q(q, n)
    rc = lib().orc_ab_step(cb, None, n, f, dt, coef, rkmts, h)
    if rc:
        raise RuntimeError(f"ab_step rc={rc}")
    return f, h.reshape(9, n)


def basis(G: Geometry, mode, u, v, w, i1):
    return lib().orc_basis(G.ptr, mode, u, v, w, i1)


def prefilter(G: Geometry, i1, i2, i3, Trad=100.0):
    return bool(lib().orc_prefilter(G.ptr, i1, i2, i3, Trad))


def seta_raw(G: Geometry, i1, i2, i3, ErrChi, ErrEps, mode=0):
    out = C.c_double()
    rc = lib().orc_seta_raw(G.ptr, mode, i1, i2, i3, ErrChi, ErrEps, C.byref(out))
    if rc:
        raise RuntimeError("SetA_DGV: number of nodes too big")
    return out.value


def seta(G: Geometry, i1, Trad=100.0, ErrChi=1e-6, ErrEps=1e-6, min_sens=0.0, mode=0,
         cap=None, nthreads=8):
    """SetA_DGV for one basis function: records (A, A_xi, A_xi1) in the reference's order."""
    if cap is None:
        cap = G.nnodes * (G.nnodes - 1) // 2
    A = np.zeros(cap)
    xi = np.zeros(cap, np.int32)
    xi1 = np.zeros(cap, np.int32)
    n = lib().orc_seta(G.ptr, mode, i1, Trad, ErrChi, ErrEps, min_sens, cap, A, xi, xi1,
                       nthreads)
    if n < 0:
        raise RuntimeError(f"orc_seta failed ({n})")
    return A[:n].copy(), xi[:n].copy(), xi1[:n].copy()


def trunc_threshold(capphi, A, A_xi, A_xi1, trhld):
    """TruncAarrsTrhls_DGV: returns new (capphi, A, A_xi, A_xi1)."""
    cap = np.array(capphi, np.int32)
    a, x, y = np.array(A, np.float64), np.array(A_xi, np.int32), np.array(A_xi1, np.int32)
    n = lib().orc_trunc_threshold(len(cap), cap, a, x, y, trhld)
    return cap, a[:n].copy(), x[:n].copy(), y[:n].copy()


def trunc_radius(G: "Geometry", capphi, A, A_xi, A_xi1, rad):
    """TruncAarrsRadius_DGV: returns new (capphi, A, A_xi, A_xi1)."""
    cap = np.array(capphi, np.int32)
    a, x, y = np.array(A, np.float64), np.array(A_xi, np.int32), np.array(A_xi1, np.int32)
    n = lib().orc_trunc_radius(G.ptr, cap, a, x, y, rad)
    return cap, a[:n].copy(), x[:n].copy(), y[:n].copy()


def eval_mixed(G: "Geometry", capphi, A, A_xi, A_xi1, f, fm, nthreads=1):
    """EvalCollisionPeriodicMixedTermsA_DGV(f, fm)."""
    fc = np.zeros(G.nnodes)
    lib().orc_eval_mixed(G.ptr, _c(capphi, np.int32), _c(A, np.float64), _c(A_xi, np.int32),
                         _c(A_xi1, np.int32), _c(f, np.float64), _c(fm, np.float64), fc, nthreads)
    return fc


def prepare_fma(G: "Geometry", capphi, A, A_xi, A_xi1, fm):
    """PrepareFMA_DGV: the dense fmA as an (N, N) array indexed [row, column]."""
    N = G.nnodes
    buf = np.zeros(N * N)
    lib().orc_prepare_fma(G.ptr, _c(capphi, np.int32), _c(A, np.float64), _c(A_xi, np.int32),
                          _c(A_xi1, np.int32), _c(fm, np.float64), buf)
    return buf.reshape(N, N).T


def eval_linear(fmA, f):
    """EvalCollisionLinear: fc = 2 f . fmA (fmA indexed [row, column])."""
    N = fmA.shape[0]
    colmajor = np.ascontiguousarray(fmA.T).reshape(-1)
    fc = np.zeros(N)
    lib().orc_eval_linear(N, colmajor, _c(f, np.float64), fc)
    return fc


class Regime:
    """CheckSolutionMode_DGV (spat_oper_mod.f90:195-272) and the run_mode
    switch of SpatialOperSH_DGV (spat_oper_mod.f90:41-60) over the oracle's
    evaluators.  ``collision(f)`` is the unscaled collision term of one
    SpatialOperSH_DGV call; pass it as ``q`` to ``rk_step`` / ``ab_step``.
    ``trace`` holds (run_mode after the check, L1_err) per call."""

    def __init__(self, G: "Geometry", capphi, A, A_xi, A_xi1, decomp_lev, linear_lev, run_mode=0):
        self.G = G
        self.A = (capphi, A, A_xi, A_xi1)
        self.decomp_lev = decomp_lev
        self.linear_lev = linear_lev
        self.run_mode = run_mode
        self.fm = None
        self.fmA = None
        self.fma_builds = 0
        self.trace = []

    def _prepare(self, fmn):
        self.fm = fmn.copy()                                   # fm = fMaxwellNew
        self.fmA = prepare_fma(self.G, *self.A, self.fm)       # PrepareFMA_DGV
        self.fma_builds += 1

    def check(self, f):
        """Returns Df (None in run_mode 0)."""
        G = self.G
        gw = G.nodes_gwts
        dens, ub, vb, wb, T = masscheck(G, f)                  # line 236
        fmn = np.array([maxwell(T, ub, vb, wb, dens, u, v, w)
                        for u, v, w in zip(G.nodes_u, G.nodes_v, G.nodes_w)])
        l1 = float(np.sum(np.abs(f - fmn) * gw) / dens)        # line 238
        df = None
        if l1 < self.decomp_lev:
            if self.run_mode == 0:
                self._prepare(fmn)
                self.run_mode = 2 if l1 < self.linear_lev else 1
            elif self.run_mode in (1, 2):
                l1_fm = float(np.sum(np.abs(self.fm - fmn) * gw) / dens)
                if l1_fm > self.linear_lev / 2.0:
                    self._prepare(fmn)
                if self.run_mode == 1 and l1 < self.linear_lev:
                    self.run_mode = 2
                elif self.run_mode == 2 and l1 > self.linear_lev:
                    self.run_mode = 1
            else:
                raise ValueError("run_mode")
            df = f - self.fm                                   # line 264
        else:
            self.run_mode = 0
        self.trace.append((self.run_mode, l1))
        return df

    def collision(self, f):
        df = self.check(f)
        if self.run_mode == 0:
            return eval_aplus(self.G, *self.A, f, nthreads=8)
        lin = eval_linear(self.fmA, df)
        if self.run_mode == 1:
            return eval_aplus(self.G, *self.A, df, nthreads=8) + lin
        return lin


def read_chunks_ofst_nrec(chunks, fstchnk, ofst, numrec):
    """ReadAarraysChnksOfstNrecDGV (readwrite.f90:1261-1449) on in-memory chunks:
    chunks[j] = (A_capphi, A, A_xi, A_xi1, A_phi) of chunk j (0-based list, the
    routine's chunk numbers are 1-based).  Returns the arrays as the routine
    leaves them in commvar: (A_capphi(mm), A, A_xi, A_xi1, A_phi) of length
    numrec (shorter if the chunks run out).  Raises where it prints and stops.
    Plain loops with the routine's 1-based indices."""
    nchnks = len(chunks)
    cap1, a1, x1, y1, p1 = chunks[fstchnk - 1]
    mm, nn = len(cap1), len(a1)
    A_capphi = [0] * (mm + 1)                        # 1-based
    A, A_xi, A_xi1, A_phi = [], [], [], []
    Act = 0
    Act_need = min(nn - ofst, numrec - Act)
    A += list(a1[ofst:ofst + Act_need])
    A_xi += list(x1[ofst:ofst + Act_need])
    A_xi1 += list(y1[ofst:ofst + Act_need])
    A_phi += list(p1[ofst:ofst + Act_need])
    Act = Act_need
    csum = lambda i: int(sum(cap1[:i]))              # sum(Achnks_capphi(1:i))
    i = 1
    while i <= mm and ofst > csum(i):
        i = i + 1
    if i > mm:
        raise IndexError("no records in the chunk after offset taken")
    if csum(i) - ofst < Act_need:
        A_capphi[i] = csum(i) - ofst
        Act_need = Act_need - csum(i) + ofst
    else:
        A_capphi[i] = Act_need
        Act_need = 0
    while Act_need > 0 and i < mm:
        i = i + 1
        if cap1[i - 1] >= Act_need:
            A_capphi[i] = Act_need
            Act_need = 0
        else:
            A_capphi[i] = int(cap1[i - 1])
            Act_need = Act_need - int(cap1[i - 1])
    if i >= mm and Act_need > 0:
        raise IndexError("no records in the chunk after offset taken")
    j = fstchnk + 1
    while Act < numrec and j <= nchnks:
        capj, aj, xj, yj, pj = chunks[j - 1]
        nn = len(aj)
        Act_need = min(nn, numrec - Act)
        A += list(aj[:Act_need])
        A_xi += list(xj[:Act_need])
        A_xi1 += list(yj[:Act_need])
        A_phi += list(pj[:Act_need])
        Act = Act + Act_need
        i = 0
        while Act_need > 0 and i < mm:
            i = i + 1
            if capj[i - 1] >= Act_need:
                A_capphi[i] = A_capphi[i] + Act_need
                Act_need = 0
            else:
                A_capphi[i] = A_capphi[i] + int(capj[i - 1])
                Act_need = Act_need - int(capj[i - 1])
        if i >= mm and Act_need > 0:
            raise IndexError("no records in the chunk after offset taken")
        j = j + 1
    return (np.array(A_capphi[1:], np.int32), np.array(A, np.float64), np.array(A_xi, np.int32),
            np.array(A_xi1, np.int32), np.array(A_phi, np.int32))
