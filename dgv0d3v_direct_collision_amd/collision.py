"""Host-side mirror of the reference's collision / spatial-operator / stepper API.

The reference keeps its grid and A-arrays in the ``commvar`` module and exposes
Fortran procedures that read them (alexmalekseenko/DGV0D3V_direct_collision):

=====================================  =================================================
reference                              here (``DGV`` instance method)
=====================================  =================================================
SetNodesDGV / SetCellsUGINdsSftArrs    ``DGV(M, s, box)`` (grid + canonical cell)
ReadAarraysChnksOfstNrecDGV            ``set_aarray`` / ``read_aarray``
SetA_DGV                               ``generate_aarray``
WriteAarraysDGV                        ``write_aarray``
EvalCollisionPeriodicAPlus_DGV(f,fc)   ``EvalCollisionPeriodicAPlus_DGV(f)``
EvalCollisionPeriodicA_DGV(f,fc)       ``EvalCollisionPeriodicA_DGV(f)``
SpatialOperSH_DGV(f,fcol,dt)           ``SpatialOperSH_DGV(f, dt)``
TimeIntegratorORK_SH_DGV(f,dt)         ``TimeIntegratorORK_SH_DGV(f, dt)`` (f in place)
TimeIntegratorMTS_SH_DGV(f,dt)         ``TimeIntegratorMTS_SH_DGV(f, dt)`` (f in place)
PrepareMTS_SH_DGV(f,dt)                ``PrepareMTS_SH_DGV(f, dt)``
MassCheckRec(f,n,u,v,w,T)              ``MassCheckRec(f)``
=====================================  =================================================

Errors surface as ``DGVError`` where the reference prints and STOPs.  All
arithmetic runs in libdgv_amd.so on the GPU; there is no host fallback.
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Optional

import numpy as np

from . import _lib
from ._lib import AgenParams, DGVError, GridDesc, check, dptr, iptr, load


def grid_desc(M, s=3, box=(-1.0, 1.0)) -> GridDesc:
    if np.isscalar(M):
        M = (M, M, M)
    if np.isscalar(s):
        s = (s, s, s)
    if len(box) == 2:
        box = (box[0], box[1]) * 3
    return GridDesc(int(M[0]), int(M[1]), int(M[2]), int(s[0]), int(s[1]), int(s[2]),
                    *[float(b) for b in box])


def agen_pair(M, s, box, i2: int, i3: int, Trad=100.0, min_sens=0.0, nquad=24,
              moment_mode=0):
    """Host evaluation of one pair's Atemp_l and A_l (no GPU needed)."""
    L = load()
    g = grid_desc(M, s, box)
    D = g.su * g.sv * g.sw
    raw = np.zeros(D)
    A = np.zeros(D)
    p = AgenParams(Trad, min_sens, nquad, moment_mode)
    rc = L.dgv_agen_pair(C.byref(g), i2, i3, C.byref(p), dptr(raw), dptr(A))
    check(rc)
    return raw, A, rc == 1


def grid_geometry(M, s=3, box=(-1.0, 1.0)):
    """(u, v, w, gwts, ccell) of a grid, computed on the host (no GPU needed)."""
    L = load()
    g = grid_desc(M, s, box)
    n = C.c_int64()
    check(L.dgv_grid_geometry(C.byref(g), None, None, None, None, None, C.byref(n)))
    u, v, w, gw = (np.zeros(n.value) for _ in range(4))
    cc = C.c_int32()
    check(L.dgv_grid_geometry(C.byref(g), dptr(u), dptr(v), dptr(w), dptr(gw), C.byref(cc),
                              C.byref(n)))
    return u, v, w, gw, cc.value


def chunk_name(M, s, box, directory="", base="", chunk=0, k_c=2, N_x=3) -> str:
    """MakeBaseNameAoperChnksDGV file name (readwrite.f90:2224-2281) + "_Aarrs.dat"."""
    g = grid_desc(M, s, box)
    buf = C.create_string_buffer(1024)
    check(load().dgv_aarray_chunk_name(C.byref(g), directory.encode(), base.encode(), chunk,
                                       k_c, N_x, buf, 1024))
    return buf.value.decode()


def _lptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def aarray_file_header(path: str):
    """(mm, nn) of an A-array file: record 1 of WriteAarraysDGV's layout."""
    mm, nn = C.c_int64(), C.c_int64()
    check(load().dgv_aarray_file_header(path.encode(), C.byref(mm), C.byref(nn)))
    return mm.value, nn.value


def aarray_file_read(path: str):
    """ReadAarraysDGV on host arrays (readwrite.f90:1066-1118): returns
    (A_capphi, A, A_xi, A_xi1, A_phi); records above 2^31 - 9 bytes in either
    framing (Intel / gfortran subrecords, or LLVM Flang's single record)."""
    mm, nn = aarray_file_header(path)
    cap = np.zeros(mm, np.int32)
    A = np.zeros(nn)
    xi, xi1, phi = (np.zeros(nn, np.int32) for _ in range(3))
    check(load().dgv_aarray_file_read(path.encode(), iptr(cap), mm, dptr(A), iptr(xi), iptr(xi1), iptr(phi), nn))
    return cap, A, xi, xi1, phi


def aarray_file_write(path: str, capphi, A, A_xi, A_xi1, A_phi, framing: int = _lib.DGV_FRAMING_SUBRECORDS):
    """WriteAarraysDGV's two records (readwrite.f90:1459-1484) from host arrays."""
    cap = np.ascontiguousarray(capphi, np.int32)
    A = np.ascontiguousarray(A, np.float64)
    xi, xi1, phi = (np.ascontiguousarray(x, np.int32) for x in (A_xi, A_xi1, A_phi))
    check(load().dgv_aarray_file_write(path.encode(), iptr(cap), len(cap), dptr(A), iptr(xi), iptr(xi1), iptr(phi),
                                       len(A), framing))


def set_fortran_subrecord_limit(nbytes: int = 2147483639):
    """Largest subrecord payload of the Intel / gfortran framing (a test hook)."""
    check(load().dgv_set_fortran_subrecord_limit(nbytes))


def ScanAarrsChnks4Acapphi(M, s, box, nchnks: int, directory="", base="", k_c=2, N_x=3):
    """Records per chunk of a chunk set (readwrite.f90:2292-2365)."""
    g = grid_desc(M, s, box)
    act = np.zeros(nchnks, np.int64)
    check(load().dgv_aarray_chunks_scan(C.byref(g), directory.encode(), base.encode(), k_c, N_x, nchnks, _lptr(act)))
    return act


def ReadAarraysChnksOfstNrecDGV(M, s, box, fstchnk: int, ofst: int, numrec: int, nchnks: int, directory="",
                                base="", k_c=2, N_x=3):
    """numrec records of a chunk set from chunk fstchnk (1-based) on, after the
    first ofst records of that chunk (readwrite.f90:1261-1449): returns
    (A_capphi, A, A_xi, A_xi1, A_phi) as the reference's arrays hold them."""
    g = grid_desc(M, s, box)
    mm, _ = aarray_file_header(chunk_name(M, s, box, directory, base, fstchnk - 1, k_c, N_x))
    cap = np.zeros(mm, np.int32)
    A = np.zeros(numrec)
    xi, xi1, phi = (np.zeros(numrec, np.int32) for _ in range(3))
    nread = C.c_int64()
    check(load().dgv_aarray_chunks_read(C.byref(g), directory.encode(), base.encode(), k_c, N_x, fstchnk, ofst,
                                        numrec, nchnks, iptr(cap), mm, dptr(A), iptr(xi), iptr(xi1), iptr(phi),
                                        numrec, C.byref(nread)))
    n = nread.value
    return cap, A[:n], xi[:n], xi1[:n], phi[:n]


def ReadAarraysChnksDGV(M, s, box, nchnks: int, directory="", base="", k_c=2, N_x=3):
    """Every record of a chunk set (readwrite.f90:1130-1243)."""
    total = int(ScanAarrsChnks4Acapphi(M, s, box, nchnks, directory, base, k_c, N_x).sum())
    return ReadAarraysChnksOfstNrecDGV(M, s, box, 1, 0, total, nchnks, directory, base, k_c, N_x)


def sol_file_name(M, s, box, suff, directory="", base="", k_c=1, N_x=1) -> str:
    """WriteSol_DGV / RestartSolutionDGV file name: MakeBaseNameDGV
    (readwrite.f90:1505-1560) + "_time" + suff + "_sol.dat"."""
    g = grid_desc(M, s, box)
    buf = C.create_string_buffer(1024)
    check(load().dgv_sol_file_name(C.byref(g), directory.encode(), base.encode(), k_c, N_x, suff.encode(),
                                   buf, 1024))
    return buf.value.decode()


def log_file_name(M, s, box, suffix, directory="", base="", k_c=1, N_x=1) -> str:
    """MakeBaseNameDGV + suffix ("_macroerr.txt", "_L1err.txt")."""
    g = grid_desc(M, s, box)
    buf = C.create_string_buffer(1024)
    check(load().dgv_log_file_name(C.byref(g), directory.encode(), base.encode(), k_c, N_x, suffix.encode(),
                                   buf, 1024))
    return buf.value.decode()


def write_sol_file(path, e_time, dt, rkmts, f, hist=None):
    """WriteSol_DGV (readwrite.f90:1756-1796); hist = (9, n) f1..f4, frhs1..frhs5."""
    f = np.ascontiguousarray(f, np.float64)
    h = None if hist is None else np.ascontiguousarray(hist, np.float64)
    check(load().dgv_sol_write_file(str(path).encode(), e_time, dt, rkmts, len(f), dptr(f),
                                    None if h is None else dptr(h)))


def write_sol_init_file(path, e_time, dt, rkmts, f):
    """WriteSol_init_DGV (readwrite.f90:1810-1850): the solution file with f alone."""
    f = np.ascontiguousarray(f, np.float64)
    check(load().dgv_sol_write_init_file(str(path).encode(), e_time, dt, rkmts, len(f), dptr(f)))


def read_sol_file(path, n):
    """RestartSolutionDGV (readwrite.f90:2389-2513) -> (e_time, dt, rkmts, f, hist)."""
    f = np.zeros(n)
    hist = np.zeros((9, n))
    e, dt, k = C.c_double(), C.c_double(), C.c_int32()
    check(load().dgv_sol_read_file(str(path).encode(), C.byref(e), C.byref(dt), C.byref(k), n, dptr(f),
                                   dptr(hist)))
    return e.value, dt.value, k.value, f, hist


class DGV:
    """One velocity grid, its A-array and the evaluator state on one GPU."""

    def __init__(self, M, s=3, box=(-1.0, 1.0), device: int = 0, mol_diam=1.0, L_inf=1.0,
                 N_inf=1.0, rkmts: int = 2):
        self._L = load()
        self.desc = grid_desc(M, s, box)
        p = C.c_void_p()
        rc = self._L.dgv_ctx_create(C.byref(self.desc), device, C.byref(p))
        check(rc)
        self._ctx = p
        n = C.c_int64()
        check(self._L.dgv_ctx_num_nodes(self._ctx, C.byref(n)), self._ctx)
        self.nnodes = n.value
        self.D = self.desc.su * self.desc.sv * self.desc.sw
        self.M = (self.desc.Mu, self.desc.Mv, self.desc.Mw)
        cc = C.c_int32()
        check(self._L.dgv_ctx_canonical_cell(self._ctx, C.byref(cc)), self._ctx)
        self.ccell = cc.value
        self.nodes_u = np.zeros(self.nnodes)
        self.nodes_v = np.zeros(self.nnodes)
        self.nodes_w = np.zeros(self.nnodes)
        self.nodes_gwts = np.zeros(self.nnodes)
        check(self._L.dgv_ctx_get_nodes(self._ctx, dptr(self.nodes_u), dptr(self.nodes_v),
                                        dptr(self.nodes_w), dptr(self.nodes_gwts)), self._ctx)
        # (mol_diam/L_inf)**2*N_inf of SpatialOperSH_DGV (spat_oper_mod.f90:65)
        self.coef = (mol_diam / L_inf) ** 2 * N_inf
        self.rkmts = rkmts
        self.frhs1 = np.zeros(self.nnodes)
        self.hist = np.zeros((9, self.nnodes))
        self.e_time = 0.0

    # -- lifetime ----------------------------------------------------------------
    def _register_dependent(self, obj):
        """Batches and surrogates built on this context; closed before it."""
        if not hasattr(self, "_dependents"):
            self._dependents = weakref.WeakSet()
        self._dependents.add(obj)

    def close(self):
        for obj in list(getattr(self, "_dependents", ())):
            obj.close()
        if getattr(self, "_ctx", None):
            self._L.dgv_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def ctx(self):
        return self._ctx

    def _chk(self, rc):
        check(rc, self._ctx)

    def _f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        if f.shape != (self.nnodes,):
            raise DGVError(-1, f"array of shape {f.shape}, expected ({self.nnodes},)")
        return f

    # -- geometry ------------------------------------------------------------------
    def shifts(self):
        n = self.nnodes
        ph, du, dv, dw = (np.zeros(n, np.int32) for _ in range(4))
        self._chk(self._L.dgv_ctx_get_shifts(self._ctx, iptr(ph), iptr(du), iptr(dv), iptr(dw)))
        return ph, du, dv, dw

    # -- A-array -------------------------------------------------------------------
    def set_aarray(self, capphi, A, A_xi, A_xi1, index_base: int = 1):
        capphi = np.ascontiguousarray(capphi, np.int32)
        A = np.ascontiguousarray(A, np.float64)
        xi = np.ascontiguousarray(A_xi, np.int32)
        xi1 = np.ascontiguousarray(A_xi1, np.int32)
        if not (len(A) == len(xi) == len(xi1)):
            raise DGVError(-1, "A, A_xi and A_xi1 must have equal length")
        self._chk(self._L.dgv_ctx_set_aarray(self._ctx, iptr(capphi), len(capphi), dptr(A),
                                             iptr(xi), iptr(xi1), len(A), index_base))

    def generate_aarray(self, Trad=100.0, min_sens=0.0, nquad=24, moment_mode=0):
        p = AgenParams(Trad, min_sens, nquad, moment_mode)
        self._chk(self._L.dgv_ctx_generate_aarray(self._ctx, C.byref(p)))

    def aarray_size(self):
        n = C.c_int64()
        per = np.zeros(self.D, np.int64)
        self._chk(self._L.dgv_ctx_aarray_size(self._ctx, C.byref(n),
                                              per.ctypes.data_as(C.POINTER(C.c_int64))))
        return n.value, per

    def workload(self):
        """dict(records, pairs, visits, pair_visits) of one Q evaluation."""
        out = np.zeros(4, np.int64)
        self._chk(self._L.dgv_ctx_workload(self._ctx, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return dict(records=int(out[0]), pairs=int(out[1]), visits=int(out[2]),
                    pair_visits=int(out[3]))

    def get_aarray_basis(self, l: int, index_base: int = 1):
        n = C.c_int64()
        self._chk(self._L.dgv_ctx_get_aarray_basis(self._ctx, l, None, None, None, 0,
                                                   C.byref(n), index_base))
        A = np.zeros(n.value)
        xi = np.zeros(n.value, np.int32)
        xi1 = np.zeros(n.value, np.int32)
        self._chk(self._L.dgv_ctx_get_aarray_basis(self._ctx, l, dptr(A), iptr(xi), iptr(xi1),
                                                   n.value, C.byref(n), index_base))
        return A, xi, xi1

    def get_aarray_bases(self, ls, index_base: int = 1):
        """Records of the canonical basis functions ``ls`` (0-based), each list in
        the reference's serial order, from one pass over the device layout:
        returns (counts, A, A_xi, A_xi1) with the lists back to back."""
        ls = np.ascontiguousarray(ls, np.int32)
        counts = np.zeros(len(ls), np.int64)
        self._chk(self._L.dgv_ctx_get_aarray_bases(self._ctx, iptr(ls), len(ls), None, None, None, 0,
                                                   counts.ctypes.data_as(C.POINTER(C.c_int64)), index_base))
        n = int(counts.sum())
        A, xi, xi1 = np.zeros(n), np.zeros(n, np.int32), np.zeros(n, np.int32)
        self._chk(self._L.dgv_ctx_get_aarray_bases(self._ctx, iptr(ls), len(ls), dptr(A), iptr(xi), iptr(xi1), n,
                                                   counts.ctypes.data_as(C.POINTER(C.c_int64)), index_base))
        return counts, A, xi, xi1

    def get_aarray(self, index_base: int = 1):
        """The whole A-array in the reference layout (capphi, A, A_xi, A_xi1)."""
        capphi = np.zeros(self.nnodes, np.int32)
        As, xs, x1s = [], [], []
        for l in range(self.D):
            A, xi, xi1 = self.get_aarray_basis(l, index_base)
            capphi[(self.ccell - 1) * self.D + l] = len(A)
            As.append(A)
            xs.append(xi)
            x1s.append(xi1)
        return capphi, np.concatenate(As), np.concatenate(xs), np.concatenate(x1s)

    def write_aarray(self, path: str):
        self._chk(self._L.dgv_aarray_write_file(self._ctx, path.encode()))

    def read_aarray(self, path: str):
        self._chk(self._L.dgv_ctx_read_aarray_file(self._ctx, path.encode()))

    def write_aarray_chunks(self, directory: str, base: str, max_records: int, k_c: int = 2, N_x: int = 3,
                            framing: int = _lib.DGV_FRAMING_SUBRECORDS) -> int:
        """The A-array as a chunk set of at most max_records records per chunk,
        named by MakeBaseNameAoperChnksDGV; returns the number of chunks."""
        n = C.c_int32()
        self._chk(self._L.dgv_ctx_write_aarray_chunks(self._ctx, directory.encode(), base.encode(), k_c, N_x,
                                                      max_records, framing, C.byref(n)))
        return n.value

    def read_aarray_chunks(self, nchnks: int, directory: str, base: str, k_c: int = 2, N_x: int = 3):
        """ReadAarraysChnksDGV(nchnks) (readwrite.f90:1130-1243) into this context."""
        self._chk(self._L.dgv_ctx_read_aarray_chunks(self._ctx, directory.encode(), base.encode(), k_c, N_x, nchnks))

    def TruncAarrsTrhls_DGV(self, trhld: float):
        """Drops the records with |A| <= trhld (dgvtools_mod.f90:1313-1362)."""
        self._chk(self._L.dgv_ctx_trunc_aarray_threshold(self._ctx, trhld))

    def TruncAarrsRadius_DGV(self, rad: float):
        """Drops the records whose nodes lie more than rad apart
        (dgvtools_mod.f90:1375-1429)."""
        self._chk(self._L.dgv_ctx_trunc_aarray_radius(self._ctx, rad))

    def aarray_abs(self):
        """Replaces every coefficient by |A| (diagnostic: Q then becomes the
        scale sum 2/w |A| f f of the terms of Q)."""
        self._chk(self._L.dgv_ctx_aarray_abs(self._ctx))

    def set_kernel(self, kernel: int):
        self._chk(self._L.dgv_ctx_set_kernel(self._ctx, kernel))

    def set_projection(self, on: bool):
        """Project every Q onto the complement of the collision invariants (FP64)."""
        self._chk(self._L.dgv_ctx_set_projection(self._ctx, int(on)))

    def set_tile(self, pairs_per_item: int):
        self._chk(self._L.dgv_ctx_set_tile(self._ctx, pairs_per_item))

    def autotune(self):
        best = C.c_int32()
        ms = C.c_double()
        self._chk(self._L.dgv_ctx_autotune(self._ctx, C.byref(best), C.byref(ms)))
        return best.value, ms.value

    # -- the reference's procedures ---------------------------------------------
    def EvalCollisionPeriodicAPlus_DGV(self, f) -> np.ndarray:
        """collision_mod.f90:143-229"""
        f = self._f(f)
        fc = np.zeros(self.nnodes)
        self._chk(self._L.dgv_eval_q(self._ctx, dptr(f), dptr(fc), self.nnodes))
        return fc

    def EvalCollisionPeriodicA_DGV(self, f) -> np.ndarray:
        """collision_mod.f90:29-129 (same quantity as the APlus variant)"""
        f = self._f(f)
        fc = np.zeros(self.nnodes)
        self._chk(self._L.dgv_eval_q_a(self._ctx, dptr(f), dptr(fc), self.nnodes))
        return fc

    def EvalCollisionPeriodicMixedTermsA_DGV(self, f, fm) -> np.ndarray:
        """Cross term 2/w sum A (fm(xi) f(xi1) + f(xi) fm(xi1)) (collision_mod.f90:244-332)."""
        f = self._f(f)
        fm = self._f(fm)
        fc = np.zeros(self.nnodes)
        self._chk(self._L.dgv_eval_mixed(self._ctx, dptr(f), dptr(fm), dptr(fc), self.nnodes))
        return fc

    def IphiDGV(self, f) -> np.ndarray:
        """dgvtools_mod.f90:847-877: 2/w_i sum A f(xi) f(xi1) for every node's basis
        function, times (mol_diam/L_inf)^2 N_inf; the plain sums, whatever the
        projection setting."""
        f = self._f(f)
        out = np.zeros(self.nnodes)
        self._chk(self._L.dgv_iphi(self._ctx, dptr(f), dptr(out), self.nnodes, self.coef))
        return out

    def IphiDGV_decomp(self, f) -> np.ndarray:
        """dgvtools_mod.f90:896-927: IphiDGV with f split at its local Maxwellian
        fM: 2/w_i sum A (fM Df + Df fM + Df Df), Df = f - fM."""
        f = self._f(f)
        out = np.zeros(self.nnodes)
        self._chk(self._L.dgv_iphi_decomp(self._ctx, dptr(f), dptr(out), self.nnodes, self.coef))
        return out

    def PrepareFMA_DGV(self, fm=None):
        """collision_mod.f90:377-464: the dense linearised operator at the
        Maxwellian fm (the context's last one when fm is None), kept on the GPU."""
        if fm is None:
            fm = self.fm
        fm = self._f(fm)
        self._chk(self._L.dgv_prepare_fma(self._ctx, dptr(fm), self.nnodes))

    def EvalCollisionLinear(self, f) -> np.ndarray:
        """collision_mod.f90:346-359: fc = 2 MATMUL(f, fmA)."""
        f = self._f(f)
        fc = np.zeros(self.nnodes)
        self._chk(self._L.dgv_eval_linear(self._ctx, dptr(f), dptr(fc), self.nnodes))
        return fc

    def eval_linear_dev(self, d_f: int, d_fc: int, nvec: int = 1, ldf: int = 0, ldq: int = 0):
        n = self.nnodes
        self._chk(self._L.dgv_eval_linear_dev(self._ctx, C.c_void_p(int(d_f)), ldf or n, C.c_void_p(int(d_fc)),
                                              ldq or n, n, nvec))

    @property
    def fmA(self) -> np.ndarray:
        """fmA indexed [row, column] as the reference's fmA(row, column)."""
        n = self.nnodes
        buf = np.zeros(n * n)
        self._chk(self._L.dgv_ctx_get_fma(self._ctx, dptr(buf), n))
        return buf.reshape(n, n).T

    @property
    def fm(self) -> np.ndarray:
        """The Maxwellian of the linearised operator (commvar fm)."""
        out = np.zeros(self.nnodes)
        self._chk(self._L.dgv_ctx_get_fm(self._ctx, dptr(out), self.nnodes))
        return out

    def set_regime(self, run_mode: int = 0, decomp_lev: float = 0.0, linear_lev: float = 0.0):
        """commvar run_mode, decomp_lev, linear_lev.  decomp_lev > 0 turns on
        CheckSolutionMode_DGV before every spatial operator."""
        self._chk(self._L.dgv_ctx_set_regime(self._ctx, run_mode, decomp_lev, linear_lev))

    def regime(self) -> dict:
        """run_mode, the last L1 deviation from the local Maxwellian, the last
        |fm - f_M|_1 / n, and the numbers of checks and fmA builds."""
        m = C.c_int32()
        e = C.c_double()
        efm = C.c_double()
        cnt = (C.c_int64 * 2)()
        self._chk(self._L.dgv_ctx_get_regime(self._ctx, C.byref(m), C.byref(e), C.byref(efm), cnt))
        return {"run_mode": m.value, "l1_err": e.value, "l1_err_fm": efm.value, "checks": cnt[0],
                "fma_builds": cnt[1]}

    @property
    def run_mode(self) -> int:
        return self.regime()["run_mode"]

    def CheckSolutionMode_DGV(self, f) -> np.ndarray:
        """spat_oper_mod.f90:195-272: regime switch; returns Df = f - fm
        (zeros when the solution is in run_mode 0)."""
        f = self._f(f)
        df = np.zeros(self.nnodes)
        self._chk(self._L.dgv_check_solution_mode(self._ctx, dptr(f), dptr(df), self.nnodes))
        return df

    def SpatialOperSH_DGV(self, f, dt: float) -> np.ndarray:
        """spat_oper_mod.f90:28-68: run_mode 0, 1 or 2 after the regime check
        when it is on (set_regime)."""
        f = self._f(f)
        fcol = np.zeros(self.nnodes)
        self._chk(self._L.dgv_spatial_oper(self._ctx, dptr(f), dptr(fcol), self.nnodes, dt,
                                           self.coef))
        return fcol

    def TimeIntegratorORK_SH_DGV(self, f: np.ndarray, dt: float) -> np.ndarray:
        """time_integr_mod.f90:26-152: advances f in place (f must be float64,
        contiguous), leaves frhs1 = first stage / dt."""
        if f.dtype != np.float64 or not f.flags.c_contiguous or f.shape != (self.nnodes,):
            raise DGVError(-1, "f must be a contiguous float64 array of length nnodes")
        self._chk(self._L.dgv_rk_step(self._ctx, dptr(f), self.nnodes, dt, self.coef,
                                      self.rkmts, dptr(self.frhs1)))
        return f

    def TimeIntegratorMTS_SH_DGV(self, f: np.ndarray, dt: float) -> np.ndarray:
        """time_integr_mod.f90:314-364 (Adams-Bashforth rkmts), advances e_time."""
        if f.dtype != np.float64 or not f.flags.c_contiguous or f.shape != (self.nnodes,):
            raise DGVError(-1, "f must be a contiguous float64 array of length nnodes")
        self.hist[4] = self.frhs1
        self._chk(self._L.dgv_ab_step(self._ctx, dptr(f), self.nnodes, dt, self.coef, self.rkmts,
                                      dptr(self.hist)))
        self.frhs1 = self.hist[4].copy()
        self.e_time += dt
        return f

    def PrepareMTS_SH_DGV(self, f: np.ndarray, dt: float) -> np.ndarray:
        """time_integr_mod.f90:443-569: rkmts-1 Runge-Kutta steps of order rkmts,
        pushing f and frhs1 down the multistep history before each (lines 536-548)."""
        k = self.rkmts
        h = self.hist  # f1..f4 = h[0..3], frhs1..frhs5 = h[4..8]
        for _ in range(2, k + 1):
            h[4] = self.frhs1
            if k == 2:
                h[0] = f
            elif k == 3:
                h[1] = h[0]; h[0] = f; h[5] = h[4]
            elif k == 4:
                h[2] = h[1]; h[1] = h[0]; h[0] = f; h[6] = h[5]; h[5] = h[4]
            elif k == 5:
                h[3] = h[2]; h[2] = h[1]; h[1] = h[0]; h[0] = f
                h[7] = h[6]; h[6] = h[5]; h[5] = h[4]
            self.TimeIntegratorORK_SH_DGV(f, dt)
            h[4] = self.frhs1
            self.e_time += dt
        return f

    def MassCheckRec(self, f):
        """dgvtools_mod.f90:1437-1458 -> (n, ubar, vbar, wbar, temp)"""
        f = self._f(f)
        out = np.zeros(5)
        self._chk(self._L.dgv_moments(self._ctx, dptr(f), self.nnodes, dptr(out)))
        return tuple(out)

    def MassCheckRecHighCMoments(self, f):
        """dgvtools_mod.f90:1501-1549 -> (n, ubar, vbar, wbar, temp, temp_u, temp_v,
        temp_w, mom3_u, mom3_v, mom3_w, mom4_*, mom5_*, mom6_*)"""
        f = self._f(f)
        out = np.zeros(20)
        self._chk(self._L.dgv_moments_high(self._ctx, dptr(f), self.nnodes, dptr(out)))
        return tuple(out)

    def RecordMomDerivRelaxTimes_DGV(self, log, curr_time: float) -> bool:
        """dgvtools_mod.f90:1589-2023 into a driver_io.MomDerivLog: the moments'
        derivatives and relaxation times from the stepper history, f1 (the
        solution of the previous step) and frhs1 (its right-hand side), as the
        driver calls it after each multistep step (DGVblzmPL.f90:330)."""
        return log.record(curr_time, self.hist[0], self.frhs1, self)

    def WriteSol_DGV(self, path, f, dt: float):
        """readwrite.f90:1756-1796: e_time, dt, rkmts, f and the multistep
        history (hist, with frhs1) to a Fortran-unformatted file."""
        self.hist[4] = self.frhs1
        write_sol_file(path, self.e_time, dt, self.rkmts, self._f(f), self.hist)

    def WriteSol_init_DGV(self, path, f, dt: float):
        """readwrite.f90:1810-1850: e_time, dt, rkmts and f alone (the driver's
        initial-data file)."""
        write_sol_init_file(path, self.e_time, dt, self.rkmts, self._f(f))

    def RestartSolutionDGV(self, path):
        """readwrite.f90:2389-2513: restores e_time, rkmts and the history;
        returns (f, dt)."""
        e, dt, k, f, hist = read_sol_file(path, self.nnodes)
        self.e_time, self.rkmts = e, k
        self.hist = hist
        self.frhs1 = hist[4].copy()
        return f, dt

    # -- additions --------------------------------------------------------------------
    def collision_moments(self, q):
        """sum_i w_i psi_k(v_i) q_i for psi = 1, u, v, w, |u|^2"""
        q = self._f(q)
        out = np.zeros(5)
        self._chk(self._L.dgv_collision_moments(self._ctx, dptr(q), self.nnodes, dptr(out)))
        return out

    def project_invariants(self, q) -> np.ndarray:
        q = self._f(q).copy()
        self._chk(self._L.dgv_project_invariants(self._ctx, dptr(q), self.nnodes))
        return q

    # -- device-resident use (bench, graphs) -----------------------------------------
    def device_alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self._chk(self._L.dgv_device_alloc(self._ctx, nbytes, C.byref(p)))
        return p.value

    def device_free(self, ptr: int):
        self._chk(self._L.dgv_device_free(self._ctx, C.c_void_p(int(ptr))))

    def htod(self, dst: int, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self._chk(self._L.dgv_memcpy_htod(self._ctx, C.c_void_p(int(dst)), a.ctypes.data_as(C.c_void_p),
                                          a.nbytes))

    def dtoh(self, a: np.ndarray, src: int):
        self._chk(self._L.dgv_memcpy_dtoh(self._ctx, a.ctypes.data_as(C.c_void_p), C.c_void_p(int(src)),
                                          a.nbytes))

    def device_copy(self, dst: int, src: int, nbytes: int):
        self._chk(self._L.dgv_device_copy(self._ctx, C.c_void_p(int(dst)), C.c_void_p(int(src)), nbytes))

    def eval_q_dev(self, d_f: int, d_q: int):
        self._chk(self._L.dgv_eval_q_dev(self._ctx, C.c_void_p(int(d_f)), C.c_void_p(int(d_q)), self.nnodes))

    def rk_run_dev(self, d_f: int, dt: float, nsteps: int, rkmts: Optional[int] = None,
                   use_graph: bool = True):
        self._chk(self._L.dgv_rk_run_dev(self._ctx, C.c_void_p(int(d_f)), self.nnodes, dt, self.coef,
                                         rkmts or self.rkmts, nsteps, int(use_graph)))

    def synchronize(self):
        self._chk(self._L.dgv_ctx_synchronize(self._ctx))

    def time_eval_q_dev(self, d_f: int, d_q: int, niter: int):
        a = C.c_double()
        b = C.c_double()
        self._chk(self._L.dgv_time_eval_q_dev(self._ctx, C.c_void_p(int(d_f)), C.c_void_p(int(d_q)), niter,
                                              C.byref(a), C.byref(b)))
        return a.value, b.value


class DGVBatch:
    """Many VDFs on one or several grids (each a ``DGV`` with its A-array),
    evaluated together in one persistent launch.  precision: DGV_FP64 or
    DGV_FP32_COMP (FP32 products with two-sum compensated sums, FP64 final sum)."""

    def __init__(self, grids, grid_of_instance, precision=_lib.DGV_FP64):
        self._L = load()
        self.grids = list(grids)
        self.grid_of_instance = np.ascontiguousarray(grid_of_instance, np.int32)
        arr = (C.c_void_p * len(self.grids))(*[g.ctx for g in self.grids])
        p = C.c_void_p()
        check(self._L.dgv_batch_create(arr, len(self.grids), iptr(self.grid_of_instance),
                                       len(self.grid_of_instance), precision, C.byref(p)))
        self._b = p
        for g in self.grids:
            g._register_dependent(self)
        self.ninstances = len(self.grid_of_instance)

    def close(self):
        if getattr(self, "_b", None):
            self._L.dgv_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        check(rc, None)

    def set_f(self, i: int, f):
        f = np.ascontiguousarray(f, np.float64)
        self._chk(self._L.dgv_batch_set_f(self._b, i, dptr(f), len(f)))

    def get_q(self, i: int) -> np.ndarray:
        n = self.grids[self.grid_of_instance[i]].nnodes
        q = np.zeros(n)
        self._chk(self._L.dgv_batch_get_q(self._b, i, dptr(q), n))
        return q

    def get_f(self, i: int) -> np.ndarray:
        """The instance's current f (after stepping)."""
        off, stride, total = C.c_int64(), C.c_int64(), C.c_int64()
        self._chk(self._L.dgv_batch_layout(self._b, i, C.byref(off), C.byref(stride), C.byref(total)))
        df, dq = C.c_void_p(), C.c_void_p()
        self._chk(self._L.dgv_batch_device_buffers(self._b, C.byref(df), C.byref(dq)))
        g = self.grids[self.grid_of_instance[i]]
        allf = np.zeros(total.value)
        g.dtoh(allf, df.value)
        return allf[off.value::stride.value][:g.nnodes].copy()

    def set_f_rows(self, d_rows: int, first: int, count: int, ld: Optional[int] = None):
        """Instances first .. first+count-1 (one grid) from device rows."""
        n = self.grids[self.grid_of_instance[first]].nnodes
        self._chk(self._L.dgv_batch_set_f_rows_dev(self._b, C.c_void_p(int(d_rows)), ld or n, first, count))

    def get_q_rows(self, d_rows: int, first: int, count: int, ld: Optional[int] = None):
        n = self.grids[self.grid_of_instance[first]].nnodes
        self._chk(self._L.dgv_batch_get_q_rows_dev(self._b, C.c_void_p(int(d_rows)), ld or n, first, count))

    def get_f_rows(self, d_rows: int, first: int, count: int, ld: Optional[int] = None):
        n = self.grids[self.grid_of_instance[first]].nnodes
        self._chk(self._L.dgv_batch_get_f_rows_dev(self._b, C.c_void_p(int(d_rows)), ld or n, first, count))

    def eval_q(self):
        self._chk(self._L.dgv_batch_eval_q(self._b))

    def rk_run(self, dt: float, nsteps: int, rkmts: int = 2, coef: float = 1.0, use_graph=True):
        self._chk(self._L.dgv_batch_rk_run(self._b, dt, coef, rkmts, nsteps, int(use_graph)))

    def project_invariants(self):
        """FP64 projection of every instance's Q onto the complement of the
        collision invariants (in place on the device)."""
        self._chk(self._L.dgv_batch_project_invariants(self._b))

    def set_probe_cells(self, cells):
        """Cells (0-based) evaluated by ``eval_q_probe``."""
        c = np.ascontiguousarray(cells, np.int32)
        self._chk(self._L.dgv_batch_set_probe_cells(self._b, iptr(c), len(c)))

    def eval_q_probe(self):
        """Q of every instance on the probe cells only."""
        self._chk(self._L.dgv_batch_eval_q_probe(self._b))

    def autotune(self):
        best = C.c_int32()
        ms = C.c_double()
        self._chk(self._L.dgv_batch_autotune(self._b, C.byref(best), C.byref(ms)))
        return best.value, ms.value

    def set_tile(self, pairs_per_segment: int, waves_per_cu: int = 0):
        self._chk(self._L.dgv_batch_set_tile(self._b, pairs_per_segment, waves_per_cu))

    def set_projection(self, on: bool):
        """Project every evaluation's Q onto the complement of the collision
        invariants (FP64), also inside the RK stages."""
        self._chk(self._L.dgv_batch_set_projection(self._b, int(on)))

    def time_eval(self, niter: int = 3, main_kernel: bool = False):
        """Milliseconds per batched evaluation (HIP events on the batch's
        stream); with main_kernel, also those of the persistent kernel alone."""
        a = C.c_double()
        m = C.c_double()
        self._chk(self._L.dgv_batch_time_eval(self._b, niter, C.byref(a), C.byref(m) if main_kernel else None))
        return (a.value, m.value) if main_kernel else a.value

    def device_buffers(self):
        """(device pointer of packed f, of packed q, offset of instance 0, stride)."""
        df, dq = C.c_void_p(), C.c_void_p()
        self._chk(self._L.dgv_batch_device_buffers(self._b, C.byref(df), C.byref(dq)))
        off, stride, total = C.c_int64(), C.c_int64(), C.c_int64()
        self._chk(self._L.dgv_batch_layout(self._b, 0, C.byref(off), C.byref(stride), C.byref(total)))
        return df.value, dq.value, off.value, stride.value
