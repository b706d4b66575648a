"""Learned surrogate of Q(f) on one grid (libdgv_amd.so, ``dgv_sur_*``).

An MLP widths[0] -> ... -> widths[-1] with GELU between layers maps
f / f_scale to Q(f) / q_scale; BF16 weights and activations on the matrix
cores with FP32 accumulation, FP32 master weights and Adam.  Its output is
projected in FP64 onto the complement of the collision invariants
(1, u, v, w, |u|^2), and ``eval`` uses the direct evaluator on a set of probe
cells as judge: rows whose probe residual exceeds ``tol`` get the direct Q.

Rows are VDFs in reference node order; device arrays hold ``nrows`` rows of
``ld`` doubles (``ld`` = number of nodes unless given).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from ._lib import DGVError, check, fptr, iptr, load
from .collision import DGV


class Surrogate:
    def __init__(self, dgv: DGV, widths: Sequence[int], batch: int = 256, seed: int = 0):
        self._L = load()
        self.dgv = dgv
        self.widths = np.ascontiguousarray(widths, np.int32)
        if self.widths[0] != dgv.nnodes or self.widths[-1] != dgv.nnodes:
            raise DGVError(-1, "first and last widths must equal the number of nodes")
        self.nlayers = len(self.widths) - 1
        self.batch = batch
        p = C.c_void_p()
        check(self._L.dgv_sur_create(dgv.ctx, iptr(self.widths), self.nlayers, batch, seed, C.byref(p)), dgv.ctx)
        self._s = p
        dgv._register_dependent(self)
        self.f_scale = 1.0
        self.q_scale = 1.0

    def close(self):
        if getattr(self, "_s", None):
            self._L.dgv_sur_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        check(rc, self.dgv.ctx)

    def set_scales(self, f_scale: float, q_scale: float):
        self._chk(self._L.dgv_sur_set_scales(self._s, f_scale, q_scale))
        self.f_scale, self.q_scale = f_scale, q_scale

    # -- parameters ------------------------------------------------------------------
    def get_layer(self, l: int):
        W = np.zeros((self.widths[l + 1], self.widths[l]), np.float32)
        b = np.zeros(self.widths[l + 1], np.float32)
        self._chk(self._L.dgv_sur_get_layer(self._s, l, fptr(W), fptr(b)))
        return W, b

    def set_layer(self, l: int, W, b):
        W = np.ascontiguousarray(W, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        if W.shape != (self.widths[l + 1], self.widths[l]) or b.shape != (self.widths[l + 1],):
            raise DGVError(-1, "layer shape mismatch")
        self._chk(self._L.dgv_sur_set_layer(self._s, l, fptr(W), fptr(b)))

    def get_grad(self, l: int):
        W = np.zeros((self.widths[l + 1], self.widths[l]), np.float32)
        b = np.zeros(self.widths[l + 1], np.float32)
        self._chk(self._L.dgv_sur_get_grad(self._s, l, fptr(W), fptr(b)))
        return W, b

    # -- device-pointer API ---------------------------------------------------------
    def train_step(self, d_f: int, d_q: int, nrows: int, lr: float, ld: Optional[int] = None) -> float:
        loss = C.c_double()
        self._chk(self._L.dgv_sur_train_step(self._s, C.c_void_p(int(d_f)), C.c_void_p(int(d_q)), ld or self.dgv.nnodes,
                                             nrows, lr, C.byref(loss)))
        return loss.value

    def predict(self, d_f: int, nrows: int, d_q: int, project: bool = True, ld: Optional[int] = None):
        self._chk(self._L.dgv_sur_predict(self._s, C.c_void_p(int(d_f)), ld or self.dgv.nnodes, nrows,
                                          C.c_void_p(int(d_q)), int(project)))

    def set_decomposition(self, on: bool = True):
        """Decomposition mode: the network maps Df = f - fm to Q(Df) and the
        surrogate's Q(f) is L(Df) + net(Df), with fm and L the context's
        prepared linearised operator (PrepareFMA_DGV); the reference's run_mode 1
        split (spat_oper_mod.f90:28-68).  Train with ``decomposition_rows``."""
        self._chk(self._L.dgv_sur_set_decomposition(self._s, int(bool(on))))

    def set_probe_cells(self, cells):
        c = np.ascontiguousarray(cells, np.int32)
        self._chk(self._L.dgv_sur_set_probe_cells(self._s, iptr(c), len(c)))

    def eval(self, d_f: int, nrows: int, d_q: int, tol: float, ld: Optional[int] = None):
        """Surrogate Q with the direct evaluator as judge and fallback; returns
        (residual per row, fallback flag per row, number of fallbacks)."""
        res = np.zeros(nrows)
        fb = np.zeros(nrows, np.int32)
        nfb = C.c_int64()
        self._chk(self._L.dgv_sur_eval(self._s, C.c_void_p(int(d_f)), ld or self.dgv.nnodes, nrows, C.c_void_p(int(d_q)),
                                       tol, res.ctypes.data_as(C.POINTER(C.c_double)), iptr(fb), C.byref(nfb)))
        return res, fb.astype(bool), nfb.value

    def time_predict(self, d_f: int, nrows: int, d_q: int, niter: int = 5, ld: Optional[int] = None) -> float:
        ms = C.c_double()
        self._chk(self._L.dgv_sur_time_predict(self._s, C.c_void_p(int(d_f)), ld or self.dgv.nnodes, nrows,
                                               C.c_void_p(int(d_q)), niter, C.byref(ms)))
        return ms.value

    # -- host-array conveniences (small sizes, tests) --------------------------------
    def _to_device(self, rows: np.ndarray) -> int:
        rows = np.ascontiguousarray(rows, np.float64)
        p = self.dgv.device_alloc(rows.nbytes)
        self.dgv.htod(p, rows)
        return p

    def train_step_host(self, f_rows, q_rows, lr: float) -> float:
        df, dq = self._to_device(f_rows), self._to_device(q_rows)
        try:
            return self.train_step(df, dq, len(f_rows), lr)
        finally:
            self.dgv.device_free(df)
            self.dgv.device_free(dq)

    def predict_host(self, f_rows, project: bool = True) -> np.ndarray:
        f_rows = np.ascontiguousarray(f_rows, np.float64)
        out = np.zeros_like(f_rows)
        df, dq = self._to_device(f_rows), self._to_device(out)
        try:
            self.predict(df, len(f_rows), dq, project)
            self.dgv.dtoh(out, dq)
        finally:
            self.dgv.device_free(df)
            self.dgv.device_free(dq)
        return out

    def eval_host(self, f_rows, tol: float):
        f_rows = np.ascontiguousarray(f_rows, np.float64)
        out = np.zeros_like(f_rows)
        df, dq = self._to_device(f_rows), self._to_device(out)
        try:
            res, fb, nfb = self.eval(df, len(f_rows), dq, tol)
            self.dgv.dtoh(out, dq)
        finally:
            self.dgv.device_free(df)
            self.dgv.device_free(dq)
        return out, res, fb, nfb


def decomposition_rows(dgv: DGV, f_rows, q_rows=None, chunk: int = 256):
    """Training rows of the decomposition mode, as host arrays.  With Df = f - fm
    and s the root mean square of a row's Df: X = Df / s and, when q_rows (Q(f))
    is given, Y = (Q(f) - Q(fm) - L(Df)) / s^2, the quadratic part Q(Df) / s^2;
    fm and L = 2 MATMUL(., fmA) are the context's prepared linearised operator
    (collision_mod.f90:346-359).  Returns (X, Y, s); Y is None without q_rows."""
    f_rows = np.asarray(f_rows, np.float64)
    fm = dgv.fm
    Df = np.ascontiguousarray(f_rows - fm[None, :])
    rs = np.sqrt(np.mean(Df * Df, axis=1))
    inv = np.where(rs > 0, 1.0 / np.where(rs > 0, rs, 1.0), 0.0)
    X = Df * inv[:, None]
    if q_rows is None:
        return X, None, rs
    N = dgv.nnodes
    k = max(1, min(chunk, len(Df)))
    Lr = np.zeros_like(Df)
    d_in = dgv.device_alloc(8 * N * k)
    d_out = dgv.device_alloc(8 * N * k)
    try:
        tmp = np.zeros((k, N))
        for r0 in range(0, len(Df), k):
            n = min(k, len(Df) - r0)
            dgv.htod(d_in, np.ascontiguousarray(Df[r0:r0 + n]))
            dgv.eval_linear_dev(d_in, d_out, nvec=n)
            dgv.dtoh(tmp[:n], d_out)
            Lr[r0:r0 + n] = tmp[:n]
    finally:
        dgv.device_free(d_in)
        dgv.device_free(d_out)
    qfm = dgv.EvalCollisionPeriodicAPlus_DGV(fm)
    Y = (np.asarray(q_rows, np.float64) - qfm[None, :] - Lr) * (inv * inv)[:, None]
    return X, Y, rs
