"""Text logs of the reference's driver, host side.

``MacroLog``  WriteErrHighCMoments_DGV (readwrite.f90:2025-2060): the moment
              history of MassCheckRecHighCMoments, ``<base>_macroerr.txt``.
``L1Track``   KeepTrackL1_err_DGV (readwrite.f90:2071-2128): the relative L1
              deviation from the local Maxwellian over time, sampled at most
              1000 times on [t_L, t_R] and rewritten num_save_solution times,
              ``<base>_L1err.txt``.
``MomDerivLog`` RecordMomDerivRelaxTimes_DGV (dgvtools_mod.f90:1589-2023) and
              WriteMomDerivRelaxTimes_DGV (readwrite.f90:2528-2609): time
              derivatives of the 20 moments from the last right-hand side and
              the relaxation times they imply, ``<base>_dermom.txt`` and
              ``<base>_relaxtimes.txt``.

Numbers are written with the reference's edit descriptors (F18.14 for fields
1, 10 and 16 as its format lists them, F18.15 for the others, E18.6 in the
derivative files, separated by A2 ","); a value too wide for its field prints
as asterisks, as Fortran does.
"""
from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np


def fortran_f(x: float, w: int, d: int) -> str:
    """Fortran Fw.d output of one value."""
    s = f"{x:{w}.{d}f}"
    return "*" * w if len(s) > w else s


def fortran_e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d output of one value: 0.ddd...E+ee (the exponent letter is
    dropped for exponents of three digits), right-justified in w."""
    x = float(x)
    if np.isnan(x):
        s = "NaN"
    elif np.isinf(x):
        s = "-Infinity" if x < 0 else "Infinity"
    else:
        if x == 0.0:
            digits, e = "0" * d, 0
        else:
            m, ex = f"{abs(x):.{d - 1}e}".split("e")
            digits, e = m.replace(".", ""), int(ex) + 1
        es = f"E{e:+03d}" if abs(e) <= 99 else f"{e:+04d}"
        s = ("-" if x < 0 else "") + "0." + digits + es
    return "*" * w if len(s) > w else s.rjust(w)


def list_directed(text: str) -> str:
    """A list-directed WRITE (*) of character items (concatenated in ``text``) as
    the LLVM Flang runtime lays it out: records of at most 79 columns, each a
    blank and the next 78 characters (the reference's own readwrite.f90
    routines, built with amdflang, write their headers so)."""
    lines = []
    rest = text
    while True:
        lines.append(" " + rest[:78])
        rest = rest[78:]
        if not rest:
            return "\n".join(lines)


# field formats of WriteErrHighCMoments_DGV: time, then 20 moments
_MACRO_D = [14] + [15] * 8 + [14] + [15] * 5 + [14] + [15] * 5
_MACRO_HEAD = (list_directed("This is the file containing the values of density, average velocity and "
                             "temperature, and the time stamp:"),
               list_directed("Time , Ndens , Ubar , Vbar , Wbar , T , Tx , Ty , Tz , (u-ubar)^3 ,(v-vbar)^3 , "
                             "(w-wbar)^3 , (u-ubar)^4 , (v-vbar)^4 , (w-wbar)^4, (u-ubar)^5 , (v-vbar)^5 , "
                             "(w-wbar)^5 , (u-ubar)^6 , (v-vbar)^6 , (w-wbar)^6"))


def macro_line(time: float, moments: Sequence[float]) -> str:
    vals = [time] + list(moments)
    if len(vals) != 21:
        raise ValueError("21 fields: the time and the 20 values of MassCheckRecHighCMoments")
    return " ,".join(fortran_f(v, 18, d) for v, d in zip(vals, _MACRO_D))


class MacroLog:
    """Rows of (time, n, ubar, vbar, wbar, T, Tx, Ty, Tz, mom3..6 per direction),
    at most 999 as in the driver (err_count < 1000)."""

    capacity = 999

    def __init__(self, path: str):
        self.path = path
        self.rows: List[tuple] = []

    def record(self, time: float, moments: Sequence[float]) -> bool:
        """the 20 values of MassCheckRecHighCMoments; False when full."""
        if len(self.rows) >= self.capacity:
            return False
        self.rows.append((time, tuple(moments)))
        return True

    def write(self) -> None:
        with open(self.path, "w") as fh:
            for h in _MACRO_HEAD:
                fh.write(h + "\n")
            for t, m in self.rows:
                fh.write(macro_line(t, m) + "\n")


class L1Track:
    """KeepTrackL1_err_DGV's schedule: the first call records and sets the
    periods; later calls record when e_time reaches the next evaluation time
    (at most 1000 entries) and rewrite the file when it reaches the next
    record time."""

    capacity = 1000

    def __init__(self, path: str, t_L: float, t_R: float, num_save_solution: int):
        self.path = path
        self.t_L, self.t_R, self.nsave = t_L, t_R, num_save_solution
        self.t: List[float] = []
        self.a: List[float] = []
        self.next_eval = self.next_record = None

    def __call__(self, l1_err: float, e_time: float) -> None:
        if self.next_eval is None:
            self.t, self.a = [e_time], [l1_err]
            self.record_period = (self.t_R - self.t_L) / float(self.nsave)
            self.next_record = e_time + self.record_period
            self.eval_period = (self.t_R - self.t_L) / float(self.capacity)
            self.next_eval = e_time + self.eval_period
        if e_time >= self.next_eval and len(self.a) < self.capacity:
            self.next_eval = e_time + self.eval_period
            self.t.append(e_time)
            self.a.append(l1_err)
        if e_time >= self.next_record:
            self.next_record = e_time + self.record_period
            self.write()

    def write(self) -> None:
        with open(self.path, "w") as fh:
            fh.write(list_directed("This file contains the values of relative L1 error of deviation from local "
                                   "maxwellian and the time stamp:") + "\n")
            fh.write(list_directed("Time\t\tL1_err") + "\n")
            for t, a in zip(self.t, self.a):
                fh.write(fortran_f(t, 18, 14) + " ," + fortran_f(a, 18, 15) + "\n")


_DERIV_HEAD = ("This file contains values of time derivatives of density, average velocity and temperature, "
               "and the time stamp:",
               "This file contains values of relaxation times for density, average velocity and temperature, "
               "and the time stamp:")


def moment_derivatives(f1: np.ndarray, frhs1: np.ndarray, u: np.ndarray, v: np.ndarray, w: np.ndarray,
                       gw: np.ndarray) -> np.ndarray:
    """The 40 values that RecordMomDerivRelaxTimes_DGV records after the time
    stamp (dgvtools_mod.f90:1635-1963): dn, dubar, dvbar, dwbar, dT, dT_u,
    dT_v, dT_w, d(mom3..mom6) per direction from f1 and its right-hand side
    frhs1, then the relaxation times tau = -d/m of the same 20 quantities (dn/m
    for the density), where m is the moment's distance from its value for the
    local Maxwellian; tau = 0 where |m| <= 0.01 |dn|.  The formulas are the
    reference's, including its temperature derivative, whose second term sums
    f1 * 2 (u - ubar) . dubar without the quadrature weights."""
    f1, frhs1 = np.asarray(f1, np.float64), np.asarray(frhs1, np.float64)
    n = np.sum(f1 * gw)
    ub, vb, wb = np.sum(f1 * gw * u) / n, np.sum(f1 * gw * v) / n, np.sum(f1 * gw * w) / n
    cu, cv, cw = u - ub, v - vb, w - wb
    c2 = cu ** 2 + cv ** 2 + cw ** 2
    temp = np.sum(f1 * gw * c2) / n / 3.0 * 2.0
    dn = np.sum(frhs1 * gw)
    du = (np.sum(frhs1 * gw * u) - dn * ub) / n
    dv = (np.sum(frhs1 * gw * v) - dn * vb) / n
    dw = (np.sum(frhs1 * gw * w) - dn * wb) / n
    dtemp = (np.sum(frhs1 * gw * c2 - f1 * 2.0 * (cu * du + cv * dv + cw * dw)) / 3.0 * 2.0 - dn * temp) / n
    d = [dn, du, dv, dw, dtemp]
    d += [np.sum(frhs1 * gw * c ** 2) / 3.0 * 2.0 / n for c in (cu, cv, cw)]
    for k in (3, 4, 5, 6):
        d += [np.sum(frhs1 * gw * c ** k) / n for c in (cu, cv, cw)]
    fm = n * np.exp(-c2 / temp) / (np.pi * temp) ** 1.5   # maxwelveldist, distributions_mod.f90:83-94
    g = f1 - fm
    m = [np.sum(g * gw), np.sum(g * gw * u) / n, np.sum(g * gw * v) / n, np.sum(g * gw * w) / n,
         np.sum(g * gw * c2) / n / 3.0 * 2.0]
    m += [np.sum(f1 * gw * c ** 2) / n / 3.0 * 2.0 - temp / 3.0 for c in (cu, cv, cw)]
    gauss = {3: 0.0, 4: 3.0 * temp ** 2 / 4.0, 5: 0.0, 6: 15.0 * temp ** 3 / 8.0}
    for k in (3, 4, 5, 6):
        m += [np.sum(f1 * gw * c ** k) / n - gauss[k] for c in (cu, cv, cw)]
    tresh = abs(dn * 0.01)
    tau = [(d[0] / m[0]) if abs(m[0]) > tresh else 0.0]
    tau += [(-dk / mk) if abs(mk) > tresh else 0.0 for dk, mk in zip(d[1:], m[1:])]
    return np.array(d + tau, np.float64)


class MomDerivLog:
    """RecordMomDerivRelaxTimes_DGV's store and WriteMomDerivRelaxTimes_DGV.
    Room for num_eval_error + 3 + rkmts + 1 records, as the reference
    allocates; a record past that is dropped with the reference's message, and
    a record at a time past t_R releases the store (the reference deallocates
    it), after which writing is an error until the next record."""

    def __init__(self, base: str, num_eval_error: int, rkmts: int, t_R: float):
        self.base = base
        self.capacity = int(num_eval_error) + 3 + int(rkmts) + 1
        self.t_R = float(t_R)
        self.rows: Optional[List[np.ndarray]] = None

    def record(self, curr_time: float, f1, frhs1, grid) -> bool:
        """grid: anything with nodes_u, nodes_v, nodes_w and nodes_gwts."""
        if self.rows is None:
            self.rows = []
        stored = len(self.rows) < self.capacity
        if stored:
            vals = moment_derivatives(f1, frhs1, grid.nodes_u, grid.nodes_v, grid.nodes_w, grid.nodes_gwts)
            self.rows.append(np.concatenate([[curr_time], vals]))
        else:
            print("RecordMomDerivRelaxTimes_DGV: Not enough memory in (MomDerivRelaxT) to store moments derivs. "
                  "and relax. times.")
        if curr_time > self.t_R:
            self.rows = None
        return stored

    def write(self) -> None:
        if self.rows is None:
            raise RuntimeError("WriteMomDerivRelaxTimes_DGV: the flag for array MomDerivRelaxT indicates array "
                               "does not exist. stop")
        for suffix, head, cols in (("_dermom.txt", _DERIV_HEAD[0], range(0, 21)),
                                   ("_relaxtimes.txt", _DERIV_HEAD[1], [0] + list(range(21, 41)))):
            with open(self.base + suffix, "w") as fh:
                fh.write(list_directed(head) + "\n")
                fh.write(_MACRO_HEAD[1] + "\n")
                for r in self.rows:
                    fh.write(" ,".join(fortran_e(r[c], 18, 6) for c in cols) + "\n")
