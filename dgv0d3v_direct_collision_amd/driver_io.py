"""Text logs of the reference's driver, host side.

``MacroLog``  WriteErrHighCMoments_DGV (readwrite.f90:2025-2060): the moment
              history of MassCheckRecHighCMoments, ``<base>_macroerr.txt``.
``L1Track``   KeepTrackL1_err_DGV (readwrite.f90:2071-2128): the relative L1
              deviation from the local Maxwellian over time, sampled at most
              1000 times on [t_L, t_R] and rewritten num_save_solution times,
              ``<base>_L1err.txt``.

Numbers are written with the reference's edit descriptors (F18.14 for fields
1, 10 and 16 as its format lists them, F18.15 for the others, separated by
A2 ","); a value too wide for its field prints as asterisks, as Fortran does.
"""
from __future__ import annotations

from typing import List, Sequence


def fortran_f(x: float, w: int, d: int) -> str:
    """Fortran Fw.d output of one value."""
    s = f"{x:{w}.{d}f}"
    return "*" * w if len(s) > w else s


# field formats of WriteErrHighCMoments_DGV: time, then 20 moments
_MACRO_D = [14] + [15] * 8 + [14] + [15] * 5 + [14] + [15] * 5
_MACRO_HEAD = (" This is the file containing the values of density, average velocity and temperature, "
               "and the time stamp:",
               " Time , Ndens , Ubar , Vbar , Wbar , T , Tx , Ty , Tz , (u-ubar)^3 ,(v-vbar)^3 , (w-wbar)^3 , "
               "(u-ubar)^4 , (v-vbar)^4 , (w-wbar)^4, (u-ubar)^5 , (v-vbar)^5 , (w-wbar)^5 , (u-ubar)^6 , "
               "(v-vbar)^6 , (w-wbar)^6")


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
            fh.write(" This file contains the values of relative L1 error of deviation from local maxwellian "
                     "and the time stamp:\n")
            fh.write(" Time\t\tL1_err\n")
            for t, a in zip(self.t, self.a):
                fh.write(fortran_f(t, 18, 14) + " ," + fortran_f(a, 18, 15) + "\n")
