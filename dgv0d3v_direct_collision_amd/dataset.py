"""Training data for the learned surrogate: (f, Q(f)) pairs streamed to disk.

A set is a directory:

  meta.json   format, grid (M, s, box), node count, A-array fingerprint and
              workload, evaluator precision, seed base, VDF family, row count
  f.npy       float64 (rows, N)  VDFs, reference node order
  q.npy       float64 (rows, N)  Q(f) (EvalCollisionPeriodicAPlus_DGV)
  seeds.npy   uint64  (rows, 2)  (seed base, row index): the row's generator seed

Rows are produced in chunks.  While the batched evaluator works on one chunk
on the GPU, a host thread draws the next chunk's VDFs, and the previous
chunk's rows go to the memory-mapped files.  Every f can be drawn again from
its seed (``vdf_from_seed``), and ``verify`` re-evaluates sampled rows with
the single-VDF FP64 evaluator, so a set can be checked against the A-array
it claims to come from.
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

FORMAT = 1


def _maxwellian(u, v, w, n, ub, T):
    return n * np.exp(-((u - ub[0]) ** 2 + (v - ub[1]) ** 2 + (w - ub[2]) ** 2) / T) / (np.pi * T) ** 1.5


FAMILIES = ("mixtures", "perturbed_maxwellian")


def equilibrium(grid, T: float) -> np.ndarray:
    """The perturbed-Maxwellian family's fm: zero bulk velocity, temperature T,
    unit discrete density."""
    u, v, w, gw = grid.nodes_u, grid.nodes_v, grid.nodes_w, grid.nodes_gwts
    fm = _maxwellian(u, v, w, 1.0, (0.0, 0.0, 0.0), T)
    return fm / np.sum(gw * fm)


def vdf_from_seed(grid, seed_base: int, row: int, family: str = "mixtures", T: float = 0.8,
                  amplitude=(1e-3, 1e-1)) -> np.ndarray:
    """The row's VDF.  Family "mixtures", normalised to unit density: even rows
    1-4 Maxwellians (density U[0.2,1], bulk velocity U[-1,1]^3, temperature
    U[0.3,1]); odd rows two Maxwellians plus a non-negative 5 % perturbation on a
    random tenth of the cells.  Family "perturbed_maxwellian" (the decomposition
    regime): equilibrium(grid, T) plus an invariant-free polynomial perturbation
    of log-uniform amplitude in the given range (workloads.perturbed_maxwellian_vdf)."""
    rng = np.random.default_rng([seed_base, row])
    u, v, w, gw = grid.nodes_u, grid.nodes_v, grid.nodes_w, grid.nodes_gwts
    if family == "perturbed_maxwellian":
        from .workloads import perturbed_maxwellian_vdf
        a = float(np.exp(rng.uniform(np.log(amplitude[0]), np.log(amplitude[1]))))
        return perturbed_maxwellian_vdf(u, v, w, gw, equilibrium(grid, T), T, rng, a)
    if family != "mixtures":
        raise ValueError(f"unknown VDF family {family!r}")
    if row % 2 == 0:
        f = np.zeros(len(u))
        for _ in range(int(rng.integers(1, 5))):
            f += _maxwellian(u, v, w, rng.uniform(0.2, 1.0), rng.uniform(-1, 1, 3), rng.uniform(0.3, 1.0))
    else:
        n1 = rng.uniform(0.3, 0.7)
        f = (_maxwellian(u, v, w, n1, rng.uniform(-1, 1, 3), rng.uniform(0.3, 1.0)) +
             _maxwellian(u, v, w, 1 - n1, rng.uniform(-1, 1, 3), rng.uniform(0.3, 1.0)))
        D = grid.D
        ncell = len(f) // D
        fmax = f.max()
        for c in rng.choice(ncell, size=max(1, ncell // 10), replace=False):
            f[c * D:(c + 1) * D] += 0.05 * rng.uniform(0, 1, D) * fmax
    return f / np.sum(gw * f)


def aarray_fingerprint(grid) -> str:
    """SHA-256 over the grid description, the A-array's record and pair counts
    and Q of a fixed VDF (rounded to 10 significant digits, so that the
    evaluator's tiling does not change it)."""
    u, v, w = grid.nodes_u, grid.nodes_v, grid.nodes_w
    f = _maxwellian(u, v, w, 1.0, (0.31, -0.17, 0.07), 0.8) + _maxwellian(u, v, w, 0.5, (-0.4, 0.2, 0.1), 0.5)
    q = grid.EvalCollisionPeriodicAPlus_DGV(f)
    scale = float(np.max(np.abs(q))) or 1.0
    h = hashlib.sha256()
    d = grid.desc
    h.update(json.dumps([d.Mu, d.Mv, d.Mw, d.su, d.sv, d.sw, d.u_L, d.u_R, d.v_L, d.v_R, d.w_L, d.w_R]).encode())
    wl = grid.workload()
    h.update(json.dumps([wl["records"], wl["pairs"]]).encode())
    h.update(np.round(q / scale, 10).tobytes())
    return h.hexdigest()


class TrainingSetWriter:
    """Streams `nrows` seeded (f, Q(f)) pairs of one grid into `path`."""

    def __init__(self, path: str, grid, nrows: int, precision=None, batch: int = 256, seed_base: int = 0,
                 tile: int = 0, family: str = "mixtures", T: float = 0.8, amplitude=(1e-3, 1e-1)):
        from . import DGV_FP64
        if family not in FAMILIES:
            raise ValueError(f"unknown VDF family {family!r}")
        self.family, self.T, self.amplitude = family, float(T), (float(amplitude[0]), float(amplitude[1]))
        self.path, self.grid, self.nrows = path, grid, int(nrows)
        self.precision = DGV_FP64 if precision is None else precision
        self.batch, self.seed_base, self.tile = min(int(batch), self.nrows), int(seed_base), tile
        os.makedirs(path, exist_ok=True)

    def run(self, log=None, log_every: int = 1) -> dict:
        """Writes the set; `log(msg)` gets a progress line every `log_every` chunks."""
        from . import PRECISION_NAMES, DGVBatch
        from numpy.lib.format import open_memmap
        g, N, B = self.grid, self.grid.nnodes, self.batch
        t0 = time.perf_counter()
        fp = open_memmap(os.path.join(self.path, "f.npy"), mode="w+", dtype=np.float64, shape=(self.nrows, N))
        qp = open_memmap(os.path.join(self.path, "q.npy"), mode="w+", dtype=np.float64, shape=(self.nrows, N))
        seeds = np.zeros((self.nrows, 2), np.uint64)
        seeds[:, 0] = self.seed_base
        seeds[:, 1] = np.arange(self.nrows, dtype=np.uint64)
        np.save(os.path.join(self.path, "seeds.npy"), seeds)
        b = DGVBatch([g], np.zeros(B, np.int32), self.precision)
        if self.tile:
            b.set_tile(self.tile)
        d_f = g.device_alloc(8 * N * B)
        d_q = g.device_alloc(8 * N * B)
        draw = lambda r0, k: np.stack([vdf_from_seed(g, self.seed_base, r, self.family, self.T, self.amplitude)
                                       for r in range(r0, r0 + k)])
        t_gpu = 0.0
        try:
            with ThreadPoolExecutor(max_workers=1) as pool:
                nxt = pool.submit(draw, 0, min(B, self.nrows))
                for r0 in range(0, self.nrows, B):
                    k = min(B, self.nrows - r0)
                    F = nxt.result()
                    if r0 + k < self.nrows:
                        nxt = pool.submit(draw, r0 + k, min(B, self.nrows - r0 - k))
                    tg = time.perf_counter()
                    g.htod(d_f, F)
                    b.set_f_rows(d_f, 0, k)
                    b.eval_q()
                    b.get_q_rows(d_q, 0, k)
                    Q = np.zeros((k, N))
                    g.dtoh(Q, d_q)
                    t_gpu += time.perf_counter() - tg
                    fp[r0:r0 + k] = F
                    qp[r0:r0 + k] = Q
                    if log and ((r0 // B) % max(1, log_every) == 0 or r0 + k == self.nrows):
                        log(f"dataset: {r0 + k}/{self.nrows} rows")
        finally:
            b.close()
            g.device_free(d_f)
            g.device_free(d_q)
        fp.flush()
        qp.flush()
        del fp, qp
        secs = time.perf_counter() - t0
        d = g.desc
        meta = {"format": FORMAT, "rows": self.nrows, "nodes": N,
                "grid": {"M": [d.Mu, d.Mv, d.Mw], "s": [d.su, d.sv, d.sw],
                         "box": [d.u_L, d.u_R, d.v_L, d.v_R, d.w_L, d.w_R]},
                "aarray_fingerprint": aarray_fingerprint(g), "workload": g.workload(),
                "precision": PRECISION_NAMES[self.precision],
                "seed_base": self.seed_base, "vdf_family": self.family,
                "family_params": {"T": self.T, "amplitude": list(self.amplitude)}, "batch": B,
                "seconds": secs, "gpu_seconds": t_gpu}
        with open(os.path.join(self.path, "meta.json"), "w") as fh:
            json.dump(meta, fh, indent=1)
        return meta


def load(path: str):
    """(meta, f, q, seeds) with f and q memory-mapped read-only."""
    with open(os.path.join(path, "meta.json")) as fh:
        meta = json.load(fh)
    if meta.get("format") != FORMAT:
        raise ValueError(f"{path}: unknown training-set format {meta.get('format')}")
    f = np.load(os.path.join(path, "f.npy"), mmap_mode="r")
    q = np.load(os.path.join(path, "q.npy"), mmap_mode="r")
    seeds = np.load(os.path.join(path, "seeds.npy"))
    return meta, f, q, seeds


def verify(path: str, grid, nsample: int = 8, seed: int = 0) -> dict:
    """Checks a set against `grid`: the A-array fingerprint, f of sampled rows
    drawn again from their seeds (bit for bit), and their Q re-evaluated by the
    single-VDF FP64 evaluator (relative L2)."""
    meta, f, q, seeds = load(path)
    fp_ok = meta["aarray_fingerprint"] == aarray_fingerprint(grid)
    rng = np.random.default_rng(seed)
    rows = rng.choice(meta["rows"], size=min(nsample, meta["rows"]), replace=False)
    f_same, err = True, 0.0
    fam = meta.get("vdf_family", "mixtures")
    fp = meta.get("family_params", {"T": 0.8, "amplitude": [1e-3, 1e-1]})
    for r in rows:
        sb, ri = int(seeds[r, 0]), int(seeds[r, 1])
        f_same &= bool(np.array_equal(vdf_from_seed(grid, sb, ri, fam, fp["T"], tuple(fp["amplitude"])), f[r]))
        qd = grid.EvalCollisionPeriodicAPlus_DGV(np.asarray(f[r]))
        err = max(err, float(np.linalg.norm(qd - q[r]) / np.linalg.norm(qd)))
    return {"fingerprint_matches": fp_ok, "f_from_seeds": f_same, "q_rel_l2_max": err, "rows_checked": len(rows)}
