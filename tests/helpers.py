"""Test-side helpers: synthetic A-arrays in the reference layout, seeded VDFs.

A synthetic A-array has the record structure of SetA_DGV (all node pairs
xi < xi1 that pass its geometric prefilter, dgvtools_mod.f90:484-530) with
random values; the evaluator is a bilinear form in f for any given A, so parity
of the GPU evaluator with the oracle does not depend on how A was computed.
"""
from __future__ import annotations

import numpy as np


def prefilter_pairs(G, Trad=100.0):
    """All (i2, i3), 1-based, i2 < i3, passing the prefilter for the canonical cell."""
    P = np.stack([G.nodes_u, G.nodes_v, G.nodes_w], 1)
    cc = G.ccell - 1
    lo = np.array([G.cells_lu[cc], G.cells_lv[cc], G.cells_lw[cc]])
    hi = np.array([G.cells_ru[cc], G.cells_rv[cc], G.cells_rw[cc]])
    dphi = np.sqrt(((hi - lo) ** 2).sum())
    out_i, out_j = [], []
    N = len(P)
    for i in range(N - 1):
        j = np.arange(i + 1, N)
        d = P[i] - P[j]
        dsph = np.sqrt((d * d).sum(1))
        s = P[i] + P[j] - lo - hi
        dist2 = np.sqrt((s * s).sum(1))
        keep = ~((dist2 > dsph + dphi) | (dist2 + dphi < dsph) | (dsph > Trad))
        out_i.append(np.full(keep.sum(), i + 1, np.int32))
        out_j.append((j[keep] + 1).astype(np.int32))
    return np.concatenate(out_i), np.concatenate(out_j)


def synthetic_aarray(G, seed=0, drop=0.1, Trad=100.0):
    """Reference-layout A-array (capphi, A, A_xi, A_xi1) with 1-based ids and
    random values; a fraction `drop` of records is removed per basis function so
    that record sets differ between basis functions."""
    rng = np.random.default_rng(seed)
    i2, i3 = prefilter_pairs(G, Trad)
    D = int(np.prod(G.s))
    capphi = np.zeros(G.nnodes, np.int32)
    As, xs, x1s = [], [], []
    for l in range(D):
        keep = rng.random(len(i2)) >= drop
        n = int(keep.sum())
        capphi[(G.ccell - 1) * D + l] = n
        As.append(rng.standard_normal(n) * 1e-3)
        xs.append(i2[keep])
        x1s.append(i3[keep])
    return capphi, np.concatenate(As), np.concatenate(xs), np.concatenate(x1s)


def maxwellian(u, v, w, n, ub, T):
    b = np.sqrt(np.pi * T) * (np.pi * T)
    return n * np.exp(-((u - ub[0]) ** 2 + (v - ub[1]) ** 2 + (w - ub[2]) ** 2) / T) / b


def bimodal(u, v, w, seed):
    """Seeded sum of two Maxwellians: densities U[0.3,0.7] summing to 1, bulk
    velocities U[-1,1]^3, temperatures U[0.3,1.0]."""
    rng = np.random.default_rng(seed)
    n1 = rng.uniform(0.3, 0.7)
    u1, u2 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
    T1, T2 = rng.uniform(0.3, 1.0, 2)
    return maxwellian(u, v, w, n1, u1, T1) + maxwellian(u, v, w, 1 - n1, u2, T2)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
