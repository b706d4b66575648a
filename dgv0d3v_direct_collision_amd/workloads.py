"""Seeded synthetic VDFs of the benchmark configurations.

Maxwellian convention n exp(-|v - u|^2 / T) / (pi T)^{3/2} (maxwelveldist,
distributions_mod.f90:83-94).  Every generator takes the grid's node
coordinates and quadrature weights and a numpy Generator, so a workload is
reproduced from its seed.
"""
from __future__ import annotations

import numpy as np


def maxwellian(u, v, w, n, ub, T):
    return n * np.exp(-((u - ub[0]) ** 2 + (v - ub[1]) ** 2 + (w - ub[2]) ** 2) / T) / (np.pi * T) ** 1.5


def bimodal_vdf(u, v, w, rng):
    """Sum of two Maxwellians: densities U[0.3,0.7] summing to 1, bulk
    velocities U[-1,1]^3, temperatures U[0.3,1.0] (the plain and single-VDF
    configurations)."""
    n1 = rng.uniform(0.3, 0.7)
    ub1, ub2 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
    T1, T2 = rng.uniform(0.3, 1.0, 2)
    return maxwellian(u, v, w, n1, ub1, T1) + maxwellian(u, v, w, 1 - n1, ub2, T2)


def perturbed_bimodal_vdf(u, v, w, gwts, M, rng):
    """The single-VDF configuration: the bimodal VDF plus a 5% non-negative
    perturbation 0.05 U[0,1] max f on a seeded random 10% of the cells,
    renormalised to sum(w f) = 1."""
    f = bimodal_vdf(u, v, w, rng)
    ncell = M ** 3
    D = len(f) // ncell
    cells = rng.choice(ncell, size=max(1, ncell // 10), replace=False)
    pert = np.zeros(len(f))
    fmax = f.max()
    for c in cells:
        pert[c * D:(c + 1) * D] = 0.05 * rng.uniform(0, 1, D) * fmax
    f = f + pert
    return f / np.sum(gwts * f)


def mixture_vdf(u, v, w, gwts, rng, kmax=3):
    """Mixture of 1..kmax Maxwellians (densities U[0.2,1], bulk velocities
    U[-1,1]^3, temperatures U[0.3,1]), unit density (the batching configuration)."""
    f = np.zeros(len(u))
    for _ in range(rng.integers(1, kmax + 1)):
        f += maxwellian(u, v, w, rng.uniform(0.2, 1.0), rng.uniform(-1, 1, 3), rng.uniform(0.3, 1.0))
    return f / np.sum(gwts * f)


def near_equilibrium_vdf(u, v, w, gwts, rng, amplitude):
    """The precision configuration: a Maxwellian (bulk velocity U[-0.3,0.3]^3,
    temperature U[0.6,1]) times (1 + a U(-1,1)) node by node, unit density."""
    f0 = maxwellian(u, v, w, 1.0, rng.uniform(-0.3, 0.3, 3), rng.uniform(0.6, 1.0))
    f = f0 * (1.0 + amplitude * rng.uniform(-1, 1, len(u)))
    return f / np.sum(gwts * f)
