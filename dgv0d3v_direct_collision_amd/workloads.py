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


def invariant_free(u, v, w, gwts, fm, delta):
    """delta minus fm * (a + b.u + c |u|^2) with the five coefficients chosen so
    that the result has no density, momentum or energy (sum_i w_i psi_k(i) x_i = 0
    for psi = 1, u, v, w, |u|^2): fm + result keeps fm's moments."""
    psi = np.stack([np.ones(len(u)), u, v, w, u * u + v * v + w * w])
    G = (psi * (gwts * fm)) @ psi.T
    lam = np.linalg.solve(G, psi @ (gwts * delta))
    return delta - fm * (lam @ psi)


def perturbed_maxwellian_vdf(u, v, w, gwts, fm, T, rng, amplitude, degree=4):
    """fm + delta for the decomposition regime: delta = fm p(u) with p a random
    polynomial in (u, v, w) / sqrt(T) of total degree 1..degree (N(0,1)
    coefficients on the monomials), made invariant-free (invariant_free) and
    scaled to max|delta| = amplitude max fm.  The rows of one fm share its
    density, bulk velocity and temperature, so f - fm is the Df of the
    reference's run_mode 1 split at that Maxwellian."""
    s = 1.0 / np.sqrt(T)
    x, y, z = u * s, v * s, w * s
    p = np.zeros(len(u))
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            for c in range(degree + 1 - a - b):
                if 1 <= a + b + c:
                    p += rng.standard_normal() * x ** a * y ** b * z ** c
    d = invariant_free(u, v, w, gwts, fm, fm * p)
    return fm + d * (amplitude * np.max(fm) / np.max(np.abs(d)))
