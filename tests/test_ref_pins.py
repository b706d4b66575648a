"""The oracle against the reference's own routines: dgvtools_mod.f90 and
distributions_mod.f90 (with the modules they use) compiled unmodified into
oracle/_ref by oracle/Makefile's "ref" target, driven through oracle/ref.py.
Skipped where that library is not built (the reference sources are absent).

On a 3 x 3 x 3-cell grid with 2 Gauss nodes per dimension and a synthetic
A-array replicated to every node (the full A that IphiDGV walks):
  * IphiDGV (dgvtools_mod.f90:847-877) equals the oracle's orc_iphi bit for bit,
    at unit and non-unit (mol_diam/L_inf)^2 N_inf;
  * IphiDGV_decomp (:896-927) equals the oracle's restatement, fed the oracle's
    MassCheckRec moments and Maxwellian, bit for bit;
  * MassCheckRec (:1437-1458) and MassCheckRecHighCMoments (:1501-1549) equal
    orc_masscheck / orc_masscheck_high bit for bit;
  * maxwelveldist (distributions_mod.f90:83-94) equals orc_maxwell bit for bit;
  * TruncAarrsTrhls_DGV and TruncAarrsRadius_DGV (:1313-1433) give exactly the
    oracle's truncated record arrays.
The GPU paths are held to these oracle functions by the GPU tests, so this pins
them to the reference's code as well.
"""
import numpy as np
import pytest

import oracle
from oracle import ref
from tests.helpers import bimodal, synthetic_aarray

pytestmark = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built (reference sources absent)")
M, S, BOX = 3, 2, (-1.5, 1.5)


@pytest.fixture(scope="module")
def case():
    G = oracle.Geometry(M, S, BOX)
    A = synthetic_aarray(G, seed=21, drop=0.3)
    full = oracle.full_aarray(G, *A)
    f = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, seed=8)
    return G, A, full, f


def nodes(G):
    return G.nodes_u, G.nodes_v, G.nodes_w, G.nodes_gwts


@pytest.mark.parametrize("coefs", [(1.0, 1.0, 1.0), (0.7, 1.3, 2.5)])
def test_iphi_is_the_reference_bit_for_bit(case, coefs):
    G, A, full, f = case
    d, l, n = coefs
    got = ref.iphi(*nodes(G), *full, f, d, l, n)
    want = oracle.iphi(*full, G.nodes_gwts, f, (d / l) ** 2 * n)
    assert np.array_equal(got, want), np.max(np.abs(got - want))
    assert np.count_nonzero(got) > G.nnodes // 2


def test_iphi_decomp_is_the_reference_bit_for_bit(case):
    G, A, full, f = case
    got = ref.iphi(*nodes(G), *full, f, decomp=True)
    n, ub, vb, wb, T = oracle.masscheck(G, f)
    fM = np.array([oracle.maxwell(T, ub, vb, wb, n, a, b, c) for a, b, c in zip(G.nodes_u, G.nodes_v, G.nodes_w)])
    want = oracle.iphi_decomp(*full, G.nodes_gwts, f, fM)
    assert np.array_equal(got, want), np.max(np.abs(got - want))


def test_moments_are_the_reference_bit_for_bit(case):
    G, A, full, f = case
    got, want = ref.masscheck(*nodes(G), f), oracle.masscheck(G, f)
    assert np.array_equal(got, want), (got, want)
    got, want = ref.masscheck_high(*nodes(G), f), oracle.masscheck_high(G, f)
    assert np.array_equal(got, want), np.max(np.abs(got - want))


def test_maxwellian_is_the_reference_bit_for_bit():
    rng = np.random.default_rng(4)
    for _ in range(200):
        T, n = rng.uniform(0.05, 2.0), rng.uniform(0.1, 3.0)
        u0, v0, w0, u, v, w = rng.uniform(-2, 2, 6)
        assert ref.maxwell(T, u0, v0, w0, n, u, v, w) == oracle.maxwell(T, u0, v0, w0, n, u, v, w)
    # the temperature floor max(T, 1e-7)
    assert ref.maxwell(0.0, 0, 0, 0, 1.0, 1e-4, 0, 0) == oracle.maxwell(0.0, 0, 0, 0, 1.0, 1e-4, 0, 0)


def test_truncations_are_the_reference(case):
    G, A, full, f = case
    t = float(np.quantile(np.abs(full[1]), 0.4))
    for got, want in ((ref.trunc_threshold(*full, t), oracle.trunc_threshold(*full, t)),
                      (ref.trunc_radius(*nodes(G), *full, 1.1), oracle.trunc_radius(G, *full, 1.1))):
        assert 0 < len(want[1]) < len(full[1])
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
