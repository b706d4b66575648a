"""The oracle against the reference's own outputs stored in
tests/golden/ref_dgvtools.npz (written by tests/golden/make_ref_vectors.py from
the reference's dgvtools_mod / distributions_mod routines compiled unmodified):
IphiDGV at unit and non-unit scale, IphiDGV_decomp, MassCheckRec,
MassCheckRecHighCMoments, maxwelveldist on 200 argument sets, and the two
A-array truncations, all bit for bit.  Runs without the reference sources."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_dgvtools.npz")


@pytest.fixture(scope="module")
def gold():
    z = np.load(GOLD)   # plain arrays, no pickles
    G = oracle.Geometry(int(z["grid"][0]), int(z["grid"][1]), (float(z["grid"][2]), float(z["grid"][3])))
    for k in ("nodes_u", "nodes_v", "nodes_w", "nodes_gwts"):
        assert np.array_equal(getattr(G, k), z[k])   # the inputs were the oracle's geometry
    return z, G


def test_iphi_and_decomposition(gold):
    z, G = gold
    full = (z["capphi"], z["A"], z["A_xi"], z["A_xi1"])
    assert np.array_equal(oracle.iphi(*full, z["nodes_gwts"], z["f"]), z["iphi"])
    d, l, n = z["coef"]
    assert np.array_equal(oracle.iphi(*full, z["nodes_gwts"], z["f"], (d / l) ** 2 * n), z["iphi_coef"])
    m = oracle.masscheck(G, z["f"])
    fM = np.array([oracle.maxwell(m[4], m[1], m[2], m[3], m[0], a, b, c)
                   for a, b, c in zip(G.nodes_u, G.nodes_v, G.nodes_w)])
    assert np.array_equal(oracle.iphi_decomp(*full, z["nodes_gwts"], z["f"], fM), z["iphi_decomp"])


def test_moments_and_maxwellian(gold):
    z, G = gold
    assert np.array_equal(oracle.masscheck(G, z["f"]), z["masscheck"])
    assert np.array_equal(oracle.masscheck_high(G, z["f"]), z["masscheck_high"])
    got = np.array([oracle.maxwell(*r) for r in z["maxwell_in"]])
    assert np.array_equal(got, z["maxwell_out"])


def test_truncations(gold):
    z, G = gold
    full = (z["capphi"], z["A"], z["A_xi"], z["A_xi1"])
    for got, name in ((oracle.trunc_threshold(*full, float(z["trunc_threshold"][0])), "trunc_threshold"),
                      (oracle.trunc_radius(G, *full, float(z["trunc_radius"][0])), "trunc_radius")):
        keep = z[name + "_keep"]
        assert 0 < keep.sum() < len(keep)
        assert np.array_equal(got[0], z[name + "_capphi"])
        for g, a in zip(got[1:], full[1:]):
            assert np.array_equal(g, a[keep])
