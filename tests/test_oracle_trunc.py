"""Pins of the oracle's A-array truncations (TruncAarrsTrhls_DGV,
TruncAarrsRadius_DGV) by their defining properties, no GPU: the survivors are
exactly the records passing the test, in their original order (a subsequence),
per-basis counts add up, a second truncation changes nothing, a looser
criterion keeps a superset, and the extreme criteria keep all or nothing."""
import numpy as np

import oracle
from tests.helpers import synthetic_aarray


def _subsequence(small, big):
    it = iter(big)
    return all(any(x == y for y in it) for x in small)


def test_threshold_truncation_properties():
    G = oracle.Geometry(3, 2, (-1.5, 1.5))
    capphi, A, xi, xi1 = synthetic_aarray(G, seed=4)
    t = float(np.quantile(np.abs(A), 0.4))
    cap2, A2, x2, y2 = oracle.trunc_threshold(capphi, A, xi, xi1, t)
    assert np.all(np.abs(A2) > t)
    assert len(A2) == int(np.sum(np.abs(A) > t)) == int(cap2.sum())
    recs = list(zip(A, xi, xi1))
    kept = list(zip(A2, x2, y2))
    assert _subsequence(kept, recs)
    # per basis function: the count of its records above the threshold
    off = np.concatenate([[0], np.cumsum(capphi)])
    for k in np.nonzero(capphi)[0]:
        assert cap2[k] == np.sum(np.abs(A[off[k]:off[k + 1]]) > t)
    again = oracle.trunc_threshold(cap2, A2, x2, y2, t)
    assert all(np.array_equal(a, b) for a, b in zip(again, (cap2, A2, x2, y2)))
    loose = oracle.trunc_threshold(capphi, A, xi, xi1, t / 2)
    assert _subsequence(kept, list(zip(*loose[1:])))
    assert len(oracle.trunc_threshold(capphi, A, xi, xi1, 0.0)[1]) == int(np.sum(A != 0))
    assert len(oracle.trunc_threshold(capphi, A, xi, xi1, np.abs(A).max())[1]) == 0


def test_radius_truncation_properties():
    G = oracle.Geometry(3, 2, (-1.5, 1.5))
    capphi, A, xi, xi1 = synthetic_aarray(G, seed=6)
    P = np.stack([G.nodes_u, G.nodes_v, G.nodes_w], 1)
    d = np.linalg.norm(P[xi - 1] - P[xi1 - 1], axis=1)
    rad = float(np.median(d))
    cap2, A2, x2, y2 = oracle.trunc_radius(G, capphi, A, xi, xi1, rad)
    d2 = np.linalg.norm(P[x2 - 1] - P[y2 - 1], axis=1)
    assert np.all(d2 <= rad * (1 + 1e-15))
    assert len(A2) == int(cap2.sum())
    assert abs(len(A2) - int(np.sum(d <= rad))) <= int(np.sum(np.abs(d - rad) <= 1e-14 * rad))
    assert _subsequence(list(zip(A2, x2, y2)), list(zip(A, xi, xi1)))
    again = oracle.trunc_radius(G, cap2, A2, x2, y2, rad)
    assert all(np.array_equal(a, b) for a, b in zip(again, (cap2, A2, x2, y2)))
    assert len(oracle.trunc_radius(G, capphi, A, xi, xi1, 2 * d.max())[1]) == len(A)
    assert len(oracle.trunc_radius(G, capphi, A, xi, xi1, 0.5 * d.min())[1]) == 0
