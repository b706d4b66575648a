"""Host side of the training-data writer, no GPU: the seeded VDF family is
reproducible row by row, normalised to unit density, non-negative, and rows
differ from each other."""
from types import SimpleNamespace

import numpy as np

import oracle
from dgv0d3v_direct_collision_amd.dataset import vdf_from_seed


def test_vdf_from_seed_is_reproducible_and_normalised():
    G = oracle.Geometry(3, 3, (-3.0, 3.0))
    g = SimpleNamespace(nodes_u=G.nodes_u, nodes_v=G.nodes_v, nodes_w=G.nodes_w, nodes_gwts=G.nodes_gwts, D=27)
    rows = [vdf_from_seed(g, 7, r) for r in range(6)]
    for r, f in enumerate(rows):
        assert np.array_equal(f, vdf_from_seed(g, 7, r))
        assert abs(np.sum(G.nodes_gwts * f) - 1.0) < 1e-14
        assert np.all(f >= 0.0)
    assert not any(np.array_equal(rows[i], rows[j]) for i in range(6) for j in range(i))
    assert not np.array_equal(vdf_from_seed(g, 8, 0), rows[0])
