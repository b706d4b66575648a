"""A-array truncation on the device (TruncAarrsTrhls_DGV, TruncAarrsRadius_DGV)
against the oracle's restatement applied to the exported records: the records
left must be identical, in the reference's order, and Q on the truncated
A-array must match the oracle's (relative L2 <= 1e-12)."""
import numpy as np
import pytest

import oracle
from tests.helpers import bimodal, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,s,box", [(3, 3, (-1.5, 1.5)), (4, 2, (-2.0, 2.0))], ids=["M3s3", "M4s2"])
def test_truncations_match_oracle(M, s, box):
    import dgv0d3v_direct_collision_amd as dgv
    print(f"\n[trunc] M={M} s={s}", flush=True)
    d = dgv.DGV(M, s, box)
    d.generate_aarray(min_sens=0.0, nquad=24)
    G = oracle.Geometry(M, s, box)
    cap, A, xi, xi1 = d.get_aarray()
    t = float(np.quantile(np.abs(A), 0.3))
    P = np.stack([G.nodes_u, G.nodes_v, G.nodes_w], 1)
    rad = float(np.quantile(np.linalg.norm(P[xi - 1] - P[xi1 - 1], axis=1), 0.7))
    ref = oracle.trunc_threshold(cap, A, xi, xi1, t)
    ref = oracle.trunc_radius(G, *ref, rad)
    d.TruncAarrsTrhls_DGV(t)
    n1, _ = d.aarray_size()
    d.TruncAarrsRadius_DGV(rad)
    got = d.get_aarray()
    n2, per = d.aarray_size()
    print(f"[trunc] {len(A)} -> {n1} -> {n2} records", flush=True)
    assert n2 == len(ref[1]) == per.sum() and n2 < n1 < len(A)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    f = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, 7)
    q = d.EvalCollisionPeriodicAPlus_DGV(f)
    qo = oracle.eval_aplus(G, ref[0], ref[1], ref[2], ref[3], f)
    assert rel_l2(q, qo) <= 1e-12
    d.close()
