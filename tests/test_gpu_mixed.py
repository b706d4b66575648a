"""EvalCollisionPeriodicMixedTermsA_DGV on the device against the oracle
(relative L2 <= 1e-12) for every pair-block kernel, and Mixed(f, f) = 2 Q(f)
bit for bit on the device."""
import numpy as np
import pytest

import oracle
from tests.helpers import bimodal, rel_l2, synthetic_aarray

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,s,box", [(3, 3, (-1.5, 1.5)), (4, 2, (-2.0, 2.0)), (5, 1, (-2.5, 2.5))],
                         ids=["M3s3", "M4s2", "M5s1"])
def test_mixed_term_matches_oracle(M, s, box):
    import dgv0d3v_direct_collision_amd as dgv
    print(f"\n[mixed] M={M} s={s}", flush=True)
    G = oracle.Geometry(M, s, box)
    cap, A, xi, xi1 = synthetic_aarray(G, seed=M)
    f = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, 4)
    fm = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, 9)
    ref = oracle.eval_mixed(G, cap, A, xi, xi1, f, fm)
    d = dgv.DGV(M, s, box)
    d.set_aarray(cap, A, xi, xi1)
    kernels = [dgv.DGV_KERNEL_PAIRBLOCK] + ([dgv.DGV_KERNEL_MFMA] if s > 1 else [])
    for k in kernels:
        d.set_kernel(k)
        m = d.EvalCollisionPeriodicMixedTermsA_DGV(f, fm)
        assert rel_l2(m, ref) <= 1e-12, k
        q = d.EvalCollisionPeriodicAPlus_DGV(f)
        assert np.array_equal(d.EvalCollisionPeriodicMixedTermsA_DGV(f, f), 2.0 * q), k
    d.close()
