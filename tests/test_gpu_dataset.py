"""The training-data writer on the GPU: 300 rows in chunks of 128 (the last
one partial) from the batched FP64 evaluator; every stored Q equals the
single-VDF evaluator's to relative L2 1e-12, f is drawn again from its seed
bit for bit, and the A-array fingerprint identifies the grid's A-array (a
truncated A-array gives another)."""
import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def test_training_set_round_trip(tmp_path):
    import dgv0d3v_direct_collision_amd as dgv
    from dgv0d3v_direct_collision_amd import dataset
    print("\n[dataset] M=3 s=3, 300 rows", flush=True)
    g = dgv.DGV(3, 3, (-3.0, 3.0))
    g.generate_aarray()
    meta = dataset.TrainingSetWriter(str(tmp_path / "set"), g, 300, batch=128, seed_base=5).run()
    assert meta["rows"] == 300 and meta["precision"] == "FP64"
    m2, f, q, seeds = dataset.load(str(tmp_path / "set"))
    assert f.shape == (300, g.nnodes) and q.shape == (300, g.nnodes) and seeds.shape == (300, 2)
    for r in (0, 127, 128, 255, 256, 299):
        assert np.array_equal(f[r], dataset.vdf_from_seed(g, 5, r))
        assert rel_l2(q[r], g.EvalCollisionPeriodicAPlus_DGV(np.asarray(f[r]))) <= 1e-12
    chk = dataset.verify(str(tmp_path / "set"), g, nsample=10)
    assert chk["fingerprint_matches"] and chk["f_from_seeds"] and chk["q_rel_l2_max"] <= 1e-12
    fp = meta["aarray_fingerprint"]
    assert dataset.aarray_fingerprint(g) == fp
    A = g.get_aarray()[1]
    g.TruncAarrsTrhls_DGV(float(np.quantile(np.abs(A), 0.5)))     # drops half of the records
    assert dataset.aarray_fingerprint(g) != fp
    g.close()
