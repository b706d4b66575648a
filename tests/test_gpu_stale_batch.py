"""A batch holds device pointers into its grids' A-arrays and copies made from
them.  Once a grid's A-array is truncated, replaced or changed in place, every
entry point of the batch that would launch on them returns DGV_ERR_STATE, and a
batch created afterwards runs on the new A-array."""
import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

DGV_ERR_STATE = -6


def _raises_state(fn, *a):
    import dgv0d3v_direct_collision_amd as dgv
    with pytest.raises(dgv.DGVError) as e:
        fn(*a)
    assert e.value.code == DGV_ERR_STATE, str(e.value)
    assert "changed after the batch was created" in str(e.value)


@pytest.mark.parametrize("mutation", ["truncate", "abs", "regenerate", "set_aarray"])
def test_batch_refuses_a_changed_aarray(mutation):
    import dgv0d3v_direct_collision_amd as dgv
    print(f"\n[stale] M=3 mutation={mutation}", flush=True)
    g = dgv.DGV(3, 3, (-1.5, 1.5))
    try:
        g.generate_aarray()
        rng = np.random.default_rng(3)
        f = np.exp(-(g.nodes_u ** 2 + g.nodes_v ** 2 + g.nodes_w ** 2) / 0.6) * (1 + 0.1 * rng.uniform(-1, 1, g.nnodes))
        for prec in ("DGV_FP64", "DGV_FP32_COMP", "DGV_INT8_SPLIT"):
            b = dgv.DGVBatch([g], np.zeros(3, np.int32), getattr(dgv, prec))
            for i in range(3):
                b.set_f(i, f)
            b.eval_q()
            q0 = b.get_q(0)
            assert rel_l2(q0, g.EvalCollisionPeriodicAPlus_DGV(f)) <= (1e-12 if prec == "DGV_FP64" else 1e-6)
            if mutation == "truncate":
                g.TruncAarrsTrhls_DGV(1e-6 if prec == "DGV_FP64" else 1e-5)
            elif mutation == "abs":
                g.aarray_abs()
            elif mutation == "regenerate":
                g.generate_aarray(nquad=16)
            else:
                g.set_aarray(*g.get_aarray())
            _raises_state(b.eval_q)
            _raises_state(b.rk_run, 0.01, 1, 2, 1.0, True)
            _raises_state(b.rk_run, 0.01, 1, 2, 1.0, False)
            _raises_state(b.set_tile, 1 << 18)
            _raises_state(b.set_probe_cells, [0, 1])
            b.close()
        # a new batch sees the current A-array
        b = dgv.DGVBatch([g], np.zeros(1, np.int32), dgv.DGV_FP64)
        b.set_f(0, f)
        b.eval_q()
        assert rel_l2(b.get_q(0), g.EvalCollisionPeriodicAPlus_DGV(f)) <= 1e-12
        b.close()
    finally:
        g.close()
