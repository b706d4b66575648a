"""The batching configuration's workload in one persistent launch: instances on
the 15^3 (5 cells on [-3.5,3.5]^3), 21^3 (7 cells on [-4,4]^3) and 27^3 (9 cells
on [-4,4]^3) grids in the ratio 1:1:2, interleaved, each a distinct seeded
mixture of 1-3 Maxwellians (96 instances instead of 512, to keep the test short;
the launch structure is the same).  FP64: sampled instances of every grid equal
the single-VDF evaluator to 1e-12.  FP32 modes: every sampled instance within
4 * 2^-24 of its term scale.  INT8_SPLIT: within 2^-28 of the term scale and
1e-6 relative L2 on every sampled instance."""
import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

GRIDS = ((5, 3.5), (7, 4.0), (9, 4.0))


def test_mixed_grid_configuration_one_launch():
    import dgv0d3v_direct_collision_amd as dgv
    from dgv0d3v_direct_collision_amd.workloads import mixture_vdf
    print("\n[config mixed] A-arrays for 15^3, 21^3, 27^3", flush=True)
    gs = [dgv.DGV(M, 3, (-L, L)) for M, L in GRIDS]
    try:
        for g in gs:
            g.generate_aarray(min_sens=0.0, nquad=24)
        rng = np.random.default_rng(20261204)
        which = np.array([0] * 24 + [1] * 24 + [2] * 48, np.int32)
        rng.shuffle(which)
        fs = [mixture_vdf(gs[k].nodes_u, gs[k].nodes_v, gs[k].nodes_w, gs[k].nodes_gwts, rng, 3) for k in which]
        sample = [int(np.flatnonzero(which == k)[j]) for k in range(3) for j in (0, 7, 19)]
        qs = {}
        for prec in ("FP64", "FP32_COMP", "FP32_F64ACC", "INT8_SPLIT"):
            print(f"[config mixed] {len(which)} instances, {prec}", flush=True)
            b = dgv.DGVBatch(gs, which, getattr(dgv, "DGV_" + prec))
            for i, f in enumerate(fs):
                b.set_f(i, f)
            b.eval_q()
            qs[prec] = {i: b.get_q(i) for i in sample}
            b.close()
        ref = {i: gs[which[i]].EvalCollisionPeriodicAPlus_DGV(fs[i]) for i in sample}
        for i in sample:
            assert rel_l2(qs["FP64"][i], ref[i]) <= 1e-12, i
        print("[config mixed] term scales on |A| copies", flush=True)
        absq = {}
        for k, (M, L) in enumerate(GRIDS):
            a = dgv.DGV(M, 3, (-L, L))
            a.generate_aarray(min_sens=0.0, nquad=24)
            a.aarray_abs()
            for i in sample:
                if which[i] == k:
                    absq[i] = np.linalg.norm(a.EvalCollisionPeriodicAPlus_DGV(fs[i]))
            a.close()
        for prec, bound in (("FP32_COMP", 4.0 * 2.0 ** -24), ("FP32_F64ACC", 4.0 * 2.0 ** -24),
                            ("INT8_SPLIT", 2.0 ** -28)):
            worst = max(np.linalg.norm(qs[prec][i] - ref[i]) / absq[i] for i in sample)
            rel = max(rel_l2(qs[prec][i], ref[i]) for i in sample)
            print(f"[config mixed] {prec}: max error / term scale {worst / 2.0 ** -24:.4f} x 2^-24, "
                  f"max rel L2 {rel:.3e}", flush=True)
            assert worst <= bound
            if prec == "INT8_SPLIT":
                assert rel <= 1e-6
    finally:
        for g in gs:
            g.close()
