"""A 33^3-node grid (M=11, 3 nodes per dimension, box [-4.5, 4.5]^3), whose
pair-coefficient array holds more than 2^32 elements (155M pairs x 28 values).

Element-wise kernels over that array must reach every element: HIP limits one
grid dimension to fewer than 2^32 work-items, so a one-thread-per-element launch
would silently cover only the first part. The instances are distinct
VDFs of the precision configuration (a Maxwellian times 1 + a U(-1,1) node by
node, a = 1e-3, 1e-2, 1e-1) and a three-Maxwellian mixture.  Checked at full
size through size-independent properties:
 * FP64 batch == single-VDF evaluator for every instance (relative L2 <= 1e-12);
 * both FP32 batches (two-sum compensated, and FP64-accumulated) against FP64:
   ||dq|| within 4 * 2^-24 of ||2/w sum |A| f f||, the terms' scale (their FP32
   coefficients are converted element-wise);
 * INT8_SPLIT against FP64: ||dq|| within 2^-28 of the terms' scale and every
   VDF within 1e-6 relative L2;
 * the |A| context (element-wise fabs) obeys the triangle inequality
   |Q(f)| <= Q_|A|(f) at every node for f >= 0.
"""
import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def test_33cubed_grid_past_2_32_coefficients():
    import dgv0d3v_direct_collision_amd as dgv
    print("\n[large] grid M=11 s=3 on [-4.5,4.5]^3", flush=True)
    g = dgv.DGV(11, 3, (-4.5, 4.5))
    g.generate_aarray(min_sens=0.0)
    wl = g.workload()
    print(f"[large] pairs {wl['pairs']}, coefficient elements {wl['pairs'] * 28}", flush=True)
    assert wl["pairs"] * 28 > 2 ** 32
    rng = np.random.default_rng(33)
    u, v, w, gw = g.nodes_u, g.nodes_v, g.nodes_w, g.nodes_gwts
    from dgv0d3v_direct_collision_amd import workloads
    fs = []
    f = np.zeros(g.nnodes)
    for k in range(3):
        n, T = rng.uniform(0.3, 1.0), rng.uniform(0.4, 1.0)
        ub = rng.uniform(-0.8, 0.8, 3)
        f += n * np.exp(-((u - ub[0]) ** 2 + (v - ub[1]) ** 2 + (w - ub[2]) ** 2) / T) / (np.pi * T) ** 1.5
    fs.append(f)
    for a in (1e-3, 1e-2, 1e-1, 1e-3, 1e-2, 1e-1, 3e-3):
        fs.append(workloads.near_equilibrium_vdf(u, v, w, gw, rng, a))
    q_single = [g.EvalCollisionPeriodicAPlus_DGV(x) for x in fs]
    qs = {}
    for prec in ("FP64", "FP32_COMP", "FP32_F64ACC", "INT8_SPLIT"):
        print(f"[large] batch of {len(fs)} distinct VDFs, {prec}", flush=True)
        b = dgv.DGVBatch([g], np.zeros(len(fs), np.int32), getattr(dgv, "DGV_" + prec))
        for i, x in enumerate(fs):
            b.set_f(i, x)
        b.eval_q()
        qs[prec] = [b.get_q(i) for i in range(len(fs))]
        b.close()
    for i in range(len(fs)):
        assert rel_l2(qs["FP64"][i], q_single[i]) <= 1e-12
    g.close()
    a = dgv.DGV(11, 3, (-4.5, 4.5))
    a.generate_aarray(min_sens=0.0)
    a.aarray_abs()
    qabs = [a.EvalCollisionPeriodicAPlus_DGV(x) for x in fs]
    a.close()
    slack = min(np.min(qa - np.abs(q)) / np.max(qa) for qa, q in zip(qabs, q_single))
    print(f"[large] triangle slack min {slack:.3e}", flush=True)
    assert slack >= -1e-12
    for prec, bound in (("FP32_COMP", 4.0 * 2.0 ** -24), ("FP32_F64ACC", 4.0 * 2.0 ** -24), ("INT8_SPLIT", 2.0 ** -28)):
        for i in range(len(fs)):
            dq = qs[prec][i] - qs["FP64"][i]
            ratio = np.linalg.norm(dq) / np.linalg.norm(qabs[i])
            node = np.max(np.abs(dq) / np.maximum(qabs[i], 1e-300))
            rel = rel_l2(qs[prec][i], qs["FP64"][i])
            print(f"[large] {prec} instance {i}: rel L2 {rel:.3e}, ||dq|| / term scale {ratio / 2.0 ** -24:.4f} "
                  f"x 2^-24 (worst node {node / 2.0 ** -24:.2f} x 2^-24)", flush=True)
            assert ratio <= bound
            if prec == "INT8_SPLIT":
                assert rel <= 1e-6
