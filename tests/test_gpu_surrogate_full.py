"""The surrogate configuration's network at its real size: 27^3 nodes in and
out, three hidden layers of 2048 (19683 -> 2048 -> 2048 -> 2048 -> 19683, GELU,
BF16 operands on the matrix cores, FP32 sums).  The forward output and every
weight and bias gradient of one training step on 8 rows of the training-data
family (1-4 Maxwellian mixtures with a perturbation), against the CPU reference
(oracle/surrogate_ref.py) with BF16 rounding at the GPU's points and float64
sums: relative L2 <= 2e-3 per tensor."""
import numpy as np
import pytest

from oracle import surrogate_ref as sr
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def test_full_size_network_forward_and_gradients():
    import dgv0d3v_direct_collision_amd as dgv
    from dgv0d3v_direct_collision_amd.dataset import vdf_from_seed
    print("\n[surrogate full] 27^3 grid, 19683-2048x3-19683", flush=True)
    g = dgv.DGV(9, 3, (-4.0, 4.0))
    try:
        N = g.nnodes
        widths = [N, 2048, 2048, 2048, N]
        rng = np.random.default_rng(5)
        Ws = [(rng.standard_normal((widths[l + 1], widths[l])) / np.sqrt(widths[l])).astype(np.float32)
              for l in range(4)]
        bs = [(0.05 * rng.standard_normal(widths[l + 1])).astype(np.float32) for l in range(4)]
        s = dgv.Surrogate(g, widths, batch=256, seed=5)
        for l in range(4):
            s.set_layer(l, Ws[l], bs[l])
        f = np.stack([vdf_from_seed(g, 20261207, r) for r in range(8)])
        q = rng.standard_normal((8, N)) * 0.02
        fs, qs = float(np.sqrt(np.mean(f ** 2))), 0.02
        s.set_scales(fs, qs)
        loss = s.train_step_host(f, q, lr=0.0)
        W64 = [W.astype(np.float64) for W in Ws]
        b64 = [b.astype(np.float64) for b in bs]
        rloss, dWs, dbs, _ = sr.forward_backward(W64, b64, f, q, fs, qs, bf16_path=True)
        print(f"[surrogate full] loss gpu {loss:.8e} cpu {rloss:.8e}", flush=True)
        assert abs(loss - rloss) <= 1e-3 * rloss
        for l in range(4):
            gW, gb = s.get_grad(l)
            eW, eb = rel_l2(gW, dWs[l]), rel_l2(gb, dbs[l])
            print(f"[surrogate full] layer {l} ({widths[l]} -> {widths[l + 1]}): dW rel L2 {eW:.2e}, "
                  f"db rel L2 {eb:.2e}", flush=True)
            assert eW <= 2e-3 and eb <= 2e-3
        p = s.predict_host(f, project=False)
        _, _, _, y = sr.forward_backward(W64, b64, f, np.zeros_like(f), fs, qs, bf16_path=True)
        e = rel_l2(p, qs * y)
        print(f"[surrogate full] forward rel L2 {e:.2e}", flush=True)
        assert e <= 2e-3
        s.close()
    finally:
        g.close()
