"""The learned surrogate's GPU kernels against the CPU reference
(oracle/surrogate_ref.py), through the C ABI:

 * forward output and every weight / bias gradient of one training step, with
   BF16 rounding at the GPU's points and float64 sums on the CPU: relative L2
   <= 2e-3 per tensor (the GPU accumulates in FP32);
 * one Adam step against the textbook update from the same gradients, and the
   BF16 weight copies it writes against those made from its master weights;
 * the FP64 projection: sum_i w_i psi_k q_i <= 1e-13 of sum_i |w_i psi_k q_i|;
 * training on (f, Q(f)) pairs from the direct evaluator: the held-out error
   falls far below that of the method without a surrogate (Q taken as 0);
 * the direct evaluator as judge and fallback: residuals on the probe cells
   (of the surrogate's Q before the projection), fallback rows equal to the
   direct Q bit for bit, the others to the projected surrogate.
"""
import numpy as np
import pytest

from oracle import surrogate_ref as sr
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grid():
    import dgv0d3v_direct_collision_amd as dgv
    print("\n[surrogate] grid M=3 s=3 on [-3,3]^3", flush=True)
    g = dgv.DGV(3, 3, (-3.0, 3.0))
    g.generate_aarray()
    yield g
    g.close()


def vdf_rows(g, rng, n):
    """Seeded mixtures of 1-4 Maxwellians plus a small non-negative perturbation,
    normalised to unit density."""
    out = np.zeros((n, g.nnodes))
    for r in range(n):
        f = np.zeros(g.nnodes)
        for _ in range(rng.integers(1, 5)):
            dens, T = rng.uniform(0.2, 1.0), rng.uniform(0.3, 1.2)
            ub = rng.uniform(-0.8, 0.8, 3)
            f += dens * np.exp(-((g.nodes_u - ub[0]) ** 2 + (g.nodes_v - ub[1]) ** 2 +
                                 (g.nodes_w - ub[2]) ** 2) / T) / (np.pi * T) ** 1.5
        f += 0.02 * f.max() * rng.uniform(0, 1, g.nnodes)
        out[r] = f / np.sum(g.nodes_gwts * f)
    return out


def random_layers(widths, rng):
    Ws = [(rng.standard_normal((widths[l + 1], widths[l])) / np.sqrt(widths[l])).astype(np.float32)
          for l in range(len(widths) - 1)]
    bs = [(0.1 * rng.standard_normal(widths[l + 1])).astype(np.float32) for l in range(len(widths) - 1)]
    return Ws, bs


def test_gradients_and_forward_match_cpu(grid):
    import dgv0d3v_direct_collision_amd as dgv
    rng = np.random.default_rng(3)
    widths = [grid.nnodes, 200, 136, grid.nnodes]     # padded to 768, 256, 256, 768
    s = dgv.Surrogate(grid, widths, batch=100, seed=3)
    Ws, bs = random_layers(widths, rng)
    for l in range(len(Ws)):
        s.set_layer(l, Ws[l], bs[l])
    f = vdf_rows(grid, rng, 77)                         # 77 of 100 rows: masking
    q = rng.standard_normal((77, grid.nnodes)) * 0.05
    s.set_scales(0.2, 0.05)
    loss = s.train_step_host(f, q, lr=0.0)
    rloss, dWs, dbs, y = sr.forward_backward([W.astype(np.float64) for W in Ws], [b.astype(np.float64) for b in bs],
                                             f, q, 0.2, 0.05, bf16_path=True)
    print(f"[surrogate] loss gpu {loss:.8e} cpu {rloss:.8e}", flush=True)
    assert abs(loss - rloss) <= 1e-4 * rloss
    for l in range(len(Ws)):
        gW, gb = s.get_grad(l)
        eW, eb = rel_l2(gW, dWs[l]), rel_l2(gb, dbs[l])
        print(f"[surrogate] layer {l}: dW rel L2 {eW:.2e}, db rel L2 {eb:.2e}", flush=True)
        assert eW <= 2e-3 and eb <= 2e-3
        W2, b2 = s.get_layer(l)                          # lr = 0: unchanged
        assert np.array_equal(W2, Ws[l]) and np.array_equal(b2, bs[l])
    # forward over more rows than one batch (two passes), unprojected
    f2 = vdf_rows(grid, rng, 150)
    p = s.predict_host(f2, project=False)
    _, _, _, y2 = sr.forward_backward([W.astype(np.float64) for W in Ws], [b.astype(np.float64) for b in bs],
                                      f2, np.zeros_like(f2), 0.2, 0.05, bf16_path=True)
    e = rel_l2(p, 0.05 * y2)
    print(f"[surrogate] forward rel L2 {e:.2e}", flush=True)
    assert e <= 2e-3
    s.close()


def test_adam_step_matches_cpu(grid):
    import dgv0d3v_direct_collision_amd as dgv
    rng = np.random.default_rng(8)
    widths = [grid.nnodes, 128, grid.nnodes]
    s = dgv.Surrogate(grid, widths, batch=64, seed=1)
    Ws, bs = random_layers(widths, rng)
    for l in range(len(Ws)):
        s.set_layer(l, Ws[l], bs[l])
    f = vdf_rows(grid, rng, 64)
    q = rng.standard_normal((64, grid.nnodes)) * 0.05
    s.set_scales(0.2, 0.05)
    s.train_step_host(f, q, lr=0.0)
    grads = [s.get_grad(l) for l in range(len(Ws))]
    s2 = dgv.Surrogate(grid, widths, batch=64, seed=1)
    for l in range(len(Ws)):
        s2.set_layer(l, Ws[l], bs[l])
    s2.set_scales(0.2, 0.05)
    s2.train_step_host(f, q, lr=1e-3)
    for l in range(len(Ws)):
        W1, b1 = s2.get_layer(l)
        eW, _, _ = sr.adam(Ws[l].astype(np.float64), 0.0, 0.0, grads[l][0].astype(np.float64), 1, 1e-3)
        eb, _, _ = sr.adam(bs[l].astype(np.float64), 0.0, 0.0, grads[l][1].astype(np.float64), 1, 1e-3)
        assert np.max(np.abs(W1 - eW)) <= 2e-6 and np.max(np.abs(b1 - eb)) <= 2e-6
    # the BF16 copies the Adam step wrote (forward [O][I] and transposed [I][O])
    # equal those made from the updated master weights by the conversion kernel:
    # forward outputs and every gradient bit for bit
    s3 = dgv.Surrogate(grid, widths, batch=64, seed=1)
    for l in range(len(Ws)):
        s3.set_layer(l, *s2.get_layer(l))
    s3.set_scales(0.2, 0.05)
    np.testing.assert_array_equal(s2.predict_host(f, project=False), s3.predict_host(f, project=False))
    s2.train_step_host(f, q, lr=0.0)
    s3.train_step_host(f, q, lr=0.0)
    for l in range(len(Ws)):
        for a, b in zip(s2.get_grad(l), s3.get_grad(l)):
            np.testing.assert_array_equal(a, b)
    s.close()
    s2.close()
    s3.close()


def test_projection_conserves_to_1e13(grid):
    import dgv0d3v_direct_collision_amd as dgv
    rng = np.random.default_rng(4)
    widths = [grid.nnodes, 256, grid.nnodes]
    s = dgv.Surrogate(grid, widths, batch=32, seed=2)
    s.set_scales(0.2, 0.05)
    f = vdf_rows(grid, rng, 40)
    raw = s.predict_host(f, project=False)
    p = s.predict_host(f, project=True)
    u, v, w, gw = grid.nodes_u, grid.nodes_v, grid.nodes_w, grid.nodes_gwts
    psi = np.stack([np.ones_like(u), u, v, w, u * u + v * v + w * w])
    mom = (p * gw) @ psi.T
    scale = np.abs(p * gw) @ np.abs(psi.T)
    print(f"[surrogate] projected moments / scale max {np.max(np.abs(mom) / scale):.2e}", flush=True)
    assert np.all(np.abs(mom) <= 1e-13 * scale)
    assert rel_l2(p, sr.project_invariants(raw, u, v, w, gw)) <= 1e-12
    s.close()


def test_training_gain_and_direct_fallback(grid):
    import dgv0d3v_direct_collision_amd as dgv
    rng = np.random.default_rng(20261207)
    N = grid.nnodes
    ntr, nte = 1024, 128
    f = vdf_rows(grid, rng, ntr + nte)
    d_f = grid.device_alloc(f.nbytes)
    d_q = grid.device_alloc(f.nbytes)
    grid.htod(d_f, f)
    for r in range(ntr + nte):                          # direct FP64 Q of every row
        grid.eval_q_dev(d_f + 8 * N * r, d_q + 8 * N * r)
    grid.synchronize()
    q = np.zeros_like(f)
    grid.dtoh(q, d_q)
    s = dgv.Surrogate(grid, [N, 512, 512, 512, N], batch=128, seed=7)
    s.set_scales(float(np.sqrt(np.mean(f[:ntr] ** 2))), float(np.sqrt(np.mean(q[:ntr] ** 2))))
    losses = []
    for it in range(600):
        idx = rng.integers(0, ntr // 128) * 128
        lr = 1e-3 if it < 400 else 3e-4
        losses.append(s.train_step(d_f + 8 * N * idx, d_q + 8 * N * idx, 128, lr))
    print(f"[surrogate] loss {losses[0]:.3e} -> {np.mean(losses[-20:]):.3e}", flush=True)
    f_te, q_te = f[ntr:], q[ntr:]
    p = s.predict_host(f_te, project=True)
    err = np.array([rel_l2(p[i], q_te[i]) for i in range(nte)])
    print(f"[surrogate] held-out rel L2: mean {err.mean():.3e}, max {err.max():.3e} "
          f"(without the surrogate, Q = 0: 1.0)", flush=True)
    assert np.mean(losses[-20:]) < 0.2 * losses[0]
    assert err.mean() < 0.4
    # judge and fallback
    s.set_probe_cells([0, 13, 26, 4])
    out, res, fb, nfb = s.eval_host(f_te, tol=np.inf)
    assert nfb == 0 and np.array_equal(out, p)
    probe_nodes = np.concatenate([np.arange(c * 27, c * 27 + 27) for c in (0, 4, 13, 26)])
    p_raw = s.predict_host(f_te, project=False)     # the judge compares the Q before the projection
    r_host = np.array([np.linalg.norm(p_raw[i, probe_nodes] - q_te[i, probe_nodes]) /
                       np.linalg.norm(q_te[i, probe_nodes]) for i in range(nte)])
    assert np.max(np.abs(res - r_host) / r_host) <= 1e-9
    tol = float(np.median(res))
    out, res2, fb, nfb = s.eval_host(f_te, tol=tol)
    assert np.array_equal(res2, res) and nfb == int(np.sum(res > tol)) and 0 < nfb < nte
    for i in range(nte):
        if fb[i]:
            assert np.array_equal(out[i], q_te[i])        # the direct FP64 evaluator, bit for bit
        else:
            assert np.array_equal(out[i], p[i])
    out, _, fb, nfb = s.eval_host(f_te, tol=0.0)
    assert nfb == nte and np.array_equal(out, q_te)
    s.close()
    grid.device_free(d_f)
    grid.device_free(d_q)
