"""Batched evaluation (many VDFs, mixed grids, one persistent launch) against
the single-VDF evaluator, whose parity with the oracle is tested elsewhere.
FP64: relative L2 <= 1e-12 for every VDF.  FP32 products (two-sum compensated
FP32 sums, or FP64 sums of the FP32 partial sums): error <= 4 * 2^-24 of the
summed terms' scale for every VDF, i.e. relative L2 <= 4 * 2^-24 * kappa with
kappa = ||sum |A f f| || / ||Q||; a VDF with kappa above about 17 can miss 1e-6,
and the test reports how many do.  INT8_SPLIT (exact integer digit products):
error <= 2^-28 of the terms' scale and relative L2 <= 1e-6 for every VDF."""
import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def vdf(g, rng, k):
    """Seeded mixture of k Maxwellians on grid g."""
    f = np.zeros(g.nnodes)
    for _ in range(k):
        n, T = rng.uniform(0.2, 1.0), rng.uniform(0.2, 0.8)
        ub = rng.uniform(-0.6, 0.6, 3)
        f += n * np.exp(-((g.nodes_u - ub[0]) ** 2 + (g.nodes_v - ub[1]) ** 2 + (g.nodes_w - ub[2]) ** 2) / T) \
            / (np.pi * T) ** 1.5
    return f


@pytest.fixture(scope="module")
def grids():
    import dgv0d3v_direct_collision_amd as dgv
    print("\n[batch] grids M=3 and M=5", flush=True)
    g0 = dgv.DGV(3, 3, (-1.5, 1.5))
    g1 = dgv.DGV(5, 3, (-3.5, 3.5))
    g0.generate_aarray()
    g1.generate_aarray()
    yield [g0, g1]
    g0.close()
    g1.close()


@pytest.fixture(scope="module")
def absq(grids):
    """Per grid: f -> 2/w sum |A| f f (the scale of the terms summed into Q)."""
    import dgv0d3v_direct_collision_amd as dgv
    ctxs = []
    for g in grids:
        c = dgv.DGV(g.M[0], 3, (g.desc.u_L, g.desc.u_R))   # both grids have 3 nodes per dimension
        c.set_aarray(*g.get_aarray())
        c.aarray_abs()
        ctxs.append(c)
    yield [c.EvalCollisionPeriodicAPlus_DGV for c in ctxs]
    for c in ctxs:
        c.close()


@pytest.mark.parametrize("precision", ["FP64", "FP32_COMP", "FP32_F64ACC", "INT8_SPLIT"])
def test_mixed_grid_batch_matches_single(grids, absq, precision):
    import dgv0d3v_direct_collision_amd as dgv
    rng = np.random.default_rng(11)
    which = np.array([0] * 70 + [1] * 10 + [0] * 3, np.int32)   # uneven, interleaved
    rng.shuffle(which)
    b = dgv.DGVBatch(grids, which, getattr(dgv, "DGV_" + precision))
    fs = []
    for i, g in enumerate(which):
        f = vdf(grids[g], rng, 1 + i % 3)
        fs.append(f)
        b.set_f(i, f)
    b.eval_q()
    errs, scaled, kappa = [], [], []
    for i, g in enumerate(which):
        ref = grids[g].EvalCollisionPeriodicAPlus_DGV(fs[i])
        q = b.get_q(i)
        errs.append(rel_l2(q, ref))
        if precision != "FP64":   # error against the scale of the summed terms
            qa = np.linalg.norm(absq[g](fs[i]))
            scaled.append(np.linalg.norm(q - ref) / qa)
            kappa.append(qa / np.linalg.norm(ref))
    errs = np.array(errs)
    print(f"[batch] {precision}: max rel L2 {errs.max():.3e}, {np.mean(errs <= 1e-6):.3f} of VDFs within 1e-6",
          flush=True)
    if precision == "FP64":
        assert errs.max() <= 1e-12
    elif precision == "INT8_SPLIT":
        print(f"[batch] {precision}: max ||dq|| / ||sum |A f f| || = {max(scaled):.3e} "
              f"({max(scaled) * 2 ** 30:.3f} x 2^-30)", flush=True)
        assert max(scaled) <= 2.0 ** -28
        assert errs.max() <= 1e-6
    else:
        # FP32 products carry a relative error of a few 2^-24 each; after the
        # compensated sums the error is that times the terms' scale, not Q's
        print(f"[batch] {precision}: max ||dq|| / ||sum |A f f| || = {max(scaled):.3e}; kappa max {max(kappa):.1f}",
              flush=True)
        assert max(scaled) <= 4 * 2.0 ** -24
        assert np.all(errs <= 4 * 2.0 ** -24 * np.array(kappa))
    b.close()


@pytest.mark.parametrize("precision,tol", [("FP64", 1e-12), ("INT8_SPLIT", 1e-8)])
def test_batch_rk_graph_matches_single(grids, precision, tol):
    import dgv0d3v_direct_collision_amd as dgv
    rng = np.random.default_rng(5)
    which = np.array([0, 1, 0, 1, 1], np.int32)
    b = dgv.DGVBatch(grids, which, getattr(dgv, "DGV_" + precision))
    fs = [vdf(grids[g], rng, 2) for g in which]
    for i, f in enumerate(fs):
        b.set_f(i, f)
    b.rk_run(0.02, 2, rkmts=2, use_graph=True)
    b.rk_run(0.02, 1, rkmts=2, use_graph=True)   # replay after eager work
    for i, g in enumerate(which):
        ref = fs[i].copy()
        grids[g].rkmts, grids[g].coef = 2, 1.0
        for _ in range(3):
            grids[g].TimeIntegratorORK_SH_DGV(ref, 0.02)
        assert rel_l2(b.get_f(i), ref) <= tol
    b.close()


def test_batch_probe_cells_and_projection(grids):
    import dgv0d3v_direct_collision_amd as dgv
    rng = np.random.default_rng(9)
    which = np.array([1] * 70 + [0] * 5, np.int32)     # two 64-instance chunks on grid 1
    b = dgv.DGVBatch(grids, which, dgv.DGV_FP32_COMP)
    fs = [vdf(grids[g], rng, 1 + i % 4) for i, g in enumerate(which)]
    for i, f in enumerate(fs):
        b.set_f(i, f)
    b.eval_q()
    full = [b.get_q(i) for i in range(len(which))]
    # probe cells: only their nodes are rewritten, with the same values
    cells = [0, 7, 13, 26, 7]                          # valid on both grids
    for i in range(len(which)):
        b.set_f(i, 2.0 * fs[i])                          # Q(2f) = 4 Q(f)
    b.set_probe_cells(cells)
    b.eval_q_probe()
    for i, g in enumerate(which):
        D = grids[g].D
        nodes = np.concatenate([np.arange(c * D, c * D + D) for c in sorted(set(cells))])
        q = b.get_q(i)
        other = np.setdiff1d(np.arange(grids[g].nnodes), nodes)
        assert np.array_equal(q[other], full[i][other])
        assert rel_l2(q[nodes], 4.0 * full[i][nodes]) <= 1e-6
    # FP64 projection of every instance: moments vanish to rounding
    b.eval_q()
    raw = [b.get_q(i) for i in range(len(which))]
    b.project_invariants()
    for i, g in enumerate(which):
        G = grids[g]
        q = b.get_q(i)
        psi = np.stack([np.ones(G.nnodes), G.nodes_u, G.nodes_v, G.nodes_w,
                        G.nodes_u ** 2 + G.nodes_v ** 2 + G.nodes_w ** 2])
        mom = (q * G.nodes_gwts) @ psi.T
        scale = np.abs(q * G.nodes_gwts) @ np.abs(psi.T)
        assert np.all(np.abs(mom) <= 1e-13 * scale)
        assert rel_l2(q, G.project_invariants(raw[i])) <= 1e-13
    b.close()


def test_batch_projection_in_rk_conserves_moments(grids):
    import dgv0d3v_direct_collision_amd as dgv
    rng = np.random.default_rng(21)
    which = np.array([0, 1, 1, 0, 1], np.int32)
    b = dgv.DGVBatch(grids, which, dgv.DGV_FP32_COMP)
    b.set_projection(True)
    fs = [vdf(grids[g], rng, 2) for g in which]
    for i, f in enumerate(fs):
        b.set_f(i, f)
    b.rk_run(0.02, 3, rkmts=3, use_graph=True)
    for i, g in enumerate(which):
        G = grids[g]
        psi = np.stack([np.ones(G.nnodes), G.nodes_u, G.nodes_v, G.nodes_w,
                        G.nodes_u ** 2 + G.nodes_v ** 2 + G.nodes_w ** 2])
        m0 = (fs[i] * G.nodes_gwts) @ psi.T
        m1 = (b.get_f(i) * G.nodes_gwts) @ psi.T
        scale = np.abs(fs[i] * G.nodes_gwts) @ np.abs(psi.T)
        assert np.all(np.abs(m1 - m0) <= 1e-13 * scale), (i, m1 - m0)
    b.close()


def test_batch_device_rows_match_per_instance_copies(grids):
    import dgv0d3v_direct_collision_amd as dgv
    rng = np.random.default_rng(33)
    which = np.array([1] * 70, np.int32)
    g = grids[1]
    b = dgv.DGVBatch(grids, which, dgv.DGV_FP64)
    F = np.stack([vdf(g, rng, 1 + i % 3) for i in range(70)])
    d_rows = g.device_alloc(F.nbytes)
    d_q = g.device_alloc(F.nbytes)
    g.htod(d_rows, F)
    b.set_f_rows(d_rows, 0, 70)
    for i in (0, 37, 69):
        assert np.array_equal(b.get_f(i), F[i])
    b.eval_q()
    b.get_q_rows(d_q, 0, 70)
    Q = np.zeros_like(F)
    g.dtoh(Q, d_q)
    for i in (0, 1, 63, 64, 69):
        assert np.array_equal(Q[i], b.get_q(i))
    g.device_free(d_rows)
    g.device_free(d_q)
    b.close()


@pytest.mark.parametrize("precision,tol", [("FP64", 1e-12), ("FP32_COMP", 1e-5), ("FP32_F64ACC", 1e-5),
                                           ("INT8_SPLIT", 1e-6)])
def test_two_nodes_per_dimension_batch_matches_single(precision, tol):
    """D = 8 (s = 2) grids, synthetic A-arrays with uneven blocks: the batched
    kernels' D = 8 instantiation against the single-VDF evaluator."""
    import dgv0d3v_direct_collision_amd as dgv
    import oracle
    from tests.helpers import synthetic_aarray
    print(f"\n[batch] s=2 grids M=4 and M=5, {precision}", flush=True)
    gs = []
    for M, box, seed in ((4, (-2.0, 2.0), 5), (5, (-2.5, 2.5), 6)):
        G = oracle.Geometry(M, 2, box)
        g = dgv.DGV(M, 2, box)
        g.set_aarray(*synthetic_aarray(G, seed=seed))
        gs.append(g)
    rng = np.random.default_rng(12)
    which = np.array([0, 1] * 40 + [1] * 7, np.int32)
    b = dgv.DGVBatch(gs, which, getattr(dgv, "DGV_" + precision))
    fs = [vdf(gs[g], rng, 1 + i % 3) for i, g in enumerate(which)]
    for i, f in enumerate(fs):
        b.set_f(i, f)
    b.eval_q()
    worst = max(rel_l2(b.get_q(i), gs[g].EvalCollisionPeriodicAPlus_DGV(fs[i])) for i, g in enumerate(which))
    print(f"[batch] s=2 {precision}: worst rel L2 {worst:.2e}", flush=True)
    assert worst <= tol
    b.close()
    for g in gs:
        g.close()


@pytest.mark.parametrize("precision", ["FP64", "INT8_SPLIT"])
def test_autotune_keeps_q(grids, precision):
    """dgv_batch_autotune times segment sizes 2^14..2^21 and keeps one of them;
    Q after it agrees with Q from the default plan (sums regrouped: 1e-12)."""
    import dgv0d3v_direct_collision_amd as dgv
    rng = np.random.default_rng(13)
    which = np.array([0, 1] * 20, np.int32)
    b = dgv.DGVBatch(grids, which, getattr(dgv, "DGV_" + precision))
    for i, g in enumerate(which):
        b.set_f(i, vdf(grids[g], rng, 2))
    b.eval_q()
    q0 = [b.get_q(i) for i in range(len(which))]
    best, ms = b.autotune()
    print(f"[batch] {precision} autotune: {best} pairs per segment, {ms:.3f} ms", flush=True)
    assert best in [1 << k for k in range(14, 22)] and ms > 0
    b.eval_q()
    for i in range(len(which)):
        assert rel_l2(b.get_q(i), q0[i]) <= 1e-12
    b.close()
