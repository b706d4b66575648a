"""dgv_eval_q_dev_noalloc: the single-VDF evaluator with its partial sums in
the library's static device array instead of memory the library allocates.
Through the C ABI, against dgv_eval_q_dev (the same kernels and sum order, so
bit-identical), on two calls in a row -- a 15^3 grid, then a smaller 9^3 grid
sharing the same static array -- with and without the FP64 projection, and a
plan whose partial sums exceed the array (21^3 with one block per work item)
is refused before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    import dgv0d3v_direct_collision_amd as dgv
    return dgv.load().dgv_eval_q_dev_noalloc


def _f(g, seed):
    rng = np.random.default_rng(seed)
    from dgv0d3v_direct_collision_amd.workloads import mixture_vdf
    return mixture_vdf(g.nodes_u, g.nodes_v, g.nodes_w, g.nodes_gwts, rng, 3)


def test_noalloc_entry_point_equals_eval_q_dev():
    import dgv0d3v_direct_collision_amd as dgv
    fn = _lib()
    print("\n[noalloc] grids 15^3 then 9^3", flush=True)
    grids = [dgv.DGV(5, 3, (-3.5, 3.5)), dgv.DGV(3, 3, (-1.5, 1.5))]
    try:
        for g in grids:
            g.generate_aarray(nquad=16)
        for project in (False, True):
            for k, g in enumerate(grids):
                g.set_projection(project)
                f = _f(g, 10 + k)
                n = g.nnodes
                d_f, d_q1, d_q2 = g.device_alloc(8 * n), g.device_alloc(8 * n), g.device_alloc(8 * n)
                g.htod(d_f, f)
                g.eval_q_dev(d_f, d_q1)
                rc = fn(g.ctx, d_f, d_q2, n)
                assert rc == 0, dgv.load().dgv_last_error(g.ctx)
                g.synchronize()
                q1, q2 = np.zeros(n), np.zeros(n)
                g.dtoh(q1, d_q1)
                g.dtoh(q2, d_q2)
                print(f"[noalloc] N={n} projection={project}: max |dq| {np.max(np.abs(q1 - q2)):.3e}", flush=True)
                assert np.array_equal(q1, q2)
                for p in (d_f, d_q1, d_q2):
                    g.device_free(p)
            for g in grids:
                g.set_projection(False)
    finally:
        for g in grids:
            g.close()
    # a plan larger than the static array (21^3, one block per item) is refused
    # before any launch: nothing is written
    g = dgv.DGV(7, 3, (-4.0, 4.0))
    try:
        g.generate_aarray(nquad=16)
        g.set_tile(1)
        n = g.nnodes
        d_f, d_q = g.device_alloc(8 * n), g.device_alloc(8 * n)
        g.htod(d_f, _f(g, 3))
        sentinel = np.full(n, 7.0)
        g.htod(d_q, sentinel)
        assert fn(g.ctx, d_f, d_q, n) == -8
        g.synchronize()
        out = np.zeros(n)
        g.dtoh(out, d_q)
        assert np.array_equal(out, sentinel)
        assert fn(g.ctx, d_f, d_q, n - 1) == -1
        g.device_free(d_f)
        g.device_free(d_q)
    finally:
        g.close()
