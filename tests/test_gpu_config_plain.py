"""The plain CPU-runnable workload end to end on the GPU: M = 5, s = 3 on
[-3.5, 3.5]^3 (N = 3375), the A-array from the GPU generator, the seeded
bimodal VDF f = n1 M(T1, u1) + (1 - n1) M(T2, u2), one Q evaluation and 100
RK2 steps at dt = 0.01 (201 Q evaluations, TimeIntegratorORK_SH_DGV with
rkmts = 2, time_integr_mod.f90:55-59, and (d/L_inf)^2 N_inf = 1).

 * Q(f0) against the oracle (EvalCollisionPeriodicAPlus_DGV restated,
   collision_mod.f90:143-229) on the exported A-array: relative L2 <= 1e-12;
 * the first two RK2 steps against the oracle's RK2 on the oracle evaluator:
   relative L2 <= 1e-12 after each;
 * all 100 steps: a second eager run equals the first bit for bit, and the
   hipGraph run (rk_run_dev) agrees to 1e-13;
 * Q at the state after 100 steps against the oracle: relative L2 <= 1e-12, so
   the evaluator is pinned along the whole trajectory, not only at its start.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import bimodal, rel_l2

pytestmark = pytest.mark.gpu

M, S, BOX = 5, 3, (-3.5, 3.5)
SEED = 20261202
DT, STEPS = 0.01, 100


@pytest.fixture(scope="module")
def plain():
    import dgv0d3v_direct_collision_amd as dgv
    print(f"\n[plain] M={M} s={S} box {BOX}: generating the A-array", flush=True)
    d = dgv.DGV(M, S, BOX)
    d.generate_aarray(min_sens=0.0, nquad=24)
    d.rkmts, d.coef = 2, 1.0
    A = d.get_aarray()
    G = oracle.Geometry(M, S, BOX)
    print(f"[plain] {len(A[1])} records", flush=True)
    f0 = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, SEED)
    q = lambda x: oracle.eval_aplus(G, *A, x, nthreads=16)
    yield d, G, q, f0
    d.close()


def test_first_q_and_rk2_steps_match_oracle(plain):
    d, G, q, f0 = plain
    assert rel_l2(d.EvalCollisionPeriodicAPlus_DGV(f0), q(f0)) <= 1e-12
    fo, fg = f0.copy(), f0.copy()
    for k in range(2):
        fo, _ = oracle.rk_step(q, fo, DT, 1.0, 2)
        d.TimeIntegratorORK_SH_DGV(fg, DT)
        print(f"[plain] RK2 step {k + 1}: rel L2 vs oracle {rel_l2(fg, fo):.2e}", flush=True)
        assert rel_l2(fg, fo) <= 1e-12, k


def test_hundred_rk2_steps(plain):
    d, G, q, f0 = plain
    runs = []
    for _ in range(2):
        f = f0.copy()
        for _ in range(STEPS):
            d.TimeIntegratorORK_SH_DGV(f, DT)
        runs.append(f)
    assert np.array_equal(runs[0], runs[1])
    f100 = runs[0]
    assert np.all(np.isfinite(f100)) and rel_l2(f100, f0) > 1e-4
    n = len(f0)
    ptr = d.device_alloc(8 * n)
    try:
        d.htod(ptr, f0)
        d.rk_run_dev(ptr, DT, STEPS, rkmts=2, use_graph=True)
        d.synchronize()
        fgraph = np.zeros(n)
        d.dtoh(fgraph, ptr)
    finally:
        d.device_free(ptr)
    assert rel_l2(fgraph, f100) <= 1e-13
    # the evaluator at the end of the trajectory
    assert rel_l2(d.EvalCollisionPeriodicAPlus_DGV(f100), q(f100)) <= 1e-12
    m0, m1 = (np.round(np.array(d.MassCheckRec(x), float), 6).tolist() for x in (f0, f100))
    print(f"[plain] (n, u, v, w, T) {m0} -> {m1} after {STEPS} steps", flush=True)
