"""PrepareMTS_SH_DGV (time_integr_mod.f90:443-569) and the multistep step after
it, against the oracle: the rkmts - 1 Runge-Kutta steps with their history
pushes (time_integr_mod.f90:536-548) driven here on the oracle's rk_step, then one
TimeIntegratorMTS_SH_DGV step (time_integr_mod.f90:314-364) from that history on
the oracle's ab_step.  The solution and all nine history rows (f1..f4,
frhs1..frhs5) agree to 1e-12 relative after the preparation and after the step,
for rkmts = 2..5, and e_time has advanced by rkmts - 1 steps.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import bimodal, rel_l2, synthetic_aarray

pytestmark = pytest.mark.gpu
M, S, BOX = 3, 3, (-1.5, 1.5)


@pytest.fixture(scope="module")
def env():
    import dgv0d3v_direct_collision_amd as dgv
    print("\n[prepare mts] M=3 s=3", flush=True)
    G = oracle.Geometry(M, S, BOX)
    A = synthetic_aarray(G, seed=11)
    d = dgv.DGV(M, S, BOX)
    d.set_aarray(*A)
    f = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, seed=5)
    q = lambda x: oracle.eval_aplus(G, *A, x, nthreads=8)
    yield d, f, q
    d.close()


def push(h, k, f):
    """The history push before each preparation step (time_integr_mod.f90:536-548);
    rows f1..f4 = h[0..3], frhs1..frhs5 = h[4..8]."""
    if k == 2:
        h[0] = f
    elif k == 3:
        h[1] = h[0]; h[0] = f; h[5] = h[4]
    elif k == 4:
        h[2] = h[1]; h[1] = h[0]; h[0] = f; h[6] = h[5]; h[5] = h[4]
    elif k == 5:
        h[3] = h[2]; h[2] = h[1]; h[1] = h[0]; h[0] = f
        h[7] = h[6]; h[6] = h[5]; h[5] = h[4]


def same_rows(got, want):
    for r in range(9):
        if np.any(want[r]):
            assert rel_l2(got[r], want[r]) <= 1e-12, r
        else:
            assert not np.any(got[r]), r


@pytest.mark.parametrize("rkmts", [2, 3, 4, 5])
def test_prepare_mts_then_multistep_step(env, rkmts):
    d, f0, q = env
    print(f"[prepare mts] rkmts {rkmts}", flush=True)
    dt, coef = 0.02, 0.9
    d.rkmts, d.coef = rkmts, coef
    d.hist[:] = 0.0
    d.frhs1[:] = 0.0
    d.e_time = 0.0
    f = f0.copy()
    d.PrepareMTS_SH_DGV(f, dt)

    h = np.zeros((9, len(f0)))
    fo = f0.copy()
    for _ in range(2, rkmts + 1):
        push(h, rkmts, fo.copy())
        fo, h[4] = oracle.rk_step(q, fo, dt, coef, rkmts)
    assert rel_l2(f, fo) <= 1e-12
    same_rows(d.hist, h)
    assert rel_l2(d.frhs1, h[4]) <= 1e-12
    assert abs(d.e_time - (rkmts - 1) * dt) <= 1e-15

    d.TimeIntegratorMTS_SH_DGV(f, dt)
    fo2, h2 = oracle.ab_step(q, fo, h, dt, coef, rkmts)
    print(f"[prepare mts] rkmts {rkmts}: multistep step {rel_l2(f, fo2):.2e}", flush=True)
    assert rel_l2(f, fo2) <= 1e-12
    same_rows(d.hist, h2)
    assert abs(d.e_time - rkmts * dt) <= 1e-15
