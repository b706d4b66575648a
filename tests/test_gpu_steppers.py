"""GPU steppers, moments and projection against the oracle's restatement of
TimeIntegratorORK_SH_DGV / TimeIntegratorMTS_SH_DGV / MassCheckRec."""
import numpy as np
import pytest

import oracle
from tests.helpers import bimodal, rel_l2, synthetic_aarray

pytestmark = pytest.mark.gpu
M, S, BOX = 3, 3, (-1.5, 1.5)


@pytest.fixture(scope="module")
def env():
    import dgv0d3v_direct_collision_amd as dgv
    print("\n[steppers] M=3 s=3", flush=True)
    G = oracle.Geometry(M, S, BOX)
    A = synthetic_aarray(G, seed=7)
    d = dgv.DGV(M, S, BOX)
    d.set_aarray(*A)
    f = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, seed=3)
    q = lambda x: oracle.eval_aplus(G, *A, x, nthreads=8)
    yield G, A, d, f, q
    d.close()


@pytest.mark.parametrize("rk", [1, 2, 3, 4, 5])
def test_rk_step(env, rk):
    G, A, d, f, q = env
    dt, coef = 0.05, 0.8
    fr, r1 = oracle.rk_step(q, f, dt, coef, rk)
    d.rkmts, d.coef = rk, coef
    fg = f.copy()
    d.TimeIntegratorORK_SH_DGV(fg, dt)
    assert rel_l2(fg, fr) <= 1e-12
    assert rel_l2(d.frhs1, r1) <= 1e-12


@pytest.mark.parametrize("rk", [1, 2, 3, 4, 5])
def test_ab_step(env, rk):
    G, A, d, f, q = env
    n = len(f)
    rng = np.random.default_rng(rk)
    hist = np.zeros((9, n))
    hist[:4] = f * (1 + 0.01 * rng.standard_normal((4, n)))
    hist[4:] = 1e-3 * rng.standard_normal((5, n))
    fr, hr = oracle.ab_step(q, f, hist, 0.02, 1.0, rk)
    d.rkmts, d.coef = rk, 1.0
    d.hist = hist.copy()
    d.frhs1 = hist[4].copy()
    fg = f.copy()
    d.TimeIntegratorMTS_SH_DGV(fg, 0.02)
    assert rel_l2(fg, fr) <= 1e-12
    assert rel_l2(d.hist[:rk], hr[:rk]) <= 1e-12 if rk > 1 else True
    assert rel_l2(d.frhs1, hr[4]) <= 1e-12


def test_rk_graph_replay_matches_eager(env):
    G, A, d, f, q = env
    n = len(f)
    d.coef = 1.0
    eager = f.copy()
    d.rkmts = 3
    for _ in range(4):
        d.TimeIntegratorORK_SH_DGV(eager, 0.03)
    ptr = d.device_alloc(8 * n)
    try:
        d.htod(ptr, f)
        d.rk_run_dev(ptr, 0.03, 2, rkmts=3, use_graph=True)
        d.synchronize()
        mid = np.zeros(n)
        d.dtoh(mid, ptr)                      # eager work between replays
        d.rk_run_dev(ptr, 0.03, 2, rkmts=3, use_graph=True)
        d.synchronize()
        out = np.zeros(n)
        d.dtoh(out, ptr)
    finally:
        d.device_free(ptr)
    assert rel_l2(out, eager) <= 1e-13


def test_masscheck(env):
    G, A, d, f, q = env
    ref = oracle.masscheck(G, f)
    got = np.array(d.MassCheckRec(f))
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-15)


def test_projection_zeroes_moments(env):
    G, A, d, f, q = env
    Q = d.EvalCollisionPeriodicAPlus_DGV(f)
    P = d.project_invariants(Q)
    psi = [np.ones_like(f), G.nodes_u, G.nodes_v, G.nodes_w,
           G.nodes_u ** 2 + G.nodes_v ** 2 + G.nodes_w ** 2]
    for p in psi:
        scale = np.sum(np.abs(G.nodes_gwts * p * Q))
        assert abs(np.sum(G.nodes_gwts * p * P)) <= 1e-13 * scale
    # projection is idempotent and only removes span{psi}
    P2 = d.project_invariants(P)
    assert rel_l2(P2, P) <= 1e-13
    mom = d.collision_moments(P)
    assert np.all(np.abs(mom) <= 1e-13 * np.sum(np.abs(G.nodes_gwts * Q)) * 10)


def test_projected_rk_graph_conserves_moments(env):
    # with the conservation-moment correction every Q has zero moments, so the
    # RK combinations keep density, momentum and energy to rounding
    G, A, d, f, q = env
    d.set_projection(True)
    try:
        N = len(f)
        d_f = d.device_alloc(8 * N)
        d.htod(d_f, f)
        d.rk_run_dev(d_f, 0.02, 4, rkmts=3, use_graph=True)
        d.synchronize()
        fr = np.zeros(N)
        d.dtoh(fr, d_f)
        d.device_free(d_f)
    finally:
        d.set_projection(False)
    psi = np.stack([np.ones(N), G.nodes_u, G.nodes_v, G.nodes_w, G.nodes_u ** 2 + G.nodes_v ** 2 + G.nodes_w ** 2])
    m0 = (f * G.nodes_gwts) @ psi.T
    m1 = (fr * G.nodes_gwts) @ psi.T
    scale = np.abs(f * G.nodes_gwts) @ np.abs(psi.T)
    assert np.all(np.abs(m1 - m0) <= 1e-13 * scale), m1 - m0
    assert rel_l2(fr, f) > 1e-6            # the VDF did change


def test_high_moments_match_oracle(env):
    G, A, d, f, q = env
    ref = oracle.masscheck_high(G, f)
    got = np.array(d.MassCheckRecHighCMoments(f))
    assert np.all(np.abs(got - ref) <= 1e-13 * np.maximum(np.abs(ref), 1e-3))
    assert np.allclose(got[:5], d.MassCheckRec(f), rtol=1e-14, atol=0)


def test_moment_derivative_record_reads_the_step_history(env):
    # after a multistep step, f1 is the solution the step started from and
    # frhs1 its right-hand side (DGVblzmPL.f90:326-330); the record made from
    # the device stepper's history equals the one made from the oracle's step
    from dgv0d3v_direct_collision_amd.driver_io import MomDerivLog, moment_derivatives
    G, A, d, f, q = env
    n = len(f)
    rng = np.random.default_rng(11)
    hist = np.zeros((9, n))
    hist[:4] = f * (1 + 0.01 * rng.standard_normal((4, n)))
    hist[4:] = 1e-3 * rng.standard_normal((5, n))
    fr, hr = oracle.ab_step(q, f, hist, 0.02, 1.0, 3)
    d.rkmts, d.coef = 3, 1.0
    d.hist = hist.copy()
    d.frhs1 = hist[4].copy()
    fg = f.copy()
    d.TimeIntegratorMTS_SH_DGV(fg, 0.02)
    log = MomDerivLog("unused", num_eval_error=10, rkmts=3, t_R=1.0)
    assert d.RecordMomDerivRelaxTimes_DGV(log, 0.0)
    ref = moment_derivatives(hr[0], hr[4], G.nodes_u, G.nodes_v, G.nodes_w, G.nodes_gwts)
    assert np.array_equal(hr[0], f)
    assert np.allclose(log.rows[0][1:21], ref[:20], rtol=1e-9, atol=1e-14)
