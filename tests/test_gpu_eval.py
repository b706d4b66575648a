"""GPU parity of the collision evaluator against the oracle (restated
EvalCollisionPeriodicAPlus_DGV / EvalCollisionPeriodicA_DGV, collision_mod.f90).

Tolerance: relative L2 <= 1e-12 (FP64; the GPU sums in a different order).
"""
import numpy as np
import pytest

import oracle
from tests.helpers import bimodal, rel_l2, synthetic_aarray

pytestmark = pytest.mark.gpu
TOL = 1e-12

CASES = [(3, 3, (-1.5, 1.5)), (4, 3, (-2.0, 2.0)), (5, 2, (-3.5, 3.5)), (3, 1, (-1.5, 1.5)),
         (5, 3, (-3.5, 3.5))]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"M{c[0]}s{c[1]}")
def case(request):
    import dgv0d3v_direct_collision_amd as dgv
    M, s, box = request.param
    print(f"\n[case] M={M} s={s} box={box}", flush=True)
    G = oracle.Geometry(M, s, box)
    A = synthetic_aarray(G, seed=10 * M + s)
    d = dgv.DGV(M, s, box)
    d.set_aarray(*A)
    f = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, seed=M + 100 * s)
    ref = oracle.eval_aplus(G, *A, f, nthreads=16)
    yield G, A, d, f, ref
    d.close()


def test_geometry_matches_oracle(case):
    G, A, d, f, ref = case
    assert d.ccell == G.ccell
    assert np.array_equal(d.nodes_gwts, G.nodes_gwts)
    assert np.array_equal(d.nodes_u, G.nodes_u)
    ph, du, dv, dw = d.shifts()
    assert np.array_equal(ph, G.nodes_phican)
    assert np.array_equal(du, G.nodes_dui) and np.array_equal(dw, G.nodes_dwi)


@pytest.mark.parametrize("kernel", ["PAIRBLOCK", "MFMA"])
def test_pairblock_kernels(case, kernel):
    import dgv0d3v_direct_collision_amd as dgv
    G, A, d, f, ref = case
    d.set_kernel(getattr(dgv, "DGV_KERNEL_" + kernel))
    q = d.EvalCollisionPeriodicAPlus_DGV(f)
    d.set_kernel(dgv.DGV_KERNEL_AUTO)
    assert rel_l2(q, ref) <= TOL, rel_l2(q, ref)


def test_record_kernel(case):
    import dgv0d3v_direct_collision_amd as dgv
    G, A, d, f, ref = case
    d.set_kernel(dgv.DGV_KERNEL_RECORD)
    q = d.EvalCollisionPeriodicAPlus_DGV(f)
    d.set_kernel(dgv.DGV_KERNEL_AUTO)
    assert rel_l2(q, ref) <= TOL, rel_l2(q, ref)


def test_eval_a_variant(case):
    G, A, d, f, ref = case
    ref_a = oracle.eval_a(G, *A, f, nthreads=16)
    q = d.EvalCollisionPeriodicA_DGV(f)
    assert rel_l2(q, ref_a) <= TOL


def test_tiles_and_determinism(case):
    G, A, d, f, ref = case
    q0 = d.EvalCollisionPeriodicAPlus_DGV(f)
    q1 = d.EvalCollisionPeriodicAPlus_DGV(f)
    assert np.array_equal(q0, q1)  # fixed summation order
    for t in (64, 500, 100000):
        d.set_tile(t)
        q = d.EvalCollisionPeriodicAPlus_DGV(f)
        assert rel_l2(q, ref) <= TOL
    d.set_tile(2048)


def test_spatial_operator_scaling(case):
    G, A, d, f, ref = case
    d.coef = 0.37
    fcol = d.SpatialOperSH_DGV(f, 0.01)
    d.coef = 1.0
    assert rel_l2(fcol, ref * (0.37 * 0.01)) <= TOL


def test_homogeneity(case):
    # Q(alpha f) = alpha^2 Q(f): bilinear form
    G, A, d, f, ref = case
    q2 = d.EvalCollisionPeriodicAPlus_DGV(3.0 * f)
    assert rel_l2(q2, 9.0 * ref) <= TOL
