"""Pins of the CPU oracle (no GPU).  The reference ships no tests or golden
vectors, so every restated function is checked against what the mathematics
fixes: closed forms, invariants, brute force on tiny inputs, and each
evaluator variant against an independent formulation.
"""
import math

import numpy as np
import pytest

import oracle
from tests.helpers import bimodal, synthetic_aarray


# ---- GauLeg (gaussian_mod.f90:4-53) -----------------------------------------------
def test_gauleg_three_point_closed_form():
    x, w = oracle.gauleg(3)
    assert np.allclose(x, [-math.sqrt(0.6), 0.0, math.sqrt(0.6)], atol=1e-15, rtol=0)
    assert np.allclose(w, [5 / 9, 8 / 9, 5 / 9], atol=1e-14, rtol=0)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 8])
def test_gauleg_exact_for_degree_2n_minus_1(n):
    x, w = oracle.gauleg(n, -1.0, 2.0)
    for k in range(2 * n):
        exact = (2.0 ** (k + 1) - (-1.0) ** (k + 1)) / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) <= 1e-13 * max(1, abs(exact))


# ---- geometry (miscset.f90:1112-1311) ------------------------------------------------
@pytest.mark.parametrize("M,s,box", [(5, 3, (-3.5, 3.5)), (4, 2, (-2.0, 2.0)), ((3, 5, 7), 3, (-1, 1, -2, 2, -3, 3))])
def test_geometry_invariants(M, s, box):
    G = oracle.Geometry(M, s, box)
    Ms = (M,) * 3 if np.isscalar(M) else M
    b = box if len(box) == 6 else box * 3
    vol = (b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4])
    assert abs(G.nodes_gwts.sum() - vol) <= 1e-12 * vol
    # node order: cell u outer ... w inner, then local u,v,w
    D = int(np.prod(G.s))
    assert np.all(np.diff(G.nodes_pcell) >= 0)
    cell = np.arange(G.nnodes) // D
    assert np.array_equal(G.nodes_pcell - 1, cell)
    cu = cell // (Ms[1] * Ms[2])
    assert np.array_equal(G.cells_ugi[cell] - 1, cu)
    # canonical cell contains the origin; phican / shift formulas (lines 1298-1307)
    cc = G.ccell - 1
    assert G.cells_lu[cc] <= 0 <= G.cells_ru[cc] and G.cells_lw[cc] <= 0 <= G.cells_rw[cc]
    loc = np.arange(G.nnodes) % D
    assert np.array_equal(G.nodes_phican - 1, cc * D + loc)
    assert np.array_equal(G.nodes_dui, G.cells_ugi[cell] - G.cells_ugi[cc])
    # symmetric box: first moments of the weights vanish
    if len(box) == 2:
        assert abs(np.sum(G.nodes_gwts * G.nodes_u)) <= 1e-12 * vol


def test_canonical_cell_even_grid_is_right_of_origin():
    G = oracle.Geometry(4, 3, (-2.0, 2.0))
    cc = G.ccell - 1
    assert G.cells_lu[cc] == 0.0 and G.cells_lv[cc] == 0.0 and G.cells_lw[cc] == 0.0


def test_product_geometry_is_bit_identical():
    import dgv0d3v_direct_collision_amd as dgv
    for M, s, box in [(5, 3, (-3.5, 3.5)), (9, 3, (-4.0, 4.0)), (4, 2, (-2.0, 2.0))]:
        G = oracle.Geometry(M, s, box)
        u, v, w, gw, cc = dgv.grid_geometry(M, s, box)
        assert cc == G.ccell
        assert np.array_equal(u, G.nodes_u) and np.array_equal(w, G.nodes_w)
        assert np.array_equal(gw, G.nodes_gwts)


# ---- the evaluator (collision_mod.f90:29-229) ---------------------------------------
@pytest.fixture(scope="module")
def small():
    M, s, box = 3, 3, (-1.5, 1.5)
    G = oracle.Geometry(M, s, box)
    A = synthetic_aarray(G, seed=1)
    f = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, seed=2)
    return G, A, f


def test_single_record_closed_form():
    G = oracle.Geometry(3, 2, (-1.5, 1.5))
    D = 8
    cc = G.ccell - 1
    l = 5
    xi, xi1, a = cc * D + 1 + 2, (cc + 1) * D + 1 + 6, 0.37   # 1-based nodes
    capphi = np.zeros(G.nnodes, np.int32)
    capphi[cc * D + l] = 1
    f = 1.0 + np.arange(G.nnodes) / G.nnodes
    Q = oracle.eval_aplus(G, capphi, [a], [xi], [xi1], f)
    M = 3
    for i in range(G.nnodes):
        c, li = divmod(i, D)
        if li != l:
            assert Q[i] == 0.0
            continue
        cu, cv, cw = c // 9, (c // 3) % 3, c % 3
        d = np.array([cu, cv, cw]) - np.array([cc // 9, (cc // 3) % 3, cc % 3])
        ok, idx = True, []
        for n in (xi - 1, xi1 - 1):
            nc, nl = divmod(n, D)
            p = np.array([nc // 9, (nc // 3) % 3, nc % 3]) + d
            if np.any(p < 0) or np.any(p >= M):
                ok = False
            idx.append(((p[0] * 3 + p[1]) * 3 + p[2]) * D + nl)
        expect = 2 * f[idx[0]] * f[idx[1]] * a / G.nodes_gwts[i] if ok else 0.0
        assert Q[i] == pytest.approx(expect, rel=1e-15, abs=0)


def test_homogeneity_bitwise(small):
    G, A, f = small
    q1 = oracle.eval_aplus(G, *A, f)
    q2 = oracle.eval_aplus(G, *A, 2.0 * f)
    assert np.array_equal(q2, 4.0 * q1)


def test_shift_equivariance_bitwise():
    # f supported away from the boundary, translated by one cell in w:
    # Q of the translated f is Q translated (same records, same order)
    M, s, box = 5, 2, (-2.5, 2.5)
    G = oracle.Geometry(M, s, box)
    A = synthetic_aarray(G, seed=4, drop=0.0)
    D = 8
    cells = np.arange(G.nnodes) // D
    cu, cv, cw = cells // 25, (cells // 5) % 5, cells % 5
    inner = (cu >= 1) & (cu <= 3) & (cv >= 1) & (cv <= 3) & (cw >= 1) & (cw <= 2)
    f = np.where(inner, 1.0 + np.sin(np.arange(G.nnodes)), 0.0)
    fs = np.zeros_like(f)
    fs[D:] = f[:-D]                                 # cell (u,v,w) -> (u,v,w+1)
    q = oracle.eval_aplus(G, *A, f)
    qs = oracle.eval_aplus(G, *A, fs)
    interior = (cu >= 1) & (cu <= 3) & (cv >= 1) & (cv <= 3) & (cw >= 1) & (cw <= 3)
    idx = np.where(interior)[0]
    # nodes of equal local index share the weight, so Q shifts exactly
    assert np.array_equal(qs[idx], q[idx - D])


def test_full_a_equivalence_with_iphi(small):
    # IphiDGV (dgvtools_mod.f90:847-877) over the canonical records replicated
    # to every cell (shift + bounds test applied once) equals the shifted evaluator
    G, (capphi, A, xi, xi1), f = small
    fullcap, fa, fx, fx1 = oracle.full_aarray(G, capphi, A, xi, xi1)
    q_full = oracle.iphi(fullcap, fa, fx, fx1, G.nodes_gwts, f)
    q = oracle.eval_aplus(G, capphi, A, xi, xi1, f)
    # same records in the same order, products grouped differently: rounding only
    assert np.linalg.norm(q_full - q) <= 1e-13 * np.linalg.norm(q)


def test_iphi_decomp_is_the_split_of_iphi(small):
    # IphiDGV_decomp (dgvtools_mod.f90:896-927) sums A (fM Df + Df fM + Df Df) per
    # record: with fM = 0 it is IphiDGV of f, and in general IphiDGV(f) - IphiDGV(fM)
    # (Q is a symmetric bilinear form); a dropped or doubled term fails one of them
    G, (capphi, A, xi, xi1), f = small
    full = oracle.full_aarray(G, capphi, A, xi, xi1)
    gw = G.nodes_gwts
    fM = 0.3 * np.cos(np.arange(G.nnodes)) ** 2
    d0 = oracle.iphi_decomp(*full, gw, f, np.zeros_like(f), 2.0)
    assert np.linalg.norm(d0 - oracle.iphi(*full, gw, f, 2.0)) <= 1e-14 * np.linalg.norm(d0)
    d1 = oracle.iphi_decomp(*full, gw, f, fM, 2.0)
    diff = oracle.iphi(*full, gw, f, 2.0) - oracle.iphi(*full, gw, fM, 2.0)
    assert np.linalg.norm(d1 - diff) <= 1e-12 * np.linalg.norm(oracle.iphi(*full, gw, f, 2.0))


def test_swap_and_variant_agreement(small):
    G, (capphi, A, xi, xi1), f = small
    q = oracle.eval_aplus(G, capphi, A, xi, xi1, f)
    q_sw = oracle.eval_aplus(G, capphi, A, xi1, xi, f)
    q_a = oracle.eval_a(G, capphi, A, xi, xi1, f)
    q_c, ab = oracle.eval_aplus_comp(G, capphi, A, xi, xi1, f)
    sc = np.abs(q).max()
    assert np.abs(q_sw - q).max() <= 1e-13 * sc
    assert np.abs(q_a - q).max() <= 1e-13 * sc
    assert np.abs(q_c - q).max() <= 1e-13 * sc
    assert np.all(ab >= np.abs(q_c) * (1 - 1e-12))


def test_openmp_order_independent(small):
    G, A, f = small
    assert np.array_equal(oracle.eval_aplus(G, *A, f, nthreads=1), oracle.eval_aplus(G, *A, f, nthreads=8))


# ---- moments and Maxwellian (dgvtools_mod.f90:1437-1458, distributions_mod.f90:83-94) ----
def test_maxwellian_moments_converge():
    # MassCheckRec of a Maxwellian returns (n, ubar, T) up to the O(h^6)
    # error of the 3-point Gauss rule per cell
    n, ub, T = 1.3, np.array([0.2, -0.3, 0.1]), 0.7
    errs = []
    for M, L in ((9, 4.5), (18, 4.5)):
        G = oracle.Geometry(M, 3, (-L, L))
        b = math.sqrt(math.pi * T) * (math.pi * T)
        f = n * np.exp(-((G.nodes_u - ub[0]) ** 2 + (G.nodes_v - ub[1]) ** 2 + (G.nodes_w - ub[2]) ** 2) / T) / b
        errs.append(np.abs(oracle.masscheck(G, f) - np.array([n, *ub, T])).max())
    assert errs[1] < 3e-6 and errs[0] / errs[1] > 30


def test_maxwell_closed_form():
    v = oracle.maxwell(0.5, 0.1, 0.0, 0.0, 2.0, 0.3, 0.2, -0.1)
    expect = 2.0 * math.exp(-((0.2) ** 2 + 0.2 ** 2 + 0.1 ** 2) / 0.5) / (math.pi * 0.5) ** 1.5
    assert v == pytest.approx(expect, rel=1e-15)


# ---- steppers (time_integr_mod.f90:26-152, 314-364) --------------------------------
@pytest.mark.parametrize("rk,order", [(1, 1), (2, 2), (3, 3), (4, 4), (5, 5)])
def test_rk_amplification_factor(rk, order):
    lam = -1.3
    f0 = np.array([1.0, 2.0])
    errs = []
    for dt in (0.1, 0.05):
        f1, r1 = oracle.rk_step(lambda x: lam * x, f0, dt, 1.0, rk)
        z = lam * dt
        errs.append(np.abs(f1 / f0 - math.exp(z)).max())
        # frhs1 = first stage / dt = lam f0
        assert np.allclose(r1, lam * f0, rtol=1e-14)
    rate = math.log2(errs[0] / errs[1])
    assert rate > order + 0.7, (rate, order)


def test_rk_stage_polynomials_exact():
    # the truncated series of exp(z) for RK2/RK3 (Ralston coefficients)
    lam, dt = -0.7, 0.3
    z = lam * dt
    f0 = np.array([1.0])
    f2, _ = oracle.rk_step(lambda x: lam * x, f0, dt, 1.0, 2)
    f3, _ = oracle.rk_step(lambda x: lam * x, f0, dt, 1.0, 3)
    assert f2[0] == pytest.approx(1 + z + z * z / 2, rel=1e-15)
    assert f3[0] == pytest.approx(1 + z + z * z / 2 + z ** 3 / 6, rel=1e-15)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_ab_integrates_polynomials_exactly(k):
    # y' = t^(k-1): the k-step Adams-Bashforth rule is exact.  Before the step
    # frhs1..frhs4 hold the derivatives at t0-dt, t0-2dt, ... (line 362 leaves
    # them divided by dt); the spatial operator returns dt*y'(t0).
    dt, t0 = 0.1, 1.0
    deg = k - 1
    yp = lambda t: t ** deg
    hist = np.zeros((9, 1))
    for j in range(5):
        hist[4 + j] = yp(t0 - (j + 1) * dt)
    y0 = np.array([0.5])
    y1, h1 = oracle.ab_step(lambda x: np.array([yp(t0)]), y0, hist, dt, 1.0, k)
    exact = y0[0] + ((t0 + dt) ** (deg + 1) - t0 ** (deg + 1)) / (deg + 1)
    assert y1[0] == pytest.approx(exact, rel=1e-13)
    assert h1[4][0] == pytest.approx(yp(t0), rel=1e-15)   # frhs1 = increment / dt
