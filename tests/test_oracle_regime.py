"""Pins of the oracle's restatement of CheckSolutionMode_DGV and of the
run_mode switch of SpatialOperSH_DGV (oracle.Regime), no GPU.

 * far from equilibrium (L1 >= decomp_lev) the term is Q(f) and fmA is never built;
 * run_mode 1 returns Q(Df) + L(Df), which equals Q(f) - Q(fm) because Q is a
   quadratic form and L(Df) = Mixed(Df, fm) is its derivative at fm;
 * run_mode 2 returns L(Df) = Mixed(Df, fm), by the independent mixed-term route;
 * a scripted sequence of VDFs with known L1 deviations walks the transitions
   0 -> 1 -> 2 -> 1 -> 0 in the reference's order, and fmA is rebuilt exactly
   when the local Maxwellian moves by more than linear_lev / 2 in L1.
"""
import numpy as np

import oracle
from tests.helpers import bimodal, rel_l2, synthetic_aarray


def _setup():
    G = oracle.Geometry(3, 2, (-2.5, 2.5))
    A = synthetic_aarray(G, seed=11)
    return G, A


def _maxwellian(G, dens, u0, T):
    return np.array([oracle.maxwell(T, u0[0], u0[1], u0[2], dens, u, v, w)
                     for u, v, w in zip(G.nodes_u, G.nodes_v, G.nodes_w)])


def _l1_dev(G, f):
    """relative L1 distance of f from its own local Maxwellian"""
    dens, ub, vb, wb, T = oracle.masscheck(G, f)
    fmn = _maxwellian(G, dens, (ub, vb, wb), T)
    return float(np.sum(np.abs(f - fmn) * G.nodes_gwts) / dens), fmn


def _perturbed(G, eps, u0=(0.1, -0.2, 0.05), seed=4):
    rng = np.random.default_rng(seed)
    base = _maxwellian(G, 1.0, u0, 0.9)
    pert = rng.uniform(-1.0, 1.0, G.nnodes) * base
    return base + eps * pert


def test_far_from_equilibrium_is_plain_q():
    G, A = _setup()
    f = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, 3)
    l1, _ = _l1_dev(G, f)
    r = oracle.Regime(G, *A, decomp_lev=0.5 * l1, linear_lev=0.1 * l1)
    c = r.collision(f)
    assert r.run_mode == 0 and r.fma_builds == 0 and r.fmA is None
    assert np.array_equal(c, oracle.eval_aplus(G, *A, f, nthreads=8))


def test_decomposition_mode_is_q_of_f_minus_q_of_fm():
    G, A = _setup()
    f = _perturbed(G, 0.2)
    l1, _ = _l1_dev(G, f)
    r = oracle.Regime(G, *A, decomp_lev=2.0 * l1, linear_lev=0.5 * l1)
    c = r.collision(f)
    assert r.run_mode == 1 and r.fma_builds == 1
    qf = oracle.eval_aplus(G, *A, f)
    qm = oracle.eval_aplus(G, *A, r.fm)
    scale = np.linalg.norm(qf) + np.linalg.norm(qm)
    assert np.linalg.norm(c - (qf - qm)) <= 1e-12 * scale


def test_linear_mode_is_the_mixed_term():
    G, A = _setup()
    f = _perturbed(G, 0.01)
    l1, fmn = _l1_dev(G, f)
    r = oracle.Regime(G, *A, decomp_lev=4.0 * l1, linear_lev=2.0 * l1)
    c = r.collision(f)
    assert r.run_mode == 2 and r.fma_builds == 1
    assert np.array_equal(r.fm, fmn)
    df = f - fmn
    assert rel_l2(c, oracle.eval_mixed(G, *A, df, fmn)) <= 1e-12


def test_transition_sequence_and_fma_refresh():
    G, A = _setup()
    eps = [0.5, 0.2, 0.2, 0.02, 0.02, 0.2, 0.6]
    l1s = [_l1_dev(G, _perturbed(G, e))[0] for e in eps]
    assert all(np.diff(l1s[:4]) <= 0)
    # decomp between the 0.5 and 0.2 deviations, linear between 0.2 and 0.02
    decomp = 0.5 * (l1s[0] + l1s[1])
    linear = 0.5 * (l1s[2] + l1s[3])
    r = oracle.Regime(G, *A, decomp_lev=decomp, linear_lev=linear)
    modes = []
    for e in eps:
        r.collision(_perturbed(G, e))
        modes.append(r.run_mode)
    assert modes == [0, 1, 1, 2, 2, 1, 0]
    # the same Maxwellian throughout: built once on entering run_mode 1
    assert r.fma_builds == 1
    # a Maxwellian that moves by more than linear_lev / 2 rebuilds fmA
    r2 = oracle.Regime(G, *A, decomp_lev=decomp, linear_lev=linear)
    r2.collision(_perturbed(G, 0.02))
    assert r2.run_mode == 2 and r2.fma_builds == 1
    fm0 = r2.fm.copy()
    r2.collision(_perturbed(G, 0.02, u0=(0.1, -0.2, 0.06)))      # tiny shift: kept
    assert r2.fma_builds == 1 and np.array_equal(r2.fm, fm0)
    r2.collision(_perturbed(G, 0.02, u0=(0.6, -0.2, 0.05)))      # large shift: rebuilt
    assert r2.fma_builds == 2 and not np.array_equal(r2.fm, fm0)
