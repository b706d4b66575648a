"""Pins of the oracle's decomposition-mode routines, no GPU:
EvalCollisionPeriodicMixedTermsA_DGV, PrepareFMA_DGV and EvalCollisionLinear.

 * Mixed(f, f) = 2 Q(f) bit for bit (both sums see 2 f(xi) f(xi1) A per record);
 * Mixed is symmetric bit for bit and satisfies the polarisation identity
   Mixed(f, g) = Q(f + g) - Q(f) - Q(g) (Q is a quadratic form);
 * the dense linearised operator at fm applied to f equals Mixed(f, fm) -- the
   derivative of Q at fm in direction f -- computed by an independent route."""
import numpy as np

import oracle
from tests.helpers import bimodal, rel_l2, synthetic_aarray


def _setup(M=3, s=2, box=(-1.5, 1.5), seed=3):
    G = oracle.Geometry(M, s, box)
    cap, A, xi, xi1 = synthetic_aarray(G, seed=seed)
    f = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, 1)
    g = bimodal(G.nodes_u, G.nodes_v, G.nodes_w, 2)
    return G, cap, A, xi, xi1, f, g


def test_mixed_of_f_with_itself_is_twice_q():
    G, cap, A, xi, xi1, f, g = _setup()
    q = oracle.eval_aplus(G, cap, A, xi, xi1, f)
    m = oracle.eval_mixed(G, cap, A, xi, xi1, f, f)
    assert np.array_equal(m, 2.0 * q)


def test_mixed_symmetric_and_polarisation():
    G, cap, A, xi, xi1, f, g = _setup(M=3, s=3)
    m1 = oracle.eval_mixed(G, cap, A, xi, xi1, f, g)
    m2 = oracle.eval_mixed(G, cap, A, xi, xi1, g, f)
    assert np.array_equal(m1, m2)
    pol = (oracle.eval_aplus(G, cap, A, xi, xi1, f + g) - oracle.eval_aplus(G, cap, A, xi, xi1, f)
           - oracle.eval_aplus(G, cap, A, xi, xi1, g))
    assert rel_l2(m1, pol) <= 1e-12


def test_linearised_operator_is_the_mixed_term():
    G, cap, A, xi, xi1, f, fm = _setup(M=3, s=2)
    fmA = oracle.prepare_fma(G, cap, A, xi, xi1, fm)
    lin = oracle.eval_linear(fmA, f)
    assert rel_l2(lin, oracle.eval_mixed(G, cap, A, xi, xi1, f, fm)) <= 1e-12
    # at f = fm: 2 Q(fm)
    assert rel_l2(oracle.eval_linear(fmA, fm), 2.0 * oracle.eval_aplus(G, cap, A, xi, xi1, fm)) <= 1e-12
