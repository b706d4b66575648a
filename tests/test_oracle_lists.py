"""Pins of the oracle's node-list evaluators (no GPU).

eval_aplus_list is the reference loop (collision_mod.f90:186-226) restricted to
a node list: it must equal eval_aplus at those nodes bit for bit.
eval_aplus_dd_list forms each term exactly and sums in double-double: against
exact rational arithmetic (fractions.Fraction) its result is Q rounded to within
one ulp, including a construction whose sequential FP64 sum cancels to noise.
"""
from fractions import Fraction

import numpy as np

import oracle
from tests.helpers import synthetic_aarray


def _geom_and_a(M=3, s=2, seed=4):
    G = oracle.Geometry(M, s, (-1.0, 1.0))
    return G, synthetic_aarray(G, seed=seed)


def test_list_equals_full_loop_bitwise():
    G, (capphi, A, xi, xi1) = _geom_and_a()
    rng = np.random.default_rng(1)
    f = rng.uniform(0.1, 1.0, G.nnodes)
    full = oracle.eval_aplus(G, capphi, A, xi, xi1, f)
    nodes = rng.choice(G.nnodes, size=40, replace=False)
    lst = oracle.eval_aplus_list(G, capphi, A, xi, xi1, f, nodes, nthreads=3)
    assert np.array_equal(lst, full[nodes])


def _exact_q(G, capphi, A, xi, xi1, f, iphi):
    """Q at node iphi (0-based) in exact rational arithmetic, by the reference's
    shift-and-drop rule applied to the canonical records."""
    D = int(np.prod(G.s))
    l = (G.nodes_phican[iphi] - 1) % D
    off = int(np.sum(capphi[:(G.ccell - 1) * D + l]))
    n = int(capphi[(G.ccell - 1) * D + l])
    M = np.array(G.M)
    cells3 = lambda c: np.array([c // (M[1] * M[2]), (c // M[2]) % M[1], c % M[2]])
    d = np.array([G.nodes_dui[iphi], G.nodes_dvi[iphi], G.nodes_dwi[iphi]])
    acc = Fraction(0)
    for j in range(off, off + n):
        k1, k2 = xi[j] - 1, xi1[j] - 1
        c1 = cells3((k1 // D)) + d
        c2 = cells3((k2 // D)) + d
        if np.any(c1 < 0) or np.any(c1 >= M) or np.any(c2 < 0) or np.any(c2 >= M):
            continue
        n1 = ((c1[0] * M[1] + c1[1]) * M[2] + c1[2]) * D + k1 % D
        n2 = ((c2[0] * M[1] + c2[1]) * M[2] + c2[2]) * D + k2 % D
        acc += 2 * Fraction(float(f[n1])) * Fraction(float(f[n2])) * Fraction(float(A[j]))
    return acc / Fraction(float(G.nodes_gwts[iphi]))


def _ulp(x):
    return np.spacing(abs(float(x)))


def test_double_double_list_is_rounded_exact_q():
    G, (capphi, A, xi, xi1) = _geom_and_a(M=2, s=2, seed=7)
    rng = np.random.default_rng(2)
    f = rng.uniform(0.1, 1.0, G.nnodes)
    nodes = np.arange(G.nnodes)
    q, ab = oracle.eval_aplus_dd_list(G, capphi, A, xi, xi1, f, nodes)
    for t in nodes[::3]:
        ex = _exact_q(G, capphi, A, xi, xi1, f, t)
        assert abs(Fraction(float(q[t])) - ex) <= Fraction(_ulp(float(ex))), (t, q[t], float(ex))
        assert ab[t] >= abs(float(ex)) * (1 - 1e-15)


def test_double_double_survives_cancellation():
    """Records come in pairs +a, -a(1 + 2^-40) on the same node pair, so Q is a
    tiny residual of huge terms: the sequential FP64 sum loses most digits, the
    double-double sum keeps Q to an ulp."""
    G = oracle.Geometry(2, 2, (-1.0, 1.0))
    capphi, A, xi, xi1 = synthetic_aarray(G, seed=9)
    D = int(np.prod(G.s))
    rng = np.random.default_rng(3)
    A2, x2, x12 = [], [], []
    cap2 = np.zeros_like(capphi)
    off = 0
    for l in range(D):
        n = int(capphi[(G.ccell - 1) * D + l])
        a = A[off:off + n] * 1e6
        big = np.repeat(a, 2)
        big[1::2] = -a * (1 + 2.0 ** -40) + rng.uniform(-1, 1, n) * 1e-9
        A2.append(big)
        x2.append(np.repeat(xi[off:off + n], 2))
        x12.append(np.repeat(xi1[off:off + n], 2))
        cap2[(G.ccell - 1) * D + l] = 2 * n
        off += n
    A2, x2, x12 = np.concatenate(A2), np.concatenate(x2), np.concatenate(x12)
    f = rng.uniform(0.1, 1.0, G.nnodes)
    nodes = np.arange(0, G.nnodes, 5)
    qdd, ab = oracle.eval_aplus_dd_list(G, cap2, A2, x2, x12, f, nodes)
    qseq = oracle.eval_aplus_list(G, cap2, A2, x2, x12, f, nodes)
    worst_seq = 0.0
    for k, t in enumerate(nodes):
        ex = _exact_q(G, cap2, A2, x2, x12, f, t)
        assert abs(Fraction(float(qdd[k])) - ex) <= Fraction(_ulp(float(ex))), (t, qdd[k], float(ex))
        if ex != 0:
            worst_seq = max(worst_seq, abs(float(qseq[k]) - float(ex)) / abs(float(ex)))
    assert worst_seq > 1e-9    # the construction is ill-conditioned for the sequential sum
