"""SetA_DGV's records (dgvtools_mod.f90:377-681): the generator against the
oracle's restatement of the adaptive-Simpson integration (no GPU: the
generator's host evaluation shares its arithmetic with the GPU kernel, which
test_gpu_agen holds to it).

On a 2 x 2 x 2-cell grid with 2 Gauss nodes per dimension, for every canonical
basis function:
  * the same pairs carry records, in the reference's serial order (i2, then
    i3), once values below the adaptive rule's accuracy are set aside;
  * values agree to the adaptive rule's accuracy (its sub-intervals whose three
    samples vanish drop thin sphere-cell slivers, about 1e-4 of the largest);
  * each record is Atemp * w(xi) * w(xi1) * |g| / 8 with the weights and |g|
    taken from the oracle's geometry (the scaling of dgvtools_mod.f90:654-659),
    on both sides;
  * min_sens drops exactly the pairs with |Atemp| <= min_sens;
  * Q of a VDF with the oracle's A-array and with the generator's differ by
    the quadrature's accuracy only.
"""
import numpy as np
import pytest

import dgv0d3v_direct_collision_amd as dgv
import oracle

M, S, BOX = 2, 2, (-1.0, 1.0)
TOL = 1e-6


@pytest.fixture(scope="module")
def both():
    G = oracle.Geometry(M, S, BOX)
    D = S ** 3
    cc = G.ccell
    orc = [oracle.seta(G, (cc - 1) * D + 1 + l, ErrChi=TOL, ErrEps=TOL, nthreads=8) for l in range(D)]
    # the generator over the prefiltered pairs in the serial order
    pairs = [(i2, i3) for i2 in range(1, G.nnodes + 1) for i3 in range(i2 + 1, G.nnodes + 1)
             if oracle.prefilter(G, (cc - 1) * D + 1, i2, i3)]
    raw = np.zeros((len(pairs), D))
    rec = np.zeros((len(pairs), D))
    for k, (i2, i3) in enumerate(pairs):
        raw[k], rec[k], _ = dgv.agen_pair(M, S, BOX, i2, i3, nquad=32)
    return G, D, orc, np.array(pairs, np.int32), raw, rec


def _w_g(G, i2, i3):
    w = G.nodes_gwts
    d = np.stack([G.nodes_u, G.nodes_v, G.nodes_w], 1)
    g = np.linalg.norm(d[i2 - 1] - d[i3 - 1], axis=1)
    return w[i2 - 1] * w[i3 - 1] * g / 8.0


def test_record_sets_order_and_values(both):
    G, D, orc, pairs, raw, rec = both
    for l in range(D):
        A, xi, xi1 = orc[l]
        key = xi.astype(np.int64) * G.nnodes + xi1
        assert np.all(np.diff(key) > 0)                     # the serial order of SetA_DGV's loops
        amax = np.max(np.abs(rec[:, l]))
        big_g = {tuple(p) for p, a in zip(pairs, rec[:, l]) if abs(a) > 1e-3 * amax}
        big_o = {(a, b) for a, b, v in zip(xi, xi1, A) if abs(v) > 1e-3 * amax}
        all_o = {(a, b) for a, b in zip(xi, xi1)}
        all_g = {tuple(p) for p, a in zip(pairs, rec[:, l]) if a != 0.0}
        assert big_g <= all_o and big_o <= all_g, l
        idx = {tuple(p): k for k, p in enumerate(pairs)}
        common = [(k, idx[(a, b)]) for k, (a, b) in enumerate(zip(xi, xi1)) if (a, b) in idx]
        ko, kg = np.array(common).T
        diff = np.abs(A[ko] - rec[kg, l]) / amax
        assert diff.max() <= 2e-4, (l, diff.max())
        assert np.median(diff) <= 2e-6, (l, np.median(diff))


def test_record_scaling_from_the_oracle_geometry(both):
    G, D, orc, pairs, raw, rec = both
    fac = _w_g(G, pairs[:, 0], pairs[:, 1])
    nz = raw != 0
    rel = np.abs(rec[nz] - (raw * fac[:, None])[nz]) / np.abs(rec[nz])
    assert rel.max() <= 1e-15, rel.max()    # a few ulps: the product is formed in another order
    cc = G.ccell
    for l in (0, D - 1):                                 # the oracle applies the same scaling
        A, xi, xi1 = orc[l]
        for k in range(0, len(A), max(1, len(A) // 15)):
            r = oracle.seta_raw(G, (cc - 1) * D + 1 + l, int(xi[k]), int(xi1[k]), TOL, TOL)
            assert abs(A[k] - r * _w_g(G, xi[k:k + 1], xi1[k:k + 1])[0]) <= 1e-15 * abs(A[k])


def test_min_sens_drops_exactly_the_small_atemp(both):
    G, D, orc, pairs, raw, rec = both
    m = float(np.median(np.abs(raw[raw != 0])))
    for k in range(0, len(pairs), 7):
        r, a, _ = dgv.agen_pair(M, S, BOX, int(pairs[k, 0]), int(pairs[k, 1]), min_sens=m, nquad=32)
        keep = np.abs(r) > m
        assert np.array_equal(a != 0, keep & (raw[k] != 0))
        assert np.array_equal(a[keep], rec[k][keep])


def test_q_with_oracle_and_generator_aarrays(both):
    G, D, orc, pairs, raw, rec = both
    cc = G.ccell
    cap_o = np.zeros(G.nnodes, np.int32)
    cap_g = np.zeros(G.nnodes, np.int32)
    Ao, xo, yo, Ag, xg, yg = [], [], [], [], [], []
    for l in range(D):
        A, xi, xi1 = orc[l]
        cap_o[(cc - 1) * D + l] = len(A)
        Ao.append(A), xo.append(xi), yo.append(xi1)
        nz = rec[:, l] != 0
        cap_g[(cc - 1) * D + l] = int(nz.sum())
        Ag.append(rec[nz, l]), xg.append(pairs[nz, 0]), yg.append(pairs[nz, 1])
    f = np.exp(-((G.nodes_u - 0.2) ** 2 + G.nodes_v ** 2 + (G.nodes_w + 0.1) ** 2) / 0.5)
    qo = oracle.eval_aplus(G, cap_o, np.concatenate(Ao), np.concatenate(xo), np.concatenate(yo), f)
    qg = oracle.eval_aplus(G, cap_g, np.concatenate(Ag), np.concatenate(xg), np.concatenate(yg), f)
    err = np.linalg.norm(qo - qg) / np.linalg.norm(qg)
    print(f"Q with the oracle's A-array vs the generator's: rel L2 {err:.2e}")
    assert err <= 1e-3
