"""FP64 parity of the single-VDF evaluator on the single-VDF and precision
configurations' own grids: 27^3 (3 Gauss nodes x 9 cells on [-4,4]^3) with the
seeded bimodal + 5% perturbation VDF, and 33^3 (11 cells on [-4.5,4.5]^3) with a
near-equilibrium VDF.  The reference values come from the oracle at a node sample that covers
every (or every third) canonical basis function in corner, edge, face and
interior cells:
  * eval_aplus_list: the reference loop and its sequential FP64 sum
    (collision_mod.f90:186-226);
  * eval_aplus_dd_list: the same terms formed exactly and summed in
    double-double, i.e. Q rounded once.
The GPU must be within 1e-12 relative L2 of the rounded-once Q; its distance
from the sequential sum, and that sum's own distance from the rounded-once Q,
are printed (they differ by the sequential sum's rounding at the terms' scale)."""
import os

import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _sample_cells(M, rng, nrand):
    e = [0, M // 2, M - 1]
    cells = {(a, b, c) for a in e for b in e for c in e}      # corners, edges, faces, centre
    while len(cells) < 27 + nrand:
        cells.add(tuple(int(x) for x in rng.integers(0, M, 3)))
    return sorted(((a * M + b) * M + c) for a, b, c in cells)


def _check(M, L, f, bases, ncells_rand, seed, label):
    import dgv0d3v_direct_collision_amd as dgv
    import oracle
    rng = np.random.default_rng(seed)
    print(f"\n[parity {label}] M={M} box=[-{L},{L}]^3: A-array on the GPU", flush=True)
    d = dgv.DGV(M, 3, (-L, L))
    try:
        d.generate_aarray(min_sens=0.0, nquad=24)
        q = d.EvalCollisionPeriodicAPlus_DGV(f)
        D = d.D
        cells = _sample_cells(M, rng, ncells_rand)
        nodes = np.array([c * D + l for c in cells for l in bases], np.int64)
        print(f"[parity {label}] exporting the records of {len(bases)} basis functions; "
              f"{len(nodes)} nodes in {len(cells)} cells", flush=True)
        counts, A, xi, xi1 = d.get_aarray_bases(bases)
        G = oracle.Geometry(M, 3, (-L, L))
        capphi = np.zeros(d.nnodes, np.int32)
        for l, n in zip(bases, counts):
            capphi[(G.ccell - 1) * D + l] = n
        # the lists must be in the order the reference's loop consumes: xi, then xi1
        o = 0
        for n in counts:
            key = xi[o:o + n].astype(np.int64) * d.nnodes + xi1[o:o + n]
            assert np.all(np.diff(key) > 0)
            o += n
        print(f"[parity {label}] oracle: sequential sum, then double-double ({THREADS} threads)", flush=True)
        qseq = oracle.eval_aplus_list(G, capphi, A, xi, xi1, f, nodes, nthreads=THREADS)
        qdd, qabs = oracle.eval_aplus_dd_list(G, capphi, A, xi, xi1, f, nodes, nthreads=THREADS)
        del A, xi, xi1
        e_gpu = rel_l2(q[nodes], qdd)
        e_seq = rel_l2(qseq, qdd)
        e_gs = rel_l2(q[nodes], qseq)
        kappa = float(np.linalg.norm(qabs) / np.linalg.norm(qdd))
        scale = float(np.linalg.norm(qabs))
        print(f"[parity {label}] GPU vs rounded-once Q {e_gpu:.3e}; sequential FP64 sum vs rounded-once "
              f"{e_seq:.3e}; GPU vs sequential {e_gs:.3e}; kappa {kappa:.1f}; GPU error / term scale "
              f"{np.linalg.norm(q[nodes] - qdd) / scale:.2e}", flush=True)
        assert e_gpu <= 1e-12, e_gpu
        return e_gpu, e_seq
    finally:
        d.close()


def test_parity_27cubed_single_vdf_config():
    from dgv0d3v_direct_collision_amd import workloads
    import dgv0d3v_direct_collision_amd as dgv
    M, L = 9, 4.0
    u, v, w, gw, _ = dgv.grid_geometry(M, 3, (-L, L))
    f = workloads.perturbed_bimodal_vdf(u, v, w, gw, M, np.random.default_rng(20261203))
    _check(M, L, f, list(range(27)), 0, 5, "27^3")


def test_parity_33cubed_precision_config():
    from dgv0d3v_direct_collision_amd import workloads
    import dgv0d3v_direct_collision_amd as dgv
    M, L = 11, 4.5
    u, v, w, gw, _ = dgv.grid_geometry(M, 3, (-L, L))
    f = workloads.near_equilibrium_vdf(u, v, w, gw, np.random.default_rng(20261205), 1e-3)
    _check(M, L, f, [0, 4, 8, 10, 13, 16, 18, 22, 26], 0, 6, "33^3")


def test_multi_basis_export_equals_single_basis_exports():
    """The one-pass export (order by construction) against the per-basis export
    of the cached host layout (order by sort), on a small grid."""
    import dgv0d3v_direct_collision_amd as dgv
    print("\n[export] M=3 s=3: get_aarray_bases vs get_aarray_basis", flush=True)
    d = dgv.DGV(3, 3, (-1.5, 1.5))
    try:
        d.generate_aarray(nquad=16)
        ls = [26, 0, 13, 5]
        counts, A, xi, xi1 = d.get_aarray_bases(ls)
        o = 0
        for l, n in zip(ls, counts):
            a1, x1, y1 = d.get_aarray_basis(l)
            assert n == len(a1)
            assert np.array_equal(A[o:o + n], a1)
            assert np.array_equal(xi[o:o + n], x1) and np.array_equal(xi1[o:o + n], y1)
            o += n
        assert o == len(A)
    finally:
        d.close()
