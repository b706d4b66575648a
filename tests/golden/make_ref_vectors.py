"""Writes tests/golden/ref_dgvtools.npz: inputs and the reference's own outputs
for IphiDGV, IphiDGV_decomp, MassCheckRec, MassCheckRecHighCMoments,
maxwelveldist and the two A-array truncations (dgvtools_mod.f90,
distributions_mod.f90), computed by those routines compiled unmodified into
oracle/_ref (oracle/Makefile, target ref; needs the reference sources).

Inputs: a 2 x 2 x 2-cell grid with 2 Gauss nodes per dimension on [-1.5, 1.5]^3
(nodes and weights from the oracle's geometry), a synthetic A-array replicated to
every node (80 % of the prefiltered pairs dropped, seed 5), a seeded bimodal f,
200 random Maxwellian arguments.  tests/test_golden_ref.py checks the oracle
against the stored outputs wherever the reference is absent.

Run from the repository root: python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle import ref  # noqa: E402
from tests.helpers import bimodal, synthetic_aarray  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "ref_dgvtools.npz")


def kept(A, xi, xi1, out):
    """Which input records a truncation kept: its output is the input's records
    in order, a subsequence."""
    _, oa, ox, oy = out
    keep = np.zeros(len(A), bool)
    k = 0
    for j in range(len(A)):
        if k < len(oa) and A[j] == oa[k] and xi[j] == ox[k] and xi1[j] == oy[k]:
            keep[j] = True
            k += 1
    assert k == len(oa)
    return keep


def main():
    if not ref.available():
        raise SystemExit("oracle/_ref/libdgv_ref.so is not built (make -C oracle ref)")
    G = oracle.Geometry(2, 2, (-1.5, 1.5))
    cap, A, xi, xi1 = oracle.full_aarray(G, *synthetic_aarray(G, seed=5, drop=0.8))
    u, v, w, gw = G.nodes_u, G.nodes_v, G.nodes_w, G.nodes_gwts
    f = bimodal(u, v, w, seed=6)
    rng = np.random.default_rng(7)
    mx_in = np.column_stack([rng.uniform(0.05, 2.0, 200), rng.uniform(-2, 2, (200, 3)), rng.uniform(0.1, 3.0, 200),
                             rng.uniform(-2, 2, (200, 3))])   # T, u0, v0, w0, n, u, v, w
    mx_out = np.array([ref.maxwell(*r) for r in mx_in])
    t = float(np.quantile(np.abs(A), 0.4))
    tr = ref.trunc_threshold(cap, A, xi, xi1, t)
    rr = ref.trunc_radius(u, v, w, gw, cap, A, xi, xi1, 1.1)
    np.savez_compressed(
        OUT, grid=np.array([2, 2, -1.5, 1.5]), nodes_u=u, nodes_v=v, nodes_w=w, nodes_gwts=gw,
        capphi=cap, A=A, A_xi=xi, A_xi1=xi1, f=f,
        iphi=ref.iphi(u, v, w, gw, cap, A, xi, xi1, f),
        iphi_coef=ref.iphi(u, v, w, gw, cap, A, xi, xi1, f, 0.7, 1.3, 2.5), coef=np.array([0.7, 1.3, 2.5]),
        iphi_decomp=ref.iphi(u, v, w, gw, cap, A, xi, xi1, f, decomp=True),
        masscheck=ref.masscheck(u, v, w, gw, f), masscheck_high=ref.masscheck_high(u, v, w, gw, f),
        maxwell_in=mx_in, maxwell_out=mx_out,
        trunc_threshold=np.array([t]), trunc_threshold_capphi=tr[0], trunc_threshold_keep=kept(A, xi, xi1, tr),
        trunc_radius=np.array([1.1]), trunc_radius_capphi=rr[0], trunc_radius_keep=kept(A, xi, xi1, rr))
    print(f"wrote {OUT}: {len(A)} records, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
