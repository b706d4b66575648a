"""Mass balance of the shifted A-array on a truncated velocity box, on the
oracle and on the GPU with the same generated A-array (15^3 nodes: 5 cells of
width 1.4 and 3 Gauss nodes per dimension on [-3.5,3.5]^3).

EvalCollisionPeriodicAPlus_DGV (collision_mod.f90:143-229) evaluates every
node's sum from the canonical cell's records, shifted to the node's cell; a
shifted record whose nodes leave the grid is skipped (collision_mod.f90:195-211).
The canonical records themselves only hold pairs on the grid.  For an output
node near the edge, a far partner's record (its loss term) is therefore absent,
while the gain terms that land near the centre are present:

  * a VDF supported near the centre (a Maxwellian, T = 0.2) loses no record:
    sum w Q vanishes to the A-array's quadrature accuracy;
  * the same mass split into two blobs near opposite faces (u = +-2.1) makes
    sum w Q positive, by orders of magnitude more;
  * both on the oracle's restatement and on the GPU, which agree to 1e-12.

This is the mechanism behind the density growth of the plain single-VDF
configuration (27^3 on [-4,4]^3, RK3): a property of the reference's scheme,
reproduced by the GPU evaluator, not an error of it.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu
M, S, BOX = 5, 3, (-3.5, 3.5)


def test_edge_records_break_mass_balance_on_oracle_and_gpu():
    import dgv0d3v_direct_collision_amd as dgv
    print("\n[edge mass] 15^3 A-array", flush=True)
    d = dgv.DGV(M, S, BOX)
    try:
        d.generate_aarray(min_sens=0.0)
        A = d.get_aarray()
        G = oracle.Geometry(M, S, BOX)
        u, v, w, gw = G.nodes_u, G.nodes_v, G.nodes_w, G.nodes_gwts
        mx = np.vectorize(oracle.maxwell)
        centre = mx(0.2, 0.0, 0.0, 0.0, 1.0, u, v, w)
        faces = 0.5 * (mx(0.2, 2.1, 0.0, 0.0, 1.0, u, v, w) + mx(0.2, -2.1, 0.0, 0.0, 1.0, u, v, w))
        ratio = {}
        for name, f in (("centre", centre), ("faces", faces)):
            f = f / np.sum(gw * f)
            qo = oracle.eval_aplus(G, *A, f, nthreads=8)
            qg = d.EvalCollisionPeriodicAPlus_DGV(f)
            err = rel_l2(qg, qo)
            mo, mg, scale = np.sum(gw * qo), np.sum(gw * qg), np.sum(np.abs(gw * qo))
            ratio[name] = mo / scale
            print(f"[edge mass] {name}: oracle sum w Q {mo:.4e} (GPU {mg:.4e}), sum |w Q| {scale:.4e}, "
                  f"GPU vs oracle rel L2 {err:.2e}", flush=True)
            assert err <= 1e-12
            assert abs(mg - mo) <= 1e-12 * scale
        assert abs(ratio["centre"]) <= 1e-4
        assert ratio["faces"] > 0 and ratio["faces"] >= 100 * abs(ratio["centre"])
    finally:
        d.close()
