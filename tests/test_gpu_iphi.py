"""IphiDGV and IphiDGV_decomp (dgvtools_mod.f90:847-877, 896-927) on the device,
against the oracle on the full A-array (the canonical records replicated to every
node, as the reference's IphiDGV walks them):

 * IphiDGV: relative L2 <= 1e-13 (same records, products grouped differently);
 * IphiDGV_decomp: the oracle's per-record sum of the three split terms with the
   oracle's local Maxwellian (MassCheckRec, maxwelveldist), relative L2 <= 1e-12
   of the result's norm; and, as the split is exact, the difference of two
   IphiDGV evaluations, Iphi(f) - Iphi(fM);
 * the coefficient (mol_diam/L_inf)^2 N_inf scales both;
 * the conservation projection, when enabled for Q, leaves both bit-identical
   (the reference's IphiDGV is the plain sum) while it does change Q.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def test_iphi_and_decomposition_against_full_a():
    import dgv0d3v_direct_collision_amd as dgv
    print("\n[iphi] grid M=3 s=3 on [-3,3]^3, mol_diam 2 (coef 4)", flush=True)
    d = dgv.DGV(3, 3, (-3.0, 3.0), mol_diam=2.0)
    try:
        d.generate_aarray()
        G = oracle.Geometry(3, 3, (-3.0, 3.0))
        full = oracle.full_aarray(G, *d.get_aarray())
        u, v, w, gw = d.nodes_u, d.nodes_v, d.nodes_w, d.nodes_gwts
        f = (np.exp(-((u - 0.4) ** 2 + v ** 2 + (w + 0.2) ** 2) / 0.7) +
             0.5 * np.exp(-((u + 0.6) ** 2 + (v - 0.3) ** 2 + w ** 2) / 0.5))
        coef = d.coef
        assert coef == 4.0
        want = oracle.iphi(*full, gw, f, coef)
        got = d.IphiDGV(f)
        print(f"[iphi] IphiDGV: {rel_l2(got, want):.2e}", flush=True)
        assert rel_l2(got, want) <= 1e-13
        m = oracle.masscheck(G, f)          # density, ubar, vbar, wbar, T
        fM = np.array([oracle.maxwell(m[4], m[1], m[2], m[3], m[0], u[i], v[i], w[i]) for i in range(len(u))])
        want_d = oracle.iphi_decomp(*full, gw, f, fM, coef)
        got_d = d.IphiDGV_decomp(f)
        diff = oracle.iphi(*full, gw, f, coef) - oracle.iphi(*full, gw, fM, coef)
        print(f"[iphi] IphiDGV_decomp: {rel_l2(got_d, want_d):.2e}; against Iphi(f) - Iphi(fM): "
              f"{rel_l2(got_d, diff):.2e}", flush=True)
        assert rel_l2(got_d, want_d) <= 1e-12
        assert rel_l2(got_d, diff) <= 1e-10
        q_plain = d.EvalCollisionPeriodicAPlus_DGV(f)
        d.set_projection(True)
        try:
            q_proj = d.EvalCollisionPeriodicAPlus_DGV(f)
            assert rel_l2(q_proj, q_plain) > 1e-8
            np.testing.assert_array_equal(d.IphiDGV(f), got)
            np.testing.assert_array_equal(d.IphiDGV_decomp(f), got_d)
            print(f"[iphi] with the projection on: Q moves by {rel_l2(q_proj, q_plain):.2e}, Iphi unchanged",
                  flush=True)
        finally:
            d.set_projection(False)
        np.testing.assert_array_equal(d.EvalCollisionPeriodicAPlus_DGV(f), q_plain)
    finally:
        d.close()
