"""A context's A-array through the reference's file layouts and back (through
the C ABI): a chunk set written by the library and read by ReadAarraysChnksDGV,
and the single file of WriteAarraysDGV with records split into subrecords.
The restored context must hold the same records in the same order and give the
same Q bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_chunk_set_round_trip(tmp_path):
    import dgv0d3v_direct_collision_amd as dgv
    from dgv0d3v_direct_collision_amd import collision as col
    print("\n[chunks] M=3 s=3: chunk set write / ReadAarraysChnksDGV", flush=True)
    g = dgv.DGV(3, 3, (-1.5, 1.5))
    h = dgv.DGV(3, 3, (-1.5, 1.5))
    try:
        g.generate_aarray(nquad=16)
        ref = g.get_aarray()
        nrec = len(ref[1])
        d = str(tmp_path) + "/"
        n = g.write_aarray_chunks(d, "A", max_records=nrec // 3 + 7)
        assert n == 3
        act = col.ScanAarrsChnks4Acapphi(3, 3, (-1.5, 1.5), n, d, "A")
        assert act.sum() == nrec
        h.read_aarray_chunks(n, d, "A")
        _same(h.get_aarray(), ref)
        f = np.exp(-(g.nodes_u ** 2 + g.nodes_v ** 2 + g.nodes_w ** 2))
        assert np.array_equal(h.EvalCollisionPeriodicAPlus_DGV(f), g.EvalCollisionPeriodicAPlus_DGV(f))
        # a slice as an MPI rank of the reference would read it
        cap, A, xi, xi1, phi = col.ReadAarraysChnksOfstNrecDGV(3, 3, (-1.5, 1.5), 2, 5, 1000, n, d, "A")
        o = act[0] + 5
        assert np.array_equal(A, ref[1][o:o + 1000]) and np.array_equal(xi, ref[2][o:o + 1000])
        assert cap.sum() == 1000
    finally:
        h.close()
        g.close()


def test_single_file_with_subrecords_round_trip(tmp_path):
    import dgv0d3v_direct_collision_amd as dgv
    from dgv0d3v_direct_collision_amd import collision as col
    print("\n[chunks] M=3 s=3: WriteAarraysDGV with 1 MiB subrecords", flush=True)
    g = dgv.DGV(3, 3, (-1.5, 1.5))
    h = dgv.DGV(3, 3, (-1.5, 1.5))
    try:
        g.generate_aarray(nquad=16)
        p = str(tmp_path / "A_Aarrs.dat")
        try:
            col.set_fortran_subrecord_limit(1 << 20)
            g.write_aarray(p)
        finally:
            col.set_fortran_subrecord_limit()
        h.read_aarray(p)
        _same(h.get_aarray(), g.get_aarray())
    finally:
        h.close()
        g.close()
