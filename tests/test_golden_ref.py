"""The oracle against the reference's own outputs stored in
tests/golden/ref_dgvtools.npz (written by tests/golden/make_ref_vectors.py from
the reference's dgvtools_mod / distributions_mod routines compiled unmodified):
IphiDGV at unit and non-unit scale, IphiDGV_decomp, MassCheckRec,
MassCheckRecHighCMoments, maxwelveldist on 200 argument sets, and the two
A-array truncations, all bit for bit.  Runs without the reference sources."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_dgvtools.npz")


@pytest.fixture(scope="module")
def gold():
    z = np.load(GOLD)   # plain arrays, no pickles
    G = oracle.Geometry(int(z["grid"][0]), int(z["grid"][1]), (float(z["grid"][2]), float(z["grid"][3])))
    for k in ("nodes_u", "nodes_v", "nodes_w", "nodes_gwts"):
        assert np.array_equal(getattr(G, k), z[k])   # the inputs were the oracle's geometry
    return z, G


def test_iphi_and_decomposition(gold):
    z, G = gold
    full = (z["capphi"], z["A"], z["A_xi"], z["A_xi1"])
    assert np.array_equal(oracle.iphi(*full, z["nodes_gwts"], z["f"]), z["iphi"])
    d, l, n = z["coef"]
    assert np.array_equal(oracle.iphi(*full, z["nodes_gwts"], z["f"], (d / l) ** 2 * n), z["iphi_coef"])
    m = oracle.masscheck(G, z["f"])
    fM = np.array([oracle.maxwell(m[4], m[1], m[2], m[3], m[0], a, b, c)
                   for a, b, c in zip(G.nodes_u, G.nodes_v, G.nodes_w)])
    assert np.array_equal(oracle.iphi_decomp(*full, z["nodes_gwts"], z["f"], fM), z["iphi_decomp"])


def test_moments_and_maxwellian(gold):
    z, G = gold
    assert np.array_equal(oracle.masscheck(G, z["f"]), z["masscheck"])
    assert np.array_equal(oracle.masscheck_high(G, z["f"]), z["masscheck_high"])
    got = np.array([oracle.maxwell(*r) for r in z["maxwell_in"]])
    assert np.array_equal(got, z["maxwell_out"])


def test_truncations(gold):
    z, G = gold
    full = (z["capphi"], z["A"], z["A_xi"], z["A_xi1"])
    for got, name in ((oracle.trunc_threshold(*full, float(z["trunc_threshold"][0])), "trunc_threshold"),
                      (oracle.trunc_radius(G, *full, float(z["trunc_radius"][0])), "trunc_radius")):
        keep = z[name + "_keep"]
        assert 0 < keep.sum() < len(keep)
        assert np.array_equal(got[0], z[name + "_capphi"])
        for g, a in zip(got[1:], full[1:]):
            assert np.array_equal(g, a[keep])


IO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_files")


@pytest.fixture(scope="module")
def io():
    return np.load(os.path.join(IO, "inputs.npz"))


def _bytes(path):
    with open(path, "rb") as fh:
        return fh.read()


def test_file_names(io):
    import json
    from dgv0d3v_direct_collision_amd import collision as col
    for e in json.load(open(os.path.join(IO, "names.json"))):
        g = (e["M"], 3, (-1.0, 1.0))
        for t in ("0.375", "i"):
            assert col.sol_file_name(*g, t, e["dir"], e["base"], e["k_c"], e["N_x"]) == e["sol_base"] + "_time" + t + \
                "_sol.dat"
        assert col.log_file_name(*g, "_L1err.txt", e["dir"], e["base"], e["k_c"], e["N_x"]) == \
            e["sol_base"] + "_L1err.txt"
        for ch, want in zip((0, 7, 123), e["chunks"]):
            assert col.chunk_name(*g, e["dir"], e["aop"], ch, e["k_c"], e["N_x"]) == want + "_Aarrs.dat"


def test_aarray_and_solution_files(io, tmp_path):
    from dgv0d3v_direct_collision_amd import collision as col
    arr = [io["aarr_" + k] for k in ("capphi", "A", "A_xi", "A_xi1", "A_phi")]
    col.aarray_file_write(str(tmp_path / "a.dat"), *arr)
    assert _bytes(tmp_path / "a.dat") == _bytes(os.path.join(IO, "aarrs.dat"))
    for x, y in zip(col.aarray_file_read(os.path.join(IO, "aarrs.dat")), arr):
        assert np.array_equal(x, y)
    f, hist = io["sol_f"], io["sol_hist"]
    for k in range(1, 6):
        col.write_sol_file(str(tmp_path / "s.dat"), 0.375, 0.0125, k, f, hist)
        theirs = os.path.join(IO, f"sol_k{k}.dat")
        assert _bytes(tmp_path / "s.dat") == _bytes(theirs), k
        e, dt, kk, f2, h2 = col.read_sol_file(theirs, len(f))
        assert (e, dt, kk) == (0.375, 0.0125, k) and np.array_equal(f2, f)
        for r in [j for j in range(4) if j < k - 1] + [4 + j for j in range(5) if j < k - 1]:
            assert np.array_equal(h2[r], hist[r])
    col.write_sol_init_file(str(tmp_path / "i.dat"), 0.0, 0.0125, 3, f)
    assert _bytes(tmp_path / "i.dat") == _bytes(os.path.join(IO, "sol_init.dat"))


def test_logs(io, tmp_path):
    from dgv0d3v_direct_collision_amd import driver_io
    log = driver_io.MacroLog(str(tmp_path / "m.txt"))
    for r in io["macro_rows"]:
        log.record(r[0], r[1:])
    log.write()
    assert _bytes(tmp_path / "m.txt") == _bytes(os.path.join(IO, "macroerr.txt"))
    l1 = driver_io.L1Track(str(tmp_path / "l.txt"), 0.0, 0.5, 4)
    for i, e in enumerate(io["l1_err"]):
        l1(float(e), 0.00137 * i)
    assert _bytes(tmp_path / "l.txt") == _bytes(os.path.join(IO, "l1err.txt"))
    G = oracle.Geometry(3, 3, (-2.5, 2.5))
    md = driver_io.MomDerivLog(str(tmp_path / "x"), num_eval_error=6, rkmts=3, t_R=1.0)
    for i, (f1, r1) in enumerate(zip(io["md_f1"], io["md_frhs1"])):
        md.record(0.1 * i, f1, r1, G)
    want = io["md_rows"]
    got = np.array(md.rows)
    assert np.all(np.abs(got - want) <= 1e-11 * np.abs(want) + 1e-300)
    md.write()
    for suffix, name in (("_dermom.txt", "dermom.txt"), ("_relaxtimes.txt", "relaxtimes.txt")):
        assert _bytes(str(tmp_path / "x") + suffix) == _bytes(os.path.join(IO, name))
