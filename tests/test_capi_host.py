"""The C ABI library loads and exports every entry point of include/dgv.h;
host-side logic and error paths that need no GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import dgv0d3v_direct_collision_amd as dgv
from dgv0d3v_direct_collision_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "dgv.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(dgv_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers all of them
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_grid_errors_are_status_codes():
    L = dgv.load()
    n = C.c_int64()
    bad = [(dgv.grid_desc(5, 3, (1.0, 3.0)), -2),     # origin outside the box: no canonical cell
           (dgv.grid_desc(5, 9, (-1.0, 1.0)), -2),    # more than 8 nodes per cell
           (dgv.grid_desc(0, 3, (-1.0, 1.0)), -1),    # no cells
           (dgv.grid_desc(3, 3, (1.0, -1.0)), -1)]    # empty box
    for g, code in bad:
        rc = L.dgv_grid_geometry(C.byref(g), None, None, None, None, None, C.byref(n))
        assert rc == code
        assert L.dgv_last_error(None)


def test_agen_pair_index_errors():
    with pytest.raises(dgv.DGVError) as e:
        dgv.agen_pair(3, 3, (-1.5, 1.5), 0, 5)
    assert e.value.code == -3
    with pytest.raises(dgv.DGVError):
        dgv.agen_pair(3, 3, (-1.5, 1.5), 1, 10 ** 6)


def test_chunk_file_name_format():
    # MakeBaseNameAoperChnksDGV (readwrite.f90:2224-2281), uniform meshes
    name = dgv.chunk_name(9, 3, (-4.0, 4.0), "sol/", "A", chunk=7, k_c=2, N_x=3)
    assert name == "sol/Ach007_2kc3su3sv3sw3NXU9MuUU9MvVU9MwWU_Aarrs.dat"


def test_context_without_gpu_fails_loudly():
    import subprocess
    import sys
    code = ("import dgv0d3v_direct_collision_amd as d\n"
            "try:\n    d.DGV(3, 3, (-1.5, 1.5))\n    print('GPU')\n"
            "except d.DGVError as e:\n    print('ERR', e.code)\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT,
                         env={**os.environ, "HIP_VISIBLE_DEVICES": "-1"}, timeout=120).stdout
    assert out.strip() in ("ERR -5",), out


def test_host_geometry_matches_reference_order():
    u, v, w, gw, cc = dgv.grid_geometry((2, 3, 4), (2, 2, 2), (-1, 1, -1.5, 1.5, -2, 2))
    assert len(u) == 24 * 8 and cc >= 1
    # cells u outer, w inner; local u outer, w inner
    assert u[0] < u[4] and w[0] < w[1] and np.all(np.diff(u[:8][::4]) > 0)
