"""A-array files in the reference's layout, host side (no GPU).

WriteAarraysDGV (readwrite.f90:1459-1484) writes two Fortran unformatted
sequential records: (mm, nn) and (A_capphi, A, A_xi, A_xi1, A_phi).  Pins:
  * byte identity with amdflang's runtime issuing the same two write
    statements, and amdflang's program reading the library's file back;
  * the subrecord framing of records above the subrecord limit (Intel /
    gfortran convention), parsed independently here, at a small limit;
  * a file with a record above 2 GiB written by amdflang (LLVM Flang framing,
    length modulo 2^32) read back (slow; needs a few GB of disk);
  * ReadAarraysChnksOfstNrecDGV / ReadAarraysChnksDGV / ScanAarrsChnks4Acapphi
    against the oracle's line-by-line restatement, which is pinned by
    hand-derived cases.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import dgv0d3v_direct_collision_amd as dgv
from dgv0d3v_direct_collision_amd import collision as col
import oracle

FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
GRID = (3, 3, (-1.5, 1.5))


def _arrays(mm, nn, seed):
    rng = np.random.default_rng(seed)
    cap = np.zeros(mm, np.int32)
    nodes = np.sort(rng.integers(0, mm, nn))
    np.add.at(cap, nodes, 1)
    A = rng.standard_normal(nn)
    xi = rng.integers(1, 1000, nn).astype(np.int32)
    xi1 = (xi + rng.integers(1, 50, nn)).astype(np.int32)
    phi = (nodes + 1).astype(np.int32)
    return cap, A, xi, xi1, phi


FORTRAN = """
program aio
  implicit none
  integer(4) :: mm, nn, i, code
  integer(4), allocatable :: cap(:), xi(:), xi1(:), phi(:)
  real(8), allocatable :: a(:)
  character(len=256) :: path
  call get_command_argument(1, path)
  mm = %(mm)d
  nn = %(nn)d
  allocate(cap(mm), a(nn), xi(nn), xi1(nn), phi(nn))
  if (trim(path) == 'write') then
    do i = 1, mm
      cap(i) = mod(i * 7, 5)
    end do
    do i = 1, nn
      a(i) = real(i, 8) / 3.0d0 - 1.5d0
      xi(i) = i * 3
      xi1(i) = i * 3 + 1
      phi(i) = 1 + mod(i, mm)
    end do
    open (15, file='fortran.dat', position='REWIND', action='WRITE', form='UNFORMATTED', access='SEQUENTIAL')
    write (15) mm, nn
    write (15) cap, a, xi, xi1, phi
    close (15)
  else
    open (15, file=trim(path), position='REWIND', action='READ', form='UNFORMATTED', access='SEQUENTIAL')
    read (15, iostat=code) mm, nn
    if (code /= 0) stop 2
    read (15, iostat=code) cap, a, xi, xi1, phi
    if (code /= 0) stop 3
    close (15)
    do i = 1, nn
      if (a(i) /= real(i, 8) / 3.0d0 - 1.5d0 .or. xi(i) /= i * 3 .or. xi1(i) /= i * 3 + 1) stop 4
      if (phi(i) /= 1 + mod(i, mm)) stop 5
    end do
    do i = 1, mm
      if (cap(i) /= mod(i * 7, 5)) stop 6
    end do
  end if
end program
"""


def _fortran_like(mm, nn):
    i = np.arange(1, mm + 1)
    cap = (i * 7 % 5).astype(np.int32)
    k = np.arange(1, nn + 1)
    a = k.astype(np.float64) / 3.0 - 1.5
    return cap, a, (k * 3).astype(np.int32), (k * 3 + 1).astype(np.int32), (1 + k % mm).astype(np.int32)


@pytest.mark.skipif(FC is None, reason="amdflang not found")
@pytest.mark.parametrize("framing", [dgv.DGV_FRAMING_SUBRECORDS if hasattr(dgv, "DGV_FRAMING_SUBRECORDS") else 0, 1])
def test_file_is_byte_identical_to_the_fortran_writer(tmp_path, framing):
    mm, nn = 37, 1001
    src = tmp_path / "aio.f90"
    src.write_text(FORTRAN % {"mm": mm, "nn": nn})
    exe = tmp_path / "aio"
    subprocess.run([FC, "-O1", str(src), "-o", str(exe)], check=True, cwd=tmp_path)
    subprocess.run([str(exe), "write"], check=True, cwd=tmp_path)
    ours = tmp_path / "ours.dat"
    col.aarray_file_write(str(ours), *_fortran_like(mm, nn), framing=framing)
    assert (tmp_path / "fortran.dat").read_bytes() == ours.read_bytes()
    # and the Fortran runtime reads ours back
    r = subprocess.run([str(exe), str(ours)], cwd=tmp_path)
    assert r.returncode == 0
    back = col.aarray_file_read(str(tmp_path / "fortran.dat"))
    for x, y in zip(back, _fortran_like(mm, nn)):
        assert np.array_equal(x, y)


def _parse_subrecords(raw):
    """Records of a file in the Intel / gfortran subrecord convention, parsed
    here from the convention's rules: leading marker negative while the record
    continues, trailing marker negative on every subrecord after the first."""
    recs, pos = [], 0
    while pos < len(raw):
        data, first = b"", True
        while True:
            lead = struct.unpack_from("<i", raw, pos)[0]
            n = abs(lead)
            data += raw[pos + 4:pos + 4 + n]
            trail = struct.unpack_from("<i", raw, pos + 4 + n)[0]
            assert trail == (n if first else -n)
            pos += 8 + n
            first = False
            if lead >= 0:
                break
        recs.append(data)
    return recs


def test_subrecord_framing_at_a_small_limit(tmp_path):
    arrs = _arrays(11, 97, 3)
    p = str(tmp_path / "sub.dat")
    try:
        col.set_fortran_subrecord_limit(100)
        col.aarray_file_write(p, *arrs)
    finally:
        col.set_fortran_subrecord_limit()
    raw = open(p, "rb").read()
    recs = _parse_subrecords(raw)
    assert len(recs) == 2
    assert struct.unpack("<ii", recs[0]) == (11, 97)
    payload = b"".join(a.tobytes() for a in arrs)
    assert recs[1] == payload
    assert raw.count(struct.pack("<i", -100)) >= 2 * (len(payload) // 100 - 1)
    back = col.aarray_file_read(p)
    for x, y in zip(back, arrs):
        assert np.array_equal(x, y)
    # the unsigned framing of the same small file is a plain file
    q = str(tmp_path / "plain.dat")
    col.aarray_file_write(q, *arrs, framing=1)
    assert len(open(q, "rb").read()) == 16 + 8 + len(payload)


def test_reader_rejects_corrupt_files(tmp_path):
    arrs = _arrays(5, 20, 4)
    p = tmp_path / "x.dat"
    col.aarray_file_write(str(p), *arrs)
    raw = bytearray(p.read_bytes())
    bad = tmp_path / "bad.dat"
    bad.write_bytes(bytes(raw[:-9]))                      # truncated
    with pytest.raises(dgv.DGVError) as e:
        col.aarray_file_read(str(bad))
    assert e.value.code == -7
    raw2 = bytearray(raw)
    raw2[16:20] = struct.pack("<i", 12345)                # wrong record length
    bad.write_bytes(bytes(raw2))
    with pytest.raises(dgv.DGVError):
        col.aarray_file_read(str(bad))
    with pytest.raises(dgv.DGVError) as e:
        col.aarray_file_read(str(tmp_path / "missing.dat"))
    assert e.value.code == -7


@pytest.mark.slow
@pytest.mark.skipif(FC is None, reason="amdflang not found")
def test_flang_record_above_2gib_is_read(tmp_path):
    if shutil.disk_usage(tmp_path).free < 6 << 30:
        pytest.skip("needs 6 GB of free disk")
    mm, nn = 4, 107_400_000          # 4 mm + 20 nn = 2.148e9 bytes > 2^31 - 1
    src = tmp_path / "aio.f90"
    src.write_text(FORTRAN % {"mm": mm, "nn": nn})
    exe = tmp_path / "aio"
    subprocess.run([FC, "-O1", str(src), "-o", str(exe)], check=True, cwd=tmp_path)
    subprocess.run([str(exe), "write"], check=True, cwd=tmp_path)
    f = tmp_path / "fortran.dat"
    assert f.stat().st_size == 16 + 8 + 4 * mm + 20 * nn
    cap, A, xi, xi1, phi = col.aarray_file_read(str(f))
    k = np.arange(1, nn + 1)
    assert np.array_equal(cap, (np.arange(1, mm + 1) * 7 % 5).astype(np.int32))
    for sl in (slice(0, 1000), slice(nn // 2, nn // 2 + 1000), slice(nn - 1000, nn)):
        assert np.array_equal(A[sl], k[sl] / 3.0 - 1.5)
        assert np.array_equal(xi[sl], k[sl] * 3) and np.array_equal(phi[sl], 1 + k[sl] % mm)
    f.unlink()


# ---- chunk sets ------------------------------------------------------------------

def _write_chunks(tmp_path, sizes, mm=9, seed=7):
    """A chunk set whose records are ordered by owning node, as the reference
    expects; sizes = records per chunk."""
    total = sum(sizes)
    cap, A, xi, xi1, phi = _arrays(mm, total, seed)
    chunks, r = [], 0
    for j, n in enumerate(sizes):
        sl = slice(r, r + n)
        c = np.zeros(mm, np.int32)
        np.add.at(c, phi[sl] - 1, 1)
        chunk = (c, A[sl], xi[sl], xi1[sl], phi[sl])
        col.aarray_file_write(col.chunk_name(*GRID, str(tmp_path) + "/", "t", j), *chunk)
        chunks.append(chunk)
        r += n
    return chunks


def test_oracle_chunk_reader_hand_cases():
    # two chunks over 3 nodes: chunk 1 holds records of nodes 1, 1, 2; chunk 2 of 2, 3, 3, 3
    c1 = (np.array([2, 1, 0]), np.array([1., 2, 3]), np.array([1, 2, 3]), np.array([4, 5, 6]), np.array([1, 1, 2]))
    c2 = (np.array([0, 1, 3]), np.array([4., 5, 6, 7]), np.array([7, 8, 9, 10]), np.array([11, 12, 13, 14]),
          np.array([2, 3, 3, 3]))
    cap, A, xi, xi1, phi = oracle.read_chunks_ofst_nrec([c1, c2], 1, 0, 7)
    assert list(cap) == [2, 2, 3] and list(A) == [1, 2, 3, 4, 5, 6, 7]
    # skip 1 record, read 3: the second record of node 1, node 2 from chunk 1 and node 2 from chunk 2
    cap, A, _, _, _ = oracle.read_chunks_ofst_nrec([c1, c2], 1, 1, 3)
    assert list(cap) == [1, 2, 0] and list(A) == [2, 3, 4]
    # skip 2 (the offset ends exactly after node 1): node 1 gets 0, node 2 its record in
    # chunk 1 and one in chunk 2, node 3 the two records after it
    cap, A, _, _, _ = oracle.read_chunks_ofst_nrec([c1, c2], 1, 2, 4)
    assert list(cap) == [0, 2, 2] and list(A) == [3, 4, 5, 6]
    # start in chunk 2 after one record
    cap, A, _, _, _ = oracle.read_chunks_ofst_nrec([c1, c2], 2, 1, 3)
    assert list(cap) == [0, 0, 3] and list(A) == [5, 6, 7]
    with pytest.raises(IndexError):
        oracle.read_chunks_ofst_nrec([c1, c2], 1, 5, 1)


def test_chunk_set_reader_matches_the_oracle(tmp_path):
    sizes = [40, 0, 75, 31]
    chunks = _write_chunks(tmp_path, sizes)
    d = str(tmp_path) + "/"
    act = col.ScanAarrsChnks4Acapphi(*GRID, len(sizes), d, "t")
    assert list(act) == sizes
    full = col.ReadAarraysChnksDGV(*GRID, len(sizes), d, "t")
    ref = oracle.read_chunks_ofst_nrec(chunks, 1, 0, sum(sizes))
    for x, y in zip(full, ref):
        assert np.array_equal(x, y)
    assert np.array_equal(full[0], sum(c[0] for c in chunks))    # ReadAarraysChnksDGV's A_capphi sum
    rng = np.random.default_rng(1)
    cases = [(1, 0, 10), (1, 39, 50), (3, 0, 200), (3, 74, 1), (4, 5, 20), (1, 17, 146)]
    for _ in range(12):
        f = int(rng.integers(1, 5))
        n = sizes[f - 1]
        if n == 0:
            continue
        cases.append((f, int(rng.integers(0, n)), int(rng.integers(1, 150))))
    for fst, ofst, num in cases:
        got = col.ReadAarraysChnksOfstNrecDGV(*GRID, fst, ofst, num, len(sizes), d, "t")
        exp = oracle.read_chunks_ofst_nrec(chunks, fst, ofst, num)
        for x, y in zip(got, exp):
            assert np.array_equal(x, y), (fst, ofst, num)
    with pytest.raises(dgv.DGVError):
        col.ReadAarraysChnksOfstNrecDGV(*GRID, 1, 45, 3, len(sizes), d, "t")   # offset past chunk 1
