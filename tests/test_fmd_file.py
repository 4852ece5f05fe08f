"""DB.fmd, the reference's FMD-index file (src/FMDIndex.cpp:143-161), read and written by the
library on the host: msgs_fmd_read decodes a file the reference's own FMDIndex::save wrote
(oracle/_ref) to the oracle's BWT, counts, gapSA and sampled SA; a file msgs_fmd_write-style
encoding writes is loaded by the reference's own FMDIndex::load and answers like the oracle.
The writer runs on the host (msgs_fmd_write_arrays, the code path of msgs_fmd_write) over the
oracle's BWT and suffix array."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
import msgseqtk_amd as M
from msgseqtk_amd import synth

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")


def small_db(seed):
    chroms, names = synth.metagenome(5, 7000, gc=(0.3, 0.7), seed=seed)
    chroms[1] = chroms[1].copy()
    chroms[1][300:340] = 15
    return O.OracleDb(chroms, names)


def test_read_reference_written_fmd(tmp_path):
    odb = small_db(51)
    bd, bwt, sa = odb.arrays()
    rix = O.RefIndex(bd)
    path = str(tmp_path / "ref.fmd")
    assert O.ref().ref_index_save(rix.h, path.encode()) == 0
    b2, B, Cc, gap, smp = M.fmd_read(path, header=False)
    B0, C0 = odb.counts()
    assert np.array_equal(b2, bwt) and np.array_equal(B, B0) and np.array_equal(Cc, C0)
    assert np.array_equal(gap, sa[:B0[0]])
    rows = np.nonzero((np.arange(len(sa)) % 32 == 0) | (bwt == 0))[0]
    assert np.array_equal(smp, sa[rows])


def test_written_fmd_loads_in_reference(tmp_path):
    odb = small_db(52)
    bd, bwt, sa = odb.arrays()
    path = str(tmp_path / "ours.fmd")
    host_write = M.lib().msgs_fmd_write_arrays
    assert host_write(path.encode(), 1, bwt.ctypes.data_as(M.PU8), sa.ctypes.data_as(M.P64), len(bwt)) == 0
    # our reader reads it back
    b2, B, Cc, gap, smp = M.fmd_read(path, header=True)
    assert np.array_equal(b2, bwt)
    # the reference's FMDIndex::load reads the body (after the 20-byte program header)
    h = O.ref().ref_index_load(path.encode(), 20)
    assert h
    B0, C0 = odb.counts()
    rb, rc = np.zeros(256, np.int64), np.zeros(256, np.int64)
    O.ref().ref_counts(h, O.ptr(rb, C.c_int64), O.ptr(rc, C.c_int64))
    assert np.array_equal(rb, B0) and np.array_equal(rc, C0)
    for i in range(0, len(bwt), 13):
        assert O.ref().ref_access_bwt(h, i) == bwt[i]
    for i in range(0, len(sa), 29):
        assert O.ref().ref_access_sa(h, i) == sa[i]
    # and it gives the reference's SMEMs of a read
    rix = O.RefIndex(bd)
    read = bd[1000:1120].copy()
    r = np.ascontiguousarray(read)
    for idx in (rix.h, h):
        out = np.zeros((64, 5), np.int64)
        lp = np.zeros(64)
        k = O.ref().ref_find_smems(idx, O.ptr(r, C.c_uint8), len(r), 17, 0.01, 0, O.ptr(out, C.c_int64),
                                   O.ptr(lp, C.c_double), 64)
        if idx == rix.h:
            want = (k, out[:k].copy(), lp[:k].copy())
        else:
            assert k == want[0] and np.array_equal(out[:k], want[1]) and np.array_equal(lp[:k], want[2])
    O.ref().ref_index_free(h)


def test_byte_identical_to_reference_save(tmp_path):
    """Ignoring the 4 bytes after each BitStr width field (the reference writes its 4-byte int
    as 8 bytes, the other 4 being whatever memory follows it), the file is byte for byte the
    reference's own."""
    odb = small_db(53)
    bd, bwt, sa = odb.arrays()
    rix = O.RefIndex(bd)
    p1, p2 = str(tmp_path / "ref.fmd"), str(tmp_path / "ours.fmd")
    assert O.ref().ref_index_save(rix.h, p1.encode()) == 0
    host_write = M.lib().msgs_fmd_write_arrays
    assert host_write(p2.encode(), 0, bwt.ctypes.data_as(M.PU8), sa.ctypes.data_as(M.P64), len(bwt)) == 0
    a, b = open(p1, "rb").read(), open(p2, "rb").read()
    assert len(a) == len(b)
    diff = np.nonzero(np.frombuffer(a, np.uint8) != np.frombuffer(b, np.uint8))[0]
    # every differing byte lies 4..7 bytes after a 32- or 64-bit width field value
    for d in diff:
        base = d - (d % 4)
        w = np.frombuffer(b[base - 4:base], np.int32)[0] if base >= 4 else -1
        assert w in (32, 64), d
