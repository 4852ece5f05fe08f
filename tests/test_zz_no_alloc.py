"""msgs_smem_batch_noalloc (the allocation-free SMEM search over device-resident reads)
against msgs_smem_batch on the same reads: equal counts, records and logP bit patterns, on
two calls in a row, the second with a smaller input."""
import ctypes as C

import numpy as np
import pytest

import msgseqtk_amd as M
from msgseqtk_amd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

REC = np.dtype([("from", np.int32), ("to", np.int32), ("fwd", np.int64), ("rev", np.int64),
                ("size", np.int64), ("logP", np.float64)])


def noalloc(m, reads, params, cap):
    seq = np.concatenate(reads).astype(np.uint8)
    offs = np.zeros(len(reads) + 1, np.int64)
    offs[1:] = np.cumsum([len(r) for r in reads])
    dseq = torch.from_numpy(seq).cuda()
    doffs = torch.from_numpy(offs).cuda()
    dcnt = torch.zeros(len(reads), dtype=torch.int32, device="cuda")
    drec = torch.zeros(len(reads) * cap * REC.itemsize, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    mx = C.c_int32(0)
    rc = M.lib().msgs_smem_batch_noalloc(m.h, C.c_void_p(dseq.data_ptr()), C.c_void_p(doffs.data_ptr()),
                                         len(reads), C.byref(params), cap, C.c_void_p(dcnt.data_ptr()),
                                         C.c_void_p(drec.data_ptr()), C.byref(mx))
    return rc, mx.value, dcnt.cpu().numpy(), drec.cpu().numpy().view(REC).reshape(len(reads), cap)


def check(m, reads, params):
    c0, rows, lp = m.smems(reads, params)
    cap = int(c0.max()) + 1
    rc, mx, cnt, rec = noalloc(m, reads, params, cap)
    assert rc == 0, M.lib().msgs_last_error(m.h)
    assert mx == c0.max() and np.array_equal(cnt, c0)
    k = 0
    for i in range(len(reads)):
        r = rec[i, :cnt[i]]
        want = rows[k:k + cnt[i]]
        assert np.array_equal(r["from"], want[:, 0]) and np.array_equal(r["to"], want[:, 1])
        assert np.array_equal(r["fwd"], want[:, 2]) and np.array_equal(r["rev"], want[:, 3])
        assert np.array_equal(r["size"], want[:, 4])
        assert np.array_equal(r["logP"].view(np.int64), lp[k:k + cnt[i]].view(np.int64))
        k += cnt[i]
    # too few slots: the call reports the largest count
    if cap > 2:
        rc, mx, cnt, _ = noalloc(m, reads, params, 1)
        assert rc == M.MSGS_E_CAP and mx == c0.max() and np.array_equal(cnt, c0)


def test_noalloc_matches_smem_batch_twice():
    chroms, names = synth.metagenome(6, 60000, gc=(0.3, 0.7), seed=41)
    reads = synth.simulate_reads(chroms, 3000, 150, sub=0.01, indel=0.001, seed=42)[0]
    reads += [np.full(60, 15, np.uint8), np.array([1], np.uint8), synth.encode("ACGT" * 40)]
    m = M.Mapper(0).build(chroms, names)
    p = M.default_params()
    check(m, reads, p)
    # a second, smaller call (shorter reads, no e-value limit) on the same scratch
    small = [r[:90] for r in reads[:500]]
    check(m, small, M.default_params(maxEvalue=float("inf"), minLen=12))


def test_noalloc_wide_index_twice():
    """the same on a database with the wide index layout (64-bit rows: larger candidate
    records in the scratch)"""
    chroms, names = synth.metagenome(4, 50000, gc=(0.3, 0.7), seed=43, strain_frac=0.25)
    reads = synth.simulate_reads(chroms, 2000, (40, 250), sub=0.01, indel=0.001, seed=44)[0]
    m = M.Mapper(0).set("wide_index", 1).build(chroms, names)
    assert m.get("index_wide") == 1
    check(m, reads, M.default_params())
    check(m, [r[:70] for r in reads[:300]], M.default_params(minLen=12))
