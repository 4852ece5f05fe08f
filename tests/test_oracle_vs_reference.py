"""The CPU oracle against the reference's own seeding code (oracle/_ref, compiled from
the reference sources by oracle/Makefile) on the same texts and reads.  Pins the
oracle's SA/BWT/rank/backExt (A4-A8), SMEM search (A10-A13), removeRedundant (A17) and
getChains (A18) to the reference bit for bit."""
import numpy as np
import pytest

import oracle_lib as O
from msgseqtk_amd import synth

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference absent)")


def _db(seed, n=4, length=20000, strains=0.25):
    chroms, names = synth.metagenome(n, length, gc=(0.3, 0.7), seed=seed, strain_frac=strains)
    return chroms, names, O.OracleDb(chroms, names)


@pytest.mark.parametrize("seed", [1, 2])
def test_index_matches_reference(seed):
    chroms, names, db = _db(seed)
    bd, bwt, sa = db.arrays()
    ri = O.RefIndex(bd)
    B1, C1 = db.counts()
    B2, C2 = ri.counts()
    assert (B1 == B2).all() and (C1 == C2).all()
    L = O.ref()
    idx = np.random.default_rng(seed).integers(0, db.n, 3000)
    for i in idx:
        i = int(i)
        assert L.ref_access_sa(ri.h, i) == sa[i]
        assert L.ref_access_bwt(ri.h, i) == bwt[i]
        for b in (0, 1, 2, 4, 8):
            assert L.ref_rank(ri.h, b, i) == O.orc().orc_rank(db.h, b, i)
    for b in range(16):
        assert L.ref_loglik(ri.h, b) == O.orc().orc_loglik(db.h, b) or (
            np.isinf(L.ref_loglik(ri.h, b)) and np.isinf(O.orc().orc_loglik(db.h, b)))


def _shared_ends_db(seed):
    """chromosomes whose ends (and whole sequences) recur, so that many suffixes of different
    chromosomes and blocks agree up to a '\0'"""
    rng = np.random.default_rng(seed)
    chroms, names = synth.metagenome(10, 3000, gc=(0.3, 0.7), seed=seed, strain_frac=0.4, div=(0.002, 0.01))
    tail = synth.random_genome(rng, 40)
    chroms = [np.concatenate([c[:len(c) - 40], tail]) if i % 3 == 0 else c for i, c in enumerate(chroms)]
    chroms += [chroms[0].copy(), chroms[4].copy(), synth.revcom(chroms[2])]
    return chroms


@pytest.mark.parametrize("block_size", [0, 7000, 15000, 26000, 40000, 78032, 78033])
def test_block_build_order_matches_reference(block_size):
    """the oracle's suffix array in msgseqtk-build's block-by-block order (blocks indexed
    separately and merged by FMDIndex::prepend) equals the reference's merged index, row by
    row (accessBWT everywhere, accessSA everywhere but the rows below numGaps that are
    multiples of 32), for block sizes giving 1 to 13 blocks.  The exception is the
    reference's: FMDIndex::buildSA (src/FMDIndex.cpp:496-508) stores a sample for every '\0'
    row i at SAsampled[rank1(i) - 1], but only rows with i % 32 == 0 or BWT[i] == 0 are
    sampled, so those writes land on the samples of rows 0, 32, ... (a merged index only;
    the one-step constructor samples with sampleSA).  Those rows hold suffixes that start
    with '\0', which no read search reaches."""
    chroms = _shared_ends_db(11)
    lens = np.array([len(c) for c in chroms], np.int64)
    db = O.OracleDb(chroms)
    bd, _, _ = db.arrays()
    n = len(bd)
    starts = np.concatenate([[0], np.cumsum(2 * (lens + 1))]).astype(np.int64)
    ends = O.block_ends(lens, block_size)
    sa = O.block_suffix_array(bd, ends)
    if 0 < block_size <= n:
        assert len(ends) >= 2 and ends[-1] == n
    ri = O.ref().ref_index_build_blocks(O.ptr(bd, O.C.c_uint8), O.ptr(starts, O.C.c_int64), len(chroms),
                                         block_size if block_size else n + 1)
    assert ri
    try:
        L = O.ref()
        gaps = int((bd == 0).sum())
        clobbered = 0
        for i in range(n):
            assert L.ref_access_bwt(ri, i) == (0 if sa[i] == 0 else bd[sa[i] - 1]), (i, len(ends))
            if len(ends) >= 2 and i < gaps and i % 32 == 0:
                clobbered += L.ref_access_sa(ri, i) != sa[i]
                continue
            assert L.ref_access_sa(ri, i) == sa[i], (i, len(ends))
        if len(ends) < 2:
            assert clobbered == 0
    finally:
        O.ref().ref_index_free(ri)
    if len(ends) >= 2:
        assert not np.array_equal(sa, O.block_suffix_array(bd, []))   # the order differs from one block


def test_backext_matches_reference():
    chroms, names, db = _db(3)
    bd, _, _ = db.arrays()
    ri = O.RefIndex(bd)
    B, Cc = db.counts()
    rng = np.random.default_rng(7)
    for trial in range(400):
        b0 = int(rng.choice([1, 2, 4, 8]))
        comp = {1: 8, 2: 4, 4: 2, 8: 1}[b0]
        v1 = np.array([Cc[b0], Cc[comp], Cc[b0 + 1] - Cc[b0]], np.int64)
        v2 = v1.copy()
        for step in range(int(rng.integers(1, 25))):
            b = int(rng.choice([0, 1, 2, 4, 8, 15, 3]))
            fwd = bool(rng.integers(0, 2))
            f1 = O.orc().orc_fwdext if fwd else O.orc().orc_backext
            f2 = O.ref().ref_fwdext if fwd else O.ref().ref_backext
            r1 = f1(db.h, O.ptr(v1, O.C.c_int64), b)
            r2 = f2(ri.h, O.ptr(v2, O.C.c_int64), b)
            assert r1 == r2 and (v1 == v2).all()
            if v1[2] <= 0:
                break


@pytest.mark.parametrize("minLen,maxE", [(17, 0.01), (12, 1.0), (5, float("inf")), (25, 1e-6)])
def test_smems_match_reference(minLen, maxE):
    chroms, names, db = _db(4)
    bd, _, _ = db.arrays()
    ri = O.RefIndex(bd)
    seqs, _, _, _ = synth.simulate_reads(chroms, 120, (40, 200), sub=0.02, indel=0.002, seed=minLen)
    seqs += [synth.encode("ACGTNNACGTAGGATTACA" * 3), synth.encode("A"), synth.encode("ACGTACGTAC"),
             synth.encode("NNNN")]
    cnt, rows, lps = O.orc_smems(db, seqs, minLen, maxE)
    off = 0
    for i, s in enumerate(seqs):
        r2, l2 = ri.smems(s, minLen, maxE)
        r1, l1 = rows[off:off + cnt[i]], lps[off:off + cnt[i]]
        off += cnt[i]
        assert np.array_equal(r1, r2), i
        assert np.array_equal(l1.view(np.int64), l2.view(np.int64)), i


def test_mems_match_reference():
    chroms, names, db = _db(5)
    bd, _, _ = db.arrays()
    ri = O.RefIndex(bd)
    seqs, _, _, _ = synth.simulate_reads(chroms, 60, 150, sub=0.02, indel=0.002, seed=9)
    cnt, rows, lps = O.orc_smems(db, seqs, 1, float("inf"), mems=True)
    off = 0
    for i, s in enumerate(seqs):
        r2, l2 = ri.smems(s, 1, float("inf"), mems=True)
        assert np.array_equal(rows[off:off + cnt[i]], r2)
        assert np.array_equal(lps[off:off + cnt[i]].view(np.int64), l2.view(np.int64))
        off += cnt[i]


def _random_seeds(rng, n):
    rows = []
    for _ in range(n):
        f = int(rng.integers(0, 60))
        ln = int(rng.integers(1, 30))
        st = int(rng.integers(0, 80))
        rows.append([f, f + ln, st, st + ln, int(rng.integers(0, 2)), int(rng.integers(1, 3))])
    return np.array(rows, np.int64).reshape(-1, 6)


def test_remove_redundant_matches_reference():
    rng = np.random.default_rng(11)
    for trial in range(200):
        n = int(rng.integers(0, 40))
        rows = _random_seeds(rng, n)
        if n and trial % 3 == 0:   # duplicates with distinct logP exercise the unstable sort
            rows = np.concatenate([rows, rows[: n // 2]])
        lp = -rng.random(len(rows)) * 30
        a, la = rows.copy(), lp.copy()
        b, lb = rows.copy(), lp.copy()
        ka = O.orc().orc_remove_redundant(O.ptr(a, O.C.c_int64), O.ptr(la, O.C.c_double), len(a))
        kb = O.ref().ref_remove_redundant(O.ptr(b, O.C.c_int64), O.ptr(lb, O.C.c_double), len(b))
        assert ka == kb
        assert np.array_equal(a[:ka], b[:kb]) and np.array_equal(la[:ka], lb[:kb])


def test_chains_match_reference():
    rng = np.random.default_rng(12)
    for trial in range(200):
        n = int(rng.integers(0, 30))
        rows = _random_seeds(rng, n)
        order = np.lexsort((rows[:, 3], rows[:, 2], rows[:, 5], rows[:, 4], rows[:, 1], rows[:, 0]))
        rows = np.ascontiguousarray(rows[order])
        mm, ind = int(rng.integers(0, 20)), int(rng.integers(0, 10))
        o1 = np.zeros(4096, np.int64)
        o2 = np.zeros(4096, np.int64)
        k1 = O.orc().orc_get_chains(O.ptr(rows, O.C.c_int64), len(rows), mm, ind, O.ptr(o1, O.C.c_int64), 4096)
        k2 = O.ref().ref_get_chains(O.ptr(rows, O.C.c_int64), len(rows), mm, ind, O.ptr(o2, O.C.c_int64), 4096)
        assert k1 == k2 and np.array_equal(o1[:k1], o2[:k2])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_reference_index_from_suffix_array_and_baseline_sam():
    """The reference's FMDIndex assembled from its own components over a given suffix array
    (ref_index_from_sa, the CPU baseline's index) answers like the one its constructor builds
    with divsufsort; the CPU baseline (the reference's own seeding and chaining + the oracle's
    DP and records) writes the oracle's SAM, single-end and paired."""
    chroms, names = synth.metagenome(6, 40000, gc=(0.3, 0.7), seed=77)
    odb = O.OracleDb(chroms, names)
    bd, bwt, sa = odb.arrays()
    r1 = O.RefIndex(bd, sa)
    r2 = O.RefIndex(bd)
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 600, 120, sub=0.015, indel=0.002, seed=78)
    for r in seqs[:150]:
        a, b = r1.smems(r), r2.smems(r)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.int64), b[1].view(np.int64))
    for i in range(0, len(bd), 997):
        assert O.ref().ref_access_sa(r1.h, i) == sa[i]
    exp, _ = O.orc_map(odb, O.default_params(), seqs, quals, rn, nthreads=4)
    got, _ = O.ref_baseline_map(r1, odb, O.default_params(), seqs, quals, rn, nthreads=4)
    assert exp == got
    (a1, q1, n1), (a2, q2, n2) = synth.simulate_pairs_fast(chroms, 300, 120, insert=(300, 40), seed=79, offtarget=0.05)
    pp = O.pe_params(300, 40)
    exp, _ = O.orc_map(odb, pp, a1, q1, n1, mates=(a2, q2, n2), nthreads=4)
    got, _ = O.ref_baseline_map(r1, odb, pp, a1, q1, n1, mates=(a2, q2, n2), nthreads=4)
    assert exp == got
