"""HIP path vs the oracle: index (SA, BWT, counts, loglik), rank, backExt/fwdExt and
the SMEM search (SMEM_LIST::findAllSMEMS), bit-exact including logP bit patterns."""
import numpy as np
import pytest

import oracle_lib as O
import msgseqtk_amd as M
from msgseqtk_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dbs():
    chroms, names = synth.metagenome(6, 30000, gc=(0.3, 0.7), seed=21, strain_frac=0.34)
    odb = O.OracleDb(chroms, names)
    m = M.Mapper(0).build(chroms, names)
    return chroms, names, odb, m


def test_index_parity(dbs):
    chroms, names, odb, m = dbs
    bd1, bwt1, sa1 = odb.arrays()
    bd2, bwt2, sa2 = m.arrays()
    assert np.array_equal(bd1, bd2)
    assert np.array_equal(sa1, sa2)
    assert np.array_equal(bwt1, bwt2)
    B1, C1 = odb.counts()
    B2, C2 = m.counts()
    assert np.array_equal(B1, B2) and np.array_equal(C1[:17], C2[:17])
    for b in range(16):
        a, c = O.orc().orc_loglik(odb.h, b), m.loglik(b)
        assert a == c or (np.isinf(a) and np.isinf(c))


def test_rank_parity(dbs):
    chroms, names, odb, m = dbs
    rng = np.random.default_rng(1)
    pos = np.concatenate([rng.integers(0, m.n, 20000), np.arange(0, 300), [m.n - 1, m.n, -1]]).astype(np.int64)
    for b in (0, 1, 2, 4, 8):
        got = m.rank(np.full(len(pos), b, np.uint8), pos)
        exp = np.array([O.orc().orc_rank(odb.h, b, int(i)) for i in pos])
        assert np.array_equal(got, exp), b


def test_ext_parity(dbs):
    chroms, names, odb, m = dbs
    B, Cc = odb.counts()
    rng = np.random.default_rng(2)
    n = 5000
    b0 = rng.choice([1, 2, 4, 8], n)
    comp = {1: 8, 2: 4, 4: 2, 8: 1}
    pqs = np.stack([Cc[b0], [Cc[comp[b]] for b in b0], Cc[b0 + 1] - Cc[b0]], 1).astype(np.int64)
    for step in range(12):
        bases = rng.choice([0, 1, 2, 4, 8, 15, 5], n).astype(np.uint8)
        fwd = step % 3 == 1
        got, ret = m.ext(pqs, bases, fwd)
        exp = pqs.copy()
        eret = np.zeros(n, np.int64)
        for i in range(n):
            r, v = odb.backext(*exp[i], int(bases[i]), fwd=fwd)
            exp[i] = v
            eret[i] = r
        assert np.array_equal(got, exp) and np.array_equal(ret, eret)
        keep = got[:, 2] > 0
        pqs = np.where(keep[:, None], got, pqs)


def _check_smems(odb, m, seqs, minLen, maxE):
    """every device form: plain occ-block extensions (0), the unique-interval text path
    (1), lane groups of 8 and 16 per read (2; no learned index trained here), and the
    paired chains (two lanes per read) with and without the text path"""
    p = M.default_params(minLen=minLen, maxEvalue=maxE)
    c1, r1, l1 = O.orc_smems(odb, seqs, minLen, maxE)
    try:
        for form, group, pair in ((0, 0, 0), (1, 0, 0), (2, 8, 0), (2, 16, 0), (0, 0, 1), (1, 0, 1), (2, 0, 1)):
            m.set("form", form).set("pair_chains", pair)
            if form == 2:
                m.set("group_lanes", group)
            c2, r2, l2 = m.smems(seqs, p)
            assert np.array_equal(c1, c2), (form, group, pair)
            assert np.array_equal(r1, r2), (form, group, pair)
            assert np.array_equal(l1.view(np.int64), l2.view(np.int64)), (form, group, pair)
    finally:
        m.set("form", 1).set("group_lanes", 8).set("pair_chains", 0)


@pytest.mark.parametrize("minLen,maxE", [(17, 0.01), (10, 1.0), (3, float("inf")), (30, 1e-9), (17, 0.0)])
def test_smem_parity(dbs, minLen, maxE):
    chroms, names, odb, m = dbs
    seqs, _, _, _ = synth.simulate_reads(chroms, 400, (30, 300), sub=0.02, indel=0.003, seed=minLen,
                                         offtarget=0.05)
    _check_smems(odb, m, seqs, minLen, maxE)


def test_smem_pair_slot_growth(dbs):
    """paired chains with one SMEM slot per read: the merged counts exceed the slots, the
    host grows them and searches again, and the lists still equal the oracle's"""
    chroms, names, odb, m = dbs
    seqs, _, _, _ = synth.simulate_reads(chroms, 300, (30, 300), sub=0.03, indel=0.003, seed=77,
                                         offtarget=0.1)
    p = M.default_params()
    c1, r1, l1 = O.orc_smems(odb, seqs, 17, 0.01)
    try:
        m.set("form", 0).set("pair_chains", 1).set("smem_cap", 1)
        c2, r2, l2 = m.smems(seqs, p)
    finally:
        m.set("form", 1).set("pair_chains", 0).set("smem_cap", 32)
    assert c1.max() > 1
    assert np.array_equal(c1, c2) and np.array_equal(r1, r2)
    assert np.array_equal(l1.view(np.int64), l2.view(np.int64))


def test_smem_edge_reads(dbs):
    chroms, names, odb, m = dbs
    seqs = [np.zeros(0, np.uint8), synth.encode("A"), synth.encode("N"), synth.encode("ACGTN" * 20),
            synth.encode("NNNNNNNNNN"), synth.encode("RYKMSWBDHV" * 5), chroms[0][:17].copy(),
            chroms[1][100:400].copy(), synth.revcom(chroms[2][5000:5150]), chroms[3][-150:].copy(),
            chroms[0][:1].copy()]
    _check_smems(odb, m, seqs, 17, 0.01)
    _check_smems(odb, m, seqs, 1, float("inf"))


def test_smem_db_with_n():
    """the reference's SMEM_test database: chromosomes with an N, minLen 1, maxE inf"""
    c = synth.encode("ACGTCGTAGTACTACGNACGTCGTAGTACTACG")
    chroms = [c, c.copy(), c.copy(), c.copy()]
    names = ["chr1_1", "chr1_2", "chr2_1", "chr2_2"]
    odb = O.OracleDb(chroms, names)
    m = M.Mapper(0).build(chroms, names)
    bd1, bwt1, sa1 = odb.arrays()
    bd2, bwt2, sa2 = m.arrays()
    assert np.array_equal(sa1, sa2) and np.array_equal(bwt1, bwt2)
    # rank of every code, N and never-occurring IUPAC codes included
    pos = np.arange(-1, m.n + 1, dtype=np.int64)
    for b in range(16):
        got = m.rank(np.full(len(pos), b, np.uint8), pos)
        exp = np.array([O.orc().orc_rank(odb.h, b, int(i)) for i in pos])
        assert np.array_equal(got, exp), b
    seqs = [synth.encode("ACGTAGTA"), synth.encode("ACGTCGTAGTACTACGNACG"), synth.encode("TACGNAC")]
    _check_smems(odb, m, seqs, 1, float("inf"))
    _check_smems(odb, m, seqs, 4, 10.0)
