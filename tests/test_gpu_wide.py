"""The wide index layout (64-bit SA / ISA rows and bi-intervals, occ-block counts relative to
superblocks, range-by-range suffix sorting) against the oracle, on small texts.

Databases of 2^31 BD symbols or more always get this layout.  The context option
wide_index = 1 gives it to small databases too, with superblocks of 2^14 positions and
suffix-sorting ranges of 2^16 suffixes, so that every superblock and range boundary case of
the large layout occurs in texts the oracle indexes in a second: the suffix array and BWT,
ranks at superblock boundaries, backExt / fwdExt, pattern searches, accessSA, every SMEM
search kernel (64-bit row instantiations), MEM search, the learned index's exact-read path
and whole-path SAM (single-end and paired) must equal the oracle's, bit for bit."""
import numpy as np
import pytest

import oracle_lib as O
import msgseqtk_amd as M
from msgseqtk_amd import synth
from test_gpu_seeding import _check_smems

pytestmark = pytest.mark.gpu

SUPER = 1 << 14


def _say(*a):
    print(*a, flush=True)


def _diff(a, b):
    bad = [i for i, (x, y) in enumerate(zip(a, b)) if x != y]
    return len(bad), (bad[:3], [(a[i], b[i]) for i in bad[:2]])


def wide_mapper(chroms, names):
    m = M.Mapper(0).set("wide_index", 1).build(chroms, names)
    assert m.get("index_wide") == 1
    return m


@pytest.fixture(scope="module")
def wdb():
    chroms, names = synth.metagenome(8, 60000, gc=(0.3, 0.7), seed=31, strain_frac=0.25)
    _say("wide test database: BD length %d" % sum(2 * (len(c) + 1) for c in chroms))
    odb = O.OracleDb(chroms, names)
    m = wide_mapper(chroms, names)
    return chroms, names, odb, m


def test_wide_index_parity(wdb):
    chroms, names, odb, m = wdb
    bd1, bwt1, sa1 = odb.arrays()
    bd2, bwt2, sa2 = m.arrays()
    assert m.n > 16 * SUPER   # many superblocks and sorting ranges
    assert np.array_equal(bd1, bd2)
    assert np.array_equal(sa1, sa2)
    assert np.array_equal(bwt1, bwt2)
    B1, C1 = odb.counts()
    B2, C2 = m.counts()
    assert np.array_equal(B1, B2) and np.array_equal(C1[:17], C2[:17])
    # the narrow layout of the same database gives the same arrays
    n = M.Mapper(0).build(chroms, names)
    assert n.get("index_wide") == 0
    assert np.array_equal(n.arrays()[2], sa2)
    n.close()


def test_wide_rank_parity(wdb):
    chroms, names, odb, m = wdb
    rng = np.random.default_rng(3)
    edges = np.arange(SUPER, m.n, SUPER)
    near = (edges[:, None] + np.array([-129, -128, -1, 0, 1, 127, 128])[None, :]).ravel()
    pos = np.concatenate([rng.integers(0, m.n, 20000), near, np.arange(0, 300), [m.n - 1, m.n, -1]])
    pos = pos[(pos >= -1) & (pos <= m.n)].astype(np.int64)
    for b in (0, 1, 2, 4, 8):
        got = m.rank(np.full(len(pos), b, np.uint8), pos)
        exp = np.array([O.orc().orc_rank(odb.h, b, int(i)) for i in pos])
        assert np.array_equal(got, exp), b


def test_wide_ext_search_and_sa(wdb):
    chroms, names, odb, m = wdb
    B, Cc = odb.counts()
    rng = np.random.default_rng(4)
    n = 4000
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
        pqs = np.where((got[:, 2] > 0)[:, None], got, pqs)
    _, _, sa = odb.arrays()
    rows = rng.integers(0, m.n, 5000).astype(np.int64)
    assert np.array_equal(m.access_sa(rows), sa[rows])
    pats = [np.ascontiguousarray(chroms[int(rng.integers(0, len(chroms)))][k:k + 25]) for k in range(0, 4000, 97)]
    for mode in (0, 1, 2):
        got = m.search(pats, mode)
        assert got[:, 2].min() >= 0 and (mode == 0 or got[:, 2].max() >= 1)
    # counts of the patterns' occurrences equal the oracle's extension walks
    cnt = m.search(pats, 0)[:, 2]
    for p, c in zip(pats, cnt):
        q = [Cc[p[-1]], Cc[comp[int(p[-1])]], Cc[p[-1] + 1] - Cc[p[-1]]]
        for k in range(len(p) - 2, -1, -1):
            _, q = odb.backext(*q, int(p[k]), fwd=False)
            if q[2] <= 0:
                break
        assert max(q[2], 0) == c


@pytest.mark.parametrize("minLen,maxE", [(17, 0.01), (3, float("inf")), (30, 1e-9)])
def test_wide_smem_parity(wdb, minLen, maxE):
    chroms, names, odb, m = wdb
    seqs, _, _, _ = synth.simulate_reads(chroms, 400, (30, 300), sub=0.02, indel=0.003, seed=300 + minLen,
                                         offtarget=0.05)
    seqs += [np.zeros(0, np.uint8), synth.encode("ACGTN" * 20), chroms[0][:17].copy(), chroms[3][-150:].copy()]
    _check_smems(odb, m, seqs, minLen, maxE)


def test_wide_mems_parity(wdb):
    chroms, names, odb, m = wdb
    seqs, _, _, _ = synth.simulate_reads(chroms, 300, (30, 300), sub=0.02, indel=0.003, seed=41, offtarget=0.1)
    for minLen, maxE in ((17, 0.01), (1, float("inf"))):
        p = M.default_params(minLen=minLen, maxEvalue=maxE)
        c1, r1, l1 = O.orc_smems(odb, seqs, minLen, maxE, mems=True)
        c2, r2, l2, _ = m.mems(seqs, p)
        assert np.array_equal(c1, c2) and np.array_equal(r1, r2)
        assert np.array_equal(l1.view(np.int64), l2.view(np.int64))


def test_wide_sam_parity_se_learned_index(wdb):
    chroms, names, odb, m = wdb
    info = m.lix_train(segments=1024, samples=1 << 17, steps=300, batch=2048, lr=1e-3, seed=7)
    _say("learned index", info)
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 6000, (50, 300), err_range=(0.0, 0.04), seed=51,
                                              offtarget=0.03)
    exact = synth.simulate_reads(chroms, 2000, 150, sub=0.0, indel=0.0, seed=52)
    seqs, quals, rn = seqs + exact[0], quals + exact[1], rn + exact[2]
    exp, _ = O.orc_map(odb, O.default_params(), seqs, quals, rn, nthreads=16)
    base = dict(form=1, group_lanes=8, read_filter=1, text_path=1, split_lanes=0, pair_chains=0)
    choices = [dict(form=0), dict(form=1), dict(form=2, group_lanes=8), dict(form=2, group_lanes=16),
               dict(form=2, group_lanes=0, text_path=1), dict(form=2, group_lanes=0, text_path=0, split_lanes=8),
               dict(form=2, group_lanes=0, pair_chains=1), dict(form=0, pair_chains=1)]
    try:
        for ch in choices:
            for k, v in dict(base, **ch).items():
                m.set(k, v)
            got = m.map((seqs, quals, rn), M.default_params())
            n, detail = _diff(exp, got)
            _say("%s: %d of %d reads differ" % (ch, n, len(got)))
            assert n == 0, (ch, detail)
    finally:
        for k, v in base.items():
            m.set(k, v)


def test_wide_sam_parity_pairs(wdb):
    chroms, names, odb, m = wdb
    (s1, q1, n1), (s2, q2, n2) = synth.simulate_pairs_fast(chroms, 4000, 150, insert=(350, 50), sub=0.01,
                                                           indel=0.001, seed=61, offtarget=0.1)
    exp, _ = O.orc_map(odb, O.pe_params(350, 50), s1, q1, n1, mates=(s2, q2, n2), nthreads=16)
    pm = M.default_params(insMean=350, insSD=50, insMin=350 - 500, insMax=350 + 500)
    got = m.map((s1, q1, n1), pm, mates=(s2, q2, n2))
    n, detail = _diff(exp, got)
    assert n == 0, detail


def test_wide_db_with_n():
    """N in the text (the BWT's N positions) under the wide layout"""
    rng = np.random.default_rng(8)
    chroms = [synth.random_genome(rng, 40000) for _ in range(3)]
    chroms[1][1000:1020] = 15
    chroms[2][30000] = 15
    names = ["a", "b", "c"]
    odb = O.OracleDb(chroms, names)
    m = wide_mapper(chroms, names)
    assert np.array_equal(odb.arrays()[2], m.arrays()[2])
    pos = np.concatenate([np.arange(SUPER - 200, SUPER + 200), rng.integers(0, m.n, 3000)]).astype(np.int64)
    for b in (0, 1, 2, 4, 8, 15, 5):
        got = m.rank(np.full(len(pos), b, np.uint8), pos)
        exp = np.array([O.orc().orc_rank(odb.h, b, int(i)) for i in pos])
        assert np.array_equal(got, exp), b
    seqs = [np.ascontiguousarray(chroms[1][900:1100]), np.ascontiguousarray(chroms[2][29900:30100]),
            synth.encode("ACGTN" * 10)]
    _check_smems(odb, m, seqs, 1, float("inf"))
    _check_smems(odb, m, seqs, 17, 0.01)
    m.close()


def _shared_ends(seed, n, length):
    """genomes whose ends and whole sequences recur (strains at low divergence, a shared tail,
    exact copies), so that many suffixes of different blocks agree up to a '\\0'"""
    rng = np.random.default_rng(seed)
    chroms, names = synth.metagenome(n, length, gc=(0.3, 0.7), seed=seed, strain_frac=0.4, div=(0.002, 0.01))
    tail = synth.random_genome(rng, 60)
    chroms = [np.concatenate([c[:len(c) - 60], tail]) if i % 3 == 0 else c for i, c in enumerate(chroms)]
    chroms += [chroms[0].copy(), chroms[5].copy(), synth.revcom(chroms[2])]
    return chroms, ["c%03d" % i for i in range(len(chroms))]


@pytest.fixture(scope="module")
def blockdb():
    chroms, names = _shared_ends(71, 24, 45000)
    lens = np.array([len(c) for c in chroms], np.int64)
    ends = O.block_ends(lens, 1 << 20)
    odb = O.OracleDb(chroms, names)
    bd = odb.arrays()[0]
    _say("block test database: BD length %d, blocks ending at %s" % (len(bd), ends))
    return chroms, names, ends, bd, O.block_suffix_array(bd, ends), O.block_suffix_array(bd, [])


@pytest.mark.parametrize("wide", [0, 1])
def test_block_build_order(blockdb, wide):
    """msgseqtk-build's block-by-block suffix order (1 MB blocks over a 2.4 MB text): the
    device suffix array equals the oracle's block-order suffix array (itself equal to the
    reference's merged index, test_oracle_vs_reference.py), and SMEM lists and SAM equal the
    oracle's over that index"""
    chroms, names, ends, bd0, want, one_step = blockdb
    assert len(ends) >= 3
    m = M.Mapper(0).set("wide_index", wide).set("build_block_mb", 1).build(chroms, names)
    assert m.get("index_blocks") == len(ends) and m.get("index_wide") == 1
    bd, bwt, sa = m.arrays()
    assert np.array_equal(bd, bd0)
    assert np.array_equal(sa, want)
    assert not np.array_equal(sa, one_step)
    assert np.array_equal(bwt, np.where(sa == 0, 0, bd[np.maximum(sa - 1, 0)]))
    odb = O.OracleDb(chroms, names, sa=sa)
    # reads at chromosome ends (where suffixes of different blocks tie through '\0') and inside
    rng = np.random.default_rng(5)
    seqs = []
    for t in range(len(chroms)):
        for L in (40, 150):
            seqs.append(np.ascontiguousarray(chroms[t][len(chroms[t]) - L:]))
            seqs.append(synth.revcom(chroms[t][len(chroms[t]) - L:]))
    more = synth.simulate_reads(chroms, 3000, (50, 250), sub=0.01, indel=0.001, seed=72)
    seqs += more[0]
    _check_smems(odb, m, seqs, 17, 0.01)
    quals = [np.full(len(x), 30, np.uint8) for x in seqs]
    rn = ["r%d" % i for i in range(len(seqs))]
    exp, _ = O.orc_map(odb, O.default_params(), seqs, quals, rn, nthreads=16)
    got = m.map((seqs, quals, rn), M.default_params())
    n, detail = _diff(exp, got)
    assert n == 0, detail
    # the default block size builds this database in one step
    one = M.Mapper(0).build(chroms, names)
    assert one.get("index_blocks") == 1
    assert np.array_equal(one.arrays()[2], one_step)
    one.close()
    m.close()
