"""The finer-grained C-ABI entry points behind the C++ drop-in face, against the oracle and
closed forms:

  * msgs_search_batch (FMDIndex::count / locateAllFwd / locateAllRev searches) against naive
    occurrence counts and positions in the BD text, and the oracle's own extension loop;
  * msgs_sa_batch (FMDIndex::accessSA) against the oracle's suffix array;
  * msgs_chain_batch (SeedChain::getChains over caller seeds) against the oracle's getChains
    (itself checked against the reference's, tests/test_oracle_vs_reference.py);
  * msgs_dp_batch (Alignment::buildAlignments + evaluate over caller chains) against the
    hand-derived chain known answers and the oracle's alignChain on simulated chains."""
import math

import numpy as np
import pytest

import oracle_lib as O
import msgseqtk_amd as M
from msgseqtk_amd import synth
from test_oracle_pins import CHAIN_KAT, enc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    chroms, names = synth.metagenome(5, 3000, gc=(0.3, 0.7), seed=91)
    chroms[2] = chroms[2].copy()
    chroms[2][100:110] = 15
    m = M.Mapper(0).build(chroms, names)
    odb = O.OracleDb(chroms, names)
    return chroms, names, m, odb


def naive_positions(text, pat):
    t = bytes(text)
    p = bytes(pat)
    out, k = [], t.find(p)
    while k >= 0:
        out.append(k)
        k = t.find(p, k + 1)
    return out


def test_search_count_and_locate(small):
    chroms, names, m, odb = small
    bd, _, sa = odb.arrays()
    rng = np.random.default_rng(3)
    pats = []
    for _ in range(400):
        L = int(rng.integers(1, 12))
        pos = int(rng.integers(0, len(bd) - L))
        p = bd[pos:pos + L].copy()
        if rng.random() < 0.2:
            p[int(rng.integers(0, L))] = rng.choice([1, 2, 4, 8])
        pats.append(p)
    pats += [np.array([15], np.uint8), np.array([2, 15], np.uint8), np.array([15, 4], np.uint8), np.array([0], np.uint8)]
    fwd = m.search(pats, 1)
    rev = m.search(pats, 2)
    cnt = m.search(pats, 0)
    for p, f, r, c in zip(pats, fwd, rev, cnt):
        amb = lambda b: b not in (1, 2, 4, 8)
        if len(p) and not amb(p[0]) and not any(amb(b) for b in p):
            occ = naive_positions(bd, p)
            assert f[2] == len(occ)
            assert sorted(sa[f[0]:f[0] + f[2]]) == occ
            assert sorted(m.access_sa(np.arange(f[0], f[0] + f[2]))) == occ
            rocc = naive_positions(bd, synth.revcom(p))
            assert r[2] == len(rocc) and sorted(sa[r[1]:r[1] + r[2]]) == rocc
            assert max(c[2], 0) == len(occ)
        # the oracle's own backward search of FMDIndex::count
        b = int(p[-1])
        B, Cc = odb.counts()
        pq = [int(Cc[b]), int(Cc[int(synth.COMP[b])]), int(Cc[b + 1] - Cc[b])]
        for i in range(len(p) - 1, 0, -1):
            s, pq = odb.backext(pq[0], pq[1], pq[2], int(p[i - 1]))
            pq = list(pq)
            if s <= 0:
                break
        assert [int(x) for x in c] == pq


def test_access_sa(small):
    chroms, names, m, odb = small
    _, _, sa = odb.arrays()
    rows = np.arange(0, len(sa), 7)
    assert np.array_equal(m.access_sa(rows), sa[rows])


def test_chain_batch_matches_oracle(small):
    chroms, names, m, odb = small
    reads = synth.simulate_reads(chroms, 300, 80, sub=0.03, indel=0.005, seed=92)[0]
    c, rows, _ = O.orc_seeds(odb, reads, minLen=10, maxE=math.inf)
    off = np.concatenate([[0], np.cumsum(c)])
    sets = [rows[off[i]:off[i + 1]] for i in range(len(reads))]
    rng = np.random.default_rng(4)
    mm = rng.integers(0, 20, len(sets))
    ind = rng.integers(0, 10, len(sets))
    got = m.chains(sets, mm, ind)
    for s, g, a, b in zip(sets, got, mm, ind):
        flat = np.zeros(1 << 14, np.int64)
        r = np.ascontiguousarray(s if len(s) else np.zeros((1, 6), np.int64), np.int64)
        k = O.orc().orc_get_chains(O.ptr(r, O.C.c_int64), len(s), int(a), int(b), O.ptr(flat, O.C.c_int64), len(flat))
        want, i = [], 0
        while i < k:
            n = int(flat[i])
            want.append([int(x) for x in flat[i + 1:i + 1 + n]])
            i += 1 + n
        assert g == want


def test_dp_batch_known_answers():
    """The hand-derived chain known answers (tests/test_oracle_pins.py) through the device DP."""
    for name, chrom, read, seeds, mode, geo, score, cigar, md, log10p in CHAIN_KAT:
        m = M.Mapper(0).build([enc(chrom)], ["c"])
        rows = [(f, t, s, e, 0, st) for (f, t, s, e, st) in seeds]
        (r10, sc, lp, cg, mdt), = m.dp([enc(read)], [np.full(len(read), 30, np.uint8)], [[rows]],
                                       M.default_params(mode=mode))
        assert (r10[2], r10[3], r10[4], r10[5], r10[6], r10[7], r10[9]) == geo, name
        assert sc == score and cg == cigar and mdt == md, (name, sc, cg, mdt)
        assert abs(lp - log10p) < 1e-9, name
        m.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_dp_batch_matches_oracle_chains(small, mode):
    chroms, names, m, odb = small
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 400, (40, 160), err_range=(0.0, 0.06), seed=93 + mode)
    per_read, want = [], []
    p = O.default_params(mode=mode)
    for r, q in zip(seqs, quals):
        _, rows, _ = O.orc_seeds(odb, [r], minLen=12, maxE=math.inf)
        chains = []
        for ch in O_chains(rows, len(r)):
            chains.append(ch)
            want.append(O.orc_align_chain(odb, p, r, ch, q))
        per_read.append(chains)
    got = m.dp(seqs, quals, per_read, M.default_params(mode=mode))
    assert len(got) == len(want) and len(got) > 300
    for (r10, sc, lp, cg, mdt), (w10, wsc, wlp, wcg, wmd) in zip(got, want):
        assert list(r10) == list(w10) and sc == wsc and cg == wcg and mdt == wmd
        assert lp == wlp or abs(lp - wlp) < 1e-12


def O_chains(rows, L):
    if len(rows) == 0:
        return []
    flat = np.zeros(1 << 14, np.int64)
    r = np.ascontiguousarray(rows, np.int64)
    k = O.orc().orc_get_chains(O.ptr(r, O.C.c_int64), len(rows), math.ceil(L * 0.2), math.ceil(L * 0.1),
                               O.ptr(flat, O.C.c_int64), len(flat))
    out, i = [], 0
    while i < k:
        n = int(flat[i])
        out.append(rows[flat[i + 1:i + 1 + n]])
        i += 1 + n
    return out


def test_dp_batch_rejects_bad_chains(small):
    chroms, names, m, odb = small
    r = chroms[0][10:60].copy()
    q = np.full(50, 30, np.uint8)
    bad = [
        [(0, 10, 10, 20, 0, 1), (5, 15, 30, 40, 0, 1)],     # overlapping on the read
        [(0, 10, 10, 20, 0, 1), (20, 30, 30, 40, 1, 1)],    # two chromosomes
        [(0, 10, 10, 21, 0, 1)],                            # unequal lengths
        [(45, 55, 10, 20, 0, 1)],                           # past the read end
        [(0, 10, 2995, 3005, 0, 1)],                        # past the chromosome end
    ]
    for ch in bad:
        with pytest.raises(RuntimeError):
            m.dp([r], [q], [[ch]], M.default_params())
