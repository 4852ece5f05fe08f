"""The learned index's exact-read path of form 2 against the CPU oracle.

A read that occurs exactly once in the BD text, whose halves read[0, L/2+1) and
read[L/2, L) occur once too, gets from SMEM_LIST::findAllSMEMS (src/SMEM.cpp:186-215,
:64-119) the single SMEM [0, L) with the bi-interval of the whole read.  Form 2
recognises such reads with two learned-index lower bounds (the key blocks of read[0, 16)
and read[L/2, L/2+16)) and text comparisons, and writes that SMEM without a search.
These tests build a database with the cases that must NOT take the path (a duplicated
segment, so the read or only its prefix occurs twice; a tandem-repeated 16-mer, so its
key block is longer than the bounded scan; an N run; reads at chromosome ends; reads
shorter than 32 bases; reads with one error) and check SMEM lists, logP bit patterns and
SAM against the oracle, plus the number of reads that took the path."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import oracle_lib as O  # noqa: E402
import msgseqtk_amd as M  # noqa: E402
from msgseqtk_amd import synth  # noqa: E402

pytestmark = pytest.mark.gpu

TRAIN = dict(segments=64, samples=1 << 16, steps=300, batch=1024, lr=3e-3, seed=5)


def _db():
    rng = np.random.default_rng(77)
    chroms, names = synth.metagenome(4, 20000, gc=(0.3, 0.7), seed=41, strain_frac=0.5)
    dup = synth.random_genome(rng, 3000)
    dup[500:1300] = chroms[0][1000:1800]          # an 800-bp exact copy of chrom 0
    tandem = synth.random_genome(rng, 2000)
    unit = synth.random_genome(rng, 16)
    tandem[1000:1000 + 16 * 40] = np.tile(unit, 40)   # key block of 40+ rows
    nrun = synth.random_genome(rng, 2000)
    nrun[990:1010] = 15
    chroms = list(chroms) + [dup, tandem, nrun]
    names = list(names) + ["dup", "tandem", "nrun"]
    return chroms, names


@pytest.fixture(scope="module")
def env():
    chroms, names = _db()
    odb = O.OracleDb(chroms, names)
    m = M.Mapper(0).build(chroms, names)
    info = m.lix_train(**TRAIN)
    print("lix info", info, flush=True)
    return chroms, names, odb, m


def _reads(chroms):
    rng = np.random.default_rng(9)
    seqs = []

    def take(t, pos, L, strand):
        f = np.ascontiguousarray(chroms[t][pos:pos + L], np.uint8)
        seqs.append(synth.revcom(f) if strand else f)

    # exact reads anywhere, both strands, lengths at and around the path's limits
    for _ in range(400):
        t = int(rng.integers(0, len(chroms)))
        L = int(rng.choice([31, 32, 33, 40, 100, 150, 251]))
        if len(chroms[t]) <= L:
            continue
        take(t, int(rng.integers(0, len(chroms[t]) - L + 1)), L, int(rng.integers(0, 2)))
    # chromosome starts and ends
    for t in range(len(chroms)):
        for L in (32, 150):
            for s in (0, 1):
                take(t, 0, L, s)
                take(t, len(chroms[t]) - L, L, s)
    # the duplicated segment: inside (read occurs twice), straddling its ends (a half
    # occurs twice), and the same in chrom 0
    for pos in (520, 600, 1100, 1150, 420, 450, 1200, 1250):
        for s in (0, 1):
            take(4, pos, 150, s)
    for pos in (1020, 900, 1700):
        take(0, pos, 150, 0)
    # the tandem repeat and the N run
    for pos in (950, 1000, 1200, 1500, 1600):
        take(5, pos, 150, 0)
    for pos in (900, 940, 1000):
        take(6, pos, 150, 0)
    # one substitution at the start, middle and end
    for k in range(30):
        t = int(rng.integers(0, 4))
        f = np.array(chroms[t][3000 + 97 * k:3000 + 97 * k + 150], np.uint8)
        j = (0, 74, 75, 76, 149)[k % 5]
        f[j] = {1: 2, 2: 4, 4: 8, 8: 1}[int(f[j])]
        seqs.append(f)
    return seqs


@pytest.mark.parametrize("group,read_filter,text_path,split,pair",
                         [(8, 1, 1, 0, 0), (0, 1, 1, 0, 0), (0, 1, 0, 0, 0), (16, 1, 1, 0, 0), (8, 0, 1, 0, 0),
                          (0, 1, 0, 8, 0), (0, 1, 1, 4, 0), (0, 0, 0, 16, 0), (0, 0, 0, 0, 1), (0, 1, 1, 0, 1)])
def test_exact_read_path_parity(env, group, read_filter, text_path, split, pair):
    """every search kernel behind the exact-read path (one lane per read with and without
    the text path, lane groups of 8 and 16, two lanes per read with paired chains), with and
    without the read filter, and the split launch (reads with no 16-mer at their middle,
    here random reads and reads with a substitution there, searched by lane groups of 4, 8
    or 16 first)"""
    chroms, names, odb, m = env
    seqs = _reads(chroms)
    if split:
        rng = np.random.default_rng(9)
        seqs += [synth.NT[rng.integers(0, 4, size=int(L))] for L in rng.integers(20, 300, 120)]
        for k in range(60):
            t = int(rng.integers(0, 4))
            f = np.array(chroms[t][5000 + 113 * k:5000 + 113 * k + 150], np.uint8)
            for j in (60 + k % 30, 90 - k % 20):
                f[j] = {1: 2, 2: 4, 4: 8, 8: 1}[int(f[j])]
            seqs.append(f)
    m.set("group_lanes", group).set("read_filter", read_filter).set("text_path", text_path)
    m.set("split_lanes", split).set("pair_chains", pair)
    try:
        for minLen, maxE in ((17, 0.01), (16, 1.0), (20, float("inf")), (40, 1e-30), (151, 0.01)):
            p = M.default_params(minLen=minLen, maxEvalue=maxE)
            c1, r1, l1 = O.orc_smems(odb, seqs, minLen, maxE)
            m.set("form", 2)
            c2, r2, l2 = m.smems(seqs, p)
            m.set("form", 1)
            assert np.array_equal(c1, c2), (minLen, maxE)
            assert np.array_equal(r1, r2), (minLen, maxE)
            assert np.array_equal(l1.view(np.int64), l2.view(np.int64)), (minLen, maxE)
    finally:
        m.set("form", 1).set("group_lanes", 8).set("read_filter", 1).set("text_path", 1).set("split_lanes", 0)
        m.set("pair_chains", 0)


def test_exact_read_path_sam_and_count(env):
    chroms, names, odb, m = env
    seqs = _reads(chroms)
    rng = np.random.default_rng(3)
    quals = [rng.integers(20, 41, size=len(s)).astype(np.uint8) for s in seqs]
    rn = ["x%d" % i for i in range(len(seqs))]
    sam1, _ = O.orc_map(odb, O.default_params(), seqs, quals, rn)
    m.set("form", 2)
    m.upload(0, (seqs, quals, rn))
    m.run(0, M.default_params())
    st = m.stats()
    sam2 = m.sam()
    m.set("form", 1)
    assert "".join(sam1) == sam2
    exact = int(st[26])
    # reads of 32+ bases that occur once with unique halves: most of the 400 random ones
    print("exact-path reads %d of %d, filtered %d" % (exact, len(seqs), int(st[27])), flush=True)
    assert 200 <= exact <= len(seqs) - 60
