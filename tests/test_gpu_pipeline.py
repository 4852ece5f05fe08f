"""HIP path vs the oracle after seeding: seeds (findSeeds incl. the loglik sort, maxNSeed
cut and removeRedundant), per-chain alignments (tStart/tEnd, aligned span, score,
CIGAR, hazard flag, DP cells) and the SAM text of whole reads and pairs."""
import numpy as np
import pytest

import oracle_lib as O
import msgseqtk_amd as M
from msgseqtk_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    chroms, names = synth.metagenome(8, 40000, gc=(0.3, 0.7), seed=41, strain_frac=0.25)
    odb = O.OracleDb(chroms, names)
    m = M.Mapper(0).build(chroms, names)
    return chroms, names, odb, m


def _orc_alignments(odb, p, seqs, quals):
    import ctypes as C
    sq, offs = O.pack_reads(seqs)
    ql, _ = O.pack_reads(quals)
    cap = 1 << 16
    counts = np.zeros(max(1, len(seqs)), np.int64)
    rows = np.zeros((cap, 10), np.int64)
    sc = np.zeros(cap)
    cig = C.create_string_buffer(1 << 24)
    k = O.orc().orc_alignments(odb.h, C.byref(p), O.ptr(sq, C.c_uint8), O.ptr(ql, C.c_uint8),
                               O.ptr(offs, C.c_int64), len(seqs), O.ptr(counts, C.c_int64),
                               O.ptr(rows, C.c_int64), O.ptr(sc, C.c_double), cig, len(cig), cap)
    return counts[:len(seqs)], rows[:k], sc[:k], cig.value.decode().split("\n")[:k]


@pytest.mark.parametrize("minLen,maxE,maxNSeed", [(17, 0.01, 500), (12, 1.0, 3), (25, 1e-6, 500)])
def test_seed_parity(env, minLen, maxE, maxNSeed):
    chroms, names, odb, m = env
    seqs, _, _, _ = synth.simulate_reads(chroms, 500, (40, 300), sub=0.02, indel=0.003, seed=minLen,
                                         offtarget=0.05)
    c1, r1, l1 = O.orc_seeds(odb, seqs, minLen, maxE, maxNSeed)
    c2, r2, l2 = m.seeds(seqs, M.default_params(minLen=minLen, maxEvalue=maxE, maxNSeed=maxNSeed))
    assert np.array_equal(c1, c2)
    assert np.array_equal(r1, r2)
    assert np.array_equal(l1.view(np.int64), l2.view(np.int64))


@pytest.mark.parametrize("mode", [0, 1])
def test_alignment_parity(env, mode):
    chroms, names, odb, m = env
    seqs, quals, _, _ = synth.simulate_reads(chroms, 600, (40, 300), err_range=(0.0, 0.06), seed=50 + mode)
    p = M.default_params(mode=mode)
    c1, r1, s1, g1 = _orc_alignments(odb, O.default_params(mode=mode), seqs, quals)
    c2, r2, s2, g2 = m.alignments(seqs, quals, p)
    assert np.array_equal(c1, c2)
    assert np.array_equal(r1, r2)
    assert np.array_equal(s1, s2)
    assert g1 == g2


def test_alignment_parity_small_seeds(env):
    """tiny minLen makes seed corner cases (adjacent seeds, step-c cells, window edges) common"""
    chroms, names, odb, m = env
    seqs, quals, _, _ = synth.simulate_reads(chroms, 300, (20, 120), err_range=(0.02, 0.1), seed=77)
    for minLen in (6, 9):
        for mode in (0, 1):
            po = O.default_params(minLen=minLen, maxEvalue=float("inf"), mode=mode, maxNSeed=40)
            pm = M.default_params(minLen=minLen, maxEvalue=float("inf"), mode=mode, maxNSeed=40)
            c1, r1, s1, g1 = _orc_alignments(odb, po, seqs, quals)
            c2, r2, s2, g2 = m.alignments(seqs, quals, pm)
            assert np.array_equal(c1, c2) and np.array_equal(r1, r2) and np.array_equal(s1, s2) and g1 == g2


def _cmp_sam(a, b):
    assert len(a) == len(b)
    for i, (x, y) in enumerate(zip(a, b)):
        assert x == y, (i, x, y)


@pytest.mark.parametrize("mode,maxReport", [(0, 1), (1, 1), (0, 0)])
def test_sam_parity_se(env, mode, maxReport):
    chroms, names, odb, m = env
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 800, 150, sub=0.01, indel=0.001, seed=90 + mode,
                                              offtarget=0.05)
    po = O.default_params(mode=mode, maxReport=maxReport)
    pm = M.default_params(mode=mode, maxReport=maxReport)
    sam1, _ = O.orc_map(odb, po, seqs, quals, rn)
    sam2 = m.map((seqs, quals, rn), pm)
    _cmp_sam(sam1, sam2)


def test_sam_parity_pe(env):
    chroms, names, odb, m = env
    (a, qa, na), (b, qb, nb) = synth.simulate_pairs(chroms, 400, 150, insert=(350, 50), seed=95, offtarget=0.05)
    po = O.pe_params(350, 50)
    pm = M.default_params(insMean=350, insSD=50, insMin=350 - 500, insMax=350 + 500)
    sam1, _ = O.orc_map(odb, po, a, qa, na, mates=(b, qb, nb))
    sam2 = m.map((a, qa, na), pm, mates=(b, qb, nb))
    _cmp_sam(sam1, sam2)


@pytest.mark.parametrize("form", [0, 1])
def test_sam_parity_pe_paired_chains(env, form):
    """the whole path with two lanes per read in the SMEM search (the autotuner's usual
    choice on the benchmark workload), both one-lane forms"""
    chroms, names, odb, m = env
    (a, qa, na), (b, qb, nb) = synth.simulate_pairs(chroms, 400, 150, insert=(350, 50), seed=96 + form,
                                                    offtarget=0.05)
    po = O.pe_params(350, 50)
    pm = M.default_params(insMean=350, insSD=50, insMin=350 - 500, insMax=350 + 500)
    sam1, _ = O.orc_map(odb, po, a, qa, na, mates=(b, qb, nb))
    try:
        m.set("form", form).set("pair_chains", 1)
        sam2 = m.map((a, qa, na), pm, mates=(b, qb, nb))
    finally:
        m.set("form", 1).set("pair_chains", 0)
    _cmp_sam(sam1, sam2)


def test_resident_batch(env):
    chroms, names, odb, m = env
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 500, 150, sub=0.01, indel=0.001, seed=99)
    pm = M.default_params()
    m.upload(0, (seqs, quals, rn))
    m.run(0, pm, True)
    s = m.stats()
    text = m.sam()
    sam1, _ = O.orc_map(odb, O.default_params(), seqs, quals, rn)
    assert text == "".join(sam1)
    mapped = sum(1 for line in text.splitlines() if not (int(line.split("\t")[1]) & 4))
    assert s[0] == 500 and s[1] == mapped


def test_pipelined_batches(env):
    """submit/wait: batch k's host finishing overlaps batch k+1's GPU stages; every batch's
    SAM (SE and PE slots mixed) equals the oracle's"""
    chroms, names, odb, m = env
    pe = synth.simulate_pairs(chroms, 150, 120, insert=(350, 50), sub=0.01, indel=0.001, seed=7)
    (s1, q1, n1), (s2, q2, n2) = pe
    se = synth.simulate_reads(chroms, 300, (60, 200), sub=0.02, indel=0.002, seed=8, offtarget=0.1)
    pm_pe = M.default_params(insMean=350, insSD=50, insMin=350 - 500, insMax=350 + 500)
    pm_se = M.default_params()
    m.upload(1, (s1, q1, n1), (s2, q2, n2))
    m.upload(2, (se[0], se[1], se[2]))
    exp_pe, _ = O.orc_map(odb, O.pe_params(350, 50), s1, q1, n1, mates=(s2, q2, n2))
    exp_se, _ = O.orc_map(odb, O.default_params(), se[0], se[1], se[2])
    m.upload(3, (se[0], se[1], se[2]))
    exp_local, _ = O.orc_map(odb, O.default_params(mode=1), se[0], se[1], se[2])
    pm_local = M.default_params(mode=1)
    for rep in range(3):
        # three batches in flight: two run their device stages at once on the two worker
        # streams, the third waits for a free one
        m.submit(1, pm_pe, True)
        m.submit(2, pm_se, True)
        m.submit(3, pm_local, True)
        m.wait(1)
        assert m.sam() == "".join(exp_pe), rep
        m.wait(2)
        assert m.sam() == "".join(exp_se), rep
        m.wait(3)
        assert m.sam() == "".join(exp_local), rep
    with pytest.raises(RuntimeError):
        m.wait(2)   # nothing pending
