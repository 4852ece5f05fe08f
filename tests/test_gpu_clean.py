"""HIP path vs the oracle for msgseqtk-clean's seeding: SMEM_LIST::findMEMS (src/SMEM.cpp:
174-184, SMEM::findMEM :42-62) per read, bit-exact including logP bit patterns, the loglik
of each read's MEM list (SMEM_LIST::loglik, src/SMEM.cpp:307-312), and the per-read and
per-pair keep decision of src/msgseqtk-clean.cpp:413-488 against two databases.  The
oracle's findMEMS is checked against the reference's own code in
test_oracle_vs_reference.py."""
import numpy as np
import pytest

import oracle_lib as O
import msgseqtk_amd as M
from msgseqtk_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def two_dbs():
    ref, rn = synth.metagenome(4, 30000, gc=(0.3, 0.7), seed=61, strain_frac=0.25)
    bg, bn = synth.metagenome(4, 30000, gc=(0.3, 0.7), seed=62)
    oref, obg = O.OracleDb(ref, rn), O.OracleDb(bg, bn)
    mref, mbg = M.Mapper(0).build(ref, rn), M.Mapper(0).build(bg, bn)
    return ref, bg, oref, obg, mref, mbg


def _reads(ref, bg, n, seed):
    a, _, _, _ = synth.simulate_reads(ref, n, (30, 300), sub=0.02, indel=0.003, seed=seed, offtarget=0.1)
    b, _, _, _ = synth.simulate_reads(bg, n, (30, 300), sub=0.02, indel=0.003, seed=seed + 1)
    edge = [np.zeros(0, np.uint8), synth.encode("A"), synth.encode("N"), synth.encode("ACGTN" * 20),
            synth.encode("NNNNNNNNNN"), synth.encode("RYKMSWBDHV" * 5), ref[0][:17].copy(), ref[1][-150:].copy(),
            synth.revcom(bg[2][5000:5150])]
    return a + b + edge


def _list_loglik(counts, logP):
    """SMEM_LIST::loglik: the list's logP summed in list order"""
    out, k = [], 0
    for c in counts:
        acc = 0.0
        for j in range(int(c)):
            acc += float(logP[k + j])
        out.append(acc)
        k += int(c)
    return np.array(out, np.float64)


@pytest.mark.parametrize("minLen,maxE", [(17, 0.01), (1, float("inf")), (12, 1.0), (30, 1e-9), (17, 0.0)])
def test_mems_parity(two_dbs, minLen, maxE):
    ref, bg, oref, obg, mref, mbg = two_dbs
    seqs = _reads(ref, bg, 200, 100 + minLen)
    p = M.default_params(minLen=minLen, maxEvalue=maxE)
    c1, r1, l1 = O.orc_smems(oref, seqs, minLen, maxE, mems=True)
    c2, r2, l2, ll2 = mref.mems(seqs, p)
    assert np.array_equal(c1, c2)
    assert np.array_equal(r1, r2)
    assert np.array_equal(l1.view(np.int64), l2.view(np.int64))
    assert np.array_equal(_list_loglik(c1, l1).view(np.int64), ll2.view(np.int64))


@pytest.mark.parametrize("paired", [False, True])
def test_clean_parity(two_dbs, paired):
    ref, bg, oref, obg, mref, mbg = two_dbs
    p = M.default_params(minLen=17, maxEvalue=0.01)
    if paired:
        (a, _, _), (b, _, _) = synth.simulate_pairs(ref, 150, 150, insert=(350, 50), seed=71, offtarget=0.1)
        (c, _, _), (d, _, _) = synth.simulate_pairs(bg, 150, 150, insert=(350, 50), seed=72)
        seqs, mates = a + c, b + d
    else:
        seqs, mates = _reads(ref, bg, 200, 80), None

    def lists(odb, reads):
        cnt, _, lp = O.orc_smems(odb, reads, 17, 0.01, mems=True)
        return _list_loglik(cnt, lp)

    rl, bl = lists(oref, seqs), lists(obg, seqs)
    if paired:
        rl, bl = rl + lists(oref, mates), bl + lists(obg, mates)
    for min_lod in (0.0, 25.0):
        keep = (-rl + bl) > min_lod
        grl, gbl, gkeep = mref.clean(mbg, seqs, p, min_lod, mates=mates)
        assert np.array_equal(rl.view(np.int64), grl.view(np.int64))
        assert np.array_equal(bl.view(np.int64), gbl.view(np.int64))
        assert np.array_equal(keep, gkeep)
    assert 0 < int(keep.sum()) < len(keep)
