"""Whole-path SAM parity (HIP path vs the oracle) at the full database sizes of the benchmark
configurations, on read samples the oracle maps in seconds:

  * 64 genomes x 4 Mb (256 Mbp, BD length 512,000,128): 20k pairs 2x150, insert N(350, 50),
    0.5% / 0.05% errors, 5% off-target pairs; and 20k single-end reads of 50-300 bp
    (log-uniform) at a per-read error rate U(0, 5%);
  * 256 genomes x 4 Mb with 20% near-duplicate strains (1 Gbp, BD length ~2.048e9):
    20k single-end 150 bp reads at 1% error, with the learned index trained on the GPU and
    every form-2 search kernel choice, and the plain forms.

The oracle's index is built from the device's suffix array; the device suffix array is
checked at full size by sampled size-independent properties (adjacent suffixes in order,
BWT = text before each suffix), and equals the oracle's own suffix sort at the sizes where
that runs (test_gpu_seeding.py).  SAM text is compared per read (pair)."""
import numpy as np
import pytest

import oracle_lib as O
import msgseqtk_amd as M
from msgseqtk_amd import synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _diff(a, b):
    bad = [i for i, (x, y) in enumerate(zip(a, b)) if x != y]
    return len(bad), (bad[:3], [(a[i], b[i]) for i in bad[:2]])


def check_suffix_order(bd, bwt, sa, samples=20000, seed=5, ends=()):
    """adjacent rows of the suffix array in byte order (0 lowest, a suffix that is a prefix of
    another first) on sampled rows whose first 64 symbols differ; BWT[i] = BD[SA[i]-1].  With
    msgseqtk-build's blocks (BD ends in `ends`): a suffix ends at its block's end, and two
    suffixes of different blocks that agree up to a shared '\0' are ordered later block first."""
    n = len(sa)
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, n - 1, samples)
    pad = np.concatenate([bd, np.zeros(64, np.uint8)])
    a, b = sa[rows], sa[rows + 1]
    ends = np.asarray(ends if len(ends) else [n], np.int64)
    blk_a, blk_b = np.searchsorted(ends, a, side="right"), np.searchsorted(ends, b, side="right")
    end_a, end_b = ends[blk_a], ends[blk_b]
    off = np.arange(64)[None, :]
    wa = pad[a[:, None] + off].astype(np.int16)
    wb = pad[b[:, None] + off].astype(np.int16)
    # past its block's end a suffix is exhausted: mark it below every symbol
    wa[a[:, None] + off >= end_a[:, None]] = -1
    wb[b[:, None] + off >= end_b[:, None]] = -1
    diff = wa != wb
    has = diff.any(axis=1)
    k = diff.argmax(axis=1)
    # a '\0' both share before the first difference: the block rule decides
    zero_before = ((wa == 0) & (off < k[:, None])).any(axis=1)
    cross = zero_before & (blk_a != blk_b)
    plain = wa[np.arange(samples), k] < wb[np.arange(samples), k]
    ok = np.where(cross, blk_a > blk_b, plain)
    assert ok[has].all(), "suffix array rows out of order"
    r = rng.integers(0, n, samples)
    want = np.where(sa[r] == 0, 0, bd[np.maximum(sa[r] - 1, 0)])
    assert np.array_equal(bwt[r], want)
    assert sa.min() == 0 and sa.max() == n - 1


@pytest.fixture(scope="module")
def db256():
    _say("database: 64 x 4 Mb")
    chroms, names = synth.config_genomes(2)
    m = M.Mapper(0).build(chroms, names)
    _say("device index: BD length %d" % m.n)
    assert m.n == 512_000_128
    bd, bwt, sa = m.arrays()
    check_suffix_order(bd, bwt, sa)
    odb = O.OracleDb(chroms, names, sa=sa)
    del bd, bwt, sa
    _say("oracle index from the device suffix array")
    yield chroms, names, m, odb
    m.close()


@pytest.mark.timeout(900)
def test_pairs_full_database(db256):
    chroms, names, m, odb = db256
    _say("case: 20k PE 2x150 on 64 x 4 Mb, insert N(350, 50), 5% off-target")
    (s1, q1, n1), (s2, q2, n2) = synth.simulate_pairs_fast(chroms, 20000, 150, insert=(350, 50), sub=0.005,
                                                           indel=0.0005, seed=0xC0FFEE + 1002, offtarget=0.05)
    exp, _ = O.orc_map(odb, O.pe_params(350, 50), s1, q1, n1, mates=(s2, q2, n2), nthreads=16)
    pm = M.default_params(insMean=350, insSD=50, insMin=350 - 500, insMax=350 + 500)
    m.lix_train(segments=1 << 16, samples=1 << 20, steps=300, batch=4096, lr=1e-3, seed=7)
    for form, pair in ((1, 0), (0, 1), (2, 0), (2, 1)):
        m.set("form", form).set("group_lanes", 0).set("pair_chains", pair)
        got = m.map((s1, q1, n1), pm, mates=(s2, q2, n2))
        n, detail = _diff(exp, got)
        _say("form %d pair_chains %d: %d of %d pairs differ" % (form, pair, n, len(got)))
        assert n == 0, (form, pair, detail)
    m.set("form", 1).set("pair_chains", 0)


@pytest.mark.timeout(900)
def test_uneven_reads_full_database(db256):
    chroms, names, m, odb = db256
    _say("case: 20k SE 50-300 bp on 64 x 4 Mb, error U(0, 5%)")
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 20000, (50, 300), err_range=(0.0, 0.05), seed=0xC0FFEE + 1003)
    exp, _ = O.orc_map(odb, O.default_params(), seqs, quals, rn, nthreads=16)
    for form in (1, 0):
        m.set("form", form)
        got = m.map((seqs, quals, rn), M.default_params())
        n, detail = _diff(exp, got)
        _say("form %d: %d of %d reads differ" % (form, n, len(got)))
        assert n == 0, (form, detail)
    m.set("form", 1)


@pytest.mark.timeout(1200)
def test_strains_1gbp_learned_index():
    _say("database: 256 x 4 Mb with 20% strains")
    chroms, names = synth.config_genomes(4)
    m = M.Mapper(0).build(chroms, names)
    _say("device index: BD length %d" % m.n)
    assert m.n == sum(2 * (len(c) + 1) for c in chroms) and m.n < 2 ** 31   # ~2.048e9 (strains carry indels)
    bd, bwt, sa = m.arrays()
    check_suffix_order(bd, bwt, sa)
    del bd, bwt
    odb = O.OracleDb(chroms, names, sa=sa)
    del sa
    _say("oracle index from the device suffix array")
    info = m.lix_train(segments=1 << 20, samples=1 << 22, steps=1000, batch=4096, lr=1e-3, seed=7)
    _say("learned index", info)
    seqs, quals, rn = synth.simulate_reads_fast(chroms, 20000, 150, sub=0.009, indel=0.001, seed=0xC0FFEE + 1004)
    exp, _ = O.orc_map(odb, O.default_params(), seqs, quals, rn, nthreads=16)
    _say("oracle mapped the sample")
    choices = [
        dict(form=2, group_lanes=8), dict(form=2, group_lanes=16), dict(form=2, group_lanes=8, read_filter=0),
        dict(form=2, group_lanes=0, text_path=1), dict(form=2, group_lanes=0, text_path=0),
        dict(form=2, group_lanes=0, text_path=1, split_lanes=4), dict(form=2, group_lanes=0, text_path=0, split_lanes=16),
        dict(form=2, group_lanes=0, text_path=1, pair_chains=1), dict(form=2, group_lanes=0, text_path=0, pair_chains=1),
        dict(form=1), dict(form=0),
    ]
    base = dict(form=1, group_lanes=8, read_filter=1, text_path=1, split_lanes=0, pair_chains=0)
    for ch in choices:
        for k, v in dict(base, **ch).items():
            m.set(k, v)
        got = m.map((seqs, quals, rn), M.default_params())
        n, detail = _diff(exp, got)
        _say("%s: %d of %d reads differ" % (ch, n, len(got)))
        assert n == 0, (ch, detail)
    m.close()
