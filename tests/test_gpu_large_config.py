"""Whole-path SAM parity (HIP path vs the oracle) on the largest benchmark database, at full
size: 1000 genomes with sizes log-uniform in 1-8 Mb (3.40 Gbp; BD text 6.80e9 symbols, past
2^32), GC U(0.3, 0.7), 30% near-duplicate strains at 0.5-5% divergence.  The index takes the
wide layout (64-bit rows, superblock counts, range-by-range suffix sorting); the learned
index is trained on the GPU over it.  A sample of 20k read pairs 2x150 (insert N(400, 80),
1% error, 10% off-target pairs) is mapped in the learned-index form and the plain forms and
compared record by record with the oracle, whose index is built from the device's suffix
array after sampled order checks (adjacent suffixes in the order of msgseqtk-build's default
block-by-block build, BWT = text before each suffix);
the device suffix sort equals the oracle's own at the sizes where that runs
(test_gpu_wide.py, test_gpu_seeding.py)."""
import numpy as np
import pytest

import oracle_lib as O
import msgseqtk_amd as M
from msgseqtk_amd import synth
from test_gpu_full_configs import check_suffix_order

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _diff(a, b):
    bad = [i for i, (x, y) in enumerate(zip(a, b)) if x != y]
    return len(bad), (bad[:3], [(a[i], b[i]) for i in bad[:2]])


@pytest.mark.timeout(1500)
def test_large_4gbp_pairs():
    _say("database: 1000 genomes, 1-8 Mb, 30% strains")
    chroms, names = synth.config_genomes(5)
    m = M.Mapper(0).build(chroms, names)
    _say("device index: BD length %d, wide layout %d" % (m.n, m.get("index_wide")))
    assert m.n == sum(2 * (len(c) + 1) for c in chroms) and m.n > 2 ** 32
    assert m.get("index_wide") == 1
    # msgseqtk-build's default 2048 MB blocks: the suffix order of its block-by-block build
    ends = O.block_ends([len(c) for c in chroms], 2048 << 20)
    _say("build blocks ending at %s" % ends)
    assert len(ends) >= 3 and m.get("index_blocks") == len(ends)
    bd, bwt, sa = m.arrays()
    check_suffix_order(bd, bwt, sa, samples=40000, ends=ends)
    del bd, bwt
    odb = O.OracleDb(chroms, names, sa=sa)
    del sa
    _say("oracle index from the device suffix array")
    info = m.lix_train(segments=1 << 20, samples=1 << 22, steps=1000, batch=4096, lr=1e-3, seed=7)
    _say("learned index", info)
    (s1, q1, n1), (s2, q2, n2) = synth.simulate_pairs_fast(chroms, 20000, 150, insert=(400, 80), sub=0.009,
                                                           indel=0.001, seed=0xC0FFEE + 1005, offtarget=0.1)
    exp, _ = O.orc_map(odb, O.pe_params(400, 80), s1, q1, n1, mates=(s2, q2, n2), nthreads=16)
    _say("oracle mapped the sample")
    pm = M.default_params(insMean=400, insSD=80, insMin=400 - 800, insMax=400 + 800)
    base = dict(form=1, group_lanes=8, read_filter=1, text_path=1, split_lanes=0, pair_chains=0)
    choices = [dict(form=2, group_lanes=8), dict(form=2, group_lanes=0, text_path=0),
               dict(form=2, group_lanes=0, text_path=1, pair_chains=1), dict(form=1), dict(form=0, pair_chains=1)]
    for ch in choices:
        for k, v in dict(base, **ch).items():
            m.set(k, v)
        got = m.map((s1, q1, n1), pm, mates=(s2, q2, n2))
        n, detail = _diff(exp, got)
        _say("%s: %d of %d pairs differ" % (ch, n, len(got)))
        assert n == 0, (ch, detail)
    m.close()
