"""Candidate mate pairs on the device (msgs_pair_batch: AlignmentPE::getPairs, the insert model's
likelihood terms, AlignmentPE::sort and AlignmentPE::filter; src/Alignment.cpp:540-581,
src/Alignment.h:446-575,635-639, src/PairingScheme.cpp:18-25) against the oracle's pairing, on
alignment lists built to stress them:

  * log10 likelihoods drawn from a few values, so that many pairs share their sort key and the
    order of equal keys (the reference's unstable std::sort) is tested, in lists of up to ~900
    candidate pairs (introsort partitions, not only the insertion sort of short lists);
  * insert sizes inside and outside [insMin, insMax], mates on other chromosomes (insert size
    -1), mates on the same strand (no pair), empty lists on either side;
  * the default insert model, the -m/-d form, a uniform model (insMean 0) and a range that is
    open on the right; every pairing filter and maxNPair.

The whole mapping path's pairs (the device pairing inside msgs_map_batch and the resident
batches) are compared as SAM records by the PE tests of test_gpu_pipeline.py, test_gpu_configs.py,
test_gpu_full_configs.py and test_gpu_dark_paths.py."""
import numpy as np
import pytest

import oracle_lib as O
import msgseqtk_amd as M

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def mapper():
    return M.Mapper(0)


def _lists(seed, n, kmax, big=()):
    """n read pairs of alignment lists: offsets (2n+1), rows4 (tid, strand, alnStart, alnEnd), log10P"""
    rng = np.random.default_rng(seed)
    ks = rng.integers(0, kmax + 1, size=2 * n)
    for i in big:
        ks[i] = ks[n + i] = 30
    ks[0] = 0            # a read with no alignment, its mate with some
    offs = np.zeros(2 * n + 1, np.int64)
    offs[1:] = np.cumsum(ks)
    tot = int(offs[-1])
    base = rng.integers(0, 5000, size=n)
    rows = np.zeros((tot, 4), np.int64)
    lp = np.zeros(tot)
    levels = np.array([-1.5, -2.25, -3.0, -0.75])
    for side in range(2):
        for i in range(n):
            for j in range(int(offs[side * n + i]), int(offs[side * n + i + 1])):
                s = base[i] + int(rng.integers(-700, 700))
                rows[j] = (int(rng.choice([0, 0, 0, 1, 2])), int(rng.integers(1, 3)), s, s + int(rng.integers(40, 220)))
                lp[j] = levels[rng.integers(0, 4)] if rng.random() < 0.8 else -rng.random() * 8
    return offs, rows, lp


MODELS = {
    "default": dict(insMean=200, insSD=20, insMin=0, insMax=400),
    "m350-d50": dict(insMean=350, insSD=50, insMin=350 - 500, insMax=350 + 500),
    "uniform": dict(insMean=0, insSD=1, insMin=0, insMax=400),
    "open": dict(insMean=300, insSD=80, insMin=-5.5, insMax=1e12),
}
FILTERS = [dict(), dict(noDiscordant=1), dict(noTailOver=1), dict(noContain=1), dict(noOverlap=1),
           dict(maxNPair=3), dict(noDiscordant=1, noOverlap=1, maxNPair=2)]


def _check(mapper, offs, rows, lp, kw):
    c1, r1, l1 = O.orc_pairs(O.default_params(**kw), offs, rows, lp)
    c2, r2, l2 = mapper.pairs(offs, rows, lp, M.default_params(**kw))
    assert np.array_equal(c1, c2)
    assert np.array_equal(r1, r2)
    assert np.array_equal(l1.view(np.int64), l2.view(np.int64))   # bit patterns, -inf included
    return int(c1.sum())


@pytest.mark.parametrize("model", list(MODELS))
def test_pairs_match_oracle(mapper, model):
    offs, rows, lp = _lists(11, 400, 7, big=(5, 77))
    for flt in FILTERS:
        _say("pairs", model, flt)
        k = _check(mapper, offs, rows, lp, dict(MODELS[model], **flt))
        if not flt and model != "default":
            assert k > 0


def test_pairs_ties_and_long_lists(mapper):
    """every read pair with 30 x 30 alignments on one chromosome, four log10P values: hundreds of
    equal sort keys per list"""
    rng = np.random.default_rng(3)
    n, k = 12, 30
    offs = np.arange(2 * n + 1, dtype=np.int64) * k
    tot = 2 * n * k
    s = rng.integers(0, 300, size=tot)
    rows = np.stack([np.zeros(tot, np.int64), rng.integers(1, 3, size=tot), s, s + 100], axis=1)
    lp = np.array([-1.0, -2.0, -1.5, -0.5])[rng.integers(0, 4, size=tot)]
    for model in ("default", "uniform"):
        _say("ties", model)
        assert _check(mapper, offs, rows, lp, MODELS[model]) > 0


def test_pairs_empty(mapper):
    offs = np.zeros(1, np.int64)
    c, r, l = mapper.pairs(offs, np.zeros((0, 4), np.int64), np.zeros(0), M.default_params())
    assert len(c) == 0 and len(r) == 0
    offs = np.zeros(7, np.int64)    # three read pairs, no alignments
    c, r, l = mapper.pairs(offs, np.zeros((0, 4), np.int64), np.zeros(0), M.default_params())
    assert list(c) == [0, 0, 0]
