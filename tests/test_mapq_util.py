"""msgs_util_mapq (the MAPQ step behind Alignment::calcMapQ of the C++ face) against the
oracle's calcMapQ, bit-exact, and the closed forms of a single and of two equal alignments."""
import math

import numpy as np

import oracle_lib as O
import msgseqtk_amd as M

LOG10E = math.log10(math.e)


def test_mapq_matches_oracle():
    rng = np.random.default_rng(5)
    for n in [1, 2, 3, 7, 33, 200]:
        for scale in [1, 30, 400]:   # 400: every loglik below the exp floor, the rescaled branch
            log10p = -rng.random(n) * scale
            pp, mq = M.calc_mapq(log10p / LOG10E)
            wpp, wmq = O.orc_calc_mapq(log10p)
            assert np.array_equal(pp.view(np.int64), wpp.view(np.int64)), (n, scale)
            assert np.array_equal(mq, wmq), (n, scale)


def test_mapq_closed_forms():
    pp, mq = M.calc_mapq([-3.0])
    assert pp[0] == 1.0 and mq[0] == 250          # 1 - postP = 0: MAPQ capped
    pp, mq = M.calc_mapq([-3.0, -3.0])
    assert pp[0] == pp[1] == 0.5 and mq[0] == mq[1] == 3   # -10 log10(0.5) = 3.01
