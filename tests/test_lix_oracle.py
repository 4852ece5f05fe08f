"""Pins the learned-index oracle (oracle/lix_oracle.py): the MLP gradient against
central finite differences, suffix keys against a per-suffix loop and the oracle's
suffix array order, bf16 rounding against torch.bfloat16, lower bounds against brute
force."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import lix_oracle as LX  # noqa: E402
import oracle_lib as O  # noqa: E402
from msgseqtk_amd import synth  # noqa: E402


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    W = rng.normal(0, 0.4, LX.N_PARAMS)
    X = rng.normal(0, 1, (37, 16))
    T = rng.normal(0, 1, 37)
    g = LX.mse_grad(W, X, T)
    h = 1e-6
    idx = np.concatenate([rng.choice(1024, 40, replace=False), 1024 + rng.choice(64, 10, replace=False),
                          1088 + rng.choice(64, 10, replace=False), [1152]])
    for k in idx:
        Wp, Wm = W.copy(), W.copy()
        Wp[k] += h
        Wm[k] -= h
        fd = (LX.mse_loss(Wp, X, T) - LX.mse_loss(Wm, X, T)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_suffix_keys_match_loop_and_sa_order():
    chroms, names = synth.metagenome(3, 2000, gc=(0.3, 0.7), seed=8)
    db = O.OracleDb(chroms, names)
    bd, _, sa = db.arrays()
    ct = LX.suffix_keys(bd, sa)
    assert np.all(ct[1:] >= ct[:-1])
    n = len(bd)
    for i in list(range(0, n, 97)) + [n - 1, n - 2, n - 16]:
        p = int(sa[i])
        k = 0
        for t in range(16):
            k = (k << 3) | (int(LX.SYM3[bd[p + t]]) if p + t < n else 0)
        assert int(ct[i]) == k


def test_bf16_matches_torch():
    import torch
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(0, 3, 5000), rng.normal(0, 1e-3, 1000), [0.0, -0.0, 1.0, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(LX.bf16(x).view(np.uint32), ref.view(np.uint32))


def test_lower_bounds_brute_force():
    rng = np.random.default_rng(6)
    ct = np.sort(rng.integers(0, 1000, 500).astype(np.uint64))
    keys = rng.integers(0, 1100, 300).astype(np.uint64)
    lb = LX.lower_bounds(ct, keys)
    for x, l in zip(keys, lb):
        assert l == sum(1 for v in ct if v < x)
