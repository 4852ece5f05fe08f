"""Learned index on the GPU against the CPU oracle (oracle/lix_oracle.py and the FMD
oracle): training gradients, determinism, bf16 MFMA inference against its VALU
restatement and the oracle, exact lower bounds / counts (FMDIndex::count,
src/FMDIndex.cpp:163-176) with the model, the root alone and a plain binary search,
the gain of the trained network over the method without it, and the form-2 read filter
(SMEMs and SAM bit-exact)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import lix_oracle as LX  # noqa: E402
import oracle_lib as O  # noqa: E402
import msgseqtk_amd as M  # noqa: E402
from msgseqtk_amd import synth  # noqa: E402

pytestmark = pytest.mark.gpu

TRAIN = dict(segments=64, samples=1 << 16, steps=400, batch=1024, lr=3e-3, seed=11)


@pytest.fixture(scope="module")
def env():
    chroms, names = synth.metagenome(6, 30000, gc=(0.3, 0.7), seed=23, strain_frac=0.34)
    odb = O.OracleDb(chroms, names)
    m = M.Mapper(0).build(chroms, names)
    info = m.lix_train(**TRAIN)
    print("lix info", info, flush=True)
    bd, _, sa = odb.arrays()
    ct = LX.suffix_keys(bd, sa)
    return chroms, names, odb, m, info, ct


def _fmd_count(odb, bases):
    """FMDIndex::count by backward search on the oracle (src/FMDIndex.cpp:163-176)"""
    B, Cc = odb.counts()
    b = int(bases[-1])
    comp = {1: 8, 2: 4, 4: 2, 8: 1}
    p, q, s = int(Cc[b]), int(Cc[comp[b]]), int(Cc[b + 1] - Cc[b])
    for x in bases[-2::-1]:
        r, v = odb.backext(p, q, s, int(x))
        p, q, s = (int(v[0]), int(v[1]), int(v[2]))
        if s <= 0:
            return 0
    return s


def test_gradient_matches_oracle(env):
    m = env[3]
    rng = np.random.default_rng(4)
    for B in (1, 300, 4096):
        W = rng.normal(0, 0.3, LX.N_PARAMS).astype(np.float32)
        X = rng.normal(0, 1, (B, 16)).astype(np.float32)
        T = rng.normal(0, 1, B).astype(np.float32)
        g, loss = m.lix_grad(W, X, T)
        gref = LX.mse_grad(W, X, T)
        # fp32 batch sums in a fixed order against float64: relative 1e-4 of the largest entry
        assert np.max(np.abs(g - gref)) <= 1e-4 * max(1.0, np.max(np.abs(gref))), B
        assert abs(loss - LX.mse_loss(W, X, T)) <= 1e-5 * max(1.0, LX.mse_loss(W, X, T))


def test_training_is_deterministic(env):
    chroms, names, odb, m, info, ct = env
    w1 = m.lix_weights()
    info2 = m.lix_train(**TRAIN)
    w2 = m.lix_weights()
    assert np.array_equal(w1.view(np.uint32), w2.view(np.uint32))
    assert info2["max_err"] == info["max_err"] and info2["loss"] == info["loss"]


def test_training_reduces_loss(env):
    info = env[4]
    assert info["loss"] < info["loss0"]


def test_mfma_inference_matches_valu_and_oracle(env):
    chroms, names, odb, m, info, ct = env
    rng = np.random.default_rng(7)
    keys = np.concatenate([ct[rng.integers(0, len(ct), 3000)], rng.integers(0, 1 << 48, 1000).astype(np.uint64)])
    p0 = m.lix_predict(keys, 0)
    p1 = m.lix_predict(keys, 1)
    kkey, krank = m.lix_knots()
    S = len(kkey) - 1
    j, t, base, width = LX.plr(kkey, krank, keys)
    W = m.lix_weights()
    ref = base + LX.infer_bf16(W, LX.features(keys, t)) * width
    # same bf16 products, fp32 sums in different orders: 1e-4 of a segment width
    assert np.max(np.abs(p0 - p1) / width) <= 1e-4
    assert np.max(np.abs(p0 - ref) / width) <= 1e-4
    assert S == TRAIN["segments"]


def test_lower_bounds_exact(env):
    chroms, names, odb, m, info, ct = env
    rng = np.random.default_rng(8)
    present = ct[rng.integers(0, len(ct), 20000)]
    absent = rng.integers(0, 1 << 48, 5000).astype(np.uint64)
    edge = np.array([0, 1, int(ct[0]), int(ct[-1]), int(ct[-1]) + 1, (1 << 48) - 1], np.uint64)
    keys = np.concatenate([present, absent, edge])
    ref = LX.lower_bounds(ct, keys)
    for use in (1, 0, -1):
        lb, probes, fell = m.lix_query(keys, use)
        assert np.array_equal(lb, ref), use
        if use >= 0:
            # present keys are inside the recorded bound by construction
            assert not fell[:len(present)].any(), use


def test_counts_match_fmd_backward_search(env):
    chroms, names, odb, m, info, ct = env
    rng = np.random.default_rng(9)
    kmers = []
    for _ in range(150):
        c = chroms[rng.integers(0, len(chroms))]
        p = rng.integers(0, len(c) - 16)
        kmers.append(np.array(c[p:p + 16], np.uint8))
    for _ in range(50):
        kmers.append(rng.choice(np.array([1, 2, 4, 8], np.uint8), 16))
    keys = np.array([LX.key_of(k) for k in kmers], np.uint64)
    lo, _, _ = m.lix_query(keys, 1)
    hi, _, _ = m.lix_query(keys + np.uint64(1), 1)
    for k, a, b in zip(kmers, lo, hi):
        assert b - a == _fmd_count(odb, k)


def test_trained_network_gains_over_root_and_binary_search(env):
    """what the network adds to the method without it: smaller prediction errors than the
    piecewise-linear root alone, hence narrower per-segment bounds and fewer key-table
    probes than the root alone or a plain binary search"""
    chroms, names, odb, m, info, ct = env
    rng = np.random.default_rng(10)
    rows = rng.integers(0, len(ct), 20000)
    keys = ct[rows]
    lb = LX.lower_bounds(ct, keys)
    err_mlp = np.abs(m.lix_predict(keys, 0) - lb)
    err_root = np.abs(m.lix_predict(keys, 2) - lb)
    _, pr_mlp, _ = m.lix_query(keys, 1)
    _, pr_root, _ = m.lix_query(keys, 0)
    _, pr_bin, _ = m.lix_query(keys, -1)
    print("mean |error| mlp %.1f root %.1f; max error mlp %d root %d; mean probes mlp %.2f root %.2f binary %.2f" % (
        err_mlp.mean(), err_root.mean(), info["max_err"], info["max_err_root"], pr_mlp.mean(), pr_root.mean(),
        pr_bin.mean()), flush=True)
    assert err_mlp.mean() < 0.8 * err_root.mean()
    assert pr_mlp.mean() < pr_root.mean() < pr_bin.mean()


def test_form2_filter_parity(env):
    chroms, names, odb, m, info, ct = env
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 400, (10, 250), err_range=(0.0, 0.05), seed=31, offtarget=0.3)
    seqs = list(seqs) + [np.array([1, 2, 15, 4] * 10, np.uint8), np.array([2] * 15, np.uint8)]
    quals = list(quals) + [np.full(40, 30, np.uint8), np.full(15, 30, np.uint8)]
    rn = list(rn) + ["nread", "short"]
    for minLen, maxE in ((17, 0.01), (16, 1.0), (12, 1.0)):
        p = M.default_params(minLen=minLen, maxEvalue=maxE)
        c1, r1, l1 = O.orc_smems(odb, seqs, minLen, maxE)
        m.set("form", 2)
        c2, r2, l2 = m.smems(seqs, p)
        m.set("form", 1)
        assert np.array_equal(c1, c2) and np.array_equal(r1, r2)
        assert np.array_equal(l1.view(np.int64), l2.view(np.int64))
    po = O.default_params()
    sam1, _ = O.orc_map(odb, po, seqs, quals, rn)
    m.set("form", 2)
    sam2 = m.map((seqs, quals, rn), M.default_params())
    m.set("form", 1)
    assert sam1 == sam2
