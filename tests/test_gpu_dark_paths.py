"""Paths the default workloads leave dark, against the oracle:

  * the serial DP fill (the class of jobs with more than 64 seeds or a rectangle wider than 512
    columns): forced for every job through the context option dp_serial, and reached on its own
    by long reads whose first several hundred bases match nowhere;
  * the pairing filters of msgseqtk-align (--no-mixed, --no-discordant, --no-tail-over,
    --no-contain, --no-overlap, --max-npair; src/Alignment.cpp:556-571) on pairs built to need
    them: short inserts (overlapping, contained and tail-over mates), mates from other pairs
    (discordant) and random mates (mixed);
  * non-default score sets in both alignment modes: integers, and multiples of 2^-k (whose double
    arithmetic is exact, computed as integers scaled by 2^k); other scores are refused;
  * the quality clamp max(q - 33, 2) of the FASTQ reader (src/QualStr.h:40-42) through the
    driver."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import msgseqtk_amd as M
from msgseqtk_amd import synth
from dbfiles import write_db, write_fastq
from test_gpu_pipeline import _orc_alignments

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "msgseqtk_amd", "msgseqtk-align-amd")


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def env():
    chroms, names = synth.metagenome(6, 60000, gc=(0.3, 0.7), seed=141, strain_frac=0.3)
    odb = O.OracleDb(chroms, names)
    m = M.Mapper(0).build(chroms, names)
    return chroms, names, odb, m


def _same_alignments(odb, m, seqs, quals, po, pm):
    c1, r1, s1, g1 = _orc_alignments(odb, po, seqs, quals)
    c2, r2, s2, g2 = m.alignments(seqs, quals, pm)
    assert np.array_equal(c1, c2)
    assert np.array_equal(r1, r2)
    assert np.array_equal(s1, s2)
    assert g1 == g2
    return len(r1)


def _same_sam(a, b):
    assert len(a) == len(b)
    for i, (x, y) in enumerate(zip(a, b)):
        assert x == y, (i, x, y)


@pytest.mark.parametrize("mode", [0, 1])
def test_serial_fill_forced(env, mode):
    chroms, names, odb, m = env
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 300, (40, 300), err_range=(0.0, 0.06), seed=150 + mode)
    _say("serial fill forced, mode", mode)
    try:
        m.set("dp_serial", 1)
        n = _same_alignments(odb, m, seqs, quals, O.default_params(mode=mode), M.default_params(mode=mode))
        small = synth.simulate_reads(chroms, 150, (20, 120), err_range=(0.02, 0.1), seed=160 + mode)
        po = O.default_params(minLen=7, maxEvalue=float("inf"), mode=mode, maxNSeed=40)
        pm = M.default_params(minLen=7, maxEvalue=float("inf"), mode=mode, maxNSeed=40)
        n += _same_alignments(odb, m, small[0], small[1], po, pm)
        sam1, _ = O.orc_map(odb, O.default_params(mode=mode), seqs, quals, rn)
        _same_sam(sam1, m.map((seqs, quals, rn), M.default_params(mode=mode)))
        m.upload(0, (seqs, quals, rn))
        m.run(0, M.default_params(mode=mode), True)
        st = m.stats()
    finally:
        m.set("dp_serial", 0)
    assert n > 300
    assert st[22] == 0 and st[32] > 0   # no job on the wavefront fill (the rest need no rectangle)


@pytest.mark.parametrize("mode", [0, 1])
def test_serial_fill_long_unseeded_heads(env, mode):
    """1-2 kb reads whose first 600-900 bases are random: the rectangle before the first seed is
    wider than 512 columns, so these jobs take the serial fill without any option; 150 bp reads
    in the same batch take the wavefront fill"""
    chroms, names, odb, m = env
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 60, (1000, 2000), sub=0.01, indel=0.001, seed=170 + mode)
    rng = np.random.default_rng(171 + mode)
    for s in seqs:
        h = int(rng.integers(600, 900))
        s[:h] = rng.choice(np.array([1, 2, 4, 8], np.uint8), h)
    short = synth.simulate_reads(chroms, 60, 150, sub=0.01, indel=0.001, seed=175 + mode, prefix="s")
    seqs, quals, rn = seqs + short[0], quals + short[1], rn + short[2]   # both fills in one batch
    _say("serial fill on long reads, mode", mode)
    _same_alignments(odb, m, seqs, quals, O.default_params(mode=mode), M.default_params(mode=mode))
    sam1, _ = O.orc_map(odb, O.default_params(mode=mode), seqs, quals, rn)
    _same_sam(sam1, m.map((seqs, quals, rn), M.default_params(mode=mode)))
    m.upload(1, (seqs, quals, rn))
    m.run(1, M.default_params(mode=mode), False)
    st = m.stats()
    assert st[32] > 0 and st[22] > 0   # both fills ran


def _hard_pairs(chroms, n, seed):
    """pairs whose mates overlap, contain each other (80 bp second mates) or tail over (inserts
    N(120, 40) for 150 bp mates), come from another pair (discordant) or are random (mixed)"""
    rng = np.random.default_rng(seed)
    a, b, qa, qb, names = [], [], [], [], []
    for i in range(n):
        c = chroms[int(rng.integers(0, len(chroms)))]
        lb = 150 if rng.random() < 0.6 else 80
        isz = int(np.clip(round(rng.normal(120, 40)), 30, 400))
        pos = int(rng.integers(lb, len(c) - max(isz, 150) - 8))
        x = synth.mutate(rng, c[pos:pos + 158], 0.005, 0.0005)[:150]
        y = synth.mutate(rng, synth.revcom(c[pos + isz - lb:pos + isz]), 0.005, 0.0005)[:lb]
        if rng.random() < 0.5:
            x, y = y, x
        a.append(np.ascontiguousarray(x, np.uint8))
        b.append(np.ascontiguousarray(y, np.uint8))
        names.append("h%d" % i)
    k = n // 5
    idx = rng.permutation(n)
    for i in idx[:k]:                      # mate of another pair
        b[i] = b[int(rng.integers(0, n))].copy()
    for i in idx[k:2 * k]:                 # random mate
        b[i] = rng.choice(np.array([1, 2, 4, 8], np.uint8), len(b[i]))
    qa = [rng.integers(20, 41, len(x)).astype(np.uint8) for x in a]
    qb = [rng.integers(20, 41, len(y)).astype(np.uint8) for y in b]
    return (a, qa, names), (b, qb, names)


FLAG_SETS = [
    dict(),
    dict(noMixed=1),
    dict(noDiscordant=1),
    dict(noTailOver=1),
    dict(noContain=1),
    dict(noOverlap=1),
    dict(maxNPair=1),
    dict(maxNPair=2),
    dict(noMixed=1, noDiscordant=1, noTailOver=1, noContain=1, noOverlap=1, maxNPair=3),
]


@pytest.mark.parametrize("flags", FLAG_SETS, ids=[",".join("%s=%d" % kv for kv in f.items()) or "none" for f in FLAG_SETS])
def test_pairing_filters(env, flags):
    chroms, names, odb, m = env
    (a, qa, na), (b, qb, nb) = _hard_pairs(chroms, 400, 180)
    _say("pairing filters", flags)
    for mode in (0, 1):
        po = O.pe_params(120, 40, mode=mode, maxReport=2, **flags)
        pm = M.default_params(insMean=120, insSD=40, insMin=120 - 400, insMax=120 + 400, mode=mode, maxReport=2,
                              **flags)
        sam1, _ = O.orc_map(odb, po, a, qa, na, mates=(b, qb, nb))
        sam2 = m.map((a, qa, na), pm, mates=(b, qb, nb))
        _same_sam(sam1, sam2)


SCORE_SETS = [
    dict(match=2, mismatch=4, gapO=6, gapE=1),
    dict(match=1, mismatch=1, gapO=0, gapE=1),
    dict(match=3, mismatch=2, gapO=10, gapE=3),
    dict(match=1, mismatch=5, gapO=2, gapE=0),
    dict(match=1, mismatch=2.5, gapO=4.5, gapE=1.5),        # multiples of 2^-k: exact in doubles
    dict(match=0.75, mismatch=3.25, gapO=5.125, gapE=2.0),
]


@pytest.mark.parametrize("scores", SCORE_SETS, ids=["m%(match)g-x%(mismatch)g-o%(gapO)g-e%(gapE)g" % s for s in SCORE_SETS])
def test_integer_score_sets(env, scores):
    chroms, names, odb, m = env
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 300, (40, 300), err_range=(0.0, 0.06), seed=190)
    _say("score set", scores)
    for mode in (0, 1):
        po = O.default_params(mode=mode, **scores)
        pm = M.default_params(mode=mode, **scores)
        _same_alignments(odb, m, seqs, quals, po, pm)
        sam1, _ = O.orc_map(odb, po, seqs, quals, rn)
        _same_sam(sam1, m.map((seqs, quals, rn), pm))


def test_non_dyadic_scores_refused(env):
    """scores that are not multiples of 1/256 would need the reference's rounding of every double
    operation: refused with an error, not approximated"""
    chroms, names, odb, m = env
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 20, 100, seed=195)
    with pytest.raises(RuntimeError, match="multiples of 1/256"):
        m.alignments(seqs, quals, M.default_params(mismatch=2.1))


def test_driver_quality_clamp(tmp_path):
    """FASTQ qualities below '#' read as Phred 2, as quality::encode does"""
    chroms, names = synth.metagenome(4, 30000, gc=(0.3, 0.7), seed=200)
    prefix = str(tmp_path / "db")
    write_db(prefix, chroms, names)
    odb = O.OracleDb(chroms, names)
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 300, 150, sub=0.02, indel=0.002, seed=201)
    rng = np.random.default_rng(202)
    raw = [rng.integers(0, 42, len(q)).astype(np.uint8) for q in quals]   # '!' .. 'J', including 0 and 1
    path = str(tmp_path / "r.fq")
    with open(path, "w") as f:
        for s, q, n in zip(seqs, raw, rn):
            f.write("@%s\n%s\n+\n%s\n" % (n, synth.decode(s), bytes(q + 33).decode()))
    clamped = [np.maximum(q, 2) for q in raw]
    exp, _ = O.orc_map(odb, O.default_params(mode=1), seqs, clamped, rn)
    _say("driver quality clamp")
    r = subprocess.run([CLI, prefix, path, "-o", str(tmp_path / "o.sam"), "--local"], capture_output=True,
                       text=True, timeout=110)
    assert r.returncode == 0, r.stderr
    body = [l for l in open(tmp_path / "o.sam").read().splitlines() if not l.startswith("@")]
    assert body == "".join(exp).splitlines()
    assert any("\t" + "#" in l for l in body)
