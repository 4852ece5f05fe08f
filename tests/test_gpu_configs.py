"""Whole-path SAM parity (HIP path vs the oracle) on the workload families of the
benchmark configurations, at sizes the oracle maps in seconds:

  * the CPU-runnable case at full size: 8 genomes x 250 kb (GC 0.5), 100k single-end
    150 bp reads, 1% substitutions, 0.1% indels;
  * uneven reads: 16 genomes x 250 kb, GC U(0.3, 0.7), 20k single-end reads with lengths
    log-uniform in 50-300 bp and per-read error U(0, 5%);
  * near-duplicate strains: 40 genomes x 100 kb, 20% strains at 1-5% divergence, 20k reads
    of 150 bp at 1% error, mapped with the learned-index form (trained on the GPU) and
    with the plain form;
  * paired reads: 16 genomes x 250 kb, 20k pairs 2x150, insert N(350, 50), 0.5% / 0.05%
    errors, 5% off-target pairs.

SAM text is compared per read (records keyed by input order, as the reference's output
order under OpenMP is not fixed)."""
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


def test_cpu_case_full_size_se():
    _say("case: 8 x 250 kb, 100k SE 150 bp")
    chroms, names = synth.metagenome(8, 250000, gc=None, seed=101)
    odb = O.OracleDb(chroms, names)
    m = M.Mapper(0).build(chroms, names)
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 100000, 150, sub=0.01, indel=0.001, seed=102)
    exp, _ = O.orc_map(odb, O.default_params(), seqs, quals, rn, nthreads=16)
    got = m.map((seqs, quals, rn), M.default_params())
    n, info = _diff(exp, got)
    assert n == 0, info
    mapped = sum(1 for t in got if "\t4\t*" not in t[:64])
    _say("mapped %d of %d" % (mapped, len(got)))
    assert mapped > 0.9 * len(got)


def test_uneven_reads_se():
    _say("case: 16 x 250 kb, 20k SE 50-300 bp, error U(0, 5%)")
    chroms, names = synth.metagenome(16, 250000, gc=(0.3, 0.7), seed=111)
    odb = O.OracleDb(chroms, names)
    m = M.Mapper(0).build(chroms, names)
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 20000, (50, 300), err_range=(0.0, 0.05), seed=112)
    exp, _ = O.orc_map(odb, O.default_params(), seqs, quals, rn, nthreads=16)
    for form in (1, 0):
        m.set("form", form)
        got = m.map((seqs, quals, rn), M.default_params())
        n, info = _diff(exp, got)
        assert n == 0, (form, info)


def test_strains_learned_index_form():
    _say("case: 40 x 100 kb with 20% strains, 20k SE 150 bp at 1%, learned-index form")
    chroms, names = synth.metagenome(40, 100000, gc=(0.3, 0.7), seed=121, strain_frac=0.2)
    odb = O.OracleDb(chroms, names)
    m = M.Mapper(0).build(chroms, names)
    info = m.lix_train(segments=4096, samples=1 << 18, steps=300, batch=4096, lr=1e-3, seed=7)
    _say("learned index", info)
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 20000, 150, sub=0.009, indel=0.001, seed=122)
    exp, _ = O.orc_map(odb, O.default_params(), seqs, quals, rn, nthreads=16)
    for form, group, split in ((2, 8, 0), (2, 0, 0), (2, 0, 8), (0, 0, 0)):
        m.set("form", form).set("group_lanes", group).set("split_lanes", split)
        got = m.map((seqs, quals, rn), M.default_params())
        n, detail = _diff(exp, got)
        assert n == 0, (form, group, split, detail)
    m.set("form", 1).set("group_lanes", 8).set("split_lanes", 0)


def test_pairs_with_offtarget():
    _say("case: 16 x 250 kb, 20k PE 2x150, insert N(350, 50), 5% off-target")
    chroms, names = synth.metagenome(16, 250000, gc=(0.3, 0.7), seed=131)
    odb = O.OracleDb(chroms, names)
    m = M.Mapper(0).build(chroms, names)
    m.lix_train(segments=4096, samples=1 << 18, steps=300, batch=4096, lr=1e-3, seed=7)
    (s1, q1, n1), (s2, q2, n2) = synth.simulate_pairs_fast(chroms, 20000, 150, insert=(350, 50), sub=0.005,
                                                           indel=0.0005, seed=132, offtarget=0.05)
    exp, _ = O.orc_map(odb, O.pe_params(350, 50), s1, q1, n1, mates=(s2, q2, n2), nthreads=16)
    pm = M.default_params(insMean=350, insSD=50, insMin=350 - 500, insMax=350 + 500)
    for form, split in ((2, 0), (2, 4), (1, 0)):
        m.set("form", form).set("group_lanes", 0 if split else 8).set("split_lanes", split)
        got = m.map((s1, q1, n1), pm, mates=(s2, q2, n2))
        n, detail = _diff(exp, got)
        assert n == 0, (form, split, detail)
    m.set("form", 1).set("group_lanes", 8).set("split_lanes", 0)
