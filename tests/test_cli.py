"""msgseqtk-align-amd: the drop-in driver (options, reference database files, FASTQ)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from msgseqtk_amd import synth
from dbfiles import write_db, write_fastq

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "msgseqtk_amd", "msgseqtk-align-amd")


def test_cli_usage():
    r = subprocess.run([CLI, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "READ-FILE" in r.stderr
    r = subprocess.run([CLI, "db", "reads.txt", "-o", "x.sam"], capture_output=True, text=True)
    assert r.returncode != 0


@pytest.mark.gpu
@pytest.mark.parametrize("paired", [False, True])
def test_cli_matches_oracle(tmp_path, paired):
    chroms, names = synth.metagenome(6, 30000, gc=(0.3, 0.7), seed=61)
    prefix = str(tmp_path / "db")
    write_db(prefix, chroms, names)
    odb = O.OracleDb(chroms, names)
    if paired:
        (a, qa, na), (b, qb, nb) = synth.simulate_pairs(chroms, 300, 150, insert=(350, 50), seed=62)
        write_fastq(str(tmp_path / "r1.fq"), a, qa, [n + "/1" for n in na])
        write_fastq(str(tmp_path / "r2.fq"), b, qb, [n + "/2" for n in nb])
        args = [CLI, prefix, str(tmp_path / "r1.fq"), str(tmp_path / "r2.fq"), "-o", str(tmp_path / "o.sam"),
                "-m", "350", "-d", "50", "--batch", "128"]
        exp, _ = O.orc_map(odb, O.pe_params(350, 50), a, qa, na, mates=(b, qb, nb))
    else:
        seqs, quals, rn, _ = synth.simulate_reads(chroms, 500, 150, sub=0.01, indel=0.001, seed=63)
        write_fastq(str(tmp_path / "r.fq"), seqs, quals, rn)
        args = [CLI, prefix, str(tmp_path / "r.fq"), "-o", str(tmp_path / "o.sam"), "--batch", "200"]
        exp, _ = O.orc_map(odb, O.default_params(), seqs, quals, rn)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = open(tmp_path / "o.sam").read().splitlines()
    head = [l for l in lines if l.startswith("@")]
    body = [l for l in lines if not l.startswith("@")]
    assert sum(1 for l in head if l.startswith("@SQ")) == len(chroms)
    assert body == "".join(exp).splitlines()


def _int_type(v):
    """bam_aux_update_int's type for a new integer tag (strict bounds)"""
    return "i" if v < -32768 else "s" if v < -128 else "c" if v < 0 else "C" if v < 255 else "S" if v < 65535 else "I"


@pytest.mark.gpu
@pytest.mark.parametrize("paired", [False, True])
def test_cli_bam_output(tmp_path, paired):
    """-o x.bam: the header and records of -o x.sam, in BGZF blocks of at most 0xff00 bytes
    with the end-of-file marker; integer tags in bam_aux_update_int's types, float tags as
    floats; the header text is the reference's @HD / @PG lines (its @SQ lines are the
    binary target list)"""
    import bamfile
    chroms, names = synth.metagenome(5, 30000, gc=(0.3, 0.7), seed=71)
    prefix = str(tmp_path / "db")
    write_db(prefix, chroms, names)
    if paired:
        (a, qa, na), (b, qb, nb) = synth.simulate_pairs(chroms, 3000, 150, insert=(350, 50), seed=72,
                                                        offtarget=0.05)
        write_fastq(str(tmp_path / "r1.fq"), a, qa, [n + "/1" for n in na])
        write_fastq(str(tmp_path / "r2.fq"), b, qb, [n + "/2" for n in nb])
        ins = [str(tmp_path / "r1.fq"), str(tmp_path / "r2.fq")]
        extra = ["-m", "350", "-d", "50", "-k", "0"]
    else:
        seqs, quals, rn, _ = synth.simulate_reads(chroms, 5000, 150, sub=0.01, indel=0.001, seed=73, offtarget=0.05)
        write_fastq(str(tmp_path / "r.fq"), seqs, quals, rn)
        ins = [str(tmp_path / "r.fq")]
        extra = ["--local", "-k", "3"]
    outs = {}
    for ext in ("sam", "bam"):
        o = str(tmp_path / ("o." + ext))
        r = subprocess.run([CLI, prefix] + ins + ["-o", o, "--batch", "1000"] + extra, capture_output=True,
                           text=True, timeout=110)
        assert r.returncode == 0, r.stderr
        outs[ext] = o
    lines = open(outs["sam"]).read().splitlines()
    head = [l for l in lines if l.startswith("@")]
    body = [l for l in lines if not l.startswith("@")]
    text, refs, recs, bl = bamfile.read(outs["bam"])
    assert text.replace("o.bam", "o.sam").splitlines() == [l for l in head if not l.startswith("@SQ")]
    assert head[0] == "@HD\tVN:1.6\tSQ:unsorted" and head[1].startswith("@PG\tID:MSGseqTK VN:v1.2.3 PN:")
    assert ["@SQ\tSN:%s\tLN:%d" % r for r in refs] == [l for l in head if l.startswith("@SQ")]
    assert [bamfile.sam_line(x, refs) for x in recs] == body
    assert bl[-1][0] == bamfile.EOF_BLOCK and all(len(d) <= 0xff00 for _, d in bl)
    assert len(bl) > 3
    # the header fills the first block; records never straddle blocks (each fits one)
    import itertools
    starts = set(itertools.accumulate([len(x) for x in recs], initial=0))
    ends = list(itertools.accumulate(len(d) for _, d in bl[1:-1]))
    assert ends[-1] == sum(len(x) for x in recs) and all(e in starts for e in ends)
    for x in recs:
        for tg, ty, v in bamfile.tags(x):
            if tg in ("AS", "NH", "XN"):
                assert ty == _int_type(v), (tg, ty, v)
            elif tg in ("XH", "XP"):
                assert ty == "f"
            else:
                assert tg == "MD" and ty == "Z"
