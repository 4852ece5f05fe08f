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
