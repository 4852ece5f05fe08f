"""The oracle's dense DP (Alignment, src/Alignment.cpp) pinned by the reference's own
test cases (test/Alignment_test.cpp: invariants cigar2QLen == read length and
mdTag2alnQLen == alnQLen, plus expected values derived by hand), and by exhaustive
enumeration of alignments on tiny inputs (the maximum score)."""
import ctypes as C
import itertools
import re

import numpy as np
import pytest

import oracle_lib as O
from msgseqtk_amd import synth

# (read, target, mode, alnFrom, alnTo, alnStart, alnEnd, score, CIGAR, MD) -- the five cases of
# test/Alignment_test.cpp:21-160; values worked out by hand from the recurrences.
KAT = [
    ("ATCG", "AAAATCGCCC", 1, 0, 4, 3, 7, 4, "4M", "4"),
    ("TTTATCGAAA", "AAAATCGCCC", 1, 3, 7, 3, 7, 4, "3S4M3S", "4"),
    ("AAAATCGATCGATCGATCGTTT", "GGGATCGATCAATCGATCGCCC", 1, 3, 19, 3, 19, 12, "3S16M3S", "7A8"),
    ("AAAATCGATCAATCGATCGTTT", "AAAATCGATCATCGATCGTTT", 0, 0, 22, 0, 21, 14, "10M1I11M", "10^A11"),
    ("AAAATCGATCGTCGATCGTTT", "AAAATCGATCGATCGATCGTTT", 0, 0, 21, 0, 22, 14, "11M1D10M", "21"),
]


def align_full(q, t, mode, **kw):
    p = O.default_params(**kw)
    qa, ta = np.ascontiguousarray(q, np.uint8), np.ascontiguousarray(t, np.uint8)
    res = np.zeros(5, np.int64)
    sc = np.zeros(1)
    cig, md = C.create_string_buffer(512), C.create_string_buffer(512)
    O.orc().orc_align_full(C.byref(p), O.ptr(qa, C.c_uint8), len(qa), O.ptr(ta, C.c_uint8), len(ta), mode,
                           O.ptr(res, C.c_int64), O.ptr(sc, C.c_double), cig, md, 512)
    return res, sc[0], cig.value.decode(), md.value.decode()


def cigar_qlen(cig):
    return sum(int(n) for n, op in re.findall(r"(\d+)([MIDNSHP=X])", cig) if op in "MIS=X")


def md_qlen(md):
    """mdTag2alnQLen (src/Alignment.cpp:402-418)"""
    m = re.match(r"^\d+", md)
    if not m:
        return 0
    n = int(m.group(0))
    for op, num in re.findall(r"([A-Z]|\^[A-Z]+)(\d+)", md[m.end():]):
        n += 1 if len(op) == 1 else len(op) - 1
        n += int(num)
    return n


@pytest.mark.parametrize("case", KAT)
def test_reference_alignment_cases(case):
    q, t, mode, af, at, ast, ae, score, cig, md = case
    res, sc, c, m = align_full(synth.encode(q), synth.encode(t), mode)
    assert list(res[:4]) == [af, at, ast, ae] and res[4] == 0
    assert sc == score and c == cig and m == md
    assert cigar_qlen(c) == len(q)                   # Alignment_test invariant 1
    assert md_qlen(m) == at - af                     # Alignment_test invariant 2


def brute_max(q, t, mode, match=1, mis=3, go=5, ge=2):
    """Maximum over every alignment written out op by op: query fully aligned (GLOBAL) or a
    query substring (LOCAL), a target substring, affine gaps go + ge*len, no adjacent I/D."""
    L, T = len(q), len(t)
    best = -1e18

    def s(a, b):
        return match if (a & b) and a != 15 and b != 15 else -mis

    def rec(i, j, last, acc, qs):
        nonlocal best
        if mode == 0:
            if i == L and last != "D":
                best = max(best, acc)
        else:
            if last == "M":
                best = max(best, acc)
        if i < L and j < T:
            rec(i + 1, j + 1, "M", acc + s(q[i], t[j]), qs)
        if i < L and last != "D":
            rec(i + 1, j, "I", acc - (ge if last == "I" else go + ge), qs)
        if j < T and last not in ("I", None) and not (mode == 0 and last is None):
            rec(i, j + 1, "D", acc - (ge if last == "D" else go + ge), qs)

    if mode == 0:
        for j0 in range(T + 1):
            rec(0, j0, None, 0, 0)
    else:
        best = 0
        for i0 in range(L):
            for j0 in range(T):
                rec(i0, j0, None, 0, 0)
    return best


@pytest.mark.parametrize("mode", [0, 1])
def test_full_rect_max_score_bruteforce(mode):
    rng = np.random.default_rng(5 + mode)
    for trial in range(60):
        L, T = int(rng.integers(1, 6)), int(rng.integers(1, 7))
        q = synth.NT[rng.integers(0, 4, L)]
        t = synth.NT[rng.integers(0, 4, T)]
        res, sc, c, m = align_full(q, t, mode)
        assert sc == brute_max(list(q), list(t), mode), (synth.decode(q), synth.decode(t), mode, sc)
        assert cigar_qlen(c) == L


def test_chain_alignment_invariants():
    chroms, names = synth.metagenome(4, 40000, gc=(0.3, 0.7), seed=31)
    db = O.OracleDb(chroms, names)
    seqs, quals, rn, truth = synth.simulate_reads(chroms, 150, (50, 250), sub=0.02, indel=0.004, seed=32)
    p = O.default_params()
    sq, offs = O.pack_reads(seqs)
    ql, _ = O.pack_reads(quals)
    cap = 4096
    counts = np.zeros(len(seqs), np.int64)
    rows = np.zeros((cap, 10), np.int64)
    sc = np.zeros(cap)
    cig = C.create_string_buffer(1 << 20)
    k = O.orc().orc_alignments(db.h, C.byref(p), O.ptr(sq, C.c_uint8), O.ptr(ql, C.c_uint8), O.ptr(offs, C.c_int64),
                               len(seqs), O.ptr(counts, C.c_int64), O.ptr(rows, C.c_int64), O.ptr(sc, C.c_double),
                               cig, len(cig), cap)
    cigars = cig.value.decode().split("\n")[:k]
    j = 0
    for i in range(len(seqs)):
        for _ in range(counts[i]):
            if rows[j, 8] == 0:
                assert cigar_qlen(cigars[j]) == len(seqs[i])
                assert rows[j, 5] == len(seqs[i]) and rows[j, 4] == 0      # GLOBAL: whole query
            j += 1
    assert j == k and k > 0
