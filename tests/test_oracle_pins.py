"""Pins of the oracle functions that the reference's seeding stack alone does not fix:
the nt16 alphabet and read types (A1/A2), the BD text (A3), getSeeds / findSeeds
(A14/A15), chaining invariants (A18), the chain DP with its quirks (A20-A25),
Alignment::evaluate (A28), calcMapQ (A29) and AlignmentPE's pairs (A30).

Each check is one of: the reference's own known-answer tests (test/DNAseq_test.cpp,
test/PrimarySeq_test.cpp), the reference's own invariants (test/SMEM_test.cpp
isValidSeed, test/SeedChain_test.cpp isValidChain, the assertion
seeds.size() == size of src/SMEM.cpp:240,267), the reference's own alphabet and read
code compiled from its sources (oracle/_ref), or values worked out by hand from the
reference's recurrences (the derivations are written next to each case)."""
import math

import numpy as np
import pytest

import oracle_lib as O
from msgseqtk_amd import synth

NT = {"A": 1, "C": 2, "G": 4, "T": 8, "N": 15}


def enc(s):
    return np.array([NT[c] for c in s], np.uint8)


def revcom_codes(a):
    """4-bit reversal is the nt16 complement (A=1<->T=8, C=2<->G=4)."""
    a = np.asarray(a, np.uint8)
    c = ((a & 1) << 3) | ((a & 2) << 1) | ((a & 4) >> 1) | ((a & 8) >> 3)
    return c[::-1].astype(np.uint8)


# ---------------------------------------------------------------- A1 / A2

def test_alphabet_closed_forms():
    e, d, comp, basic, toint, amb, qenc = O.alphabet(O.orc(), "orc")
    for b in range(16):
        assert comp[b] == (((b & 1) << 3) | ((b & 2) << 1) | ((b & 4) >> 1) | ((b & 8) >> 3))
        assert toint[b] == {0: 0, 1: 1, 2: 2, 4: 3, 8: 4}.get(b, 0)
        assert amb[b] == (0 if b in (1, 2, 4, 8) else 1)
        # toBasic keeps N and gaps, maps other IUPAC codes to their highest base bit
        assert basic[b] == (b if b in (0, 15) else 1 << (b.bit_length() - 1))
    assert d == b"$ACMGRSVTWYHKDBN"
    for c in "ACGTN":
        assert e[ord(c)] == NT[c] and e[ord(c.lower())] == NT[c]
    assert e[ord("U")] == NT["T"]
    for c in range(33, 127):
        assert qenc[c] == max(c - 33, 2)     # quality::encode (src/QualStr.h:40-42)
    assert qenc[32] == 255                  # the uint8_t difference wraps below the shift


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_alphabet_tables_equal_reference():
    mine = list(O.alphabet(O.orc(), "orc"))
    theirs = list(O.alphabet(O.ref(), "ref"))
    # DNAalphabet::encode(char) indexes its table with a (signed) char: bytes >= 128 read before
    # the table (undefined), so only ASCII is compared
    mine[0], theirs[0] = mine[0][:128], theirs[0][:128]
    for a, b in zip(mine, theirs):
        if isinstance(a, bytes):
            assert a == b
        else:
            assert np.array_equal(a, b)


def test_dnaseq_revcom_kat():
    """test/DNAseq_test.cpp:17-22: revcom of ATCGURYSWKMBDHVN is NBDHVKMWSRYACGAT."""
    e = O.alphabet(O.orc(), "orc")[0]
    src = np.array([e[ord(c)] for c in "ATCGURYSWKMBDHVN"], np.uint8)
    want = np.array([e[ord(c)] for c in "NBDHVKMWSRYACGAT"], np.uint8)
    out, q = np.zeros(16, np.uint8), np.zeros(16, np.uint8)
    O.orc().orc_revcom_read(O.ptr(src, O.C.c_uint8), O.ptr(np.zeros(16, np.uint8), O.C.c_uint8), 16,
                            O.ptr(out, O.C.c_uint8), O.ptr(q, O.C.c_uint8))
    assert np.array_equal(out, want)
    if O.ref_available():
        r = np.zeros(16, np.uint8)
        O.ref().ref_revcom(O.ptr(src, O.C.c_uint8), 16, O.ptr(r, O.C.c_uint8))
        assert np.array_equal(r, want)
        n = O.ref().ref_encode(b"ATCGURYSWKMBDHVN", O.ptr(r, O.C.c_uint8))
        assert n == 16 and np.array_equal(r, src)


def test_primaryseq_revcom_kat():
    """test/PrimarySeq_test.cpp:16-26: the read ATCGNTCGANatcgntcgan reverse-complements to
    ntcgancgatNTCGANCGAT (codes are case-blind); qualities are reversed."""
    e = O.alphabet(O.orc(), "orc")[0]
    s = "ATCGNTCGANatcgntcgan"
    qs = "".join(chr(33 + 2 + i) for i in range(20))
    src = np.array([e[ord(c)] for c in s], np.uint8)
    qual = np.array([ord(c) - 33 for c in qs], np.uint8)
    want = np.array([e[ord(c)] for c in "ntcgancgatNTCGANCGAT"], np.uint8)
    out, oq = np.zeros(20, np.uint8), np.zeros(20, np.uint8)
    O.orc().orc_revcom_read(O.ptr(src, O.C.c_uint8), O.ptr(qual, O.C.c_uint8), 20, O.ptr(out, O.C.c_uint8),
                            O.ptr(oq, O.C.c_uint8))
    assert np.array_equal(out, want) and np.array_equal(oq, qual[::-1])
    if O.ref_available():
        rs, rq = np.zeros(20, np.uint8), np.zeros(20, np.uint8)
        assert O.ref().ref_primary_revcom(s.encode(), qs.encode(), O.ptr(rs, O.C.c_uint8), O.ptr(rq, O.C.c_uint8)) == 20
        assert np.array_equal(rs, out) and np.array_equal(rq, oq)


# ---------------------------------------------------------------- A3

def test_bd_text_closed_form():
    """MetaGenome::getBDSeq (src/MetaGenome.cpp:104-113): per chromosome, toBasic(seq) + 0 +
    revcom(toBasic(seq)) + 0, chromosomes in order; IUPAC codes other than N become basic."""
    rng = np.random.default_rng(11)
    chroms = []
    for L in (37, 1, 64, 5):
        c = rng.choice(np.array([1, 2, 4, 8], np.uint8), L)
        c[rng.random(L) < 0.1] = 15
        c[rng.random(L) < 0.05] = 5        # R = A|G -> G
        c[rng.random(L) < 0.05] = 3        # M = A|C -> C
        chroms.append(c)
    db = O.OracleDb(chroms)
    bd = db.arrays()[0]
    basic = O.alphabet(O.orc(), "orc")[3]
    want = []
    for c in chroms:
        b = basic[c]
        want += [b, [0], revcom_codes(b), [0]]
    want = np.concatenate(want).astype(np.uint8)
    assert np.array_equal(bd, want)
    if O.ref_available():
        for c in chroms:
            b = basic[c]
            r = np.zeros(len(b), np.uint8)
            O.ref().ref_revcom(O.ptr(np.ascontiguousarray(b), O.C.c_uint8), len(b), O.ptr(r, O.C.c_uint8))
            assert np.array_equal(r, revcom_codes(b))


# ---------------------------------------------------------------- A14 / A15 / A18

def seed_spells_read(seed, read, chroms):
    """isValidSeed (test/SMEM_test.cpp:86-97): the target segment equals the read segment
    (FWD) or the reverse complement's segment (REV)."""
    frm, to, start, end, tid, strand = (int(x) for x in seed)
    q = read if strand == 1 else revcom_codes(read)
    t = np.concatenate([chroms[tid], [0]])
    return end - start == to - frm and np.array_equal(q[frm:to], t[start:end])


def test_smem_test_literal_case():
    """test/SMEM_test.cpp: two genomes x two copies of ACGTCGTAGTACTACGNACGTCGTAGTACTACG,
    read ACGTAGTA, minLen 1, no e-value limit: every seed of every MEM/SMEM spells the read."""
    ch = enc("ACGTCGTAGTACTACGNACGTCGTAGTACTACG")
    chroms = [ch, ch, ch, ch]
    db = O.OracleDb(chroms)
    read = enc("ACGTAGTA")
    cnt, sz, seeds = O.orc_raw_seeds(db, [read], minLen=1, maxE=math.inf)
    assert cnt[0] > 0 and len(seeds) > 0
    assert np.array_equal(sz[:, 0], sz[:, 1])
    assert all(seed_spells_read(s, read, chroms) for s in seeds)
    c, rows, _ = O.orc_seeds(db, [read], minLen=1, maxE=math.inf)
    assert c[0] > 0 and all(seed_spells_read(s, read, chroms) for s in rows)


def is_valid_chain(chain, read, chroms):
    """isValidChain (test/SeedChain_test.cpp:54-70) with SeedPair::isCompatitable
    (src/SeedPair.h:156-158): consecutive seeds ordered on both sequences, same tid and
    strand, and every seed spells the read."""
    for a, b in zip(chain, chain[1:]):
        if not (a[1] <= b[0] and a[3] <= b[2] and a[4] == b[4] and a[5] == b[5]):
            return False
    return all(seed_spells_read(s, read, chroms) for s in chain)


def chains_of(rows, L, trunc=False):
    mm = int(L * 0.2) if trunc else math.ceil(L * 0.2)
    ind = int(L * 0.1) if trunc else math.ceil(L * 0.1)
    flat = np.zeros(1 << 16, np.int64)
    r = np.ascontiguousarray(rows, np.int64)
    k = O.orc().orc_get_chains(O.ptr(r, O.C.c_int64), len(r), mm, ind, O.ptr(flat, O.C.c_int64), len(flat))
    out, i = [], 0
    while i < k:
        n = int(flat[i])
        out.append([rows[j] for j in flat[i + 1:i + 1 + n]])
        i += 1 + n
    return out


def test_seedchain_test_literal_case():
    """test/SeedChain_test.cpp: chromosomes ACGTCGTAACGTGGTAACGTTGTA x2, read ACGTAGTA,
    findSeeds(minLen 4, no e-value limit), getChains(seeds, 8*0.2, 8*0.1) truncated to int."""
    ch = enc("ACGTCGTAACGTGGTAACGTTGTA")
    chroms = [ch, ch]
    db = O.OracleDb(chroms)
    read = enc("ACGTAGTA")
    c, rows, _ = O.orc_seeds(db, [read], minLen=4, maxE=math.inf)
    assert c[0] > 0
    chains = chains_of(rows, len(read), trunc=True)
    assert chains and all(is_valid_chain(ch_, read, chroms) for ch_ in chains)
    # chains are made of the read's seeds only
    assert {tuple(s) for ch_ in chains for s in ch_} <= {tuple(s) for s in rows}


def test_seed_invariants_on_simulated_reads():
    """On simulated reads with errors over a small metagenome (N runs included): every SMEM
    gives exactly `size` seeds (src/SMEM.cpp:240,267), every raw and every final seed spells
    the read, and every chain is valid."""
    chroms, names = synth.metagenome(4, 20000, gc=(0.3, 0.7), seed=21)
    chroms[1] = chroms[1].copy()
    chroms[1][500:530] = 15
    reads = synth.simulate_reads(chroms, 300, 100, sub=0.02, indel=0.003, seed=22)[0]
    db = O.OracleDb(chroms, names)
    cnt, sz, raw = O.orc_raw_seeds(db, reads)
    assert cnt.sum() > 300 and np.array_equal(sz[:, 0], sz[:, 1])
    k = 0
    for i, r in enumerate(reads):
        m = int(sz[k:k + cnt[i], 1].sum())
        k += int(cnt[i])
        # raw seeds of read i are the next m rows
        for s in raw[:m]:
            assert seed_spells_read(s, r, chroms)
        raw = raw[m:]
    c, rows, _ = O.orc_seeds(db, reads)
    off = np.concatenate([[0], np.cumsum(c)])
    nchains = 0
    for i, r in enumerate(reads):
        rs = rows[off[i]:off[i + 1]]
        assert all(seed_spells_read(s, r, chroms) for s in rs)
        for ch_ in chains_of(rs, len(r)):
            assert is_valid_chain(ch_, r, chroms)
            nchains += 1
    assert nchains >= 250


# ---------------------------------------------------------------- A20-A25, A28: chain DP known answers

# Chromosome of the chain cases (30 bases; indexes used below):
# 0 A 1 C 2 G 3 T 4 T 5 G 6 C 7 A 8 A 9 G 10 C 11 T 12 A 13 G 14 G 15 C 16 T 17 T 18 A 19 C
# 20 G 21 A 22 T 23 C 24 C 25 G 26 T 27 A 28 G 29 C
T30 = "ACGTTGCAAGCTAGGCTTACGATCCGTAGC"
# Chromosome of the matchMax case
T24 = "GAGACGCTACGTGACCAGTTCCAT"

M30 = math.log10(1 - 10 ** (-30 / 10))    # evaluate(): a matched base at quality 30
X30, GAP_OPEN30 = -3.0, -3.0 - 5           # a mismatch; a gap opening (q/-10 - gapO), plus -gapE per base

# (name, chromosome, read, seeds (from, to, start, end, strand), mode,
#  (tStart, tEnd, alnFrom, alnTo, alnStart, alnEnd, cells), score, CIGAR, MD, log10P)
CHAIN_KAT = [
    # Two seeds touching on both sequences.  Seed 1 ends with M(10,10)=10 and its step-c cell;
    # seed 2's block resets that cell and starts from max(M(10,10)+1, ..., 0) = 11.  Last row:
    # only M(20,20)=20.  Cells: 10+1 (step c) + 10.
    ("adjacent seeds", T30, T30[5:25], [(0, 10, 5, 15, 1), (10, 20, 15, 25, 1)], 0,
     (5, 25, 0, 20, 5, 25, 21), 20, "20M", "20", 20 * M30),
    # Query-adjacent seeds with one target base between them (T[10] deleted from the read).
    # The gap rectangle [10,10)x[10,11) is empty, so the only link is the column-D overwrite
    # D.col(11) = M.col(10) - 7: D(10,11) = 3.  Seed 2's first cell M(11,12) = D(10,11)+1 = 4,
    # last M(20,21) = 13.  Backtrace at (11,12): matchMax(-inf, -inf, 3) = D; at (10,11):
    # delMax(M(10,10)-7 = 3, -inf) = M.  The step-c cell M(11,11)=11 survives unused.
    ("column-D overwrite (deletion)", T30, T30[0:10] + T30[11:21], [(0, 10, 0, 10, 1), (10, 20, 11, 21, 1)], 0,
     (0, 21, 0, 20, 0, 21, 21), 13, "10M1D10M", "20", 20 * M30 + GAP_OPEN30 - 2),
    # Target-adjacent seeds with one read base (G) between them.  The gap rectangle
    # [10,11)x[10,10) has no columns; the row-I overwrite I.row(11) = M.row(10) - 7 gives
    # I(11,10) = 3, seed 2 starts at M(12,11) = 4 and ends at M(21,20) = 13.  Backtrace at
    # (12,11): matchMax(-inf, 3, -inf) = I; at (11,10): insMax(M(10,10)-7 = 3, -inf) = M.
    ("row-I overwrite (insertion)", T30, T30[0:10] + "G" + T30[10:20], [(0, 10, 0, 10, 1), (11, 21, 10, 20, 1)], 0,
     (0, 20, 0, 21, 0, 20, 21), 13, "10M1I10M", "10^G10", 20 * M30 + GAP_OPEN30 - 2),
    # LOCAL: seed 1 (0,10)@0 adds the step-c cell M(11,11) = M(10,10) + match = 11 whatever the
    # bases (read[10] = T[11] = T against target T[10] = C, a mismatch).  Seed 2 (10,12)@11 is
    # query-adjacent, so no rectangle covers (11,11).  Seed 2 reaches 5 (from D(10,11) = 3), the
    # 3' rectangle rows 12..15 x cols 13..16 at most 5+4 = 9; the maximum of M is the step-c cell:
    # 11 M steps back to (0,0), 5 bases clipped.  Cells: 11 + 3 + 16.
    ("step-c cell (LOCAL)", T30, T30[0:10] + T30[11:13] + "AAAA", [(0, 10, 0, 10, 1), (10, 12, 11, 13, 1)], 1,
     (0, 17, 0, 11, 0, 11, 30), 11, "11M5S", "10C0", 10 * M30 + X30 - 4 * 5),
    # GLOBAL: three read bases before a seed that starts at the chromosome start.  No 5'
    # rectangle (tStart = start = 0); the seed's first cell takes max(I(3,0)+1 = -10, 0.0) = 0
    # (the 0.0 floor applies in GLOBAL mode too), so the score is 19, not 20 - gapPenalty(3) = 9.
    # Backtrace at (4,1): matchMax(-inf, I(3,0) = -11, -inf) = I, then I down column 0.
    ("0.0 floor at a seed start (GLOBAL)", T30, "GGG" + T30[0:20], [(3, 23, 0, 20, 1)], 0,
     (0, 20, 0, 23, 0, 20, 20), 19, "3I20M", "0^GGG20", 20 * M30 + GAP_OPEN30 - 2 * 3),
    # The read runs one base past the chromosome end: tEnd = min(len+1, 30 + 1.1) = 31 admits the
    # '\0' terminator as target base 30; the 3' rectangle's one cell scores A against '\0' as a
    # mismatch: M(11,11) = 10 - 3 = 7 (over the step-c value 11).  The row-I overwrite makes
    # I(11,10) = 3, which the last-row scan passes first; M(11,11) = 7 wins.  MD prints the
    # terminator as '$'.
    ("'\\0' target base", T30, T30[20:30] + "A", [(0, 10, 20, 30, 1)], 0,
     (20, 31, 0, 11, 20, 31, 12), 7, "11M", "10$0", 10 * M30 + X30),
    # A mismatch between two seeds: the 1x1 gap rectangle overwrites the step-c cell with
    # M(11,11) = 10 - 3 = 7, seed 2 continues from 8 to 17.
    ("mismatch between seeds", T30, T30[0:10] + "A" + T30[11:21], [(0, 10, 0, 10, 1), (11, 21, 11, 21, 1)], 0,
     (0, 21, 0, 21, 0, 21, 22), 17, "21M", "10C10", 20 * M30 + X30),
    # matchMax quirk (src/Alignment.h:577-589): with ins > match it sets max = del, so D can only
    # win over INS by beating itself.  Read GAGCGCT, seed CGC = read[3,6) at T24[4,7).
    # tStart 0, tEnd 8.  5' rectangle rows 1..3 x cols 1..4 (GAG against GAGA), column-major:
    #   col 4: M(3,4) = max(M(2,3) = -6, D(2,3) = -5, I(2,3) = -6) - 3 = -8,
    #          I(3,4) = max(M(2,4) - 7 = -5, I(2,4) - 2 = -11) = -5,
    #          D(3,4) = max(M(3,3) - 7 = -4, D(3,3) - 2 = -18) = -4.
    # Seed: M(4,5) = max(-7, -3, -4, 0) = 0 ... M(6,7) = 2; 3' cell M(7,8) = 3 (T against T).
    # Backtrace at (4,5): matchMax(-8, -5, -4): ins > match, max := del = -4, del > max fails: I
    # (a correct maximum would pick D and give 3M1D4M from target 0).  At (3,4): insMax(-5, -11)
    # = M; (2,4) and (1,3) are matches.  Cells: 12 + 3 + 1 (step c) + 1.
    ("matchMax quirk", T24, "GAGCGCT", [(3, 6, 4, 7, 1)], 0,
     (0, 8, 0, 7, 2, 8, 17), 3, "2M1I4M", "2^G4", 6 * M30 + GAP_OPEN30 - 2),
]


def rc_str(s):
    return "".join({"A": "T", "C": "G", "G": "C", "T": "A"}[c] for c in reversed(s))


@pytest.mark.parametrize("case", CHAIN_KAT, ids=[c[0] for c in CHAIN_KAT])
def test_chain_dp_known_answers(case):
    name, chrom, read, seeds, mode, geo, score, cigar, md, log10p = case
    db = O.OracleDb([enc(chrom)])
    rows = [(f, t, s, e, 0, st) for (f, t, s, e, st) in seeds]
    res, sc, lp, cg, m = O.orc_align_chain(db, O.default_params(mode=mode), enc(read), rows)
    assert (res[2], res[3], res[4], res[5], res[6], res[7], res[9]) == geo, name
    assert res[0] == 0 and res[1] == seeds[0][4] and res[8] == 0
    assert sc == score and cg == cigar and m == md
    assert abs(lp - log10p) < 1e-9


def test_chain_dp_reverse_strand():
    """A REV chain aligns the reverse complement: the read is revcom(T[5:25]), its seed (0,20)
    on the reverse complement maps to target [5,25)."""
    db = O.OracleDb([enc(T30)])
    read = enc(rc_str(T30[5:25]))
    res, sc, lp, cg, m = O.orc_align_chain(db, O.default_params(), read, [(0, 20, 5, 25, 0, 2)])
    assert res[1] == 2 and (res[4], res[5], res[6], res[7]) == (0, 20, 5, 25)
    assert sc == 20 and cg == "20M" and m == "20" and abs(lp - 20 * M30) < 1e-9


def test_evaluate_quality_closed_forms():
    """Alignment::evaluate (src/Alignment.cpp:446-489): a matched base adds
    log10(1 - 10^(-q/10)), a mismatch -q/10, a gap opening -q/10 - gapO, every gap base
    -gapE, every clipped base -clip; qualities vary along the read."""
    db = O.OracleDb([enc(T30)])
    q = np.array([2, 10, 20, 30, 40] * 4, np.uint8)
    read = T30[0:10] + "A" + T30[11:20]          # mismatch at read position 10 (target C)
    res, sc, lp, cg, m = O.orc_align_chain(db, O.default_params(), enc(read), [(0, 10, 0, 10, 0, 1), (11, 20, 11, 20, 0, 1)], q)
    assert cg == "20M" and m == "10C9"
    want = sum(math.log10(1 - 10 ** (-int(q[i]) / 10)) for i in range(20) if i != 10) - int(q[10]) / 10
    assert abs(lp - want) < 1e-9


# ---------------------------------------------------------------- A29

def test_calc_mapq_closed_forms():
    """Alignment::calcMapQ (src/Alignment.cpp:491-512): postP = 10^l_i / sum 10^l_j,
    mapQ = min(round(-10 log10(1 - postP)), 250)."""
    pp, mq = O.orc_calc_mapq([-3.0])
    assert pp[0] == 1.0 and mq[0] == 250                 # unique alignment
    pp, mq = O.orc_calc_mapq([-5.0, -5.0])
    assert pp[0] == 0.5 and list(mq) == [3, 3]           # -10 log10(0.5) = 3.01
    pp, mq = O.orc_calc_mapq([-2.0, -3.0])
    assert abs(pp[0] - 10 / 11) < 1e-15 and list(mq) == [10, 0]
    pp, mq = O.orc_calc_mapq([-7.0, -7.0, -7.0])
    assert abs(pp[0] - 1 / 3) < 1e-15 and list(mq) == [2, 2, 2]
    # far below exp's range: the scale lifts the largest loglik to -256 first
    pp, mq = O.orc_calc_mapq([-500.0, -500.0])
    assert pp[0] == 0.5 and list(mq) == [3, 3]
    pp, mq = O.orc_calc_mapq([-500.0, -501.0])
    assert abs(pp[0] - 10 / 11) < 1e-12 and list(mq) == [10, 0]


def test_pairing_known_answers():
    """AlignmentPE (src/Alignment.cpp:540-581, src/Alignment.h:446-575,635-639) and
    PairingScheme::pr (src/PairingScheme.cpp:18-25) on hand-built lists: only opposite strands
    pair; the insert size is max(end) - min(start) on one chromosome and -1 across two; pr is the
    normal pdf inside [insMin, insMax] (so log10lik adds log10(1 / (sd sqrt(2 pi))) at the mean),
    0 outside (log10lik -inf, then removed by the range filter) and 1 when insMean is 0; pairs sort
    by log10lik, largest first; each filter removes exactly the pairs its predicate names."""
    P = O.default_params()                          # N(200, 20) on [0, 400]
    c0 = math.log10(1 / (20 * math.sqrt(2 * math.pi)))
    # read: FWD at [1000, 1150) (log10P -2), REV at [1100, 1250) (-1), FWD on chrom 1 (-3)
    # mate: REV at [1050, 1200) (-1), FWD at [900, 1000) (-2)
    rows = [(0, 1, 1000, 1150), (0, 2, 1100, 1250), (1, 1, 1000, 1150),
            (0, 2, 1050, 1200), (0, 1, 900, 1000)]
    lp = [-2.0, -1.0, -3.0, -1.0, -2.0]
    offs = [0, 3, 5]
    cnt, r3, l2 = O.orc_pairs(P, offs, rows, lp)
    # candidates: (0,3) isize 200, (1,4) isize 350, (2,3) isize -1 (removed: -1 < 0)
    assert list(cnt) == [2]
    by = {(int(a), int(b)): (int(s), x, y) for (a, b, s), (x, y) in zip(r3, l2)}
    assert set(by) == {(0, 3), (1, 4)}
    assert by[(0, 3)][0] == 200 and by[(1, 4)][0] == 350
    assert math.isclose(by[(0, 3)][1], -3.0 + c0, rel_tol=1e-14)
    z = (350 - 200) / 20
    assert math.isclose(by[(1, 4)][1], -3.0 + c0 - z * z / 2 * math.log10(math.e), rel_tol=1e-13)
    assert math.isclose(by[(0, 3)][2], (-3.0 + c0) * math.log(10), rel_tol=1e-13)
    assert [tuple(x[:2]) for x in r3] == [(0, 3), (1, 4)]        # larger log10lik first
    # insert size 400 is in range, 401 is not
    c, r, _ = O.orc_pairs(P, [0, 1, 2], [(0, 1, 0, 100), (0, 2, 300, 400)], [-1.0, -1.0])
    assert list(c) == [1] and r[0][2] == 400
    c, r, _ = O.orc_pairs(P, [0, 1, 2], [(0, 1, 0, 100), (0, 2, 300, 401)], [-1.0, -1.0])
    assert list(c) == [0]
    # same strand: no candidate at all
    c, _, _ = O.orc_pairs(P, [0, 1, 2], [(0, 1, 0, 100), (0, 1, 100, 200)], [-1.0, -1.0])
    assert list(c) == [0]
    # uniform model: log10lik is the sum of the mates' values; -1 is kept when the range holds it
    U = O.default_params(insMean=0, insSD=1, insMin=-1, insMax=400)
    c, r, l = O.orc_pairs(U, offs, rows, lp)
    assert list(c) == [3] and [int(x[2]) for x in r] == [200, 350, -1]
    assert [x[0] for x in l] == [-3.0, -3.0, -4.0]
    # filters: mates at [0, 100) and [50, 150) overlap (tail-over); [0, 200) contains [50, 150)
    # (which is also tail-over: one starts first and ends past the other's start); [0, 100) and
    # [100, 200) only touch
    tail = [(0, 1, 0, 100), (0, 2, 50, 150)]
    cont = [(0, 1, 0, 200), (0, 2, 50, 150)]
    far = [(0, 1, 0, 100), (0, 2, 100, 200)]
    for flags, kept in [(dict(noTailOver=1), (0, 0, 1)), (dict(noContain=1), (1, 0, 1)),
                        (dict(noOverlap=1), (0, 0, 1)), (dict(noDiscordant=1), (1, 1, 1))]:
        got = tuple(int(O.orc_pairs(O.default_params(insMin=0, insMax=1000, **flags), [0, 1, 2], x,
                                    [-1.0, -1.0])[0][0]) for x in (tail, cont, far))
        assert got == kept, flags
    # maxNPair keeps the first pairs of the sorted list
    c, r, _ = O.orc_pairs(O.default_params(insMax=1000, maxNPair=1), offs, rows, lp)
    assert list(c) == [1] and tuple(r[0][:2]) == (0, 3)


def test_record_mate_fields_and_unmapped_layout():
    """Records are built by the BAM constructor with mtid -1, mpos 0 and, for unmapped reads,
    tid -1, pos 0, MAPQ 0 and no CIGAR (src/SAMtools/BAM.h:61-75; AlignmentPE::exportFwdBAM /
    exportRevBAM change only MAPQ, ISIZE and flags, src/Alignment.h:512-529); sam_format1 prints
    RNEXT '*' for mtid < 0, PNEXT as mpos + 1 and POS as pos + 1.  So every record reads
    '*', '1' in columns 7-8, and an unmapped one '*', '1', '0', '*' in columns 3-6; TLEN is 0
    for single-end records and +-ISIZE for pairs."""
    chroms, names = synth.metagenome(3, 20000, gc=(0.4, 0.6), seed=5)
    db = O.OracleDb(chroms, names)
    seqs, quals, rn, _ = synth.simulate_reads(chroms, 60, 100, sub=0.01, indel=0.0, seed=6, offtarget=0.2)
    sam, _ = O.orc_map(db, O.default_params(), seqs, quals, rn)
    lines = "".join(sam).splitlines()
    unmapped = [l.split("\t") for l in lines if int(l.split("\t")[1]) & 4]
    assert unmapped and all(f[2:6] == ["*", "1", "0", "*"] for f in unmapped)
    assert all(l.split("\t")[6:9] == ["*", "1", "0"] for l in lines)
    (a, qa, na), (b, qb, nb) = synth.simulate_pairs(chroms, 40, 100, insert=(250, 30), seed=7)
    pe, _ = O.orc_map(db, O.pe_params(250, 30), a, qa, na, mates=(b, qb, nb))
    pl = [l.split("\t") for l in "".join(pe).splitlines()]
    assert all(f[6:8] == ["*", "1"] for f in pl)
    proper = [f for f in pl if int(f[1]) & 2]
    assert proper and all(abs(int(f[8])) >= 100 for f in proper)
