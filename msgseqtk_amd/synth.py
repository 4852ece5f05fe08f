"""Deterministic synthetic metagenomes and reads for the mapping workloads.

Genomes are i.i.d. bases with a per-genome GC fraction; strains copy a parent
with substitutions and 1-3 bp indels.  Reads are sampled uniformly over
genomes, positions and both strands, then mutated with substitutions and
1-3 bp indels (insertion:deletion 1:1); Phred scores are uniform in a range.
All sequences are nt16 codes (A=1, C=2, G=4, T=8).
"""
import numpy as np

NT = np.array([1, 2, 4, 8], dtype=np.uint8)          # A C G T
COMP = np.array([0, 8, 4, 12, 2, 10, 6, 14, 1, 9, 5, 13, 3, 11, 7, 15], dtype=np.uint8)
SYM = np.frombuffer(b"=ACMGRSVTWYHKDBN", dtype=np.uint8)


def random_genome(rng, length, gc=0.5):
    """i.i.d. bases with P(A) = P(T) = (1 - gc) / 2 and P(C) = P(G) = gc / 2"""
    u = rng.random(length)
    c1, c2, c3 = (1 - gc) / 2, 0.5, (1 + gc) / 2
    return NT[(u >= c1).astype(np.uint8) + (u >= c2) + (u >= c3)]


def mutate(rng, seq, sub, indel, maxlen=3):
    """Apply substitutions (rate sub) and indels (rate indel per base, length U{1..maxlen})."""
    seq = seq.copy()
    n = len(seq)
    m = rng.random(n) < sub
    if m.any():
        cur = np.searchsorted(NT, seq[m]) % 4
        seq[m] = NT[(cur + rng.integers(1, 4, size=int(m.sum()))) % 4]
    k = rng.binomial(n, indel) if indel > 0 else 0
    if k == 0:
        return seq
    pos = np.sort(rng.choice(n, size=k, replace=False))
    out, last = [], 0
    for p in pos:
        ln = int(rng.integers(1, maxlen + 1))
        out.append(seq[last:p])
        if rng.random() < 0.5:          # insertion before p
            out.append(NT[rng.integers(0, 4, size=ln)])
            last = p
        else:                           # deletion of ln bases at p
            last = min(n, p + ln)
    out.append(seq[last:])
    return np.concatenate(out)


def metagenome(n_genomes, length, gc=None, seed=0, strain_frac=0.0, div=(0.01, 0.05)):
    """Return (chroms, names).  length: int or (lo, hi) for log-uniform sizes; gc: None (0.5),
    a float, or (lo, hi) for per-genome U(lo, hi)."""
    rng = np.random.default_rng(seed)
    n_str = int(round(n_genomes * strain_frac))
    n_ind = n_genomes - n_str
    chroms = []
    for i in range(n_ind):
        ln = length if np.isscalar(length) else int(round(np.exp(rng.uniform(np.log(length[0]), np.log(length[1])))))
        g = 0.5 if gc is None else (gc if np.isscalar(gc) else rng.uniform(gc[0], gc[1]))
        chroms.append(random_genome(rng, int(ln), g))
    for i in range(n_str):
        parent = chroms[int(rng.integers(0, n_ind))]
        d = rng.uniform(div[0], div[1])
        chroms.append(mutate(rng, parent, 0.9 * d, 0.1 * d))
    names = ["g%04d" % i for i in range(n_genomes)]
    return chroms, names


def revcom(seq):
    return COMP[seq[::-1]]


def simulate_reads(chroms, n, length=150, sub=0.01, indel=0.001, qual=(20, 40), seed=1,
                   err_range=None, offtarget=0.0, prefix="r"):
    """Single-end reads.  length: int or (lo, hi) log-uniform.  err_range=(lo, hi): per-read error
    rate e ~ U(lo, hi) split 90% substitutions / 10% indels (overrides sub/indel).
    Returns (seqs, quals, names, truth) with truth rows (tid, pos, strand) (tid -1 = off-target)."""
    rng = np.random.default_rng(seed)
    seqs, quals, names, truth = [], [], [], []
    for i in range(n):
        L = length if np.isscalar(length) else int(round(np.exp(rng.uniform(np.log(length[0]), np.log(length[1])))))
        if err_range is not None:
            e = rng.uniform(err_range[0], err_range[1])
            s_rate, i_rate = 0.9 * e, 0.1 * e
        else:
            s_rate, i_rate = sub, indel
        if offtarget > 0 and rng.random() < offtarget:
            frag = NT[rng.integers(0, 4, size=L)]
            t = (-1, 0, 1)
        else:
            tid = int(rng.integers(0, len(chroms)))
            c = chroms[tid]
            span = L + 8
            pos = int(rng.integers(0, max(1, len(c) - span)))
            frag = c[pos:pos + span]
            strand = int(rng.integers(0, 2))
            if strand:
                frag = revcom(frag)
            frag = mutate(rng, frag, s_rate, i_rate)[:L]
            t = (tid, pos, 2 if strand else 1)
        seqs.append(np.ascontiguousarray(frag, dtype=np.uint8))
        quals.append(rng.integers(qual[0], qual[1] + 1, size=len(frag)).astype(np.uint8))
        names.append("%s%d" % (prefix, i))
        truth.append(t)
    return seqs, quals, names, np.array(truth, dtype=np.int64).reshape(-1, 3)


def simulate_pairs(chroms, n, length=150, insert=(350, 50), sub=0.005, indel=0.0005, qual=(20, 40),
                   seed=1, offtarget=0.0, prefix="p"):
    """Paired-end reads (mate 1 forward of the fragment, mate 2 reverse-complement)."""
    rng = np.random.default_rng(seed)
    m1, m2, q1, q2, names = [], [], [], [], []
    for i in range(n):
        isz = max(length, int(round(rng.normal(insert[0], insert[1]))))
        if offtarget > 0 and rng.random() < offtarget:
            frag = NT[rng.integers(0, 4, size=isz + 8)]
        else:
            c = chroms[int(rng.integers(0, len(chroms)))]
            pos = int(rng.integers(0, max(1, len(c) - isz - 8)))
            frag = c[pos:pos + isz + 8]
            if rng.integers(0, 2):
                frag = revcom(frag)
        a = mutate(rng, frag[:length + 8], sub, indel)[:length]
        b = mutate(rng, revcom(frag[:isz])[:length + 8], sub, indel)[:length]
        m1.append(np.ascontiguousarray(a, np.uint8))
        m2.append(np.ascontiguousarray(b, np.uint8))
        q1.append(rng.integers(qual[0], qual[1] + 1, size=len(a)).astype(np.uint8))
        q2.append(rng.integers(qual[0], qual[1] + 1, size=len(b)).astype(np.uint8))
        names.append("%s%d" % (prefix, i))
    return (m1, q1, names), (m2, q2, names)


def encode(s):
    """ASCII DNA string -> nt16 codes (the reference's sym2base table for A/C/G/T/N/IUPAC)."""
    table = np.zeros(256, np.uint8)
    for ch, v in {"A": 1, "C": 2, "G": 4, "T": 8, "U": 8, "R": 5, "Y": 10, "S": 6, "W": 9, "K": 12,
                  "M": 3, "B": 14, "D": 13, "H": 11, "V": 7, "N": 15}.items():
        table[ord(ch)] = v
        table[ord(ch.lower())] = v
    for i, v in enumerate([1, 2, 4, 8, 15]):
        table[ord("0") + i] = v
    return table[np.frombuffer(s.encode(), np.uint8)]


def decode(seq):
    return bytes(np.frombuffer(b"$ACMGRSVTWYHKDBN", np.uint8)[np.asarray(seq) & 15]).decode()


def simulate_pairs_fast(chroms, n, length=150, insert=(350, 50), sub=0.005, indel=0.0005, qual=(20, 40),
                        seed=1, offtarget=0.0, prefix="p"):
    """Vectorized paired-end simulation for large batches (same model as simulate_pairs:
    fragment ~ N(insert), mate 1 = fragment start, mate 2 = revcom of fragment end,
    uniform genome / position / strand, substitutions and 1-3 bp indels).  Returns
    ((seqs1, quals1, names), (seqs2, quals2, names)) with fixed-length numpy rows."""
    rng = np.random.default_rng(seed)
    lens = np.array([len(c) for c in chroms], np.int64)
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
    allseq = np.concatenate(chroms)
    span = length + 8
    isz = np.maximum(length, np.round(rng.normal(insert[0], insert[1], n)).astype(np.int64))
    g = rng.integers(0, len(chroms), n)
    pos = (rng.random(n) * np.maximum(1, lens[g] - isz - span)).astype(np.int64)
    strand = rng.integers(0, 2, n)
    off = rng.random(n) < offtarget
    mates = []
    for m in range(2):
        if m == 0:
            a = starts[g] + pos                         # fragment start
        else:
            a = starts[g] + pos + isz - span            # window ending at the fragment end
        a = np.clip(a, 0, len(allseq) - span)
        idx = a[:, None] + np.arange(span)[None, :]
        w = allseq[idx]
        if m == 1:
            w = COMP[w[:, ::-1]]                        # mate 2 reads the reverse strand
        mates.append(w)
    # a fragment drawn from the reverse strand swaps the roles of the two windows
    w1 = np.where(strand[:, None] == 0, mates[0], mates[1])
    w2 = np.where(strand[:, None] == 0, mates[1], mates[0])
    rnd = NT[rng.integers(0, 4, (n, span))]
    w1 = np.where(off[:, None], rnd, w1)
    rnd2 = NT[rng.integers(0, 4, (n, span))]
    w2 = np.where(off[:, None], rnd2, w2)
    out = []
    for w in (w1, w2):
        w = w.copy()
        subm = rng.random(w.shape) < sub
        cur = np.searchsorted(NT, w[subm]) % 4
        w[subm] = NT[(cur + rng.integers(1, 4, int(subm.sum()))) % 4]
        seqs = [row[:length] for row in w]
        k = rng.binomial(span, indel, n) if indel > 0 else np.zeros(n, np.int64)
        for i in np.nonzero(k)[0]:
            seqs[i] = mutate(rng, w[i], 0.0, indel)[:length]
            if len(seqs[i]) < length:
                seqs[i] = np.concatenate([seqs[i], w[i][len(seqs[i]):length]])
        seqs = [np.ascontiguousarray(s, np.uint8) for s in seqs]
        quals = list(rng.integers(qual[0], qual[1] + 1, (n, length)).astype(np.uint8))
        out.append((seqs, quals))
    names = ["%s%d" % (prefix, i) for i in range(n)]
    return (out[0][0], out[0][1], names), (out[1][0], out[1][1], names)


def simulate_reads_fast(chroms, n, length=150, sub=0.01, indel=0.001, qual=(20, 40), seed=1, offtarget=0.0,
                        prefix="r"):
    """Vectorized single-end simulation for large batches (same model as simulate_reads with a
    fixed length: uniform genome / position / strand, substitutions and 1-3 bp indels, Phred
    scores uniform in qual; a fraction `offtarget` of reads is uniform random sequence).
    Returns (seqs, quals, names)."""
    rng = np.random.default_rng(seed)
    lens = np.array([len(c) for c in chroms], np.int64)
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
    allseq = np.concatenate(chroms)
    span = length + 8
    g = rng.integers(0, len(chroms), n)
    pos = (rng.random(n) * np.maximum(1, lens[g] - span)).astype(np.int64)
    a = np.clip(starts[g] + pos, 0, len(allseq) - span)
    w = allseq[a[:, None] + np.arange(span)[None, :]]
    strand = rng.integers(0, 2, n).astype(bool)
    w[strand] = COMP[w[strand][:, ::-1]]
    off = rng.random(n) < offtarget
    if off.any():
        w[off] = NT[rng.integers(0, 4, (int(off.sum()), span))]
    subm = rng.random(w.shape) < sub
    cur = np.searchsorted(NT, w[subm]) % 4
    w[subm] = NT[(cur + rng.integers(1, 4, int(subm.sum()))) % 4]
    seqs = [row[:length] for row in w]
    k = rng.binomial(span, indel, n) if indel > 0 else np.zeros(n, np.int64)
    for i in np.nonzero(k)[0]:
        s = mutate(rng, w[i], 0.0, indel)[:length]
        if len(s) < length:
            s = np.concatenate([s, w[i][len(s):length]])
        seqs[i] = s
    seqs = [np.ascontiguousarray(s, np.uint8) for s in seqs]
    quals = list(rng.integers(qual[0], qual[1] + 1, (n, length)).astype(np.uint8))
    names = ["%s%d" % (prefix, i) for i in range(n)]
    return seqs, quals, names


# The five benchmark configurations' databases (genome seeds 0xC0FFEE + config number):
#   1: 8 x 250 kb, GC 0.5;  2 and 3: 64 x 4 Mb, GC U(0.3, 0.7);
#   4: 256 x 4 Mb, GC U(0.3, 0.7), 20% strains at U(1%, 5%) divergence from a parent;
#   5: 1000 genomes, lengths log-uniform in 1-8 Mb, 30% strains at U(0.5%, 5%).
def config_genomes(k):
    seed = 0xC0FFEE + k
    if k == 1:
        return metagenome(8, 250_000, gc=None, seed=seed)
    if k in (2, 3):
        return metagenome(64, 4_000_000, gc=(0.3, 0.7), seed=0xC0FFEE + 2)
    if k == 4:
        return metagenome(256, 4_000_000, gc=(0.3, 0.7), seed=seed, strain_frac=0.2, div=(0.01, 0.05))
    if k == 5:
        return metagenome(1000, (1_000_000, 8_000_000), gc=(0.3, 0.7), seed=seed, strain_frac=0.3,
                          div=(0.005, 0.05))
    raise ValueError("configurations are 1..5")
