"""ctypes bindings to the CPU oracle (oracle/liborc.so) and, when built, to the
reference's own seeding stack (oracle/_ref/libmsgs_ref.so).  Test
infrastructure only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_SO = os.path.join(ROOT, "oracle", "liborc.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libmsgs_ref.so")

P64 = C.POINTER(C.c_int64)
PD = C.POINTER(C.c_double)
PU8 = C.POINTER(C.c_uint8)


class Params(C.Structure):
    """Parameter block shared by the oracle and the product C-ABI (same layout)."""
    _fields_ = [("minLen", C.c_int64), ("maxEvalue", C.c_double), ("maxNSeed", C.c_int64),
                ("mode", C.c_int32), ("maxReport", C.c_int32),
                ("match", C.c_double), ("mismatch", C.c_double), ("gapO", C.c_double),
                ("gapE", C.c_double), ("clip", C.c_double), ("minScoreRate", C.c_double),
                ("insMean", C.c_double), ("insSD", C.c_double), ("insMin", C.c_double),
                ("insMax", C.c_double),
                ("noMixed", C.c_int32), ("noDiscordant", C.c_int32), ("noTailOver", C.c_int32),
                ("noContain", C.c_int32), ("noOverlap", C.c_int32), ("pad0", C.c_int32),
                ("maxNPair", C.c_int64)]


def default_params(**kw):
    p = Params(minLen=17, maxEvalue=0.01, maxNSeed=500, mode=0, maxReport=1, match=1, mismatch=3,
               gapO=5, gapE=2, clip=4, minScoreRate=0.65, insMean=200, insSD=20, insMin=0,
               insMax=400, maxNPair=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def pe_params(mean, sd, **kw):
    """Insert model the way msgseqtk-align sets it from -m/-d (setMean, setSD, updateRange)."""
    return default_params(insMean=mean, insSD=sd, insMin=mean - 10 * sd, insMax=mean + 10 * sd, **kw)


def ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _load(path):
    lib = C.CDLL(path)
    return lib


_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        lib = _load(ORC_SO)
        lib.orc_db_create.restype = C.c_void_p
        lib.orc_db_create.argtypes = [PU8, P64, C.c_int32, C.c_char_p, P64]
        lib.orc_db_free.argtypes = [C.c_void_p]
        lib.orc_block_ends.argtypes = [P64, C.c_int32, C.c_int64, P64, C.c_int64]
        lib.orc_block_ends.restype = C.c_int64
        lib.orc_suffix_array.argtypes = [PU8, C.c_int64, P64, C.c_int64, P64]
        lib.orc_db_len.argtypes = [C.c_void_p]
        lib.orc_db_len.restype = C.c_int64
        lib.orc_db_counts.argtypes = [C.c_void_p, P64, P64]
        lib.orc_db_copy.argtypes = [C.c_void_p, PU8, PU8, P64]
        lib.orc_rank.argtypes = [C.c_void_p, C.c_int32, C.c_int64]
        lib.orc_rank.restype = C.c_int64
        lib.orc_backext.argtypes = [C.c_void_p, P64, C.c_int32]
        lib.orc_backext.restype = C.c_int64
        lib.orc_fwdext.argtypes = [C.c_void_p, P64, C.c_int32]
        lib.orc_fwdext.restype = C.c_int64
        lib.orc_loglik.argtypes = [C.c_void_p, C.c_int32]
        lib.orc_loglik.restype = C.c_double
        lib.orc_tid.argtypes = [C.c_void_p, C.c_int64]
        lib.orc_tid.restype = C.c_int64
        lib.orc_smems.argtypes = [C.c_void_p, PU8, P64, C.c_int64, C.c_int64, C.c_double, C.c_int32,
                                  P64, P64, PD, C.c_int64]
        lib.orc_smems.restype = C.c_int64
        lib.orc_seeds.argtypes = [C.c_void_p, PU8, P64, C.c_int64, C.c_int64, C.c_double, C.c_int64,
                                  P64, P64, PD, C.c_int64]
        lib.orc_seeds.restype = C.c_int64
        lib.orc_remove_redundant.argtypes = [P64, PD, C.c_int64]
        lib.orc_remove_redundant.restype = C.c_int64
        lib.orc_get_chains.argtypes = [P64, C.c_int64, C.c_int64, C.c_int64, P64, C.c_int64]
        lib.orc_get_chains.restype = C.c_int64
        lib.orc_align_full.argtypes = [C.POINTER(Params), PU8, C.c_int64, PU8, C.c_int64, C.c_int32,
                                       P64, PD, C.c_char_p, C.c_char_p, C.c_int64]
        lib.orc_align_full.restype = C.c_int64
        lib.orc_alignments.argtypes = [C.c_void_p, C.POINTER(Params), PU8, PU8, P64, C.c_int64, P64, P64,
                                       PD, C.c_char_p, C.c_int64, C.c_int64]
        lib.orc_alignments.restype = C.c_int64
        lib.orc_map.argtypes = [C.c_void_p, C.POINTER(Params), PU8, PU8, P64, C.c_char_p, PU8, PU8, P64,
                                C.c_char_p, C.c_int64, C.c_int32, C.c_char_p, C.c_int64, P64, PD]
        lib.orc_map.restype = C.c_int64
        lib.orc_max_threads.restype = C.c_int32
        lib.orc_alphabet.argtypes = [PU8, C.c_char_p, PU8, PU8, C.POINTER(C.c_int32), PU8, PU8]
        lib.orc_revcom_read.argtypes = [PU8, PU8, C.c_int64, PU8, PU8]
        lib.orc_raw_seeds.argtypes = [C.c_void_p, PU8, P64, C.c_int64, C.c_int64, C.c_double, P64, P64,
                                      C.c_int64, P64, C.c_int64]
        lib.orc_raw_seeds.restype = C.c_int64
        lib.orc_align_chain.argtypes = [C.c_void_p, C.POINTER(Params), PU8, PU8, C.c_int64, P64, C.c_int64, P64,
                                        PD, PD, C.c_char_p, C.c_char_p, C.c_int64]
        lib.orc_align_chain.restype = C.c_int64
        lib.orc_calc_mapq.argtypes = [PD, C.c_int64, PD, PU8]
        lib.orc_pairs.argtypes = [C.POINTER(Params), C.c_int64, P64, P64, PD, P64, P64, PD, C.c_int64]
        lib.orc_pairs.restype = C.c_int64
        lib.orc_params_convert.argtypes = [C.POINTER(Params), C.c_void_p]
        lib.orc_count_ext.argtypes = [C.c_void_p, PU8, P64, C.c_int64, C.c_int64, C.c_double]
        lib.orc_count_ext.restype = C.c_int64
        _orc = lib
    return _orc


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        lib = _load(REF_SO)
        lib.ref_index_build.restype = C.c_void_p
        lib.ref_index_build.argtypes = [PU8, C.c_int64]
        lib.ref_index_free.argtypes = [C.c_void_p]
        lib.ref_length.argtypes = [C.c_void_p]
        lib.ref_length.restype = C.c_int64
        lib.ref_counts.argtypes = [C.c_void_p, P64, P64]
        lib.ref_access_bwt.argtypes = [C.c_void_p, C.c_int64]
        lib.ref_access_bwt.restype = C.c_int32
        lib.ref_rank.argtypes = [C.c_void_p, C.c_int32, C.c_int64]
        lib.ref_rank.restype = C.c_int64
        lib.ref_access_sa.argtypes = [C.c_void_p, C.c_int64]
        lib.ref_access_sa.restype = C.c_int64
        lib.ref_backext.argtypes = [C.c_void_p, P64, C.c_int32]
        lib.ref_backext.restype = C.c_int64
        lib.ref_fwdext.argtypes = [C.c_void_p, P64, C.c_int32]
        lib.ref_fwdext.restype = C.c_int64
        lib.ref_loglik.argtypes = [C.c_void_p, C.c_int32]
        lib.ref_loglik.restype = C.c_double
        lib.ref_find_smems.argtypes = [C.c_void_p, PU8, C.c_int64, C.c_int64, C.c_double, C.c_int32, P64,
                                       PD, C.c_int64]
        lib.ref_find_smems.restype = C.c_int64
        lib.ref_remove_redundant.argtypes = [P64, PD, C.c_int64]
        lib.ref_remove_redundant.restype = C.c_int64
        lib.ref_get_chains.argtypes = [P64, C.c_int64, C.c_int64, C.c_int64, P64, C.c_int64]
        lib.ref_get_chains.restype = C.c_int64
        lib.ref_alphabet.argtypes = [PU8, C.c_char_p, PU8, PU8, C.POINTER(C.c_int32), PU8, PU8]
        lib.ref_encode.argtypes = [C.c_char_p, PU8]
        lib.ref_encode.restype = C.c_int64
        lib.ref_revcom.argtypes = [PU8, C.c_int64, PU8]
        lib.ref_primary_revcom.argtypes = [C.c_char_p, C.c_char_p, PU8, PU8]
        lib.ref_primary_revcom.restype = C.c_int64
        lib.ref_index_build_blocks.restype = C.c_void_p
        lib.ref_index_build_blocks.argtypes = [PU8, P64, C.c_int32, C.c_int64]
        lib.ref_index_from_sa.restype = C.c_void_p
        lib.ref_index_from_sa.argtypes = [PU8, P64, C.c_int64]
        lib.ref_baseline_map.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, PU8, PU8, P64, C.c_char_p, PU8, PU8, P64,
                                         C.c_char_p, C.c_int64, C.c_int32, C.c_char_p, C.c_int64, P64, PD]
        lib.ref_baseline_map.restype = C.c_int64
        lib.ref_index_save.argtypes = [C.c_void_p, C.c_char_p]
        lib.ref_index_save.restype = C.c_int32
        lib.ref_index_load.argtypes = [C.c_char_p, C.c_int64]
        lib.ref_index_load.restype = C.c_void_p
        _ref = lib
    return _ref


def block_ends(lens, block_size):
    """msgseqtk-build's blocks for chromosomes of these lengths (BD ends; empty: one step)"""
    lens = np.ascontiguousarray(lens, np.int64)
    out = np.zeros(len(lens) + 1, np.int64)
    k = orc().orc_block_ends(ptr(lens, C.c_int64), len(lens), int(block_size), ptr(out, C.c_int64), len(out))
    return out[:k].copy()


def block_suffix_array(bd, ends):
    """the oracle's suffix array of a BD text in the block-by-block build order"""
    bd = np.ascontiguousarray(bd, np.uint8)
    ends = np.ascontiguousarray(ends, np.int64)
    sa = np.zeros(len(bd), np.int64)
    orc().orc_suffix_array(ptr(bd, C.c_uint8), len(bd), ptr(ends, C.c_int64), len(ends), ptr(sa, C.c_int64))
    return sa


class OracleDb:
    """Metagenome + FMD index held by the oracle."""

    def __init__(self, chroms, names=None, sa=None):
        self.chroms = [np.ascontiguousarray(c, dtype=np.uint8) for c in chroms]
        self.lens = np.array([len(c) for c in self.chroms], dtype=np.int64)
        self.names = names or ["g%04d" % i for i in range(len(chroms))]
        seqs = np.concatenate(self.chroms) if self.chroms else np.zeros(0, np.uint8)
        sap = ptr(sa, C.c_int64) if sa is not None else None
        self.h = orc().orc_db_create(ptr(seqs, C.c_uint8), ptr(self.lens, C.c_int64), len(self.chroms),
                                     "\n".join(self.names).encode(), sap)
        if not self.h:
            raise ValueError("oracle refused the database")
        self.n = orc().orc_db_len(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            orc().orc_db_free(self.h)
            self.h = None

    def counts(self):
        B = np.zeros(256, np.int64)
        Cc = np.zeros(256, np.int64)
        orc().orc_db_counts(self.h, ptr(B, C.c_int64), ptr(Cc, C.c_int64))
        return B, Cc

    def arrays(self):
        bd = np.zeros(self.n, np.uint8)
        bwt = np.zeros(self.n, np.uint8)
        sa = np.zeros(self.n, np.int64)
        orc().orc_db_copy(self.h, ptr(bd, C.c_uint8), ptr(bwt, C.c_uint8), ptr(sa, C.c_int64))
        return bd, bwt, sa

    def backext(self, p, q, s, b, fwd=False):
        v = np.array([p, q, s], np.int64)
        f = orc().orc_fwdext if fwd else orc().orc_backext
        r = f(self.h, ptr(v, C.c_int64), b)
        return int(r), tuple(int(x) for x in v)


def pack_reads(reads):
    """list of uint8 arrays -> (concat, offsets)"""
    offs = np.zeros(len(reads) + 1, np.int64)
    offs[1:] = np.cumsum([len(r) for r in reads])
    seq = np.concatenate([np.asarray(r, np.uint8) for r in reads]) if reads else np.zeros(0, np.uint8)
    if len(seq) == 0:
        seq = np.zeros(1, np.uint8)
    return np.ascontiguousarray(seq), offs


def orc_smems(db, reads, minLen=17, maxE=0.01, mems=False):
    seq, offs = pack_reads(reads)
    n = len(reads)
    counts = np.zeros(max(n, 1), np.int64)
    cap = 1 << 16
    while True:
        out = np.zeros((cap, 5), np.int64)
        lp = np.zeros(cap, np.float64)
        k = orc().orc_smems(db.h, ptr(seq, C.c_uint8), ptr(offs, C.c_int64), n, minLen, maxE, int(mems),
                            ptr(counts, C.c_int64), ptr(out, C.c_int64), ptr(lp, C.c_double), cap)
        if k <= cap:
            return counts[:n].copy(), out[:k].copy(), lp[:k].copy()
        cap = k


def orc_seeds(db, reads, minLen=17, maxE=0.01, maxNSeed=500):
    seq, offs = pack_reads(reads)
    n = len(reads)
    counts = np.zeros(max(n, 1), np.int64)
    cap = 1 << 16
    while True:
        out = np.zeros((cap, 6), np.int64)
        lp = np.zeros(cap, np.float64)
        k = orc().orc_seeds(db.h, ptr(seq, C.c_uint8), ptr(offs, C.c_int64), n, minLen, maxE, maxNSeed,
                            ptr(counts, C.c_int64), ptr(out, C.c_int64), ptr(lp, C.c_double), cap)
        if k <= cap:
            return counts[:n].copy(), out[:k].copy(), lp[:k].copy()
        cap = k


def orc_map(db, params, reads, quals, names, mates=None, nthreads=1):
    """SAM text per read (SE) or per pair (PE), in input order."""
    seq, offs = pack_reads(reads)
    qual, _ = pack_reads(quals)
    n = len(reads)
    nm = "\n".join(names).encode()
    if mates is not None:
        seq2, offs2 = pack_reads(mates[0])
        qual2, _ = pack_reads(mates[1])
        nm2 = "\n".join(mates[2]).encode()
        a2 = (ptr(seq2, C.c_uint8), ptr(qual2, C.c_uint8), ptr(offs2, C.c_int64), nm2)
    else:
        a2 = (None, None, None, None)
    secs = np.zeros(1, np.float64)
    need = orc().orc_map(db.h, C.byref(params), ptr(seq, C.c_uint8), ptr(qual, C.c_uint8),
                         ptr(offs, C.c_int64), nm, *a2, n, nthreads, None, 0, None, ptr(secs, C.c_double))
    buf = C.create_string_buffer(int(need) + 1)
    lo = np.zeros(n + 1, np.int64)
    orc().orc_map(db.h, C.byref(params), ptr(seq, C.c_uint8), ptr(qual, C.c_uint8), ptr(offs, C.c_int64),
                  nm, *a2, n, nthreads, buf, need + 1, ptr(lo, C.c_int64), ptr(secs, C.c_double))
    text = buf.raw[:need].decode()
    return [text[lo[i]:lo[i + 1]] for i in range(n)], float(secs[0])


class RefIndex:
    """The reference's own FMDIndex built by its own divsufsort from a BD text, or (sa given)
    from its own components over that suffix array (ref_index_from_sa)."""

    def __init__(self, bd, sa=None):
        bd = np.ascontiguousarray(bd, dtype=np.uint8)
        if sa is None:
            self.h = ref().ref_index_build(ptr(bd, C.c_uint8), len(bd))
        else:
            sa = np.ascontiguousarray(sa, np.int64)
            self.h = ref().ref_index_from_sa(ptr(bd, C.c_uint8), ptr(sa, C.c_int64), len(bd))
        if not self.h:
            raise ValueError("reference FMDIndex refused the text")
        self.n = len(bd)

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_index_free(self.h)
            self.h = None

    def counts(self):
        B = np.zeros(256, np.int64)
        Cc = np.zeros(256, np.int64)
        ref().ref_counts(self.h, ptr(B, C.c_int64), ptr(Cc, C.c_int64))
        return B, Cc

    def smems(self, read, minLen=17, maxE=0.01, mems=False):
        r = np.ascontiguousarray(read, np.uint8)
        cap = 4096
        out = np.zeros((cap, 5), np.int64)
        lp = np.zeros(cap, np.float64)
        k = ref().ref_find_smems(self.h, ptr(r, C.c_uint8), len(r), minLen, maxE, int(mems),
                                 ptr(out, C.c_int64), ptr(lp, C.c_double), cap)
        assert 0 <= k <= cap
        return out[:k].copy(), lp[:k].copy()


def alphabet(lib, prefix):
    """(encode[256], decode[16], complement[16], toBasic[16], toInt[16], isAmbiguous[16],
    quality encode[256]) from orc_alphabet or ref_alphabet"""
    enc = np.zeros(256, np.uint8)
    dec = C.create_string_buffer(16)
    comp = np.zeros(16, np.uint8)
    basic = np.zeros(16, np.uint8)
    toint = np.zeros(16, np.int32)
    amb = np.zeros(16, np.uint8)
    qenc = np.zeros(256, np.uint8)
    getattr(lib, prefix + "_alphabet")(ptr(enc, C.c_uint8), dec, ptr(comp, C.c_uint8), ptr(basic, C.c_uint8),
                                       ptr(toint, C.c_int32), ptr(amb, C.c_uint8), ptr(qenc, C.c_uint8))
    return enc, dec.raw[:16], comp, basic, toint, amb, qenc


def orc_raw_seeds(db, reads, minLen=17, maxE=0.01):
    """per read its SMEM count; per SMEM (size, seeds made); all seed rows"""
    seq, offs = pack_reads(reads)
    n = len(reads)
    counts = np.zeros(max(n, 1), np.int64)
    scap, cap = 1 << 16, 1 << 18
    while True:
        sz = np.zeros((scap, 2), np.int64)
        out = np.zeros((cap, 6), np.int64)
        k = orc().orc_raw_seeds(db.h, ptr(seq, C.c_uint8), ptr(offs, C.c_int64), n, minLen, maxE,
                                ptr(counts, C.c_int64), ptr(sz, C.c_int64), scap, ptr(out, C.c_int64), cap)
        m = int(counts[:n].sum())
        if k <= cap and m <= scap:
            return counts[:n].copy(), sz[:m].copy(), out[:k].copy()
        cap, scap = max(cap, k), max(scap, m)


def orc_align_chain(db, params, read, seeds, qual=None):
    """one chain of seed rows (from, to, start, end, tid, strand): (res10, score, log10P, CIGAR, MD)"""
    r = np.ascontiguousarray(read, np.uint8)
    q = np.ascontiguousarray(qual if qual is not None else np.full(len(r), 30), np.uint8)
    rows = np.ascontiguousarray(seeds, np.int64).reshape(-1, 6)
    res = np.zeros(10, np.int64)
    sc = np.zeros(1)
    lp = np.zeros(1)
    cig, md = C.create_string_buffer(4096), C.create_string_buffer(4096)
    orc().orc_align_chain(db.h, C.byref(params), ptr(r, C.c_uint8), ptr(q, C.c_uint8), len(r),
                          ptr(rows, C.c_int64), len(rows), ptr(res, C.c_int64), ptr(sc, C.c_double),
                          ptr(lp, C.c_double), cig, md, 4096)
    return res, float(sc[0]), float(lp[0]), cig.value.decode(), md.value.decode()


def orc_pairs(params, aln_offs, rows4, log10p):
    """AlignmentPE::getPairs + sort + filter of the oracle over alignment lists laid out as
    msgs_pair_batch takes them: (counts, rows3, lik2)"""
    offs = np.ascontiguousarray(aln_offs, np.int64)
    n = (len(offs) - 1) // 2
    r4 = np.ascontiguousarray(np.asarray(rows4, np.int64).reshape(-1, 4) if len(rows4) else np.zeros((1, 4), np.int64))
    lp = np.ascontiguousarray(log10p if len(log10p) else np.zeros(1), np.float64)
    cap = 1
    for i in range(n):
        cap += int((offs[i + 1] - offs[i]) * (offs[n + i + 1] - offs[n + i]))
    counts = np.zeros(max(n, 1), np.int64)
    rows3 = np.zeros((cap, 3), np.int64)
    lik2 = np.zeros((cap, 2))
    k = orc().orc_pairs(C.byref(params), n, ptr(offs, C.c_int64), ptr(r4, C.c_int64), ptr(lp, C.c_double),
                        ptr(counts, C.c_int64), ptr(rows3, C.c_int64), ptr(lik2, C.c_double), cap)
    assert k >= 0
    return counts[:n], rows3[:k], lik2[:k]


def orc_calc_mapq(log10p):
    v = np.ascontiguousarray(log10p, np.float64)
    pp = np.zeros(len(v))
    mq = np.zeros(len(v), np.uint8)
    orc().orc_calc_mapq(ptr(v, C.c_double), len(v), ptr(pp, C.c_double), ptr(mq, C.c_uint8))
    return pp, mq


def ref_baseline_map(rix, db, params, reads, quals, names, mates=None, nthreads=1):
    """msgseqtk-align's tasks with the reference's own seeding and chaining (RefIndex rix) and the
    oracle's DP and records (OracleDb db): SAM text per read (pair) and the mapping seconds."""
    seq, offs = pack_reads(reads)
    qual, _ = pack_reads(quals)
    n = len(reads)
    nm = "\n".join(names).encode()
    if mates is not None:
        seq2, offs2 = pack_reads(mates[0])
        qual2, _ = pack_reads(mates[1])
        nm2 = "\n".join(mates[2]).encode()
        a2 = (ptr(seq2, C.c_uint8), ptr(qual2, C.c_uint8), ptr(offs2, C.c_int64), nm2)
    else:
        a2 = (None, None, None, None)
    pbuf = C.create_string_buffer(1024)
    orc().orc_params_convert(C.byref(params), pbuf)
    secs = np.zeros(1, np.float64)
    args = (rix.h, db.h, pbuf, ptr(seq, C.c_uint8), ptr(qual, C.c_uint8), ptr(offs, C.c_int64), nm) + a2
    need = ref().ref_baseline_map(*args, n, nthreads, None, 0, None, ptr(secs, C.c_double))
    buf = C.create_string_buffer(int(need) + 1)
    lo = np.zeros(n + 1, np.int64)
    ref().ref_baseline_map(*args, n, nthreads, buf, need + 1, ptr(lo, C.c_int64), ptr(secs, C.c_double))
    text = buf.raw[:need].decode()
    return [text[lo[i]:lo[i + 1]] for i in range(n)], float(secs[0])


def orc_count_ext(db, reads, minLen=17, maxE=0.01):
    """FMD extensions (backExt/fwdExt with an unambiguous base) of findAllSMEMS over the reads"""
    seq, offs = pack_reads(reads)
    return int(orc().orc_count_ext(db.h, ptr(seq, C.c_uint8), ptr(offs, C.c_int64), len(reads), minLen, maxE))
