"""msgseqtk_amd -- MI355X (gfx950) read-mapping hot path of MSGseqTK.

The compute runs in libmsgs_hip.so (HIP kernels behind the C-ABI declared in
include/msgs_hip.h).  This module is thin ctypes plumbing around that library for
tests and benchmarks; there is no CPU fallback: importing the product path without
the built library raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmsgs_hip.so")

P64 = C.POINTER(C.c_int64)
PD = C.POINTER(C.c_double)
PU8 = C.POINTER(C.c_uint8)

MSGS_E_CAP = -6

# every extern "C" symbol of include/msgs_hip.h
EXPORTS = [
    "msgs_params_default", "msgs_ctx_create", "msgs_ctx_destroy", "msgs_last_error", "msgs_ctx_set",
    "msgs_db_build", "msgs_db_load", "msgs_db_nchrom", "msgs_db_chrom", "msgs_db_length", "msgs_db_counts", "msgs_db_copy", "msgs_db_loglik",
    "msgs_rank_batch", "msgs_ext_batch", "msgs_smem_batch", "msgs_mems_batch", "msgs_clean_batch", "msgs_seed_batch",
    "msgs_map_batch", "msgs_align_batch", "msgs_batch_upload", "msgs_batch_run",
    "msgs_batch_stats", "msgs_batch_sam", "msgs_util_sort_by_key",
    "msgs_batch_submit", "msgs_batch_wait", "msgs_tune", "msgs_util_calib_copy", "msgs_util_random_gather", "msgs_util_format_g6", "msgs_ctx_get", "msgs_lix_train", "msgs_lix_query", "msgs_lix_predict", "msgs_lix_grad", "msgs_lix_weights", "msgs_lix_knots",
    "msgs_smem_batch_noalloc", "msgs_search_batch", "msgs_sa_batch", "msgs_db_chrom_seq", "msgs_chain_batch",
    "msgs_dp_batch", "msgs_fmd_read", "msgs_fmd_write", "msgs_fmd_write_arrays", "msgs_util_mapq",
    "msgs_pair_batch", "msgs_util_insert_pr",
]
N_LIX_PARAMS = 16 * 64 + 64 + 64 + 1


class Params(C.Structure):
    """msgs_params (include/msgs_hip.h)."""
    _fields_ = [("minLen", C.c_int64), ("maxEvalue", C.c_double), ("maxNSeed", C.c_int64),
                ("mode", C.c_int32), ("maxReport", C.c_int32),
                ("match", C.c_double), ("mismatch", C.c_double), ("gapO", C.c_double),
                ("gapE", C.c_double), ("clip", C.c_double), ("minScoreRate", C.c_double),
                ("insMean", C.c_double), ("insSD", C.c_double), ("insMin", C.c_double),
                ("insMax", C.c_double),
                ("noMixed", C.c_int32), ("noDiscordant", C.c_int32), ("noTailOver", C.c_int32),
                ("noContain", C.c_int32), ("noOverlap", C.c_int32), ("pad0", C.c_int32),
                ("maxNPair", C.c_int64)]


class Reads(C.Structure):
    """msgs_reads (include/msgs_hip.h)."""
    _fields_ = [("seq", PU8), ("qual", PU8), ("offs", P64), ("names", C.c_char_p), ("n", C.c_int64)]


_lib = None


def lib():
    """Load libmsgs_hip.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmsgs_hip.so is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.msgs_params_default.argtypes = [C.POINTER(Params)]
    L.msgs_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.msgs_ctx_destroy.argtypes = [vp]
    L.msgs_last_error.argtypes = [vp]
    L.msgs_last_error.restype = C.c_char_p
    L.msgs_ctx_set.argtypes = [vp, C.c_char_p, C.c_int64]
    L.msgs_db_build.argtypes = [vp, PU8, P64, C.c_int32, C.c_char_p]
    L.msgs_db_load.argtypes = [vp, C.c_char_p]
    L.msgs_db_nchrom.argtypes = [vp]
    L.msgs_db_nchrom.restype = C.c_int32
    L.msgs_db_chrom.argtypes = [vp, C.c_int32, C.c_char_p, C.c_int64, P64]
    L.msgs_db_length.argtypes = [vp]
    L.msgs_db_length.restype = C.c_int64
    L.msgs_db_counts.argtypes = [vp, P64, P64]
    L.msgs_db_copy.argtypes = [vp, PU8, PU8, P64]
    L.msgs_db_loglik.argtypes = [vp, C.c_int32]
    L.msgs_db_loglik.restype = C.c_double
    L.msgs_rank_batch.argtypes = [vp, PU8, P64, P64, C.c_int64]
    L.msgs_ext_batch.argtypes = [vp, P64, PU8, C.c_int32, P64, C.c_int64]
    L.msgs_smem_batch.argtypes = [vp, C.POINTER(Reads), C.POINTER(Params), P64, P64, PD, C.c_int64, P64]
    L.msgs_seed_batch.argtypes = [vp, C.POINTER(Reads), C.POINTER(Params), P64, P64, PD, C.c_int64, P64]
    L.msgs_mems_batch.argtypes = [vp, C.POINTER(Reads), C.POINTER(Params), P64, P64, PD, PD, C.c_int64, P64]
    L.msgs_clean_batch.argtypes = [vp, vp, C.POINTER(Reads), C.POINTER(Reads), C.POINTER(Params), C.c_double,
                                   PD, PD, C.POINTER(C.c_uint8)]
    L.msgs_map_batch.argtypes = [vp, C.POINTER(Reads), C.POINTER(Reads), C.POINTER(Params), C.c_char_p,
                                 C.c_int64, P64, P64]
    L.msgs_align_batch.argtypes = [vp, C.POINTER(Reads), C.POINTER(Params), P64, P64, PD, C.c_char_p,
                                   C.c_int64, C.c_int64, P64]
    L.msgs_batch_upload.argtypes = [vp, C.c_int32, C.POINTER(Reads), C.POINTER(Reads)]
    L.msgs_batch_run.argtypes = [vp, C.c_int32, C.POINTER(Params), C.c_int32]
    L.msgs_batch_stats.argtypes = [vp, PD, C.c_int32]
    L.msgs_batch_sam.argtypes = [vp, C.c_char_p, C.c_int64, P64]
    L.msgs_util_sort_by_key.argtypes = [PD, C.POINTER(C.c_int32), C.c_int64]
    L.msgs_tune.argtypes = [vp, C.c_int32, C.POINTER(Params), C.POINTER(C.c_int32), C.c_int32,
                            C.POINTER(C.c_int32), PD]
    L.msgs_util_calib_copy.argtypes = [vp, C.c_int64, PD]
    L.msgs_ctx_get.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
    L.msgs_util_format_g6.argtypes = [C.POINTER(C.c_float), C.c_int64, C.c_char_p]
    L.msgs_util_format_g6.restype = None
    L.msgs_util_random_gather.argtypes = [vp, C.c_int64, C.c_int64, PD]
    L.msgs_batch_submit.argtypes = [vp, C.c_int32, C.POINTER(Params), C.c_int32]
    L.msgs_batch_wait.argtypes = [vp, C.c_int32]
    PF = C.POINTER(C.c_float)
    PU64 = C.POINTER(C.c_uint64)
    L.msgs_lix_train.argtypes = [vp, C.c_int32, C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_uint64, PD]
    L.msgs_lix_query.argtypes = [vp, PU64, C.c_int64, C.c_int32, P64, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.msgs_lix_predict.argtypes = [vp, PU64, C.c_int64, C.c_int32, PD]
    L.msgs_lix_grad.argtypes = [vp, PF, PF, PF, C.c_int32, PF, PF]
    L.msgs_lix_weights.argtypes = [vp, PF, PF]
    L.msgs_lix_knots.argtypes = [vp, PU64, P64, C.POINTER(C.c_int32)]
    L.msgs_smem_batch_noalloc.argtypes = [vp, vp, vp, C.c_int64, C.POINTER(Params), C.c_int32, vp, vp,
                                          C.POINTER(C.c_int32)]
    L.msgs_search_batch.argtypes = [vp, PU8, P64, C.c_int64, C.c_int32, P64]
    L.msgs_sa_batch.argtypes = [vp, P64, C.c_int64, P64]
    L.msgs_db_chrom_seq.argtypes = [vp, C.c_int32, PU8, C.c_int64, P64]
    L.msgs_chain_batch.argtypes = [vp, P64, P64, C.c_int64, P64, P64, P64, P64, C.c_int64, P64]
    L.msgs_dp_batch.argtypes = [vp, C.POINTER(Reads), P64, P64, P64, C.POINTER(Params), P64, PD, PD, C.c_char_p,
                                C.c_int64, C.c_char_p, C.c_int64, P64]
    L.msgs_fmd_read.argtypes = [C.c_char_p, C.c_int32, P64, PU8, P64, P64, P64, P64, P64, P64]
    L.msgs_fmd_write.argtypes = [vp, C.c_char_p, C.c_int32]
    L.msgs_fmd_write_arrays.argtypes = [C.c_char_p, C.c_int32, PU8, P64, C.c_int64]
    L.msgs_pair_batch.argtypes = [vp, C.c_int64, P64, P64, PD, C.POINTER(Params), P64, P64, PD, C.c_int64, P64]
    L.msgs_util_mapq.argtypes = [PD, C.c_int64, C.c_int32, PD, PU8]
    L.msgs_util_mapq.restype = None
    L.msgs_util_insert_pr.argtypes = [C.POINTER(Params), C.c_double]
    L.msgs_util_insert_pr.restype = C.c_double
    _lib = L
    return L


def default_params(**kw):
    p = Params()
    lib().msgs_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


_RETRIES = 6   # capacity regrows of one call (each sizes from the reported need)


class ReadBatch:
    """Reads packed for the C-ABI; keeps the numpy buffers alive."""

    def __init__(self, seqs, quals=None, names=None):
        self.n = len(seqs)
        self.offs = np.zeros(self.n + 1, np.int64)
        if self.n:
            self.offs[1:] = np.cumsum([len(s) for s in seqs])
        self.seq = np.ascontiguousarray(np.concatenate(seqs) if self.n else np.zeros(0, np.uint8), np.uint8)
        if self.seq.size == 0:
            self.seq = np.zeros(1, np.uint8)
        self.qual = None
        if quals is not None:
            self.qual = np.ascontiguousarray(np.concatenate(quals) if self.n else np.zeros(1, np.uint8), np.uint8)
            if self.qual.size == 0:
                self.qual = np.zeros(1, np.uint8)
        self.names = ("\n".join(names)).encode() if names is not None else None
        self.st = Reads(_p(self.seq, C.c_uint8), _p(self.qual, C.c_uint8) if self.qual is not None else None,
                        _p(self.offs, C.c_int64), self.names, self.n)


class Mapper:
    """A device context with a resident database (one GPU)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        rc = lib().msgs_ctx_create(device, C.byref(self.h))
        if rc != 0:
            raise RuntimeError("msgs_ctx_create failed (%d): no usable GPU" % rc)

    def close(self):
        if self.h:
            lib().msgs_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed (%d): %s" % (what, rc, lib().msgs_last_error(self.h).decode()))

    def set(self, key, value):
        self._chk(lib().msgs_ctx_set(self.h, key.encode(), int(value)), "msgs_ctx_set")
        return self

    def get(self, key):
        v = C.c_int64(0)
        self._chk(lib().msgs_ctx_get(self.h, key.encode(), C.byref(v)), "msgs_ctx_get")
        return v.value

    def build(self, chroms, names=None):
        chroms = [np.ascontiguousarray(c, np.uint8) for c in chroms]
        lens = np.array([len(c) for c in chroms], np.int64)
        seqs = np.ascontiguousarray(np.concatenate(chroms), np.uint8)
        nm = "\n".join(names or ["g%04d" % i for i in range(len(chroms))]).encode()
        self._chk(lib().msgs_db_build(self.h, _p(seqs, C.c_uint8), _p(lens, C.c_int64), len(chroms), nm),
                  "msgs_db_build")
        self.n = lib().msgs_db_length(self.h)
        return self

    def load(self, prefix):
        """the reference's database files prefix.mtg + prefix.mgs"""
        self._chk(lib().msgs_db_load(self.h, prefix.encode()), "msgs_db_load")
        self.n = lib().msgs_db_length(self.h)
        return self

    def counts(self):
        B = np.zeros(256, np.int64)
        Cc = np.zeros(256, np.int64)
        self._chk(lib().msgs_db_counts(self.h, _p(B, C.c_int64), _p(Cc, C.c_int64)), "msgs_db_counts")
        return B, Cc

    def arrays(self):
        bd = np.zeros(self.n, np.uint8)
        bwt = np.zeros(self.n, np.uint8)
        sa = np.zeros(self.n, np.int64)
        self._chk(lib().msgs_db_copy(self.h, _p(bd, C.c_uint8), _p(bwt, C.c_uint8), _p(sa, C.c_int64)),
                  "msgs_db_copy")
        return bd, bwt, sa

    def loglik(self, b):
        return lib().msgs_db_loglik(self.h, b)

    def rank(self, syms, pos):
        syms = np.ascontiguousarray(syms, np.uint8)
        pos = np.ascontiguousarray(pos, np.int64)
        out = np.zeros(len(pos), np.int64)
        self._chk(lib().msgs_rank_batch(self.h, _p(syms, C.c_uint8), _p(pos, C.c_int64), _p(out, C.c_int64),
                                        len(pos)), "msgs_rank_batch")
        return out

    def ext(self, pqs, bases, forward):
        pqs = np.ascontiguousarray(pqs, np.int64).copy()
        bases = np.ascontiguousarray(bases, np.uint8)
        ret = np.zeros(len(bases), np.int64)
        self._chk(lib().msgs_ext_batch(self.h, _p(pqs, C.c_int64), _p(bases, C.c_uint8), int(forward),
                                       _p(ret, C.c_int64), len(bases)), "msgs_ext_batch")
        return pqs, ret

    def smems(self, seqs, params):
        rb = ReadBatch(seqs)
        counts = np.zeros(max(rb.n, 1), np.int64)
        need = C.c_int64(0)
        cap = max(16, 8 * rb.n)
        for _ in range(_RETRIES):
            rows = np.zeros((cap, 5), np.int64)
            lp = np.zeros(cap, np.float64)
            rc = lib().msgs_smem_batch(self.h, C.byref(rb.st), C.byref(params), _p(counts, C.c_int64),
                                       _p(rows, C.c_int64), _p(lp, C.c_double), cap, C.byref(need))
            if rc == MSGS_E_CAP:
                cap = need.value
                continue
            self._chk(rc, "msgs_smem_batch")
            k = need.value
            return counts[:rb.n].copy(), rows[:k].copy(), lp[:k].copy()
        else:
            raise RuntimeError("output capacity not reached after %d tries" % _RETRIES)

    def search(self, patterns, mode):
        """bi-intervals (p, q, s) after msgs_search_batch's search (0 count, 1 fwd, 2 rev)"""
        rb = ReadBatch(patterns)
        out = np.zeros((max(rb.n, 1), 3), np.int64)
        self._chk(lib().msgs_search_batch(self.h, _p(rb.seq, C.c_uint8), _p(rb.offs, C.c_int64), rb.n, mode,
                                          _p(out, C.c_int64)), "msgs_search_batch")
        return out[:rb.n]

    def access_sa(self, rows):
        rows = np.ascontiguousarray(rows, np.int64)
        out = np.zeros(max(len(rows), 1), np.int64)
        self._chk(lib().msgs_sa_batch(self.h, _p(rows, C.c_int64), len(rows), _p(out, C.c_int64)), "msgs_sa_batch")
        return out[:len(rows)]

    def chains(self, seed_sets, max_mm, max_indel):
        """SeedChain::getChains of each list of seed rows: per set a list of index lists"""
        rows = np.concatenate([np.asarray(s, np.int64).reshape(-1, 6) for s in seed_sets]) if seed_sets else \
            np.zeros((0, 6), np.int64)
        rows = np.ascontiguousarray(rows if len(rows) else np.zeros((1, 6), np.int64))
        offs = np.zeros(len(seed_sets) + 1, np.int64)
        offs[1:] = np.cumsum([len(s) for s in seed_sets])
        mm = np.ascontiguousarray(max_mm, np.int64)
        ind = np.ascontiguousarray(max_indel, np.int64)
        counts = np.zeros(max(1, len(seed_sets)), np.int64)
        need = C.c_int64(0)
        cap = 4 * int(offs[-1]) + 16
        for _ in range(_RETRIES):
            out = np.zeros(cap, np.int64)
            rc = lib().msgs_chain_batch(self.h, _p(rows, C.c_int64), _p(offs, C.c_int64), len(seed_sets),
                                        _p(mm, C.c_int64), _p(ind, C.c_int64), _p(counts, C.c_int64),
                                        _p(out, C.c_int64), cap, C.byref(need))
            if rc == MSGS_E_CAP:
                cap = need.value + 1
                continue
            self._chk(rc, "msgs_chain_batch")
            break
        else:
            raise RuntimeError("output capacity not reached after %d tries" % _RETRIES)
        res, k = [], 0
        for i in range(len(seed_sets)):
            cs = []
            for _ in range(int(counts[i])):
                n = int(out[k])
                cs.append([int(x) for x in out[k + 1:k + 1 + n]])
                k += 1 + n
            res.append(cs)
        return res

    def pairs(self, aln_offs, rows4, log10p, params):
        """msgs_pair_batch: AlignmentPE::getPairs + sort + filter on the device over alignment
        lists (read pair i: rows aln_offs[i]..aln_offs[i+1] and aln_offs[n+i]..aln_offs[n+i+1] of
        rows4 = tid, strand, alnStart, alnEnd); returns (counts, rows3, lik2)"""
        offs = np.ascontiguousarray(aln_offs, np.int64)
        n = (len(offs) - 1) // 2
        r4 = np.ascontiguousarray(np.asarray(rows4, np.int64).reshape(-1, 4) if len(rows4) else
                                  np.zeros((1, 4), np.int64))
        lp = np.ascontiguousarray(log10p if len(log10p) else np.zeros(1), np.float64)
        counts = np.zeros(max(n, 1), np.int64)
        need = C.c_int64(0)
        cap = 16
        for _ in range(_RETRIES):
            rows3 = np.zeros((cap, 3), np.int64)
            lik2 = np.zeros((cap, 2))
            rc = lib().msgs_pair_batch(self.h, n, _p(offs, C.c_int64), _p(r4, C.c_int64), _p(lp, C.c_double),
                                       C.byref(params), _p(counts, C.c_int64), _p(rows3, C.c_int64),
                                       _p(lik2, C.c_double), cap, C.byref(need))
            if rc == MSGS_E_CAP:
                cap = need.value + 1
                continue
            self._chk(rc, "msgs_pair_batch")
            break
        else:
            raise RuntimeError("output capacity not reached after %d tries" % _RETRIES)
        k = int(counts[:n].sum())
        return counts[:n], rows3[:k], lik2[:k]

    def dp(self, reads, quals, chains_per_read, params):
        """msgs_dp_batch: chains_per_read[i] = list of chains, a chain = seed rows (from, to, start,
        end, tid, strand); returns per chain (rows10, score, log10P, CIGAR, MD)"""
        rb = ReadBatch(reads, quals)
        chain_offs = np.zeros(len(reads) + 1, np.int64)
        chain_offs[1:] = np.cumsum([len(c) for c in chains_per_read])
        flat = [np.asarray(ch, np.int64).reshape(-1, 6) for cs in chains_per_read for ch in cs]
        seed_offs = np.zeros(len(flat) + 1, np.int64)
        seed_offs[1:] = np.cumsum([len(x) for x in flat])
        rows = np.ascontiguousarray(np.concatenate(flat) if flat else np.zeros((1, 6), np.int64))
        nj = len(flat)
        r10 = np.zeros((max(nj, 1), 10), np.int64)
        sc = np.zeros(max(nj, 1))
        lp = np.zeros(max(nj, 1))
        cap = 64 * 1024 + 1200 * max(nj, 1)
        cig, md = C.create_string_buffer(cap), C.create_string_buffer(cap)
        need = C.c_int64(0)
        self._chk(lib().msgs_dp_batch(self.h, C.byref(rb.st), _p(chain_offs, C.c_int64), _p(seed_offs, C.c_int64),
                                      _p(rows, C.c_int64), C.byref(params), _p(r10, C.c_int64), _p(sc, C.c_double),
                                      _p(lp, C.c_double), cig, cap, md, cap, C.byref(need)), "msgs_dp_batch")
        cl = cig.value.decode().split("\n")[:nj]
        ml = md.value.decode().split("\n")[:nj]
        return [(r10[k], sc[k], lp[k], cl[k], ml[k]) for k in range(nj)]

    def fmd_write(self, path, header=True):
        self._chk(lib().msgs_fmd_write(self.h, path.encode(), int(header)), "msgs_fmd_write")

    def mems(self, seqs, params):
        """SMEM_LIST::findMEMS of each read: (counts, rows5, logP, per-read list loglik)"""
        rb = ReadBatch(seqs)
        counts = np.zeros(max(rb.n, 1), np.int64)
        rll = np.zeros(max(rb.n, 1), np.float64)
        need = C.c_int64(0)
        cap = max(16, 8 * rb.n)
        for _ in range(_RETRIES):
            rows = np.zeros((cap, 5), np.int64)
            lp = np.zeros(cap, np.float64)
            rc = lib().msgs_mems_batch(self.h, C.byref(rb.st), C.byref(params), _p(counts, C.c_int64),
                                       _p(rows, C.c_int64), _p(lp, C.c_double), _p(rll, C.c_double), cap,
                                       C.byref(need))
            if rc == MSGS_E_CAP:
                cap = need.value
                continue
            self._chk(rc, "msgs_mems_batch")
            k = need.value
            return counts[:rb.n].copy(), rows[:k].copy(), lp[:k].copy(), rll[:rb.n].copy()
        else:
            raise RuntimeError("output capacity not reached after %d tries" % _RETRIES)

    def clean(self, background, seqs, params, min_lod, mates=None):
        """msgseqtk-clean's decision with this mapper's database as the reference and
        `background`'s as the background: (ref loglik, bg loglik, keep) per read or pair"""
        rb = ReadBatch(seqs)
        mb = ReadBatch(mates) if mates is not None else None
        n = max(rb.n, 1)
        rl = np.zeros(n, np.float64)
        bl = np.zeros(n, np.float64)
        keep = np.zeros(n, np.uint8)
        rc = lib().msgs_clean_batch(self.h, background.h, C.byref(rb.st), C.byref(mb.st) if mb else None,
                                    C.byref(params), float(min_lod), _p(rl, C.c_double), _p(bl, C.c_double),
                                    keep.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc:
            # the failing database's context holds the message
            msg = lib().msgs_last_error(self.h).decode() or lib().msgs_last_error(background.h).decode()
            raise RuntimeError("msgs_clean_batch failed (%d): %s" % (rc, msg))
        return rl[:rb.n].copy(), bl[:rb.n].copy(), keep[:rb.n].astype(bool)

    def seeds(self, seqs, params):
        rb = ReadBatch(seqs)
        counts = np.zeros(max(rb.n, 1), np.int64)
        need = C.c_int64(0)
        cap = max(16, 8 * rb.n)
        for _ in range(_RETRIES):
            rows = np.zeros((cap, 6), np.int64)
            lp = np.zeros(cap, np.float64)
            rc = lib().msgs_seed_batch(self.h, C.byref(rb.st), C.byref(params), _p(counts, C.c_int64),
                                       _p(rows, C.c_int64), _p(lp, C.c_double), cap, C.byref(need))
            if rc == MSGS_E_CAP:
                cap = need.value
                continue
            self._chk(rc, "msgs_seed_batch")
            k = need.value
            return counts[:rb.n].copy(), rows[:k].copy(), lp[:k].copy()
        else:
            raise RuntimeError("output capacity not reached after %d tries" % _RETRIES)

    def alignments(self, seqs, quals, params):
        rb = ReadBatch(seqs, quals)
        counts = np.zeros(max(rb.n, 1), np.int64)
        need = C.c_int64(0)
        cap = max(16, 4 * rb.n)
        ccap = 64 * cap + 64
        for _ in range(_RETRIES):
            rows = np.zeros((cap, 10), np.int64)
            sc = np.zeros(cap, np.float64)
            cig = C.create_string_buffer(ccap)
            rc = lib().msgs_align_batch(self.h, C.byref(rb.st), C.byref(params), _p(counts, C.c_int64),
                                        _p(rows, C.c_int64), _p(sc, C.c_double), cig, len(cig), cap,
                                        C.byref(need))
            if rc == MSGS_E_CAP:
                if need.value > cap:      # rows
                    cap = need.value
                    ccap = max(ccap, 64 * cap + 64)
                else:                     # CIGAR text
                    ccap *= 4
                continue
            self._chk(rc, "msgs_align_batch")
            k = need.value
            cigars = cig.value.decode().split("\n")[:k]
            return counts[:rb.n].copy(), rows[:k].copy(), sc[:k].copy(), cigars
        else:
            raise RuntimeError("msgs_align_batch: output capacity not reached after %d tries" % _RETRIES)

    def map(self, reads, params, mates=None):
        """reads/mates: (seqs, quals, names).  Returns the SAM text per read (pair)."""
        rb = ReadBatch(*reads)
        mb = ReadBatch(*mates) if mates is not None else None
        need = C.c_int64(0)
        lo = np.zeros(rb.n + 1, np.int64)
        cap = 1024 * max(rb.n, 1)
        for _ in range(_RETRIES):
            buf = C.create_string_buffer(cap)
            rc = lib().msgs_map_batch(self.h, C.byref(rb.st), C.byref(mb.st) if mb else None, C.byref(params),
                                      buf, cap, _p(lo, C.c_int64), C.byref(need))
            if rc == MSGS_E_CAP:
                cap = need.value + 1
                continue
            self._chk(rc, "msgs_map_batch")
            text = buf.raw[:need.value].decode()
            return [text[lo[i]:lo[i + 1]] for i in range(rb.n)]
        else:
            raise RuntimeError("output capacity not reached after %d tries" % _RETRIES)

    def submit(self, slot, params, emit_sam=True):
        self._chk(lib().msgs_batch_submit(self.h, slot, C.byref(params), 1 if emit_sam else 0), "msgs_batch_submit")

    def wait(self, slot):
        self._chk(lib().msgs_batch_wait(self.h, slot), "msgs_batch_wait")
        return self.stats()

    def calib_copy(self, nbytes):
        us = C.c_double(0)
        self._chk(lib().msgs_util_calib_copy(self.h, nbytes, C.byref(us)), "msgs_util_calib_copy")
        return us.value

    def random_gather(self, nbytes, nloads):
        """microseconds for nloads random 64-byte line reads over a device buffer of nbytes"""
        us = C.c_double(0)
        self._chk(lib().msgs_util_random_gather(self.h, nbytes, nloads, C.byref(us)), "msgs_util_random_gather")
        return us.value

    def tune(self, slot, params, cand=None):
        """pick the fastest lanes-per-CU for the SMEM search of a resident batch"""
        c = None if cand is None else np.ascontiguousarray(cand, np.int32)
        best = C.c_int32(0)
        us = C.c_double(0)
        self._chk(lib().msgs_tune(self.h, slot, C.byref(params), None if c is None else _p(c, C.c_int32),
                                  0 if c is None else len(c), C.byref(best), C.byref(us)), "msgs_tune")
        return best.value, us.value

    # learned index
    def lix_train(self, segments=4096, samples=1 << 20, steps=2000, batch=4096, lr=1e-3, seed=7):
        info = np.zeros(5, np.float64)
        self._chk(lib().msgs_lix_train(self.h, segments, samples, steps, batch, lr, seed, _p(info, C.c_double)),
                  "msgs_lix_train")
        return {"max_err": int(info[0]), "max_err_root": int(info[1]), "loss0": info[2], "loss": info[3],
                "train_ms": info[4]}

    def lix_query(self, keys, use_mlp=1):
        keys = np.ascontiguousarray(keys, np.uint64)
        n = len(keys)
        lb = np.zeros(n, np.int64)
        pr = np.zeros(n, np.int32)
        fb = np.zeros(n, np.int32)
        self._chk(lib().msgs_lix_query(self.h, _p(keys, C.c_uint64), n, use_mlp, _p(lb, C.c_int64),
                                       _p(pr, C.c_int32), _p(fb, C.c_int32)), "msgs_lix_query")
        return lb, pr, fb

    def lix_predict(self, keys, mode=0):
        keys = np.ascontiguousarray(keys, np.uint64)
        out = np.zeros(len(keys), np.float64)
        self._chk(lib().msgs_lix_predict(self.h, _p(keys, C.c_uint64), len(keys), mode, _p(out, C.c_double)),
                  "msgs_lix_predict")
        return out

    def lix_grad(self, W, X, T):
        W = np.ascontiguousarray(W, np.float32)
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(N_LIX_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._chk(lib().msgs_lix_grad(self.h, _p(W, C.c_float), _p(X, C.c_float), _p(T, C.c_float), len(T),
                                      _p(g, C.c_float), _p(loss, C.c_float)), "msgs_lix_grad")
        return g, float(loss[0])

    def lix_knots(self):
        S = C.c_int32(0)
        self._chk(lib().msgs_lix_knots(self.h, None, None, C.byref(S)), "msgs_lix_knots")
        keys = np.zeros(S.value + 1, np.uint64)
        ranks = np.zeros(S.value + 1, np.int64)
        self._chk(lib().msgs_lix_knots(self.h, _p(keys, C.c_uint64), _p(ranks, C.c_int64), None), "msgs_lix_knots")
        return keys, ranks

    def lix_weights(self, set_to=None):
        w = np.zeros(N_LIX_PARAMS, np.float32)
        if set_to is not None:
            s = np.ascontiguousarray(set_to, np.float32)
            self._chk(lib().msgs_lix_weights(self.h, None, _p(s, C.c_float)), "msgs_lix_weights")
        self._chk(lib().msgs_lix_weights(self.h, _p(w, C.c_float), None), "msgs_lix_weights")
        return w

    # resident batches
    def upload(self, slot, reads, mates=None):
        rb = ReadBatch(*reads)
        mb = ReadBatch(*mates) if mates is not None else None
        self._chk(lib().msgs_batch_upload(self.h, slot, C.byref(rb.st), C.byref(mb.st) if mb else None),
                  "msgs_batch_upload")

    def run(self, slot, params, emit_sam=True):
        self._chk(lib().msgs_batch_run(self.h, slot, C.byref(params), int(emit_sam)), "msgs_batch_run")

    def stats(self):
        s = np.zeros(48, np.float64)
        self._chk(lib().msgs_batch_stats(self.h, _p(s, C.c_double), 48), "msgs_batch_stats")
        return s

    def sam(self):
        need = C.c_int64(0)
        rc = lib().msgs_batch_sam(self.h, None, 0, C.byref(need))
        buf = C.create_string_buffer(need.value + 1)
        self._chk(lib().msgs_batch_sam(self.h, buf, need.value + 1, C.byref(need)), "msgs_batch_sam")
        return buf.raw[:need.value].decode()


def sort_by_key(keys):
    """Permutation the reference's std::sort produces for keys compared with '<'
    (the sort restatement of the library, run on the host)."""
    keys = np.ascontiguousarray(keys, np.float64)
    perm = np.arange(len(keys), dtype=np.int32)
    lib().msgs_util_sort_by_key(_p(keys, C.c_double), _p(perm, C.c_int32), len(keys))
    return perm


def fmd_read(path, header=True):
    """DB.fmd decoded: (bwt, B, C, gapSA, SAsampled) (msgs_fmd_read, host only)"""
    n, ng, ns = C.c_int64(0), C.c_int64(0), C.c_int64(0)
    rc = lib().msgs_fmd_read(path.encode(), int(header), C.byref(n), None, None, None, None, C.byref(ng), None,
                             C.byref(ns))
    if rc != 0:
        raise RuntimeError("msgs_fmd_read failed (%d)" % rc)
    bwt = np.zeros(max(n.value, 1), np.uint8)
    B, Cc = np.zeros(256, np.int64), np.zeros(256, np.int64)
    gap, smp = np.zeros(max(ng.value, 1), np.int64), np.zeros(max(ns.value, 1), np.int64)
    rc = lib().msgs_fmd_read(path.encode(), int(header), C.byref(n), _p(bwt, C.c_uint8), _p(B, C.c_int64),
                             _p(Cc, C.c_int64), _p(gap, C.c_int64), C.byref(ng), _p(smp, C.c_int64), C.byref(ns))
    if rc != 0:
        raise RuntimeError("msgs_fmd_read failed (%d)" % rc)
    return bwt[:n.value], B, Cc, gap[:ng.value], smp[:ns.value]


def calc_mapq(loglik, pe=False):
    """Alignment::calcMapQ / AlignmentPE::calcMapQ (src/Alignment.cpp:491-540) over natural-log
    likelihoods: the posterior probabilities and MAPQ values (run on the host)."""
    ll = np.ascontiguousarray(loglik, np.float64)
    pp = np.zeros(len(ll))
    mq = np.zeros(len(ll), np.uint8)
    lib().msgs_util_mapq(_p(ll, C.c_double), len(ll), 1 if pe else 0, _p(pp, C.c_double), _p(mq, C.c_uint8))
    return pp, mq
