/*
 * msgs_smem.hip -- batched SMEM search on the FMD index (SMEM_LIST::findAllSMEMS,
 * src/SMEM.cpp:186-215, with findFwdBackSMEMS :64-117 and findBackFwdSMEMS :119-172).
 *
 * One lane owns one read at a time and pulls the next read from a device work queue
 * when it finishes (persistent grid).  The reference's nested loops are unrolled into a
 * step machine: every loop iteration of a lane performs exactly one FMD extension, so
 * all lanes of a wavefront issue their occ-block loads at the same instruction no
 * matter which phase of which search they are in.  The lock-step candidate list of a
 * search (the reference's `prev`) keeps its first entries in LDS and the rest in a
 * per-lane global scratch area (records interleaved over lanes); it is traversed in
 * reverse of its push order and compacted in place exactly like the reference's
 * reverse + erase(remove_if).
 *
 * Floating point: logP is accumulated in fp64 in the reference's order.  The e-value
 * gates N*exp(logP) <= maxE are evaluated as logP <= T, with T found on the host from
 * glibc exp (see msgs_capi.cpp: evalueThreshold); the extended-range sums of the first
 * gate add only non-positive terms, so they stop as soon as they pass T.
 */
#include "msgs_device.h"
#include "msgs_kernels.h"

namespace {

enum : int { ST_IDLE = 0, ST_FB1 = 1, ST_FB2 = 2, ST_BF1 = 3, ST_BF2 = 4, ST_WAIT = 5 };
constexpr int kNoHand = -2147483647 - 1;
enum : int { STG_A = 0, STG_C = 1, STG_D = 2 };

template <class Row>
struct Lane {
    /* read */
    int64_t rid;
    const uint8_t* s;
    const double* P;   /* loglik prefix sums of the read (NULL: sequential gates only) */
    double gD;         /* P[i1] - P[i0] of the running phase-1 search's push gate */
    double gPL;        /* |P[L]| */
    int L;
    /* findAllSMEMS variables */
    int dFrom, dTo, stage;
    /* the running search: its by-reference from/to */
    int sFrom, sTo;
    int state;
    /* phase-1 SMEM */
    int cFrom, cTo;
    Row cF, cR, cS, cX;       /* rows and sizes (int32_t for narrow indexes, int64_t for wide) */
    double cLL;
    /* candidate list */
    int nCand, lo, k, w, nValid;
    Row s0;
    /* output */
    int nOut;
    /* paired chains: role 1 = the pair's even lane (first search, then forward searches),
     * 2 = its odd lane (backward searches), 0 = unpaired; tag = the running search's number;
     * hand = the first search's final `from`, passed to the odd lane (kNoHand: none) */
    int role, tag, hand;
    /* counters */
    int64_t nExt, nFast, nLook;
};

template <class Row>
__device__ __forceinline__ uint8_t base(const Lane<Row>& l, int i) { return i < l.L ? l.s[i] : (uint8_t)0; }

template <class Row>
struct Ctx {
    const DevIndex& ix;
    const SmemArgs& a;
    int64_t lane;
    bool grouped;   /* lane groups keep each group's list contiguous (chunks of G entries) */
    int64_t qBase, qEnd;   /* this launch's queue entries [qBase, qEnd) */
};

/* the queue entries of a launch: all of them, or one side of the split */
__device__ __forceinline__ void queueRange(const SmemArgs& a, int64_t& base, int64_t& end) {
    const int64_t split = a.qSplit ? (int64_t)*a.qSplit : 0;
    base = a.qPart == 2 ? split : 0;
    end = a.qPart == 1 ? split : a.nReads;
}

template <class Row>
__device__ __forceinline__ void emit(const Ctx<Row>& c, Lane<Row>& l, int from, int to, Row F, Row R, Row S, double ll) {
    if (l.nOut < c.a.cap) {
        const size_t k = (size_t)l.rid * c.a.cap + l.nOut;
        SmemRec* o = (l.role == 2 ? c.a.outB : c.a.out) + k;
        o->from = from; o->to = to; o->fwd = F; o->rev = R; o->size = S; o->logP = ll;
        if (l.role) (l.role == 2 ? c.a.tagB : c.a.tagA)[k] = l.tag;
    }
    l.nOut++;
}

/* candidate k of a lane: interleaved over lanes, so that lanes at the same list position
 * touch neighbouring entries; of a lane group: contiguous, so that a chunk of G entries is
 * one coalesced access */
template <class Row>
__device__ __forceinline__ size_t cidx(const Ctx<Row>& c, int k) {
    return c.grouped ? (size_t)c.lane * c.a.candCap + k : (size_t)k * c.a.nLanes + c.lane;
}

/* A candidate in the global scratch: three 64-bit words, four with X (HX: the kernels whose
 * lists carry X).  Word 0 and 1 hold F, R, S and pos: 32 bits each for 32-bit rows; for 64-bit
 * rows (all below 2^34) F in bits 0-33 of word 0, R in bits 34-63 of word 0 and 0-3 of word 1,
 * S in bits 4-37 and pos in bits 38-63 of word 1 (reads are below 2^26 bases).  Word 2 is LL's
 * bit pattern, word 3 X + 1. */
template <bool HX>
constexpr int candWords() { return HX ? 4 : 3; }

template <bool HX, class Row>
__device__ __forceinline__ void candStore(const Ctx<Row>& c, int k, int pos, Row F, Row R, Row S, double ll, Row X) {
    uint64_t* const p = static_cast<uint64_t*>(c.a.cand) + cidx(c, k) * candWords<HX>();
    uint64_t w0, w1;
    if constexpr (sizeof(Row) == 8) {
        w0 = (uint64_t)F | ((uint64_t)R << 34);
        w1 = ((uint64_t)R >> 30) | ((uint64_t)S << 4) | ((uint64_t)(uint32_t)pos << 38);
    } else {
        w0 = (uint64_t)(uint32_t)F | ((uint64_t)(uint32_t)R << 32);
        w1 = (uint64_t)(uint32_t)S | ((uint64_t)(uint32_t)pos << 32);
    }
    p[0] = w0;
    p[1] = w1;
    p[2] = (uint64_t)__double_as_longlong(ll);
    if constexpr (HX) p[3] = (uint64_t)((int64_t)X + 1);
}

template <bool HX, class Row>
__device__ __forceinline__ SmemCandT<Row> candAt(const Ctx<Row>& c, int k) {
    const uint64_t* const p = static_cast<const uint64_t*>(c.a.cand) + cidx(c, k) * candWords<HX>();
    const uint64_t w0 = p[0], w1 = p[1], w2 = p[2];
    SmemCandT<Row> v;
    if constexpr (sizeof(Row) == 8) {
        constexpr uint64_t M = (1ull << 34) - 1;
        v.F = (Row)(w0 & M);
        v.R = (Row)((w0 >> 34) | ((w1 & 15u) << 30));
        v.S = (Row)((w1 >> 4) & M);
        v.pos = (int32_t)(w1 >> 38);
    } else {
        v.F = (Row)(uint32_t)w0;
        v.R = (Row)(uint32_t)(w0 >> 32);
        v.S = (Row)(uint32_t)w1;
        v.pos = (int32_t)(w1 >> 32);
    }
    v.LL = __longlong_as_double((long long)w2);
    v.X = HX ? (Row)((int64_t)p[3] - 1) : (Row)-1;
    v.pad = 0;
    return v;
}

/* The per-lane kernels keep the top W entries of each lane's list (indices nCand - W ..
 * nCand - 1) in LDS, one slot column per thread of the 256-thread block (structure of
 * arrays, slot = index % W), and the entries below them in the global scratch.  Phase 1
 * pushes at the top, evicting the entry that leaves the window to the global scratch;
 * phase 2 reads the list from the top down and stores the kept entries from the top down,
 * so its passes after the first stay in LDS while at most W entries survive.  W is what
 * fits the blocks per CU of the row width (four for 32-bit rows: at most 38 KB per block;
 * three for 64-bit rows: at most 48 KB); X is kept only by the unique-interval path. */
template <bool FAST, class Row>
struct CandSlots {
    static constexpr int W = FAST ? (sizeof(Row) == 8 ? 4 : 5) : (sizeof(Row) == 8 ? 5 : 6);
    Row F[W][256], R[W][256], S[W][256];
    Row X[FAST ? W : 1][256];
    double LL[W][256];
    int32_t pos[W][256];
};

template <bool FAST, class Row>
__device__ __forceinline__ SmemCandT<Row> slotLoad(const CandSlots<FAST, Row>& s, int q) {
    const int t = threadIdx.x;
    SmemCandT<Row> v;
    v.F = s.F[q][t]; v.R = s.R[q][t]; v.S = s.S[q][t]; v.X = FAST ? s.X[q][t] : (Row)-1;
    v.LL = s.LL[q][t]; v.pos = s.pos[q][t]; v.pad = 0;
    return v;
}

/* entry k of a list of nCand entries */
template <bool FAST, class Row>
__device__ __forceinline__ SmemCandT<Row> laneLoad(const Ctx<Row>& c, const CandSlots<FAST, Row>& s, int k, int nCand) {
    constexpr int W = CandSlots<FAST, Row>::W;
    if (k >= nCand - W) return slotLoad(s, k % W);
    return candAt<FAST>(c, k);
}

template <bool FAST, class Row>
__device__ __forceinline__ void laneStore(const Ctx<Row>& c, CandSlots<FAST, Row>& s, int k, int nCand, int pos, Row F,
                                          Row R, Row S, double ll, Row X) {
    constexpr int W = CandSlots<FAST, Row>::W;
    if (k >= nCand - W) {
        const int t = threadIdx.x, q = k % W;
        s.F[q][t] = F; s.R[q][t] = R; s.S[q][t] = S;
        if (FAST) s.X[q][t] = X;
        s.LL[q][t] = ll; s.pos[q][t] = pos;
        return;
    }
    candStore<FAST>(c, k, pos, F, R, S, ll, X);
}

/* push entry k = nCand at the top: the entry k - W leaves the window for the global scratch */
template <bool FAST, class Row>
__device__ __forceinline__ void lanePush(const Ctx<Row>& c, CandSlots<FAST, Row>& s, int k, int pos, Row F, Row R,
                                         Row S, double ll, Row X) {
    constexpr int W = CandSlots<FAST, Row>::W;
    if (k >= W) {
        const SmemCandT<Row> o = slotLoad(s, k % W);
        candStore<FAST>(c, k - W, o.pos, o.F, o.R, o.S, o.LL, o.X);
    }
    laneStore(c, s, k, k + 1, pos, F, R, S, ll, X);
}

/* start findFwdBackSMEMS / findBackFwdSMEMS at f; false if it ends at once (ambiguous base) */
template <class Row>
__device__ __forceinline__ bool startSearch(const Ctx<Row>& c, Lane<Row>& l, int f, bool fb) {
    l.sFrom = f;
    l.sTo = f + 1;
    uint8_t b = base(l, f);
    if (amb(b)) return false;
    l.cFrom = f; l.cTo = f + 1;
    l.cF = c.ix.C[b]; l.cR = c.ix.C[comp16(b)]; l.cS = c.ix.C[b + 1] - c.ix.C[b];
    l.cX = -1;
    double ll = 0;
    ll += c.ix.ll[b];
    l.cLL = ll;
    l.nCand = 0;
    l.state = fb ? ST_FB1 : ST_BF1;
    /* the push gate's read range stays fixed during phase 1: [0, from) forward, [to, L) back */
    if (l.P) l.gD = fb ? l.P[f] - l.P[0] : l.P[l.L] - l.P[f + 1];
    return true;
}

/* the backward searches of findAllSMEMS alone (a pair's odd lane): while from > 0, search
 * back from --from; a search that ends at once leaves from unchanged */
template <class Row>
__device__ void pairBackChain(const Ctx<Row>& c, Lane<Row>& l) {
    while (l.dFrom > 0) {
        --l.dFrom;
        l.tag++;
        l.stage = STG_D;
        if (startSearch(c, l, l.dFrom, false)) return;
    }
    l.state = ST_IDLE;
}

/* findAllSMEMS control flow after a search returned; ends with a running search or an idle lane.
 * After the first search, the loop's forward searches depend only on `to` and its backward
 * searches only on `from` (src/SMEM.cpp:196-203: from0 and to0 are copies), and the k-th of
 * each runs in the loop's k-th iteration; a pair runs the two chains on its two lanes and tags
 * each SMEM with k, so that the merge restores the loop's order. */
template <class Row>
__device__ void driverAfter(const Ctx<Row>& c, Lane<Row>& l) {
    if (l.role == 1) {
        if (l.stage == STG_A) { l.dFrom = l.sFrom; l.dTo = l.sTo; l.hand = l.dFrom; }
        else l.dTo = l.sTo;
        while (l.dTo < l.L) {
            l.tag++;
            l.stage = STG_C;
            if (startSearch(c, l, l.dTo, true)) return;
            l.dTo = l.sTo;   /* ended at once: to = from0 + 1 */
        }
        l.state = ST_IDLE;
        return;
    }
    if (l.role == 2) {
        l.dFrom = l.sFrom;
        pairBackChain(c, l);
        return;
    }
    int pc;  /* 0 = loop check, 1 = back-search preparation */
    if (l.stage == STG_A) { l.dFrom = l.sFrom; l.dTo = l.sTo; pc = 0; }
    else if (l.stage == STG_C) { l.dTo = l.sTo; pc = 1; }
    else { l.dFrom = l.sFrom; pc = 0; }
    while (true) {
        if (pc == 0) {
            if (!(l.dFrom > 0 || l.dTo < l.L)) { l.state = ST_IDLE; return; }
            int from0 = l.dTo;
            if (from0 < l.L) {
                l.stage = STG_C;
                if (startSearch(c, l, from0, true)) return;
                l.dTo = l.sTo;  /* ended at once: to = from0 + 1 */
            }
            pc = 1;
        } else {
            --l.dFrom;
            if (l.dFrom >= 0) {
                l.stage = STG_D;
                if (startSearch(c, l, l.dFrom, false)) return;
                /* ended at once: from unchanged */
            }
            pc = 0;
        }
    }
}

/* loglik of base code v: A/C/G/T from registers, other codes from the table */
template <class Row>
__device__ __forceinline__ double llOf(const Ctx<Row>& c, uint32_t v, double lA, double lC, double lG, double lT) {
    return v == 1 ? lA : (v == 2 ? lC : (v == 4 ? lG : (v == 8 ? lT : c.ix.ll[v])));
}

/* x0 + ll(s[i0]) + ll(s[i0+1]) + ... + ll(s[i1-1]) <= T, summed in that order; stops once
 * the running sum is <= T (every term is <= 0).  Bases are read 8 at a time from
 * aligned words (the read buffer is padded). */
template <class Row>
__device__ __forceinline__ bool gateSum(const Ctx<Row>& c, const Lane<Row>& l, double x, int i0, int i1, double T, double lA,
                                        double lC, double lG, double lT) {
    const int64_t off = l.s - c.a.seqs;
    int64_t wb = -1;
    uint64_t w = 0;
    for (int i = i0; i < i1; ++i) {
        if (x <= T) return true;
        const int64_t g = off + i;
        if ((g & ~(int64_t)7) != wb) { wb = g & ~(int64_t)7; w = *(const uint64_t*)(c.a.seqs + wb); }
        x += llOf(c, (uint32_t)(w >> (8 * (g & 7))) & 15u, lA, lC, lG, lT);
    }
    return x <= T;
}

/* The gates below compare with T a sum of m <= L+1 non-positive terms.  Two summation
 * orders of the same terms differ by at most 2 (m-1) u sum|t| (u = 2^-53; the prefix
 * difference of gateBack adds one such bound), so a prefix-sum estimate outside
 * T +- slack, slack = 1e-15 (L+2) (1 + |ll0| + |P[L]|), decides exactly; estimates inside
 * the band (and NaN estimates from -inf terms) take the sequential sum in the
 * reference's order. */
template <class Row>
__device__ __forceinline__ double gateSlack(const Lane<Row>& l, double ll0) {
    return 1e-15 * (double)(l.L + 2) * (1.0 + fabs(ll0) + l.gPL);
}

/* the phase-1 push gate: logP of the candidate plus the logliks of read[i0, i1) <= T1,
 * i.e. evalue(0, to) for a forward search ([0, from)) and evalue(from, L) for a backward
 * one ([to, L)) (src/SMEM.cpp:26-40, :81-82, :136-137); [i0, i1) is the running search's
 * fixed range, whose prefix-sum difference startSearch keeps in l.gD */
template <class Row>
__device__ __forceinline__ bool gatePush(const Ctx<Row>& c, const Lane<Row>& l, double ll0, int i0, int i1, double lA,
                                         double lC, double lG, double lT) {
    if (!c.a.gate1) return true;
    if (l.P) {
        const double v = ll0 + l.gD, sl = gateSlack(l, ll0);
        if (v < c.a.T1 - sl) return true;
        if (v > c.a.T1 + sl) return false;
    }
    double x = 0;
    x += ll0;
    return gateSum(c, l, x, i0, i1, c.a.T1, lA, lC, lG, lT);
}
template <class Row>
__device__ __forceinline__ bool gateFwd(const Ctx<Row>& c, const Lane<Row>& l, double ll0, int from, double lA, double lC,
                                        double lG, double lT) {
    return gatePush(c, l, ll0, 0, from, lA, lC, lG, lT);
}
template <class Row>
__device__ __forceinline__ bool gateBack(const Ctx<Row>& c, const Lane<Row>& l, double ll0, int to, double lA, double lC,
                                         double lG, double lT) {
    return gatePush(c, l, ll0, to, l.L, lA, lC, lG, lT);
}

__device__ __forceinline__ uint8_t byteAt(uint64_t w0, uint64_t w1, uint64_t w2, int64_t o) {
    const uint64_t w = o < 8 ? w0 : (o < 16 ? w1 : w2);
    return (uint8_t)(w >> (8 * (o & 7)));
}

constexpr int kRun = 16;

/*
 * Number (<= kRun) of consecutive successful unique-interval steps from text position X:
 * forward steps compare the complement of read base r0 + t with BD[X-1-t], backward
 * steps read base r0 - 1 - t with BD[X-1-t]; a step needs an A/C/G/T base inside the
 * read and X - t > 0.  Text and read bytes come from three aligned 8-byte words each
 * (both buffers are padded past their ends).
 */
template <class Row>
__device__ __forceinline__ int matchRun(const DevIndex& ix, const SmemArgs& a, const Lane<Row>& l, int64_t X, int r0,
                                        bool fwd, double& LL, double lA, double lC, double lG, double lT) {
    if (X <= 0) return 0;
    const int64_t tb = X > kRun ? X - kRun : 0;
    const int64_t tA = tb & ~(int64_t)7;
    const uint64_t* tw = (const uint64_t*)(ix.bd + tA);
    const uint64_t t0 = tw[0], t1 = tw[1], t2 = tw[2];
    const int64_t off = l.s - a.seqs;
    const int64_t rb = fwd ? off + r0 : off + (r0 > kRun ? r0 - kRun : 0);
    const int64_t rA = rb & ~(int64_t)7;
    const uint64_t* rw = (const uint64_t*)(a.seqs + rA);
    const uint64_t q0 = rw[0], q1 = rw[1], q2 = rw[2];
    int t = 0;
    for (; t < kRun; ++t) {
        if (X - t <= 0) break;
        const int rp = fwd ? r0 + t : r0 - 1 - t;
        if (rp < 0 || rp >= l.L) break;
        const uint8_t b = byteAt(q0, q1, q2, off + rp - rA);
        if (amb(b)) break;
        const uint8_t tx = byteAt(t0, t1, t2, X - 1 - t - tA);
        if (tx != (fwd ? comp16(b) : b)) break;
        LL += b == 1 ? lA : (b == 2 ? lC : (b == 4 ? lG : lT));
    }
    return t;
}

/* bytes [a, a+8) and [a+8, a+16) of a padded buffer, from three aligned words */
__device__ __forceinline__ void load16(const uint8_t* buf, int64_t a, uint64_t& u0, uint64_t& u1) {
    const uint64_t* w = (const uint64_t*)(buf + (a & ~(int64_t)7));
    const uint64_t w0 = w[0], w1 = w[1], w2 = w[2];
    const uint32_t sh = 8u * (uint32_t)(a & 7);
    u0 = sh ? (w0 >> sh) | (w1 << (64 - sh)) : w0;
    u1 = sh ? (w1 >> sh) | (w2 << (64 - sh)) : w1;
}

/* per byte: complement of a basic nt16 code (bit reversal of the low nibble) */
__device__ __forceinline__ uint64_t comp8(uint64_t v) {
    return ((v & 0x0101010101010101ull) << 3) | ((v & 0x0202020202020202ull) << 1) |
           ((v & 0x0404040404040404ull) >> 1) | ((v & 0x0808080808080808ull) >> 3);
}

/* leading run (<= kRun) of matching steps, 8 bases per word: text byte X-1-t against
 * read base r0+t (forward, complemented) or r0-1-t (backward); same steps and LL sums as
 * matchRun.  Falls back to matchRun near the start of the text or of the read buffer. */
template <class Row>
__device__ __forceinline__ int matchRunW(const DevIndex& ix, const SmemArgs& a, const Lane<Row>& l, int64_t X, int r0,
                                         bool fwd, double& LL, double lA, double lC, double lG, double lT) {
    const int64_t off = l.s - a.seqs;
    if (X < kRun || (!fwd && off + r0 < kRun)) return matchRun(ix, a, l, X, r0, fwd, LL, lA, lC, lG, lT);
    uint64_t t0, t1, q0, q1;
    load16(ix.bd, X - kRun, t0, t1);
    /* text in step order: t = 0 is the byte at X-1 */
    const uint64_t T[2] = {__builtin_bswap64(t1), __builtin_bswap64(t0)};
    uint64_t R[2];
    if (fwd) {
        load16(a.seqs, off + r0, q0, q1);
        R[0] = q0; R[1] = q1;
    } else {
        load16(a.seqs, off + r0 - kRun, q0, q1);
        R[0] = __builtin_bswap64(q1); R[1] = __builtin_bswap64(q0);
    }
    /* steps allowed by the read ends and the text start */
    const int64_t lim = min((int64_t)kRun, min(fwd ? (int64_t)(l.L - r0) : (int64_t)r0, X));
    int m = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (m == 8 * h) {
            const uint64_t v = R[h];
            const uint64_t bad = notBasic8(v) | (T[h] ^ (fwd ? comp8(v) : v));
            const int run = bad ? (__builtin_ctzll(bad) >> 3) : 8;
            m += run;
        }
    }
    if (m > lim) m = (int)lim;
    for (int t = 0; t < m; ++t) {
        const uint32_t b = (uint32_t)((t < 8 ? R[0] : R[1]) >> (8 * (t & 7))) & 0xffu;
        LL += b == 1 ? lA : (b == 2 ? lC : (b == 4 ? lG : lT));
    }
    return m;
}

template <class Row>
__device__ __forceinline__ void startIteration(Lane<Row>& l) {
    l.k = l.nCand - 1;
    l.w = l.nCand - 1;
    l.s0 = -1;
    l.nValid = 0;
}

/* SMEM::fwdExt / SMEM::backExt on a bi-interval: forward steps are backward steps of the
 * reverse complement with the roles of p and q swapped (src/FMDIndex.h:226-228).  One
 * inlined extension serves both directions, so a wavefront whose lanes go different
 * ways runs the rank code once. */
template <class Row>
__device__ __forceinline__ void extendOne(const DevIndex& ix, bool fwd, Row& P, Row& Q, Row& S, uint8_t b) {
    Row x = fwd ? Q : P, y = fwd ? P : Q;
    dBackExtT<Row>(ix, x, y, S, fwd ? comp16(b) : b);
    P = fwd ? y : x;
    Q = fwd ? x : y;
}

/* fetch reads until one has a running search; false when the queue is empty */
template <class Row>
__device__ bool nextRead(const Ctx<Row>& c, Lane<Row>& l) {
    while (true) {
        int64_t rid = (int64_t)atomicAdd(c.a.queue, 1ull) + c.qBase;
        if (rid >= c.qEnd) return false;
        if (c.a.order) rid = c.a.order[rid];
        l.rid = rid;
        const int64_t off = c.a.offs[rid];
        l.s = c.a.seqs + off;
        l.L = (int)(c.a.offs[rid + 1] - off);
        l.P = c.a.prefLL ? c.a.prefLL + off + rid : nullptr;
        l.nOut = 0;
        l.tag = 0;
        l.hand = kNoHand;
        const uint8_t nd = c.a.need ? c.a.need[rid] : 1;
        if (l.L > 0 && (nd & 1)) {
            if (l.P) l.gPL = fabs(l.P[l.L]);
            l.dFrom = l.L / 2;
            l.dTo = l.dFrom + 1;
            l.stage = STG_A;
            if (startSearch(c, l, l.dFrom, true)) return true;
            driverAfter(c, l);
            if (l.role == 1) {
                /* the odd lane takes the read's backward searches whatever is left here */
                if (l.state == ST_IDLE) c.a.cnt[rid] = l.nOut;
                return true;
            }
            if (l.state != ST_IDLE) return true;
        }
        if ((nd & 3) == 2) {
            c.a.out[rid * c.a.cap] = c.a.exact[rid];
            if (l.role) c.a.tagA[rid * c.a.cap] = 0;
            l.nOut = 1;
        }
        c.a.cnt[rid] = l.nOut;
    }
}

/* a pair's odd lane takes read rid (fetched by its even lane) and waits for the hand-over */
template <class Row>
__device__ void pairTake(const Ctx<Row>& c, Lane<Row>& l, int64_t rid) {
    l.rid = rid;
    const int64_t off = c.a.offs[rid];
    l.s = c.a.seqs + off;
    l.L = (int)(c.a.offs[rid + 1] - off);
    l.P = c.a.prefLL ? c.a.prefLL + off + rid : nullptr;
    if (l.P) l.gPL = fabs(l.P[l.L]);
    l.nOut = 0;
    l.tag = 0;
    l.state = ST_WAIT;
}

}  // namespace

/*
 * FAST selects the unique-interval path: while an interval has size 1, a backward step
 * by b succeeds iff the text byte before the suffix is b (BWT[p] = BD[SA[p]-1]), and
 * then p becomes LF(p) = ISA[SA[p]-1] while q is unchanged (src/FMDIndex.cpp:577-610
 * with s = 1: no split, so no q update); forward steps act on q the same way
 * (FMDIndex::fwdExt).  The moving side is carried as its text position X and turned
 * back into a row (ISA[X]) only when the SMEM is pushed or emitted.  Values are
 * identical to the occ-block path; only the memory traffic differs.
 */
template <bool FAST, bool PAIR, class Row>
__device__ __forceinline__ void smemBody(const DevIndex& ix, const SmemArgs& a) {
    __shared__ CandSlots<FAST, Row> slots;
    Lane<Row> l;
    l.nExt = 0; l.nFast = 0; l.nLook = 0;
    l.state = ST_IDLE;
    l.tag = 0;
    l.hand = kNoHand;
    Ctx<Row> c{ix, a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, false, 0, 0};
    l.role = PAIR ? ((c.lane & 1) ? 2 : 1) : 0;
    queueRange(a, c.qBase, c.qEnd);
    const int64_t minLen = a.minLen;
    const double llA = ix.ll[1], llC = ix.ll[2], llG = ix.ll[4], llT = ix.ll[8];
    uint32_t nIter = 0;
    /* the candidate of the next phase-2 step, loaded one step ahead (valid when nxOk) */
    int nxPos = 0;
    Row nxF = 0, nxR = 0, nxS = 0, nxX = -1;
    double nxLL = 0;
    bool nxOk = false;
    while (true) {
        if (PAIR) {
            /* both lanes of a pair run this exchange in every iteration: the even lane fetches
             * when both are free, then passes the read and the first search's `from` across */
            const int free = l.state == ST_IDLE ? 1 : 0;
            const int pFree = __shfl_xor(free, 1, 64);
            int got = -2;   /* -2: no fetch, -1: queue empty, else the fetched read */
            if (l.role == 1 && free && pFree) {
                got = nextRead(c, l) ? (int)l.rid : -1;
                nxOk = false;
            }
            const int gotIn = __shfl_xor(got, 1, 64);
            const int handIn = __shfl_xor(l.hand, 1, 64);
            if (l.role == 1) {
                if (got == -1) break;
                l.hand = kNoHand;
            } else {
                if (gotIn == -1) break;
                if (gotIn >= 0) pairTake(c, l, gotIn);
                if (l.state == ST_WAIT && handIn != kNoHand) {
                    l.dFrom = handIn;
                    pairBackChain(c, l);
                    nxOk = false;
                    if (l.state == ST_IDLE) a.cntB[l.rid] = l.nOut;
                }
            }
            if (l.state == ST_IDLE || l.state == ST_WAIT) continue;
        } else if (l.state == ST_IDLE) {
            if (!nextRead(c, l)) break;
            nxOk = false;
        }
        ++nIter;
        /* ---- operands of this step's extension ---- */
        const int st = l.state;
        const bool phase1 = (st == ST_FB1 || st == ST_BF1);
        const bool fwd = (st == ST_FB1 || st == ST_BF2);
        /* the base of this step comes from the lane's own state (phase 2: the pass's
         * boundary), so its load need not wait for the candidate's */
        uint8_t b = 0;
        if (!FAST) {
            const int bp = fwd ? (phase1 ? l.cTo : l.sTo) : (phase1 ? l.cFrom : l.sFrom) - 1;
            b = bp >= 0 ? base(l, bp) : (uint8_t)0;
        }
        int pos = 0;
        Row P, Q, S, X;
        double LL;
        int mFrom, mTo;
        if (phase1) {
            P = l.cF; Q = l.cR; S = l.cS; X = l.cX; LL = l.cLL; mFrom = l.cFrom; mTo = l.cTo;
        } else {
            if (nxOk) {
                pos = nxPos; P = nxF; Q = nxR; S = nxS; LL = nxLL; X = FAST ? nxX : -1;
            } else {
                const SmemCandT<Row> cd = laneLoad(c, slots, l.k, l.nCand);
                pos = cd.pos; P = cd.F; Q = cd.R; S = cd.S; LL = cd.LL;
                X = FAST ? cd.X : -1;
            }
            if (st == ST_FB2) { mFrom = l.sFrom; mTo = pos; } else { mFrom = pos; mTo = l.sTo; }
            /* load the next step's candidate now, beside this step's occ blocks: the next
             * entry of this pass (this step stores at index w >= k, above it), or, at the
             * pass's last entry, the first entry of the next pass when an earlier step of
             * this pass stored it */
            nxOk = l.k > l.lo || l.w < l.nCand - 1;
            if (nxOk) {
                const SmemCandT<Row> cd = laneLoad(c, slots, l.k > l.lo ? l.k - 1 : l.nCand - 1, l.nCand);
                nxPos = cd.pos; nxF = cd.F; nxR = cd.R; nxS = cd.S; nxLL = cd.LL; nxX = cd.X;
            }
        }
        /* unique interval with no other candidate in play: run over matching text bytes;
         * each of those steps keeps S = 1 and so pushes, emits and drops nothing */
        if (FAST && S == 1 && (phase1 || (l.k == l.lo && l.lo == l.nCand - 1))) {
            if (X < 0) { X = saRow<Row>(ix, fwd ? Q : P); l.nLook++; }
            const int m = matchRunW(ix, a, l, X, fwd ? mTo : mFrom, fwd, LL, llA, llC, llG, llT);
            if (m > 0) {
                X -= m;
                if (fwd) mTo += m; else mFrom -= m;
                l.nFast += m;
                if (st == ST_FB1) { l.cTo = mTo; l.cLL = LL; l.cX = X; }
                else if (st == ST_BF1) { l.cFrom = mFrom; l.cLL = LL; l.cX = X; }
                else if (st == ST_FB2) l.sFrom = mFrom;
                else l.sTo = mTo;
            }
        }
        if (FAST) b = fwd ? base(l, mTo) : (mFrom > 0 ? base(l, mFrom - 1) : (uint8_t)0);
        Row F0 = P, R0 = Q;
        const Row S0 = S, X0 = X;
        const double LL0 = LL;
        /* ---- the extension (SMEM::fwdExt / SMEM::backExt, src/SMEM.h:140-177) ---- */
        if (FAST && S == 1) {
            if (X < 0) { X = saRow<Row>(ix, fwd ? Q : P); l.nLook++; }
            const uint8_t bb = fwd ? comp16(b) : b;
            if (!amb(b) && X > 0 && ix.bd[X - 1] == bb) { X--; l.nFast++; }
            else S = 0;
        } else {
            l.nExt += amb(b) ? 0 : 1;
            extendOne(ix, fwd, P, Q, S, b);
            X = -1;
        }
        if (fwd) { if (mTo < l.L) LL += llOf(c, b, llA, llC, llG, llT); mTo++; }
        else { if (mFrom > 0) LL += llOf(c, b, llA, llC, llG, llT); mFrom--; }
        /* the pre-step SMEM with its moving side turned back into a row */
        auto exact0 = [&]() {
            if (FAST && X0 >= 0) {
                const Row v = isaRow<Row>(ix, X0);
                l.nLook++;
                if (fwd) R0 = v; else F0 = v;
            }
        };

        /* ---- control ----
         * (one copy of each step for both directions: lanes of a wavefront in different
         * states then share the code that runs) */
        bool after = false;   /* the search returned: findAllSMEMS goes on */
        if (phase1) {
            /* push the pre-step SMEM when its size changes (src/SMEM.cpp:77-83, :130-138) */
            const int len = fwd ? l.cTo : (l.L - l.cFrom);
            nxOk = false;
            if (S != S0 && len >= minLen &&
                gatePush(c, l, LL0, fwd ? 0 : l.cTo, fwd ? l.cFrom : l.L, llA, llC, llG, llT)) {
                exact0();
                const int pp = fwd ? l.cTo : l.cFrom;
                lanePush(c, slots, l.nCand++, pp, F0, R0, S0, LL0, (Row)-1);
                /* the last push heads the phase-2 list if phase 1 ends here */
                nxPos = pp; nxF = F0; nxR = R0; nxS = S0; nxLL = LL0; nxX = -1;
                nxOk = true;
            }
            if (fwd) l.cTo = mTo; else l.cFrom = mFrom;
            l.cF = P; l.cR = Q; l.cS = S; l.cLL = LL; l.cX = X;
            if (!(S >= 1)) {
                if (fwd) l.sTo = mTo; else l.sFrom = mFrom;
                if (l.nCand == 0) {
                    if (fwd) l.sFrom--; else l.sTo++;
                    after = true;
                } else {
                    l.state = fwd ? ST_FB2 : ST_BF2;
                    l.lo = 0;
                    startIteration(l);
                }
            }
        } else {
            const bool fb = (st == ST_FB2);
            const int len0 = fb ? (pos - l.sFrom) : (l.sTo - pos);
            if (S != S0 && S0 >= 1 && l.nValid == 0 && len0 >= minLen && LL0 <= a.T2) {
                exact0();
                emit(c, l, fb ? l.sFrom : pos, fb ? pos : l.sTo, F0, R0, S0, LL0);
            }
            if (S >= 1 && S != l.s0) {
                l.s0 = S;
                l.nValid++;
                /* a store at the top of the list heads the next pass (no earlier step of this
                 * pass stored, so nothing was loaded ahead for it) */
                if (l.w == l.nCand - 1 && l.k == l.lo) {
                    nxPos = pos; nxF = P; nxR = Q; nxS = S; nxLL = LL; nxX = X;
                    nxOk = true;
                }
                laneStore(c, slots, l.w--, l.nCand, pos, P, Q, S, LL, X);
            }
            l.k--;
            if (l.k < l.lo) {
                l.lo = l.w + 1;
                bool more;
                if (fb) { l.sFrom--; more = l.sFrom >= 0; }
                else { l.sTo++; more = l.sTo <= l.L; }
                more = more && (l.lo <= l.nCand - 1);
                if (more) startIteration(l);
                else after = true;
            }
        }
        if (after) { driverAfter(c, l); nxOk = false; }
        if (l.state == ST_IDLE) (l.role == 2 ? a.cntB : a.cnt)[l.rid] = l.nOut;
    }
    if (a.stats) {
        atomicAdd(&a.stats[0], (unsigned long long)l.nExt);
        atomicAdd(&a.stats[2], (unsigned long long)l.nFast);
        atomicAdd(&a.stats[3], (unsigned long long)l.nLook);
        /* lane iterations, and the wavefront's (its busiest lane's) iterations */
        uint32_t w = nIter;
        for (int off = 32; off >= 1; off >>= 1) w = max(w, (uint32_t)__shfl_xor((int)w, off, 64));
        atomicAdd(&a.stats[5], (unsigned long long)nIter);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&a.stats[6], (unsigned long long)w);
            atomicMax(&a.stats[7], (unsigned long long)w);
        }
    }
}


/* 4 waves per SIMD (128 VGPRs) for 32-bit rows; the 64-bit rows of wide indexes need more
 * registers than that (they spilled 52-124 bytes per lane to scratch, which the counters
 * showed as ~13 GB of writes per launch on a 6.8e9-symbol index), so 3 waves per SIMD */
template <class Row> constexpr int smemWaves() { return sizeof(Row) == 8 ? 3 : 4; }

template <bool FAST, class Row>
__global__ void __launch_bounds__(256, smemWaves<Row>()) smem_kernel_t(DevIndex ix, SmemArgs a) {
    smemBody<FAST, false, Row>(ix, a);
}

template <bool FAST, class Row>
__global__ void __launch_bounds__(256, smemWaves<Row>()) smem_pair_kernel(DevIndex ix, SmemArgs a) {
    smemBody<FAST, true, Row>(ix, a);
}

template __global__ void smem_kernel_t<false, int32_t>(DevIndex, SmemArgs);
template __global__ void smem_kernel_t<true, int32_t>(DevIndex, SmemArgs);
template __global__ void smem_pair_kernel<false, int32_t>(DevIndex, SmemArgs);
template __global__ void smem_pair_kernel<true, int32_t>(DevIndex, SmemArgs);
template __global__ void smem_kernel_t<false, int64_t>(DevIndex, SmemArgs);
template __global__ void smem_kernel_t<true, int64_t>(DevIndex, SmemArgs);
template __global__ void smem_pair_kernel<false, int64_t>(DevIndex, SmemArgs);
template __global__ void smem_pair_kernel<true, int64_t>(DevIndex, SmemArgs);

/* one read per lane: the even lane's list (tags non-decreasing) and the odd lane's list (tags
 * from 1, non-decreasing) merged by tag, the even lane's SMEMs first on equal tags: the first
 * search, then the k-th forward and the k-th backward search of findAllSMEMS's loop
 * (src/SMEM.cpp:190-203) */
__global__ void __launch_bounds__(256) smem_merge_kernel(SmemArgs a, SmemRec* out, int32_t* cnt) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < a.nReads;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t nA = a.cnt[r], nB = a.cntB[r];
        const int32_t total = nA + nB;
        cnt[r] = total;
        if (total > a.cap) continue;   /* the host grows the slots and searches again */
        const size_t base = (size_t)r * a.cap;
        int32_t i = 0, j = 0, o = 0;
        while (i < nA || j < nB) {
            const bool takeA = j >= nB || (i < nA && a.tagA[base + i] <= a.tagB[base + j]);
            out[base + o++] = takeA ? a.out[base + i++] : a.outB[base + j++];
        }
    }
}

/* per read, P[0] = 0 and P[i+1] = P[i] + loglik(base i) (the gates' estimates) */
__global__ void __launch_bounds__(256) smem_prefix_kernel(DevIndex ix, SmemArgs a, double* pref) {
    /* one read per lane, summed in read order over chunks of 16 bases loaded as two 8-byte
     * words (the read buffer is padded); each chunk's 16 sums per read are staged in LDS and
     * stored by 16 lanes at a time, so that stores fill whole 128-byte lines */
    constexpr int kChunk = 16;
    __shared__ double sv[4][64][kChunk + 1];
    __shared__ int64_t sBase[4][64];
    __shared__ int32_t sLen[4][64];
    __shared__ double sll[16];
    if (threadIdx.x < 16) sll[threadIdx.x] = ix.ll[threadIdx.x];
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t r0 = ((int64_t)blockIdx.x * 4 + w) * 64; r0 < a.nReads; r0 += (int64_t)gridDim.x * 4 * 64) {
        const int64_t r = r0 + lane;
        /* reads the search skips (form 2: exact or filtered) need no sums */
        const bool mine = r < a.nReads && (!a.need || (a.need[r] & 1));
        const int64_t o = mine ? a.offs[r] : 0;
        const int32_t L = mine ? (int32_t)(a.offs[r + 1] - o) : -1;
        sBase[w][lane] = o + r;
        sLen[w][lane] = L;
        if (mine) pref[o + r] = 0.0;   /* P[0] */
        int32_t Lmax = L;
        for (int d = 32; d >= 1; d >>= 1) Lmax = max(Lmax, __shfl_xor(Lmax, d, 64));
        double x = 0;
        for (int32_t k0 = 0; k0 < Lmax; k0 += kChunk) {
            /* bases k0 .. k0 + 15 of the read: sums P[k0 + 1] .. P[k0 + 16] */
            uint64_t q[2] = {0, 0};
            if (k0 < L) load16(a.seqs, o + k0, q[0], q[1]);
#pragma unroll
            for (int t = 0; t < kChunk; ++t) {
                const uint32_t b = (uint32_t)(q[t >> 3] >> (8 * (t & 7))) & 15u;
                if (k0 + t < L) x += sll[b];
                sv[w][lane][t] = x;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int j = lane; j < 64 * kChunk; j += 64) {
                const int rr = j / kChunk, t = j % kChunk;
                const int32_t k = k0 + t + 1;
                if (k <= sLen[w][rr]) pref[sBase[w][rr] + k] = sv[w][rr][t];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

/* ------------------------------------------------------------------ one read per lane group */

namespace {

/* a group's shuffle source: lane `src` of the wavefront (all lanes of the group active) */
__device__ __forceinline__ int64_t shfl64(int64_t v, int src) {
    const int lo = __shfl((int)(uint32_t)(uint64_t)v, src, 64);
    const int hi = __shfl((int)(uint32_t)((uint64_t)v >> 32), src, 64);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

/* the group leader takes reads from the queue until one has a running search; the group
 * follows with the same values (false when the queue is empty) */
template <int G, class Row>
__device__ bool nextReadGroup(const Ctx<Row>& c, Lane<Row>& l, int gbase, int gl) {
    while (true) {
        int64_t rid = 0;
        if (gl == 0) rid = (int64_t)atomicAdd(c.a.queue, 1ull) + c.qBase;
        rid = shfl64(rid, gbase);
        if (rid >= c.qEnd) return false;
        if (c.a.order) rid = c.a.order[rid];
        l.rid = rid;
        const int64_t off = c.a.offs[rid];
        l.s = c.a.seqs + off;
        l.L = (int)(c.a.offs[rid + 1] - off);
        l.P = c.a.prefLL ? c.a.prefLL + off + rid : nullptr;
        l.nOut = 0;
        const uint8_t nd = c.a.need ? c.a.need[rid] : 1;
        if (l.L > 0 && (nd & 1)) {
            if (l.P) l.gPL = fabs(l.P[l.L]);
            l.dFrom = l.L / 2;
            l.dTo = l.dFrom + 1;
            l.stage = STG_A;
            if (startSearch(c, l, l.dFrom, true)) return true;
            driverAfter(c, l);
            if (l.state != ST_IDLE) return true;
        }
        if ((nd & 3) == 2) {
            if (gl == 0) c.a.out[rid * c.a.cap] = c.a.exact[rid];
            l.nOut = 1;
        }
        if (gl == 0) c.a.cnt[rid] = l.nOut;
    }
}

}  // namespace

/* lane `dst` of the wavefront receives v (a permutation within the active lanes) */
__device__ __forceinline__ int permute32(int dst, int v) { return __builtin_amdgcn_ds_permute(dst << 2, v); }
__device__ __forceinline__ int64_t permute64(int dst, int64_t v) {
    const int lo = permute32(dst, (int)(uint32_t)(uint64_t)v);
    const int hi = permute32(dst, (int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ double permuteF64(int dst, double v) {
    return __longlong_as_double(permute64(dst, __double_as_longlong(v)));
}
/* rows and sizes of either width */
template <class Row> __device__ __forceinline__ Row permuteRow(int dst, Row v) {
    if constexpr (sizeof(Row) == 8) return permute64(dst, v);
    else return permute32(dst, v);
}
template <class Row> __device__ __forceinline__ Row shflRow(Row v, int src) {
    if constexpr (sizeof(Row) == 8) return shfl64(v, src);
    else return __shfl(v, src, 64);
}

/*
 * The SMEM search with one read per group of G lanes (64 / G reads per wavefront).  The
 * read's state is uniform across its group.  A phase-1 step (the growing SMEM) is one
 * extension that all lanes of the group compute alike; a phase-2 step extends up to G
 * candidates of the lock-step list at once, lane i taking list entry i.  The reference's
 * sequential rules for one pass (src/SMEM.cpp:93-113, 146-166) become per-candidate tests
 * because sizes along the list are non-decreasing (longer strings have no more
 * occurrences, and all candidates extend by the same base):
 *   keep_i = S'_i >= 1 and S'_i != S'_{i-1}          (the reference's s0 test)
 *   emit_i = S'_i != S_i and S_i >= 1 and S'_{i-1} < 1 (nValid == 0) and len, e-value,
 * with S'_{-1} = -1.  Emitted SMEMs keep the list order (ballot prefix counts) and kept
 * candidates move to the front in order, as erase(remove_if) does.  Lists of at most G
 * entries live in the group's registers (lane i holds entry i; compaction is a lane
 * permutation); longer lists live in the per-group scratch and are processed G entries
 * per step, each chunk continuing from the last S' of the previous one.  A lone unique
 * candidate walks the text as in the per-lane kernel.  Output equals smem_kernel_t's.
 */
template <int G, class Row>
__global__ void __launch_bounds__(256) smem_group_kernel(DevIndex ix, SmemArgs a) {
    static_assert(G == 4 || G == 8 || G == 16 || G == 32, "group size");
    const int lane = threadIdx.x & 63;
    const int gl = lane % G, gbase = lane - gl;
    const uint64_t gmask = ((1ull << G) - 1) << gbase;
    const uint64_t below = (1ull << lane) - 1;
    Lane<Row> l;
    l.nExt = 0; l.nFast = 0; l.nLook = 0;
    l.state = ST_IDLE;
    l.role = 0; l.tag = 0; l.hand = kNoHand;
    l.nCand = 0; l.lo = 0; l.k = 0; l.w = 0; l.s0 = -1; l.nValid = 0;
    /* this lane's list entry in register mode (l.nValid == 1) or its phase-1 push */
    int rPos = 0;
    Row rP = 0, rQ = 0, rS = 0, rX = -1;
    double rLL = 0;
    Ctx<Row> c{ix, a, ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G, true, 0, 0};
    queueRange(a, c.qBase, c.qEnd);
    const int64_t minLen = a.minLen;
    const double llA = ix.ll[1], llC = ix.ll[2], llG = ix.ll[4], llT = ix.ll[8];
    uint32_t nIter = 0;
    while (true) {
        if (l.state == ST_IDLE) {
            if (!nextReadGroup<G, Row>(c, l, gbase, gl)) break;
        }
        ++nIter;
        const int st = l.state;
        if (st == ST_FB1 || st == ST_BF1) {
            /* ---- phase 1: one extension, alike in every lane of the group ---- */
            Row P = l.cF, Q = l.cR, S = l.cS, X = l.cX;
            double LL = l.cLL;
            int mFrom = l.cFrom, mTo = l.cTo;
            const bool fwd = st == ST_FB1;
            if (S == 1) {
                if (X < 0) { X = saRow<Row>(ix, fwd ? Q : P); if (gl == 0) l.nLook++; }
                const int m = matchRunW(ix, a, l, X, fwd ? mTo : mFrom, fwd, LL, llA, llC, llG, llT);
                if (m > 0) {
                    X -= m;
                    if (fwd) mTo += m; else mFrom -= m;
                    if (gl == 0) l.nFast += m;
                    if (fwd) l.cTo = mTo; else l.cFrom = mFrom;
                    l.cLL = LL; l.cX = X;
                }
            }
            const uint8_t b = fwd ? base(l, mTo) : (mFrom > 0 ? base(l, mFrom - 1) : (uint8_t)0);
            Row F0 = P, R0 = Q;
            const Row S0 = S, X0 = X;
            const double LL0 = LL;
            if (S == 1) {
                if (X < 0) { X = saRow<Row>(ix, fwd ? Q : P); if (gl == 0) l.nLook++; }
                const uint8_t bb = fwd ? comp16(b) : b;
                if (!amb(b) && X > 0 && ix.bd[X - 1] == bb) { X--; if (gl == 0) l.nFast++; }
                else S = 0;
            } else {
                if (gl == 0) l.nExt += amb(b) ? 0 : 1;
                extendOne(ix, fwd, P, Q, S, b);
                X = -1;
            }
            if (fwd) { if (mTo < l.L) LL += llOf(c, b, llA, llC, llG, llT); mTo++; }
            else { if (mFrom > 0) LL += llOf(c, b, llA, llC, llG, llT); mFrom--; }
            const bool push = fwd ? (S != S0 && l.cTo >= minLen && gateFwd(c, l, LL0, l.cFrom, llA, llC, llG, llT))
                                  : (S != S0 && (int64_t)(l.L - l.cFrom) >= minLen &&
                                     gateBack(c, l, LL0, l.cTo, llA, llC, llG, llT));
            if (push) {
                if (X0 >= 0) {
                    const Row v = isaRow<Row>(ix, X0);
                    if (gl == 0) l.nLook++;
                    if (fwd) R0 = v; else F0 = v;
                }
                const int pp = fwd ? l.cTo : l.cFrom;
                if (l.nCand < G) {
                    if (gl == l.nCand) { rPos = pp; rP = F0; rQ = R0; rS = S0; rX = -1; rLL = LL0; }
                } else if (gl == 0) {
                    candStore<true>(c, l.nCand, pp, F0, R0, S0, LL0, (Row)-1);
                }
                l.nCand++;
            }
            if (fwd) l.cTo = mTo; else l.cFrom = mFrom;
            l.cF = P; l.cR = Q; l.cS = S; l.cLL = LL; l.cX = X;
            if (!(S >= 1)) {
                if (fwd) l.sTo = mTo; else l.sFrom = mFrom;
                if (l.nCand == 0) {
                    if (fwd) l.sFrom--; else l.sTo++;
                    driverAfter(c, l);
                } else {
                    l.state = fwd ? ST_FB2 : ST_BF2;
                    l.lo = l.nCand - 1;   /* slot of list entry 0 (the last push) */
                    l.k = 0; l.w = 0; l.s0 = -1;
                    if (l.nCand <= G) {
                        /* registers: lane i takes push nCand-1-i (list order) */
                        const int src = gbase + max(0, l.nCand - 1 - gl);
                        rPos = __shfl(rPos, src, 64);
                        rP = shflRow(rP, src); rQ = shflRow(rQ, src); rS = shflRow(rS, src); rX = shflRow(rX, src);
                        rLL = __longlong_as_double(shfl64(__double_as_longlong(rLL), src));
                        l.nValid = 1;
                    } else {
                        /* scratch: the first G pushes join the rest */
                        candStore<true>(c, gl, rPos, rP, rQ, rS, rLL, rX);
                        l.nValid = 0;
                    }
                }
            }
        } else {
            /* ---- phase 2: up to G candidates of the lock-step list, one per lane ---- */
            const bool fb = st == ST_FB2;
            const bool fwd = !fb;
            const bool regs = l.nValid == 1;
            const int i = l.k + gl;
            const bool act = i < l.nCand;
            int pos = 0;
            Row P = 0, Q = 0, S = 0, X = -1;
            double LL = 0;
            if (regs) {
                if (act) { pos = rPos; P = rP; Q = rQ; S = rS; X = rX; LL = rLL; }
            } else if (act) {
                const SmemCandT<Row> cd = candAt<true>(c, l.lo - i);
                pos = cd.pos; P = cd.F; Q = cd.R; S = cd.S; X = cd.X; LL = cd.LL;
            }
            /* a lone unique candidate walks the text; the whole group follows its boundary */
            if (l.nCand == 1 && l.k == 0) {
                int m = 0;
                if (gl == 0 && S == 1) {
                    if (X < 0) { X = saRow<Row>(ix, fwd ? Q : P); l.nLook++; }
                    m = matchRunW(ix, a, l, X, fwd ? l.sTo : l.sFrom, fwd, LL, llA, llC, llG, llT);
                    X -= m;
                    l.nFast += m;
                }
                m = __shfl(m, gbase, 64);
                if (fb) l.sFrom -= m; else l.sTo += m;
            }
            int mFrom = fb ? l.sFrom : pos, mTo = fb ? pos : l.sTo;
            const uint8_t b = fwd ? base(l, mTo) : (mFrom > 0 ? base(l, mFrom - 1) : (uint8_t)0);
            Row F0 = P, R0 = Q;
            const Row S0 = S, X0 = X;
            const double LL0 = LL;
            if (act) {
                if (S == 1) {
                    if (X < 0) { X = saRow<Row>(ix, fwd ? Q : P); l.nLook++; }
                    const uint8_t bb = fwd ? comp16(b) : b;
                    if (!amb(b) && X > 0 && ix.bd[X - 1] == bb) { X--; l.nFast++; }
                    else S = 0;
                } else {
                    l.nExt += amb(b) ? 0 : 1;
                    extendOne(ix, fwd, P, Q, S, b);
                    X = -1;
                }
                if (fwd) { if (mTo < l.L) LL += llOf(c, b, llA, llC, llG, llT); }
                else { if (mFrom > 0) LL += llOf(c, b, llA, llC, llG, llT); }
            }
            /* S' of the previous list entry */
            const Row up = shflRow(S, gl == 0 ? lane : lane - 1);
            const int64_t prev = gl == 0 ? (int64_t)l.s0 : up;
            const int len0 = fb ? (pos - l.sFrom) : (l.sTo - pos);
            const bool keep = act && S >= 1 && S != prev;
            const bool em = act && S != S0 && S0 >= 1 && prev < 1 && len0 >= minLen && LL0 <= a.T2;
            const uint64_t kMask = __ballot(keep) & gmask, eMask = __ballot(em) & gmask;
            if (em) {
                if (X0 >= 0) {
                    const Row v = isaRow<Row>(ix, X0);
                    l.nLook++;
                    if (fwd) R0 = v; else F0 = v;
                }
                const int slot = l.nOut + __popcll(eMask & below);
                if (slot < a.cap) {
                    SmemRec* o = a.out + (size_t)l.rid * a.cap + slot;
                    if (fb) { o->from = l.sFrom; o->to = pos; } else { o->from = pos; o->to = l.sTo; }
                    o->fwd = F0; o->rev = R0; o->size = S0; o->logP = LL0;
                }
            }
            const int nKeep = __popcll(kMask);
            if (regs) {
                /* compaction: kept entries to the front in order, the others behind them */
                const int rank = keep ? __popcll(kMask & below) : nKeep + __popcll(~kMask & gmask & below);
                const int dst = gbase + rank;
                rPos = permute32(dst, pos);
                rP = permuteRow(dst, P); rQ = permuteRow(dst, Q); rS = permuteRow(dst, S); rX = permuteRow(dst, X);
                rLL = permuteF64(dst, LL);
            } else if (keep) {
                candStore<true>(c, l.lo - (l.w + __popcll(kMask & below)), pos, P, Q, S, LL, X);
            }
            l.nOut += __popcll(eMask);
            l.w += nKeep;
            const int last = min(G, l.nCand - l.k) - 1;
            l.s0 = shflRow(S, gbase + last);
            l.k += G;
            if (l.k >= l.nCand) {
                l.nCand = l.w;
                bool more;
                if (fb) { l.sFrom--; more = l.sFrom >= 0; }
                else { l.sTo++; more = l.sTo <= l.L; }
                more = more && l.nCand > 0;
                if (more) {
                    l.k = 0; l.w = 0; l.s0 = -1;
                    if (!regs && l.nCand <= G) {
                        /* the list now fits the group: load it once */
                        if (gl < l.nCand) {
                            const SmemCandT<Row> cd = candAt<true>(c, l.lo - gl);
                            rPos = cd.pos; rP = cd.F; rQ = cd.R; rS = cd.S; rX = cd.X; rLL = cd.LL;
                        }
                        l.nValid = 1;
                    }
                } else {
                    driverAfter(c, l);
                }
            }
        }
        if (l.state == ST_IDLE && gl == 0) a.cnt[l.rid] = l.nOut;
    }
    if (a.stats) {
        atomicAdd(&a.stats[0], (unsigned long long)l.nExt);
        atomicAdd(&a.stats[2], (unsigned long long)l.nFast);
        atomicAdd(&a.stats[3], (unsigned long long)l.nLook);
        uint32_t w = nIter;
        for (int off = 32; off >= 1; off >>= 1) w = max(w, (uint32_t)__shfl_xor((int)w, off, 64));
        if (gl == 0) atomicAdd(&a.stats[5], (unsigned long long)nIter);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&a.stats[6], (unsigned long long)w);
            atomicMax(&a.stats[7], (unsigned long long)w);
        }
    }
}

template __global__ void smem_group_kernel<4, int32_t>(DevIndex, SmemArgs);
template __global__ void smem_group_kernel<8, int32_t>(DevIndex, SmemArgs);
template __global__ void smem_group_kernel<16, int32_t>(DevIndex, SmemArgs);
template __global__ void smem_group_kernel<4, int64_t>(DevIndex, SmemArgs);
template __global__ void smem_group_kernel<8, int64_t>(DevIndex, SmemArgs);
template __global__ void smem_group_kernel<16, int64_t>(DevIndex, SmemArgs);

/*
 * SMEM_LIST::findMEMS (src/SMEM.cpp:174-184) with SMEM::findMEM (src/SMEM.cpp:42-62), one
 * lane per read, and the list's SMEM_LIST::loglik (src/SMEM.cpp:307-312).
 *
 * findMEM extends [from, from+1) forward until the size drops to 0 and keeps the last
 * extension that succeeded only for the next start (to = mem0.to); its backward loop then
 * starts from the failed MEM itself (mem0 = mem), whose size is 0, so it ends after one
 * step with mem0 unchanged: the MEM returned is the failed forward extension (size 0, end
 * one past the last success, logP including the failing base when it lies inside the
 * read), and from is unchanged.  An extension by any base keeps (p, q) when the new size
 * is 0 (FMDIndex::backExt, src/FMDIndex.cpp:577-610).
 */
template <class Row>
__global__ void __launch_bounds__(256) mem_kernel(DevIndex ix, SmemArgs a, double* listLL) {
    for (int64_t rid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; rid < a.nReads;
         rid += (int64_t)gridDim.x * blockDim.x) {
        const int64_t off = a.offs[rid];
        const uint8_t* s = a.seqs + off;
        const int L = (int)(a.offs[rid + 1] - off);
        int nOut = 0;
        double sum = 0;
        for (int from = 0, to = 1; from < L; from = to + 1) {
            const uint8_t b = s[from];
            Row P = (Row)ix.C[b], Q = (Row)ix.C[comp16(b)], S = (Row)(ix.C[b + 1] - ix.C[b]);
            double LL = 0;
            LL += ix.ll[b];
            int mTo = from + 1, to0 = mTo;
            while (true) {
                const uint8_t nb = mTo < L ? s[mTo] : (uint8_t)0;   /* seq[L] is the terminator */
                extendOne(ix, true, P, Q, S, nb);
                if (mTo < L) LL += ix.ll[nb];
                mTo++;
                if (S > 0) to0 = mTo;
                else break;
            }
            to = to0;
            if (mTo - from >= a.minLen && LL <= a.T2) {
                if (nOut < a.cap) {
                    SmemRec* o = a.out + (size_t)rid * a.cap + nOut;
                    o->from = from; o->to = mTo; o->fwd = P; o->rev = Q; o->size = S; o->logP = LL;
                }
                nOut++;
                sum += LL;
            }
        }
        a.cnt[rid] = nOut;
        listLL[rid] = sum;
    }
}

template __global__ void mem_kernel<int32_t>(DevIndex, SmemArgs, double*);
template __global__ void mem_kernel<int64_t>(DevIndex, SmemArgs, double*);
