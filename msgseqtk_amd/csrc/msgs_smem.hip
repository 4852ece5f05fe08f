/*
 * msgs_smem.hip -- batched SMEM search on the FMD index (SMEM_LIST::findAllSMEMS,
 * src/SMEM.cpp:186-215, with findFwdBackSMEMS :64-117 and findBackFwdSMEMS :119-172).
 *
 * One lane owns one read at a time and pulls the next read from a device work queue
 * when it finishes (persistent grid).  The reference's nested loops are unrolled into a
 * step machine: every loop iteration of a lane performs exactly one FMD extension, so
 * all lanes of a wavefront issue their occ-block loads at the same instruction no
 * matter which phase of which search they are in.  The lock-step candidate list of a
 * search (the reference's `prev`) lives in a per-lane scratch area interleaved across
 * lanes; it is traversed in reverse of its push order and compacted in place exactly
 * like the reference's reverse + erase(remove_if).
 *
 * Floating point: logP is accumulated in fp64 in the reference's order.  The e-value
 * gates N*exp(logP) <= maxE are evaluated as logP <= T, with T found on the host from
 * glibc exp (see msgs_capi.cpp: evalueThreshold); the extended-range sums of the first
 * gate add only non-positive terms, so they stop as soon as they pass T.
 */
#include "msgs_device.h"
#include "msgs_kernels.h"

namespace {

enum : int { ST_IDLE = 0, ST_FB1 = 1, ST_FB2 = 2, ST_BF1 = 3, ST_BF2 = 4 };
enum : int { STG_A = 0, STG_C = 1, STG_D = 2 };

struct Lane {
    /* read */
    int64_t rid;
    const uint8_t* s;
    int L;
    /* findAllSMEMS variables */
    int dFrom, dTo, stage;
    /* the running search: its by-reference from/to */
    int sFrom, sTo;
    int state;
    /* phase-1 SMEM */
    int cFrom, cTo;
    int64_t cF, cR, cS, cX;
    double cLL;
    /* candidate list */
    int nCand, lo, k, w, nValid;
    int64_t s0;
    /* output */
    int nOut;
    /* counters */
    int64_t nExt, nFast, nLook;
};

__device__ __forceinline__ uint8_t base(const Lane& l, int i) { return i < l.L ? l.s[i] : (uint8_t)0; }

struct Ctx {
    const DevIndex& ix;
    const SmemArgs& a;
    int64_t lane;
};

__device__ __forceinline__ void emit(const Ctx& c, Lane& l, int from, int to, int64_t F, int64_t R, int64_t S, double ll) {
    if (l.nOut < c.a.cap) {
        SmemRec* o = c.a.out + (size_t)l.rid * c.a.cap + l.nOut;
        o->from = from; o->to = to; o->fwd = F; o->rev = R; o->size = S; o->logP = ll;
    }
    l.nOut++;
}

__device__ __forceinline__ size_t cidx(const Ctx& c, int k) { return (size_t)k * c.a.nLanes + c.lane; }

__device__ __forceinline__ void candStore(const Ctx& c, int k, int pos, int64_t F, int64_t R, int64_t S, double ll,
                                          int64_t X) {
    size_t i = cidx(c, k);
    c.a.candPos[i] = pos; c.a.candF[i] = F; c.a.candR[i] = R; c.a.candS[i] = S; c.a.candLL[i] = ll;
    c.a.candX[i] = X;
}

/* start findFwdBackSMEMS / findBackFwdSMEMS at f; false if it ends at once (ambiguous base) */
__device__ __forceinline__ bool startSearch(const Ctx& c, Lane& l, int f, bool fb) {
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
    return true;
}

/* findAllSMEMS control flow after a search returned; ends with a running search or an idle lane */
__device__ void driverAfter(const Ctx& c, Lane& l) {
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

__device__ __forceinline__ bool gateFwd(const Ctx& c, const Lane& l, double ll0, int from) {
    /* evalue(0, to) <= maxE of the phase-1 candidate (src/SMEM.cpp:26-40,81-82) */
    if (!c.a.gate1) return true;
    double x = 0;
    x += ll0;
    for (int i = 0; i < from; ++i) {
        if (x <= c.a.T1) return true;
        x += c.ix.ll[l.s[i] & 15];
    }
    return x <= c.a.T1;
}
__device__ __forceinline__ bool gateBack(const Ctx& c, const Lane& l, double ll0, int to) {
    /* evalue(from, L) <= maxE (src/SMEM.cpp:136-137) */
    if (!c.a.gate1) return true;
    double x = 0;
    x += ll0;
    for (int i = to; i < l.L; ++i) {
        if (x <= c.a.T1) return true;
        x += c.ix.ll[l.s[i] & 15];
    }
    return x <= c.a.T1;
}

__device__ __forceinline__ void startIteration(Lane& l) {
    l.k = l.nCand - 1;
    l.w = l.nCand - 1;
    l.s0 = -1;
    l.nValid = 0;
}

/* fetch reads until one has a running search; false when the queue is empty */
__device__ bool nextRead(const Ctx& c, Lane& l) {
    while (true) {
        int64_t rid = (int64_t)atomicAdd(c.a.queue, 1ull);
        if (rid >= c.a.nReads) return false;
        l.rid = rid;
        const int64_t off = c.a.offs[rid];
        l.s = c.a.seqs + off;
        l.L = (int)(c.a.offs[rid + 1] - off);
        l.nOut = 0;
        if (l.L > 0 && (!c.a.need || c.a.need[rid])) {
            l.dFrom = l.L / 2;
            l.dTo = l.dFrom + 1;
            l.stage = STG_A;
            if (startSearch(c, l, l.dFrom, true)) return true;
            driverAfter(c, l);
            if (l.state != ST_IDLE) return true;
        }
        c.a.cnt[rid] = l.nOut;
    }
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
template <bool FAST>
__global__ void __launch_bounds__(256) smem_kernel_t(DevIndex ix, SmemArgs a) {
    Lane l;
    l.nExt = 0; l.nFast = 0; l.nLook = 0;
    l.state = ST_IDLE;
    Ctx c{ix, a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x};
    const int64_t minLen = a.minLen;
    while (true) {
        if (l.state == ST_IDLE) {
            if (!nextRead(c, l)) break;
        }
        /* ---- operands of this step's extension ---- */
        const int st = l.state;
        const bool phase1 = (st == ST_FB1 || st == ST_BF1);
        int pos = 0;
        int64_t P, Q, S, X;
        double LL;
        int mFrom, mTo;
        if (phase1) {
            P = l.cF; Q = l.cR; S = l.cS; X = l.cX; LL = l.cLL; mFrom = l.cFrom; mTo = l.cTo;
        } else {
            size_t i = cidx(c, l.k);
            pos = a.candPos[i]; P = a.candF[i]; Q = a.candR[i]; S = a.candS[i]; LL = a.candLL[i];
            X = FAST ? a.candX[i] : -1;
            if (st == ST_FB2) { mFrom = l.sFrom; mTo = pos; } else { mFrom = pos; mTo = l.sTo; }
        }
        const bool fwd = (st == ST_FB1 || st == ST_BF2);
        const uint8_t b = fwd ? base(l, mTo) : (mFrom > 0 ? base(l, mFrom - 1) : (uint8_t)0);
        int64_t F0 = P, R0 = Q;
        const int64_t S0 = S, X0 = X;
        const double LL0 = LL;
        /* ---- the extension (SMEM::fwdExt / SMEM::backExt, src/SMEM.h:140-177) ---- */
        if (FAST && S == 1) {
            if (X < 0) { X = ix.sa[fwd ? Q : P]; l.nLook++; }
            const uint8_t bb = fwd ? comp16(b) : b;
            if (!amb(b) && X > 0 && ix.bd[X - 1] == bb) { X--; l.nFast++; }
            else S = 0;
        } else {
            l.nExt += amb(b) ? 0 : 1;
            if (fwd) dBackExt(ix, Q, P, S, comp16(b));
            else dBackExt(ix, P, Q, S, b);
            X = -1;
        }
        if (fwd) { if (mTo < l.L) LL += ix.ll[b]; mTo++; }
        else { if (mFrom > 0) LL += ix.ll[b]; mFrom--; }
        /* the pre-step SMEM with its moving side turned back into a row */
        auto exact0 = [&]() {
            if (FAST && X0 >= 0) {
                const int64_t v = ix.isa[X0];
                l.nLook++;
                if (fwd) R0 = v; else F0 = v;
            }
        };

        /* ---- control ---- */
        if (st == ST_FB1) {
            if (S != S0 && l.cTo >= minLen && gateFwd(c, l, LL0, l.cFrom)) {
                exact0();
                candStore(c, l.nCand++, l.cTo, F0, R0, S0, LL0, -1);
            }
            l.cTo = mTo; l.cF = P; l.cR = Q; l.cS = S; l.cLL = LL; l.cX = X;
            if (!(S >= 1)) {
                l.sTo = mTo;
                if (l.nCand == 0) { l.sFrom--; driverAfter(c, l); }
                else { l.state = ST_FB2; l.lo = 0; startIteration(l); }
            }
        } else if (st == ST_BF1) {
            if (S != S0 && (int64_t)(l.L - l.cFrom) >= minLen && gateBack(c, l, LL0, l.cTo)) {
                exact0();
                candStore(c, l.nCand++, l.cFrom, F0, R0, S0, LL0, -1);
            }
            l.cFrom = mFrom; l.cF = P; l.cR = Q; l.cS = S; l.cLL = LL; l.cX = X;
            if (!(S >= 1)) {
                l.sFrom = mFrom;
                if (l.nCand == 0) { l.sTo++; driverAfter(c, l); }
                else { l.state = ST_BF2; l.lo = 0; startIteration(l); }
            }
        } else {
            const bool fb = (st == ST_FB2);
            const int len0 = fb ? (pos - l.sFrom) : (l.sTo - pos);
            if (S != S0 && S0 >= 1 && l.nValid == 0 && len0 >= minLen && LL0 <= a.T2) {
                exact0();
                if (fb) emit(c, l, l.sFrom, pos, F0, R0, S0, LL0);
                else emit(c, l, pos, l.sTo, F0, R0, S0, LL0);
            }
            if (S >= 1 && S != l.s0) {
                l.s0 = S;
                l.nValid++;
                candStore(c, l.w--, pos, P, Q, S, LL, X);
            }
            l.k--;
            if (l.k < l.lo) {
                l.lo = l.w + 1;
                bool more;
                if (fb) { l.sFrom--; more = l.sFrom >= 0; }
                else { l.sTo++; more = l.sTo <= l.L; }
                more = more && (l.lo <= l.nCand - 1);
                if (more) startIteration(l);
                else driverAfter(c, l);
            }
        }
        if (l.state == ST_IDLE) a.cnt[l.rid] = l.nOut;
    }
    if (a.stats) {
        atomicAdd(&a.stats[0], (unsigned long long)l.nExt);
        atomicAdd(&a.stats[2], (unsigned long long)l.nFast);
        atomicAdd(&a.stats[3], (unsigned long long)l.nLook);
    }
}


template __global__ void smem_kernel_t<false>(DevIndex, SmemArgs);
template __global__ void smem_kernel_t<true>(DevIndex, SmemArgs);
