/*
 * msgs_dp.hip -- chain-restricted affine DP and backtrace, one lane per chain.
 *
 * The reference (Alignment(read, rcRead, mtg, chain) + calculateScores(chain) +
 * backTrace, src/Alignment.h:84-94, src/Alignment.cpp:46-276) fills three dense
 * (L+1) x (tLen+1) double matrices with -inf and then writes, in this order:
 *   init      M row 0 = 0; NW: I(i,0) = -gapPenalty(i); SW: M col 0 = 0
 *   R5        the 5' rectangle before the first seed (if both sides are non-empty)
 *   seed k    M block := -inf, the seed diagonal (first cell floored at 0), the
 *             "step c" cell one past the seed, I row to_k+1 := M row to_k + o and
 *             D col end_k+1 := M col end_k + o (whole row / whole column)
 *   gap k     the rectangle between seed k and k+1
 *   R3        the 3' rectangle after the last seed
 * Only the rectangles hold O(area) data.  Here each rectangle is a tile of int32
 * M/I/D values in HBM; every other cell is resolved by walking that write history
 * backwards (vM/vI/vD below), which yields the value the dense matrix would hold at
 * any point in time.  Rectangle recurrences read their outside neighbours that way,
 * and the backtrace reads final values that way, so the result (score, end cell,
 * path, tie-breaks, even out-of-matrix reads) is the dense reference's.
 *
 * Scores are integer-valued (the reference's defaults 1/3/5/2 and any integer
 * setting); -inf is the sentinel NEG, which absorbs additions.  Non-integer score
 * settings are refused by the host (MSGS_E_UNSUPPORTED).
 */
#include "msgs_kernels.h"

namespace {

constexpr int32_t NEG = -1073741824;
enum : int32_t { OP_M = 0, OP_I = 1, OP_D = 2, OP_S = 4, OP_EQ = 7, OP_X = 8 };
enum : int32_t { FWD = 1, REV = 2 };

__device__ __forceinline__ int32_t addn(int32_t a, int32_t b) { return a == NEG ? NEG : a + b; }
__device__ __forceinline__ int32_t max3(int32_t a, int32_t b, int32_t c) { return max(a, max(b, c)); }

struct Rect {
    int64_t r0, r1, c0, c1, off;   /* rows r0..r1, cols c0..c1 (matrix coords), tile offset */
    __device__ bool has(int64_t i, int64_t j) const { return i >= r0 && i <= r1 && j >= c0 && j <= c1; }
    __device__ int64_t idx(int64_t i, int64_t j) const { return off + (j - c0) * (r1 - r0 + 1) + (i - r0); }
    __device__ bool empty() const { return r0 > r1 || c0 > c1; }
    __device__ int64_t area() const { return empty() ? 0 : (r1 - r0 + 1) * (c1 - c0 + 1); }
};

struct Job {
    const SeedRec* seeds;     /* the read's seeds */
    const int32_t* idx;       /* chain: seed indices */
    int32_t ns;
    int64_t L, tStart, tEnd, tLen;
    const uint8_t* q;
    const uint8_t* qual;
    const uint8_t* tg;        /* target chrom (followed by its 0 terminator) */
    int64_t chromLen;
    int32_t* tM;
    int32_t* tI;
    int32_t* tD;
    int32_t* d0;              /* first diagonal value of each seed */
    int64_t* gapOff;          /* tile offset of gap rect k */
    Rect r5, r3;
    int32_t mode, match, mis, go, ge;
    int32_t hazard;

    __device__ const SeedRec& S(int k) const { return seeds[idx[k]]; }
    __device__ Rect gap(int k) const {
        const SeedRec& a = S(k);
        const SeedRec& b = S(k + 1);
        Rect r;
        r.r0 = a.to + 1; r.r1 = b.from; r.c0 = a.end - tStart + 1; r.c1 = b.start - tStart;
        r.off = gapOff[k];
        return r;
    }
    __device__ uint8_t tb(int64_t t) const { return (t >= 0 && t <= chromLen) ? tg[t] : (uint8_t)0; }
    __device__ int32_t score(uint8_t a, uint8_t b) const {
        return ((a & b) && a != 15 && b != 15) ? match : -mis;
    }
};

/* M at (i, j) after events 1..e (0 = init only) */
__device__ int32_t vM(Job& J, int64_t i, int64_t j, int e) {
    if (i < 0 || j < 0 || i > J.L || j > J.tLen) { J.hazard = 1; return NEG; }
    const int last = 2 * J.ns + 1;
    for (; e >= 1; --e) {
        if (e == 1) {
            if (!J.r5.empty() && J.r5.has(i, j)) return J.tM[J.r5.idx(i, j)];
        } else if (e == last) {
            if (!J.r3.empty() && J.r3.has(i, j)) return J.tM[J.r3.idx(i, j)];
        } else if ((e & 1) == 0) {
            const int k = (e - 2) >> 1;
            const SeedRec& s = J.S(k);
            const int64_t bi = s.from + 1, bj = s.start - J.tStart + 1, len = s.to - s.from;
            if (i >= bi && i < bi + len && j >= bj && j < bj + len)
                return (i - bi == j - bj) ? J.d0[k] + (int32_t)(i - bi) * J.match : NEG;
            if (i == bi + len && j == bj + len && i <= J.L && j <= J.tLen) return J.d0[k] + (int32_t)len * J.match;
        } else {
            const Rect g = J.gap((e - 3) >> 1);
            if (!g.empty() && g.has(i, j)) return J.tM[g.idx(i, j)];
        }
    }
    return (i == 0 || (J.mode == 1 && j == 0)) ? 0 : NEG;
}

__device__ int32_t vI(Job& J, int64_t i, int64_t j, int e) {
    if (i < 0 || j < 0 || i > J.L || j > J.tLen) { J.hazard = 1; return NEG; }
    const int last = 2 * J.ns + 1;
    const int32_t o = -(J.go + J.ge);
    for (; e >= 1; --e) {
        if (e == 1) {
            if (!J.r5.empty() && J.r5.has(i, j)) return J.tI[J.r5.idx(i, j)];
        } else if (e == last) {
            if (!J.r3.empty() && J.r3.has(i, j)) return J.tI[J.r3.idx(i, j)];
        } else if ((e & 1) == 0) {
            const int k = (e - 2) >> 1;
            const int64_t r = J.S(k).to + 1;
            if (i == r && r <= J.L) return addn(vM(J, i - 1, j, e), o);
        } else {
            const Rect g = J.gap((e - 3) >> 1);
            if (!g.empty() && g.has(i, j)) return J.tI[g.idx(i, j)];
        }
    }
    if (J.mode == 0 && j == 0) return i == 0 ? 0 : -(J.go + J.ge * (int32_t)i);
    return NEG;
}

__device__ int32_t vD(Job& J, int64_t i, int64_t j, int e) {
    if (i < 0 || j < 0 || i > J.L || j > J.tLen) { J.hazard = 1; return NEG; }
    const int last = 2 * J.ns + 1;
    const int32_t o = -(J.go + J.ge);
    for (; e >= 1; --e) {
        if (e == 1) {
            if (!J.r5.empty() && J.r5.has(i, j)) return J.tD[J.r5.idx(i, j)];
        } else if (e == last) {
            if (!J.r3.empty() && J.r3.has(i, j)) return J.tD[J.r3.idx(i, j)];
        } else if ((e & 1) == 0) {
            const int k = (e - 2) >> 1;
            const int64_t c = J.S(k).end - J.tStart + 1;
            if (j == c && c <= J.tLen) return addn(vM(J, i, j - 1, e), o);
        } else {
            const Rect g = J.gap((e - 3) >> 1);
            if (!g.empty() && g.has(i, j)) return J.tD[g.idx(i, j)];
        }
    }
    return NEG;
}

/* calculateScoresNW / calculateScoresSW over a rectangle written at event e */
__device__ int64_t fillRect(Job& J, const Rect& R, int e) {
    if (R.empty()) return 0;
    const int32_t o = -(J.go + J.ge), ex = -J.ge;
    for (int64_t j = R.c0; j <= R.c1; ++j) {
        const uint8_t t = J.tb(J.tStart + j - 1);
        for (int64_t i = R.r0; i <= R.r1; ++i) {
            const int32_t sc = J.score(J.q[i - 1], t);
            const bool up = i - 1 >= R.r0, left = j - 1 >= R.c0;
            int32_t mD, iD, dD, mU, iU, mL, dL;
            if (up && left) { int64_t x = R.idx(i - 1, j - 1); mD = J.tM[x]; iD = J.tI[x]; dD = J.tD[x]; }
            else { mD = vM(J, i - 1, j - 1, e - 1); iD = vI(J, i - 1, j - 1, e - 1); dD = vD(J, i - 1, j - 1, e - 1); }
            if (up) { int64_t x = R.idx(i - 1, j); mU = J.tM[x]; iU = J.tI[x]; }
            else { mU = vM(J, i - 1, j, e - 1); iU = vI(J, i - 1, j, e - 1); }
            if (left) { int64_t x = R.idx(i, j - 1); mL = J.tM[x]; dL = J.tD[x]; }
            else { mL = vM(J, i, j - 1, e - 1); dL = vD(J, i, j - 1, e - 1); }
            int32_t m = max3(addn(mD, sc), addn(dD, sc), addn(iD, sc));
            if (J.mode == 1) m = max(m, 0);
            const int64_t x = R.idx(i, j);
            J.tM[x] = m;
            J.tI[x] = max(addn(mU, o), addn(iU, ex));
            J.tD[x] = max(addn(mL, o), addn(dL, ex));
        }
    }
    return R.area();
}

/* matchMax / insMax / delMax (src/Alignment.h:577-597), matchMax quirk included */
__device__ __forceinline__ int32_t matchMax(int32_t match, int32_t ins, int32_t del) {
    int32_t mx = match, s = OP_M;
    if (ins > mx) { mx = del; s = OP_I; }
    if (del > mx) s = OP_D;
    return s;
}
__device__ __forceinline__ int32_t subn(int32_t a, int32_t b) { return a == NEG ? NEG : a - b; }

__device__ __forceinline__ char decodeBase(uint8_t b) {
    const char* sym = "$ACMGRSVTWYHKDBN";
    return sym[b & 15];
}

__device__ __forceinline__ int putNum(char* md, int64_t pos, uint32_t v) {
    char buf[12];
    int n = 0;
    do { buf[n++] = (char)('0' + v % 10); v /= 10; } while (v);
    for (int k = 0; k < n; ++k) md[pos + k] = buf[n - 1 - k];
    return n;
}

__device__ void setupJob(const DevIndex& ix, const DpArgs& a, int64_t job, Job& J) {
    const int32_t rid = a.jobRead[job];
    J.seeds = a.seeds + a.seedOff[rid];
    J.idx = a.chainSeeds + a.jobChainStart[job];
    J.ns = a.jobChainLen[job];
    J.L = a.offs[rid + 1] - a.offs[rid];
    const SeedRec& f = J.S(0);
    const SeedRec& l = J.S(J.ns - 1);
    const int32_t tid = f.tid;
    J.chromLen = ix.chromLen[tid];
    J.tg = ix.tgt + ix.tgtOff[tid];
    /* initTStart / initTEnd (src/Alignment.cpp:46-52) */
    const double kIndel = 1 + 0.1;
    J.tStart = max((int64_t)0, (int64_t)((double)f.start - kIndel * (double)(f.from - 0)));
    J.tEnd = min(J.chromLen + 1, (int64_t)((double)l.end + kIndel * (double)(J.L - l.to)));
    J.tLen = J.tEnd - J.tStart;
    const bool fwd = f.strand == FWD;
    J.q = (fwd ? a.seqs : a.rcSeqs) + a.offs[rid];
    J.qual = (fwd ? a.quals : a.rcQuals) + a.offs[rid];
    J.mode = a.mode; J.match = a.match; J.mis = a.mismatch; J.go = a.gapO; J.ge = a.gapE;
    J.hazard = 0;
    J.r5.r0 = 1; J.r5.r1 = f.from; J.r5.c0 = 1; J.r5.c1 = f.start - J.tStart;
    if (!(0 < f.from && J.tStart < f.start)) { J.r5.r0 = 1; J.r5.r1 = 0; }
    J.r3.r0 = l.to + 1; J.r3.r1 = J.L; J.r3.c0 = l.end - J.tStart + 1; J.r3.c1 = J.tLen;
    if (!(l.to < J.L && l.end < J.tEnd)) { J.r3.r0 = 1; J.r3.r1 = 0; }
}

}  // namespace

__global__ void __launch_bounds__(256) dp_kernel(DevIndex ix, DpArgs a) {
    for (int64_t job = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; job < a.nJobs;
         job += (int64_t)gridDim.x * blockDim.x) {
        Job J;
        setupJob(ix, a, job, J);
        const int64_t pathCap = J.L + J.tLen + 3;
        if (a.sizeOnly) {
            int64_t area = J.r5.area() + J.r3.area();
            for (int k = 0; k + 1 < J.ns; ++k) {
                const SeedRec& s0 = J.S(k);
                const SeedRec& s1 = J.S(k + 1);
                int64_t rows = s1.from - s0.to, cols = s1.start - s0.end;
                if (rows > 0 && cols > 0) area += rows * cols;
            }
            a.tileNeed[job] = area;
            a.pathNeed[job] = pathCap;
            continue;
        }
        const int64_t base = a.tileOff[job];
        J.tM = a.tileM + base; J.tI = a.tileI + base; J.tD = a.tileD + base;
        /* per-seed scratch, addressed by the chain's range in the chain-seed pool */
        J.d0 = a.auxD0 + a.jobChainStart[job];
        J.gapOff = a.auxGap + a.jobChainStart[job];
        AlnRec rec;
        rec.rid = a.jobRead[job];
        /* tiles are laid out R5, gap 0 .. gap ns-2, R3 */
        int64_t cur = J.r5.area();
        J.r5.off = 0;
        for (int k = 0; k + 1 < J.ns; ++k) {
            const SeedRec& s0 = J.S(k);
            const SeedRec& s1 = J.S(k + 1);
            int64_t rows = s1.from - s0.to, cols = s1.start - s0.end;
            J.gapOff[k] = cur;
            if (rows > 0 && cols > 0) cur += rows * cols;
        }
        J.r3.off = cur;
        int32_t* d0p = J.d0;
        /* ---- calculateScores(chain) (src/Alignment.cpp:54-70) ---- */
        int64_t cells = fillRect(J, J.r5, 1);
        for (int k = 0; k < J.ns; ++k) {
            const int e = 2 + 2 * k;
            const SeedRec& s = J.S(k);
            const int64_t bi = s.from + 1, bj = s.start - J.tStart + 1, len = s.to - s.from;
            const int32_t sm = J.match;
            d0p[k] = max(max3(addn(vM(J, bi - 1, bj - 1, e - 1), sm), addn(vD(J, bi - 1, bj - 1, e - 1), sm),
                              addn(vI(J, bi - 1, bj - 1, e - 1), sm)), 0);
            cells += len;
            if (bi + len <= J.L && bj + len <= J.tLen) cells += 1;
            if (k + 1 < J.ns) cells += fillRect(J, J.gap(k), e + 1);
        }
        cells += fillRect(J, J.r3, 2 * J.ns + 1);
        const int T = 2 * J.ns + 1;
        /* ---- backtrace (src/Alignment.cpp:173-276) ---- */
        uint8_t* path = a.path + a.pathOff[job];
        int64_t np = 0;
        int64_t alnTo, alnEnd, i, j;
        int32_t best;
        int32_t s = OP_M;
        const int32_t go = J.go, ge = J.ge;
        if (J.mode == 0) {
            best = NEG; alnEnd = 0; alnTo = J.L;
            for (int64_t c = 1; c <= J.tLen; ++c) {
                int32_t m = vM(J, alnTo, c, T);
                if (m > best) { best = m; alnEnd = c; s = OP_M; }
                int32_t x = vI(J, alnTo, c, T);
                if (x > best) { best = x; alnEnd = c; s = OP_I; }
            }
            i = alnTo; j = alnEnd;
            while (i > 0 && np < pathCap) {
                path[np++] = (uint8_t)s;
                if (s == OP_M) { s = matchMax(vM(J, i - 1, j - 1, T), vI(J, i - 1, j - 1, T), vD(J, i - 1, j - 1, T)); i--; j--; }
                else if (s == OP_I) { s = subn(subn(vM(J, i - 1, j, T), go), ge) > subn(vI(J, i - 1, j, T), ge) ? OP_M : OP_I; i--; }
                else { s = subn(subn(vM(J, i, j - 1, T), go), ge) > subn(vD(J, i, j - 1, T), ge) ? OP_M : OP_D; j--; }
                if (J.hazard && j < -1) break;
            }
        } else {
            /* M.maxCoeff: first maximum in column-major order */
            best = 0; int64_t bi = 0, bj = 0;
            auto visit = [&](int64_t ci, int64_t cj) {
                int32_t v = vM(J, ci, cj, T);
                if (v > best || (v == best && (cj < bj || (cj == bj && ci < bi)))) { best = v; bi = ci; bj = cj; }
            };
            auto visitRect = [&](const Rect& R) {
                if (R.empty()) return;
                for (int64_t cj = R.c0; cj <= R.c1; ++cj)
                    for (int64_t ci = R.r0; ci <= R.r1; ++ci) visit(ci, cj);
            };
            visitRect(J.r5);
            for (int k = 0; k < J.ns; ++k) {
                const SeedRec& sd = J.S(k);
                const int64_t b0 = sd.from + 1, b1 = sd.start - J.tStart + 1, len = sd.to - sd.from;
                for (int64_t d = 0; d < len; ++d) visit(b0 + d, b1 + d);
                if (b0 + len <= J.L && b1 + len <= J.tLen) visit(b0 + len, b1 + len);
                if (k + 1 < J.ns) visitRect(J.gap(k));
            }
            visitRect(J.r3);
            alnTo = bi; alnEnd = bj;
            i = alnTo; j = alnEnd;
            while ((s != OP_M || vM(J, i, j, T) > 0) && np < pathCap) {
                path[np++] = (uint8_t)s;
                if (s == OP_M) { s = matchMax(vM(J, i - 1, j - 1, T), vI(J, i - 1, j - 1, T), vD(J, i - 1, j - 1, T)); i--; j--; }
                else if (s == OP_I) { s = subn(subn(vM(J, i - 1, j, T), go), ge) > subn(vI(J, i - 1, j, T), ge) ? OP_M : OP_I; i--; }
                else { s = subn(subn(vM(J, i, j - 1, T), go), ge) > subn(vD(J, i, j - 1, T), ge) ? OP_M : OP_D; j--; }
                if (J.hazard && (i < -1 || j < -1)) break;
            }
        }
        /* path was collected end -> start; reverse it */
        for (int64_t x = 0, y = np - 1; x < y; ++x, --y) { uint8_t t = path[x]; path[x] = path[y]; path[y] = t; }
        const int64_t alnFrom = i, alnStart = J.tStart + j;
        rec.tid = J.S(0).tid; rec.strand = J.S(0).strand; rec.hazard = J.hazard;
        rec.tStart = J.tStart; rec.tEnd = J.tEnd;
        rec.alnFrom = alnFrom; rec.alnTo = alnTo; rec.alnStart = alnStart; rec.alnEnd = alnEnd + J.tStart;
        rec.score = best;
        rec.cells = cells;
        /* ---- getAlnCigar (src/Alignment.cpp:278-303) ---- */
        uint32_t* cg = a.cigar + a.pathOff[job] + 2 * job;
        int32_t nc = 0;
        const int32_t clip5 = (int32_t)(alnFrom - 0), clip3 = (int32_t)(J.L - alnTo);
        if (clip5 > 0) cg[nc++] = ((uint32_t)clip5 << 4) | OP_S;
        if (np > 0) {
            uint32_t op = path[0], len = 1;
            for (int64_t x = 1; x < np; ++x) {
                if (path[x] != op) { cg[nc++] = (len << 4) | op; len = 1; }
                else len++;
                op = path[x];
            }
            cg[nc++] = (len << 4) | op;
        }
        if (clip3 > 0) cg[nc++] = ((uint32_t)clip3 << 4) | OP_S;
        rec.nCigar = nc;
        /* ---- getAlnMDTag (src/Alignment.cpp:305-353) and evaluate (:446-489) ---- */
        char* md = a.md + 3 * a.pathOff[job] + 8 * job;
        int64_t ml = 0;
        uint32_t prev = 0xffffffffu, run = 0;
        double x = 0;
        for (int64_t c = 0; c < alnFrom; ++c) x += -a.clip;
        int64_t qi = alnFrom, tj = alnStart;
        for (int64_t k = 0; k < np; ++k) {
            const uint8_t qb = (qi >= 0 && qi < J.L) ? J.q[qi] : (uint8_t)0;
            const uint8_t tbv = J.tb(tj);
            const int qv = (qi >= 0 && qi < J.L) ? J.qual[qi] : 0;
            uint32_t op = path[k];
            if (op == OP_M) {
                x += (qb & tbv) ? a.qMatch[qv] : a.qMis[qv];
                op = (qb & tbv) ? OP_EQ : OP_X;
            } else {
                const bool open = (k == 0 || path[k - 1] != path[k]);
                if (open) x += a.qGapOpen[qv];
                x += -a.gapE_d;
            }
            if (op == OP_EQ) { qi++; tj++; run++; }
            else if (op == OP_X) { ml += putNum(md, ml, run); run = 0; md[ml++] = decodeBase(tbv); qi++; tj++; }
            else if (op == OP_I) {
                if (prev != OP_I) ml += putNum(md, ml, run);
                run = 0;
                if (op != prev) md[ml++] = '^';
                md[ml++] = decodeBase(qb);
                qi++;
            } else if (op == OP_D) { tj++; }
            prev = op;
        }
        ml += putNum(md, ml, run);
        for (int64_t c = alnTo; c < J.L; ++c) x += -a.clip;
        rec.mdLen = (int32_t)ml;
        rec.log10P = x;
        rec.pad = (int32_t)np;
        a.out[job] = rec;
    }
}
