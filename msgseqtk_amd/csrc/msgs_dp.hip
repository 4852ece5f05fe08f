/*
 * msgs_dp.hip -- chain-restricted affine DP and backtrace.
 *
 * The reference (Alignment(read, rcRead, mtg, chain) + calculateScores(chain) +
 * backTrace, src/Alignment.h:84-94, src/Alignment.cpp:46-276) fills three dense
 * (L+1) x (tLen+1) double matrices with -inf and then writes, in this order:
 *   event 1       R5, the 5' rectangle before the first seed (if both sides are non-empty)
 *   event 2+2k    seed k: M block := -inf, the seed diagonal (first cell floored at 0),
 *                 the "step c" cell one past the seed, I row to_k+1 := M row to_k + o and
 *                 D col end_k+1 := M col end_k + o (whole row / whole column)
 *   event 3+2k    gap k, the rectangle between seed k and k+1
 *   event 2ns+1   R3, the 3' rectangle after the last seed
 * on top of the initial values (M row 0 = 0; NW: I(i,0) = -gapPenalty(i); SW: M col 0 = 0).
 * Only the rectangles hold O(area) data.  Here each rectangle is a tile of int32
 * (M, I, D) cells in HBM, and every other cell is resolved from the write history:
 * seeds are ordered and compatible, so rows (and columns) split into alternating
 * rectangle and seed bands, and the latest writer of a cell at any event is found from
 * the band of its row (vMk / vIk / vDk).  Rectangle borders, seed diagonals and the
 * backtrace read values exactly as the dense reference would, including the order of
 * ties and the out-of-matrix reads of the backtrace (flagged as hazard, read as -inf).
 *
 * Three kernels:
 *   dp_setup_kernel  one lane per job: tStart/tEnd, seeds in matrix coordinates,
 *                    rectangle geometry and pool sizes, and the fill class;
 *   dp_fill_kernel   one wavefront per job with rectangles: each rectangle is swept in
 *                    strips of 64 rows along anti-diagonals (lane = row, up/diagonal
 *                    neighbours by lane shuffle, the strip's top row in LDS), seeds'
 *                    first diagonal values in between, rectangle summaries at the end;
 *   dp_trace_kernel  one lane per job: the serial fill for jobs the wavefront kernel does
 *                    not take, the NW last-row scan / SW maximum, the backtrace (seed
 *                    diagonal interiors are crossed in one step: their neighbours are
 *                    provably -inf), CIGAR, MD and the quality-based log10 likelihood.
 *
 * Scores are integer-valued (the reference's defaults 1/3/5/2 and any integer setting);
 * -inf is the sentinel NEG, which absorbs additions.  Non-integer score settings are
 * refused by the host (MSGS_E_UNSUPPORTED).
 */
#include "msgs_kernels.h"

#include <type_traits>

namespace {

constexpr int32_t NEG = -1073741824;
enum : int32_t { OP_M = 0, OP_I = 1, OP_D = 2, OP_S = 4, OP_EQ = 7, OP_X = 8 };
enum : int32_t { FWD = 1, REV = 2 };
/* limits of the wavefront fill: seeds per job (LDS copy) and rectangle columns; with 64
 * seeds a one-wavefront block holds 7.7 KB of LDS, so LDS allows 20 blocks per CU and the
 * kernel's registers decide (the host sizes the persistent grid by the occupancy query) */
constexpr int FILL_NS = 64;
constexpr int FILL_COLS = 512;
/* limits of the 16-lane fill (four jobs per wavefront): seeds per job and rectangle columns;
 * its LDS holds, per job, the seeds (512 B) and the row above the current strip (M, I, D and
 * the target bases: 2.8 KB), 13.2 KB per four-job block */
constexpr int NS16 = 32;
constexpr int COLS16 = 192;

__device__ __forceinline__ int32_t addn(int32_t a, int32_t b) { return a == NEG ? NEG : a + b; }
__device__ __forceinline__ int32_t subn(int32_t a, int32_t b) { return a == NEG ? NEG : a - b; }
__device__ __forceinline__ int32_t max3(int32_t a, int32_t b, int32_t c) { return max(a, max(b, c)); }
/* the value of lane - 1 (lane 0 keeps its own), by a DPP whole-wavefront shift right by one
 * lane (wave_shr:1) instead of an LDS permute; every lane of the wavefront runs it */
__device__ __forceinline__ int32_t laneUp1(int32_t v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
/* within each 16-lane DPP row: the value of lane - 1, lane 0 of the row taking its own `first`
 * (row_shr:1 with the out-of-row source keeping the old value) */
__device__ __forceinline__ int32_t rowUp1(int32_t first, int32_t v) {
    return __builtin_amdgcn_update_dpp(first, v, 0x111, 0xf, 0xf, false);
}
/* c ? a : b as one v_cndmask on the condition's lane mask (kept from becoming a branch) */
__device__ __forceinline__ int32_t selectI(bool c, int32_t a, int32_t b) {
    int32_t r;
    const uint64_t m = __builtin_amdgcn_ballot_w64(c);
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(a), "s"(m));
    return r;
}
/* acc * 2 + c: one decision bit shifted in at the bottom, by an add with carry whose carry-in
 * is the condition's lane mask (one instruction per bit) */
__device__ __forceinline__ uint32_t pushBit(uint32_t acc, bool c) {
    uint32_t r;
    uint64_t co;
    const uint64_t m = __builtin_amdgcn_ballot_w64(c);
    asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(co) : "v"(acc), "s"(m));
    return r;
}

/* Backtrace decisions of a rectangle are packed per row and per sweep step: the fill sweeps
 * a strip of `rows` (<= 64) rows and C columns in C + rows - 1 anti-diagonal steps, and the
 * lane of row l decides cell (l, jj) at step t = jj + l.  Each lane packs the decisions of 8
 * consecutive steps into one word, and the lanes store their words side by side, so that
 * every store of the strip is one contiguous run of rows words: cell (l, jj) is in word
 * ((jj + l) / 8) * rows + l.  GLOBAL: 32-bit words of 4-bit decisions, the first step of a
 * word in its top nibble (step s at bit 4 * (7 - s % 8)), as the fill shifts each decision bit
 * in at the bottom; LOCAL: 64-bit words of 8-bit decisions (bit 4 being M > 0), step s at bit
 * 8 * (s % 8).  Strips of H rows follow each other (H = 64 for jobs filled one per wavefront,
 * 16 for jobs filled 16 lanes each); a strip of `rows` rows takes ceil((C + rows - 1) / 8) *
 * rows words. */
constexpr int DEC_SH = 3;   /* log2 of steps per word */
__device__ __forceinline__ int decBytes(int mode) { return mode == 0 ? 4 : 8; }   /* bytes per word */
__device__ __forceinline__ int decShiftOf(int mode, int32_t t) {
    return mode == 0 ? 4 * (7 - (t & 7)) : 8 * (t & 7);
}
template <int MODE> using DecWord = typename std::conditional<MODE == 0, uint32_t, uint64_t>::type;
__device__ __forceinline__ int64_t stripWords(int64_t rows, int64_t cols) {
    return ((cols + rows - 1 + 7) >> DEC_SH) * rows;
}
/* hs = log2 of the strip height H */
__device__ __forceinline__ int64_t stripBytes(int64_t rows, int64_t cols, int mode, int hs) {
    const int64_t full = rows >> hs, rem = rows & ((1 << hs) - 1);
    return decBytes(mode) * (full * stripWords(1 << hs, cols) + (rem ? stripWords(rem, cols) : 0));
}


/* sequential byte reads through the aligned 8-byte word holding each byte */
struct Bytes8 {
    uint64_t w = 0;
    uintptr_t wa = ~(uintptr_t)0;
    __device__ __forceinline__ uint32_t get(const uint8_t* p) {
        const uintptr_t a = (uintptr_t)p & ~(uintptr_t)7;
        if (a != wa) { wa = a; w = *(const uint64_t*)a; }
        return (uint32_t)(w >> (8 * ((uintptr_t)p & 7))) & 0xffu;
    }
};

/* a rectangle with its tile offset and event index */
struct Rc {
    int32_t r0, r1, c0, c1, ev;
    int64_t off;               /* first cell in the job's decision (or tile) region */
    int64_t boff;              /* first cell in the job's border region */
    bool ok;
    __device__ __forceinline__ int64_t idx(int32_t i, int32_t j) const { return off + (int64_t)(j - c0) * (r1 - r0 + 1) + (i - r0); }
    __device__ __forceinline__ bool hasCol(int32_t j) const { return j >= c0 && j <= c1; }
    __device__ __forceinline__ int64_t area() const { return ok ? (int64_t)(r1 - r0 + 1) * (c1 - c0 + 1) : 0; }
    /* the word holding the decisions of cell (i, j) and their bit position (see stripBytes) */
    __device__ __forceinline__ int64_t decWord(int32_t i, int32_t j, int mode, int hs, int& shift) const {
        const int32_t H = 1 << hs;
        const int32_t ri = i - r0, jj = j - c0, s0 = ri & ~(H - 1), l = ri & (H - 1);
        const int32_t rows = min(H, r1 - r0 + 1 - s0), C = c1 - c0 + 1;
        const int32_t t = jj + l;
        shift = decShiftOf(mode, t);
        return off / decBytes(mode) + (int64_t)(s0 >> hs) * stripWords(H, C) + (int64_t)(t >> DEC_SH) * rows + l;
    }
};

struct Geo {
    const int4* sg;          /* bi, bj, len, d0 per seed */
    const int64_t* gapOff;   /* cell offset of gap rect k */
    const int64_t* gapBOff;  /* border offset of gap rect k */
    int4* tile;              /* serially filled jobs: M, I, D per rectangle cell */
    uint8_t* bits;           /* wavefront-filled jobs: backtrace decisions per rectangle cell */
    int4* border;            /* wavefront-filled jobs: last row and last column of each rectangle */
    int32_t L, tLen, ns, mode, match, mis, go, ge, o;
    int32_t r5r1, r5c1, r3r0, r3c0;
    int64_t r3off, r3boff;
    int32_t hazard;
    bool serial;             /* rectangle values kept in full (tile) */
    int32_t hs;              /* log2 of the decision layout's strip height */
    int32_t interior;        /* set if an interior value of a border-only rectangle was asked for */
};

__device__ __forceinline__ int32_t initM(const Geo& G, int32_t i, int32_t j) {
    return (i == 0 || (G.mode == 1 && j == 0)) ? 0 : NEG;
}
__device__ __forceinline__ int32_t initI(const Geo& G, int32_t i, int32_t j) {
    return (G.mode == 0 && j == 0) ? (i == 0 ? 0 : -(G.go + G.ge * i)) : NEG;
}

/* largest k with bi_k <= i (-1 if none) */
__device__ __forceinline__ int rowSeg(const Geo& G, int32_t i) {
    int lo = -1, hi = G.ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (G.sg[mid].x <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}
/* largest k with bj_k <= j (-1 if none) */
__device__ __forceinline__ int colSeg(const Geo& G, int32_t j) {
    int lo = -1, hi = G.ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (G.sg[mid].y <= j) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ Rc rectR5(const Geo& G) {
    Rc R;
    R.r0 = 1; R.r1 = G.r5r1; R.c0 = 1; R.c1 = G.r5c1; R.ev = 1; R.off = 0; R.boff = 0;
    R.ok = G.r5r1 >= 1 && G.r5c1 >= 1;
    return R;
}
__device__ __forceinline__ Rc rectR3(const Geo& G) {
    Rc R;
    R.r0 = G.r3r0; R.r1 = G.L; R.c0 = G.r3c0; R.c1 = G.tLen; R.ev = 2 * G.ns + 1; R.off = G.r3off; R.boff = G.r3boff;
    R.ok = R.r0 <= R.r1 && R.c0 <= R.c1;
    return R;
}
/* gap k between seed s = k and t = k+1 */
__device__ __forceinline__ Rc rectGap(const Geo& G, int k, const int4& s, const int4& t) {
    Rc R;
    R.r0 = s.x + s.z; R.r1 = t.x - 1; R.c0 = s.y + s.z; R.c1 = t.y - 1; R.ev = 3 + 2 * k; R.off = G.gapOff[k];
    R.boff = G.gapBOff[k];
    R.ok = R.r0 <= R.r1 && R.c0 <= R.c1;
    return R;
}
/* rectangle of the band after seed k (k = rowSeg(i) and row i past seed k) */
__device__ __forceinline__ Rc rectAfter(const Geo& G, int k, const int4& s) {
    if (k + 1 < G.ns) return rectGap(G, k, s, G.sg[k + 1]);
    return rectR3(G);
}

/* M, I, D of a cell of rectangle R after its fill.  Wavefront-filled jobs keep only each
 * rectangle's last row and last column: the backtrace takes decisions inside a rectangle
 * from the decision bytes, and a cell outside every rectangle has its neighbours inside
 * one only on that rectangle's last row or column.  Any other request is flagged. */
__device__ __forceinline__ int4 rectVal(const Geo& G, const Rc& R, int32_t i, int32_t j) {
    if (G.serial) return G.tile[R.idx(i, j)];
    if (i == R.r1) return G.border[R.boff + (j - R.c0)];
    if (j == R.c1) return G.border[R.boff + (R.c1 - R.c0 + 1) + (i - R.r0)];
    const_cast<Geo&>(G).interior = 1;
    return make_int4(NEG, NEG, NEG, 0);
}

/* M at (i, j) after events 1..e; (i, j) inside the matrix, k = rowSeg(i) */
/* RECTS = false: the caller knows that the rectangle after band k is not filled at event e
 * (it is the one about to be filled), so no rectangle value is looked up (no global reads) */
template <bool RECTS = true>
__device__ __forceinline__ int32_t vMk(const Geo& G, int32_t i, int32_t j, int e, int k) {
    if (k < 0) {
        if constexpr (RECTS) {
            const Rc R = rectR5(G);
            if (i >= 1 && R.ok && e >= 1 && R.hasCol(j)) return rectVal(G, R, i, j).x;
        }
        return initM(G, i, j);
    }
    const int4 s = G.sg[k];
    if (i < s.x + s.z) {   /* seed k's rows */
        if (e >= 2 + 2 * k && j >= s.y && j < s.y + s.z) return (i - s.x == j - s.y) ? s.w + (i - s.x) * G.match : NEG;
        if (k >= 1 && e >= 2 * k) {   /* step c of seed k-1 on the first row of seed k */
            const int4 p = G.sg[k - 1];
            if (i == p.x + p.z && j == p.y + p.z) return p.w + p.z * G.match;
        }
        return initM(G, i, j);
    }
    if constexpr (RECTS) {
        const Rc R = rectAfter(G, k, s);
        if (R.ok && e >= R.ev && R.hasCol(j)) return rectVal(G, R, i, j).x;
    }
    if (e >= 2 + 2 * k && i == s.x + s.z && j == s.y + s.z) return s.w + s.z * G.match;   /* step c */
    return initM(G, i, j);
}

template <bool RECTS = true>
__device__ __forceinline__ int32_t vIk(const Geo& G, int32_t i, int32_t j, int e, int k) {
    if (k >= 0) {
        const int4 s = G.sg[k];
        if (i >= s.x + s.z) {
            if constexpr (RECTS) {
                const Rc R = rectAfter(G, k, s);
                if (R.ok && e >= R.ev && R.hasCol(j)) return rectVal(G, R, i, j).y;
            }
            /* I row to_k + 1 := M row to_k + o at seed k's event */
            if (i == s.x + s.z && e >= 2 + 2 * k) return addn(vMk<RECTS>(G, i - 1, j, 2 + 2 * k, k), G.o);
        } else if (k >= 1 && e >= 2 * k) {
            const int4 p = G.sg[k - 1];
            if (i == p.x + p.z) return addn(vMk<RECTS>(G, i - 1, j, 2 * k, k - 1), G.o);
        }
    } else if constexpr (RECTS) {
        const Rc R = rectR5(G);
        if (i >= 1 && R.ok && e >= 1 && R.hasCol(j)) return rectVal(G, R, i, j).y;
    }
    return initI(G, i, j);
}

/* D at (i, j) after events 1..e; k = rowSeg(i), m = colSeg(j) */
template <bool RECTS = true>
__device__ __forceinline__ int32_t vDkm(const Geo& G, int32_t i, int32_t j, int e, int k, int m) {
    if constexpr (RECTS) {
        if (k < 0) {
            const Rc R = rectR5(G);
            if (i >= 1 && R.ok && e >= 1 && R.hasCol(j)) return rectVal(G, R, i, j).z;
        } else {
            const int4 s = G.sg[k];
            if (i >= s.x + s.z) {
                const Rc R = rectAfter(G, k, s);
                if (R.ok && e >= R.ev && R.hasCol(j)) return rectVal(G, R, i, j).z;
            }
        }
    }
    /* D col end_m + 1 := M col end_m + o at seed m's event; m is colSeg(j) or the one before */
    if (m >= 0) {
        const int4 s = G.sg[m];
        if (j == s.y + s.z && e >= 2 + 2 * m) return addn(vMk<RECTS>(G, i, j - 1, 2 + 2 * m, k), G.o);
        if (m >= 1) {
            const int4 p = G.sg[m - 1];
            if (j == p.y + p.z && e >= 2 * m) return addn(vMk<RECTS>(G, i, j - 1, 2 * m, k), G.o);
        }
    }
    return NEG;
}
__device__ __forceinline__ int32_t vDk(const Geo& G, int32_t i, int32_t j, int e, int k) {
    return vDkm(G, i, j, e, k, colSeg(G, j));
}

__device__ __forceinline__ bool inside(Geo& G, int32_t i, int32_t j) {
    if (i < 0 || j < 0 || i > G.L || j > G.tLen) { G.hazard = 1; return false; }
    return true;
}
__device__ __forceinline__ int32_t vM(Geo& G, int32_t i, int32_t j, int e) { return inside(G, i, j) ? vMk(G, i, j, e, rowSeg(G, i)) : NEG; }
__device__ __forceinline__ int32_t vI(Geo& G, int32_t i, int32_t j, int e) { return inside(G, i, j) ? vIk(G, i, j, e, rowSeg(G, i)) : NEG; }
__device__ __forceinline__ int32_t vD(Geo& G, int32_t i, int32_t j, int e) { return inside(G, i, j) ? vDk(G, i, j, e, rowSeg(G, i)) : NEG; }

/* the rectangle holding cell (i, j), if any (the band rules of vMk) */
__device__ __forceinline__ bool cellRect(const Geo& G, int32_t i, int32_t j, Rc& R) {
    if (i < 1 || j < 1 || i > G.L || j > G.tLen) return false;
    const int k = rowSeg(G, i);
    if (k < 0) R = rectR5(G);
    else {
        const int4 s = G.sg[k];
        if (i < s.x + s.z) return false;
        R = rectAfter(G, k, s);
    }
    return R.ok && i >= R.r0 && i <= R.r1 && R.hasCol(j);
}

/* first diagonal value of seed k (src/Alignment.cpp:146-150), from its corner at event 1+2k */
__device__ __forceinline__ int32_t seedD0(Geo& G, int k) {
    const int4 s = G.sg[k];
    const int e = 1 + 2 * k;
    const int32_t sm = G.match;
    return max(max3(addn(vM(G, s.x - 1, s.y - 1, e), sm), addn(vD(G, s.x - 1, s.y - 1, e), sm),
                    addn(vI(G, s.x - 1, s.y - 1, e), sm)), 0);
}

__device__ __forceinline__ bool swBetter(int32_t v, int32_t i, int32_t j, int32_t bv, int32_t bi, int32_t bj) {
    return v > bv || (v == bv && (j < bj || (j == bj && i < bi)));
}

/* matchMax / insMax / delMax (src/Alignment.h:577-597), matchMax quirk included */
__device__ __forceinline__ int32_t matchMax(int32_t match, int32_t ins, int32_t del) {
    int32_t mx = match, s = OP_M;
    if (ins > mx) { mx = del; s = OP_I; }
    if (del > mx) s = OP_D;
    return s;
}

/* nt16 code -> IUPAC symbol ("$ACMGRSVTWYHKDBN"), from two packed words */
__device__ __forceinline__ char decodeBase(uint8_t b) {
    constexpr uint64_t kSym0 = (uint64_t)'$' | (uint64_t)'A' << 8 | (uint64_t)'C' << 16 | (uint64_t)'M' << 24 |
                               (uint64_t)'G' << 32 | (uint64_t)'R' << 40 | (uint64_t)'S' << 48 | (uint64_t)'V' << 56;
    constexpr uint64_t kSym1 = (uint64_t)'T' | (uint64_t)'W' << 8 | (uint64_t)'Y' << 16 | (uint64_t)'H' << 24 |
                               (uint64_t)'K' << 32 | (uint64_t)'D' << 40 | (uint64_t)'B' << 48 | (uint64_t)'N' << 56;
    const uint32_t c = b & 15u;
    return (char)((c < 8 ? kSym0 : kSym1) >> (8 * (c & 7)));
}

/* decimal digits of v written at md[pos..], their count returned */
__device__ __forceinline__ int putNum(char* md, int64_t pos, uint32_t v) {
    int n = 1;
    for (uint32_t t = v; t >= 10; t /= 10) ++n;
    for (int k = n - 1; k >= 0; --k) { md[pos + k] = (char)('0' + v % 10); v /= 10; }
    return n;
}

/* 8 bytes from p (any alignment) through the aligned words that hold them: the second
 * word is read only when p is unaligned, and then it holds p[7], so no read goes past the
 * aligned word of the last byte */
__device__ __forceinline__ uint64_t load8u(const uint8_t* p) {
    const uintptr_t a = (uintptr_t)p & ~(uintptr_t)7;
    const uint32_t sh = 8u * (uint32_t)((uintptr_t)p & 7);
    const uint64_t w0 = *(const uint64_t*)a;
    if (!sh) return w0;
    const uint64_t w1 = *(const uint64_t*)(a + 8);
    return (w0 >> sh) | (w1 << (64 - sh));
}

__device__ __forceinline__ void makeGeo(Geo& G, const DpArgs& a, const JobGeo& g, const int4* sg, int64_t job) {
    G.sg = sg;
    G.gapOff = a.auxGap + g.cs;
    G.gapBOff = a.auxGapB + g.cs;
    G.serial = g.cls == CLS_SERIAL;
    G.hs = g.cls == CLS_WAVE16 ? 4 : 6;
    G.tile = a.tile + a.serialOff[job];
    G.bits = a.bits + a.tileOff[job];
    G.border = a.border + a.borderOff[job];
    G.interior = 0;
    G.L = g.L; G.tLen = g.tLen; G.ns = g.ns; G.mode = a.mode;
    G.match = a.match; G.mis = a.mismatch; G.go = a.gapO; G.ge = a.gapE; G.o = -(a.gapO + a.gapE);
    G.r5r1 = g.r5r1; G.r5c1 = g.r5c1; G.r3r0 = g.r3r0; G.r3c0 = g.r3c0; G.r3off = g.r3off; G.r3boff = g.r3boff;
    G.hazard = 0;
}

}  // namespace

/* ------------------------------------------------------------------ setup */

__global__ void __launch_bounds__(256) dp_setup_kernel(DevIndex ix, DpArgs a) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.tileNeed[a.nJobs] = 0; a.pathNeed[a.nJobs] = 0; a.borderNeed[a.nJobs] = 0; a.serialNeed[a.nJobs] = 0;
    }
    /* shape counters of the 16-lane jobs' rectangles (when a.dstat is set) */
    unsigned long long nRect = 0, nCell = 0, nLaneStep = 0, nStrip = 0;
    unsigned long long jRect = 0, jCell = 0, jLaneStep = 0, jStrip = 0;   /* of the current job */
    auto shape = [&](int64_t rows, int64_t cols) {
        const int64_t strips = (rows + 15) >> 4;
        jRect += 1;
        jCell += (unsigned long long)(rows * cols);
        jLaneStep += (unsigned long long)(16 * (strips * (cols - 1) + rows));
        jStrip += (unsigned long long)strips;
    };
    for (int64_t job = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; job < a.nJobs;
         job += (int64_t)gridDim.x * blockDim.x) {
        const int32_t rid = a.jobRead[job];
        const SeedRec* seeds = a.seeds + a.seedOff[rid];
        const int32_t cs = a.jobChainStart[job];
        const int32_t* idx = a.chainSeeds + cs;
        const int32_t ns = a.jobChainLen[job];
        const int64_t L = a.offs[rid + 1] - a.offs[rid];
        const SeedRec& f = seeds[idx[0]];
        const SeedRec& l = seeds[idx[ns - 1]];
        const int64_t chromLen = ix.chromLen[f.tid];
        /* initTStart / initTEnd (src/Alignment.cpp:46-52) */
        const double kIndel = 1 + 0.1;
        const int64_t tStart = max((int64_t)0, (int64_t)((double)f.start - kIndel * (double)(f.from - 0)));
        const int64_t tEnd = min(chromLen + 1, (int64_t)((double)l.end + kIndel * (double)(L - l.to)));
        JobGeo g;
        g.tStart = tStart; g.tEnd = tEnd;
        g.L = (int32_t)L; g.tLen = (int32_t)(tEnd - tStart); g.ns = ns; g.cs = cs;
        g.tid = f.tid; g.strand = f.strand; g.rid = rid;
        /* R5 only when both sides are non-empty (src/Alignment.cpp:57-58) */
        g.r5r1 = (int32_t)f.from; g.r5c1 = (int32_t)(f.start - tStart);
        if (!(0 < f.from && tStart < f.start)) { g.r5r1 = 0; g.r5c1 = 0; }
        g.r3r0 = (int32_t)(l.to + 1); g.r3c0 = (int32_t)(l.end - tStart + 1);
        if (!(l.to < L && l.end < tEnd)) { g.r3r0 = g.L + 1; g.r3c0 = g.tLen + 1; }
        /* the fill class from the rectangles' widths and the seed count; it sets the strip
         * height of the decision layout */
        bool anyRect = g.r5r1 >= 1 && g.r5c1 >= 1;
        int32_t maxCols = anyRect ? g.r5c1 : 0;
        jRect = 0; jCell = 0; jLaneStep = 0; jStrip = 0;
        if (anyRect) shape(g.r5r1, g.r5c1);
        for (int k = 0; k + 1 < ns; ++k) {
            const SeedRec& s = seeds[idx[k]];
            const SeedRec& t = seeds[idx[k + 1]];
            const int64_t rows = t.from - s.to, cols = t.start - s.end;
            if (rows > 0 && cols > 0) { anyRect = true; maxCols = max(maxCols, (int32_t)cols); shape(rows, cols); }
        }
        if (g.r3r0 <= g.L) { anyRect = true; maxCols = max(maxCols, g.tLen - g.r3c0 + 1); shape(g.L - g.r3r0 + 1, g.tLen - g.r3c0 + 1); }
        g.cls = !anyRect ? CLS_NONE
                : a.forceSerial ? CLS_SERIAL
                : (ns <= NS16 && maxCols <= COLS16) ? CLS_WAVE16
                : (ns <= FILL_NS && maxCols <= FILL_COLS) ? CLS_WAVE : CLS_SERIAL;
        const int hs = g.cls == CLS_WAVE16 ? 4 : 6;
        int64_t area = (g.r5r1 >= 1 && g.r5c1 >= 1) ? stripBytes(g.r5r1, g.r5c1, a.mode, hs) : 0;
        int64_t bord = (g.r5r1 >= 1 && g.r5c1 >= 1) ? (int64_t)g.r5r1 + g.r5c1 : 0;
        for (int k = 0; k < ns; ++k) {
            const SeedRec& s = seeds[idx[k]];
            a.auxSG[cs + k] = make_int4((int32_t)(s.from + 1), (int32_t)(s.start - tStart + 1), (int32_t)(s.to - s.from), 0);
            if (k + 1 < ns) {
                const SeedRec& t = seeds[idx[k + 1]];
                const int64_t rows = t.from - s.to, cols = t.start - s.end;
                a.auxGap[cs + k] = area;
                a.auxGapB[cs + k] = bord;
                if (rows > 0 && cols > 0) {
                    area += stripBytes(rows, cols, a.mode, hs);
                    bord += rows + cols;
                }
            }
        }
        g.r3off = area;
        g.r3boff = bord;
        if (g.r3r0 <= g.L) {
            const int64_t rows = g.L - g.r3r0 + 1, cols = g.tLen - g.r3c0 + 1;
            area += stripBytes(rows, cols, a.mode, hs);
            bord += rows + cols;
        }
        int64_t cellsAll = 0;
        if (g.cls == CLS_SERIAL) {
            /* a serially filled job keeps M, I, D per cell in the tile pool instead of packed
             * decisions: its rectangle offsets count cells */
            cellsAll = (g.r5r1 >= 1 && g.r5c1 >= 1) ? (int64_t)g.r5r1 * g.r5c1 : 0;
            for (int k = 0; k + 1 < ns; ++k) {
                const SeedRec& s = seeds[idx[k]];
                const SeedRec& t = seeds[idx[k + 1]];
                const int64_t rows = t.from - s.to, cols = t.start - s.end;
                a.auxGap[cs + k] = cellsAll;
                if (rows > 0 && cols > 0) cellsAll += rows * cols;
            }
            g.r3off = cellsAll;
            if (g.r3r0 <= g.L) cellsAll += (int64_t)(g.L - g.r3r0 + 1) * (g.tLen - g.r3c0 + 1);
        }
        if (a.dstat && g.cls == CLS_SERIAL) {
            atomicAdd(&a.dstat[0], 1ull);
            atomicAdd(&a.dstat[1], (unsigned long long)cellsAll);
            atomicMax(&a.dstat[2], (unsigned long long)cellsAll);
        }
        if (g.cls == CLS_WAVE16) { nRect += jRect; nCell += jCell; nLaneStep += jLaneStep; nStrip += jStrip; }
        a.geo[job] = g;
        a.jobCls[job] = g.cls;
        const bool wave = g.cls == CLS_WAVE || g.cls == CLS_WAVE16;
        a.tileNeed[job] = wave ? area : 0;
        a.pathNeed[job] = L + g.tLen + 3;
        a.borderNeed[job] = wave ? bord : 0;
        a.serialNeed[job] = cellsAll;
    }
    if (a.dstat) {
        /* per wavefront, then per block: one set of atomics per block */
        __shared__ unsigned long long red[4][4];
        for (int off = 32; off >= 1; off >>= 1) {
            nRect += __shfl_xor(nRect, off, 64);
            nCell += __shfl_xor(nCell, off, 64);
            nLaneStep += __shfl_xor(nLaneStep, off, 64);
            nStrip += __shfl_xor(nStrip, off, 64);
        }
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) { red[w][0] = nRect; red[w][1] = nCell; red[w][2] = nLaneStep; red[w][3] = nStrip; }
        __syncthreads();
        if (threadIdx.x < 4) {
            unsigned long long v = 0;
            for (int k = 0; k < (int)(blockDim.x >> 6); ++k) v += red[k][threadIdx.x];
            if (v) atomicAdd(&a.dstat[4 + threadIdx.x], v);
        }
    }
}

/* ------------------------------------------------------------------ wavefront fill */

template <int MODE>
__global__ void __launch_bounds__(64) dp_fill_kernel(DevIndex ix, DpArgs a) {
    __shared__ int4 sSG[FILL_NS];
    __shared__ int32_t sTM[FILL_COLS + 1], sTI[FILL_COLS + 1], sTD[FILL_COLS + 1];
    __shared__ uint8_t sTg[FILL_COLS];
    const int lane = threadIdx.x;
    for (int64_t w = blockIdx.x; w < a.nWaveJobs; w += gridDim.x) {
        const int64_t job = a.waveJobs[w];
        const JobGeo g = a.geo[job];
        for (int k = lane; k < g.ns; k += 64) sSG[k] = a.auxSG[g.cs + k];
        __syncthreads();
        Geo G;
        makeGeo(G, a, g, sSG, job);
        const uint8_t* q = ((g.strand == FWD) ? a.seqs : a.rcSeqs) + a.offs[g.rid];
        const uint8_t* tg = ix.tgt + ix.tgtOff[g.tid];
        const int64_t chromLen = ix.chromLen[g.tid];
        int32_t bv = NEG, bi = 0, bj = 0;               /* SW: best rectangle cell of this lane */
        int32_t nv = NEG, nc = 0, nst = OP_M;           /* NW: scan of R3's last row (its lane) */
        int32_t cm = NEG, ci = NEG, cd = NEG;           /* corner of the last filled rectangle */

        /* sweep rectangle R (event R.ev), reading its borders at event R.ev - 1 */
        auto fill = [&](const Rc& R) {
            const int32_t Rn = R.r1 - R.r0 + 1, C = R.c1 - R.c0 + 1;
            const int e = R.ev;
            for (int t = lane; t <= C; t += 64) {
                const int32_t j = R.c0 - 1 + t;
                sTM[t] = vM(G, R.r0 - 1, j, e - 1);
                sTI[t] = vI(G, R.r0 - 1, j, e - 1);
                sTD[t] = vD(G, R.r0 - 1, j, e - 1);
                if (t < C) {
                    const int64_t tp = g.tStart + (R.c0 + t) - 1;
                    sTg[t] = (tp >= 0 && tp <= chromLen) ? tg[tp] : (uint8_t)0;
                }
            }
            __syncthreads();
            for (int32_t s0 = 0; s0 < Rn; s0 += 64) {
                const int32_t rows = min(64, Rn - s0);
                const int32_t i = R.r0 + s0 + lane;
                const bool act = lane < rows;
                const bool more = s0 + rows < Rn;
                int32_t cM = NEG, cI = NEG, cD = NEG;
                if (act) {
                    cM = vM(G, i, R.c0 - 1, e - 1);
                    cI = vI(G, i, R.c0 - 1, e - 1);
                    cD = vD(G, i, R.c0 - 1, e - 1);
                }
                const int32_t bM0 = cM, bI0 = cI, bD0 = cD;
                const uint8_t qb = act ? q[i - 1] : (uint8_t)0;
                int32_t pM = NEG, pI = NEG, pD = NEG;
                const int32_t steps = C + rows - 1;
                /* this strip's decision words, packed per lane over spw steps */
                DecWord<MODE>* const dw = (DecWord<MODE>*)G.bits + R.off / (int64_t)sizeof(DecWord<MODE>) +
                                          (int64_t)(s0 >> 6) * stripWords(64, C);
                DecWord<MODE> acc = 0;
                for (int32_t t = 0; t < steps; ++t) {
                    /* up = (i-1, j): row above, one step ago; diagonal = (i-1, j-1): two steps ago */
                    int32_t uM = laneUp1(cM), uI = laneUp1(cI), uD = laneUp1(cD);
                    int32_t dM = pM, dI = pI, dD = pD;
                    const int32_t jj = t - lane;
                    if (lane == 0) {
                        const int32_t x = min(jj, C - 1);
                        uM = sTM[x + 1]; uI = sTI[x + 1]; uD = sTD[x + 1];
                        dM = sTM[x]; dI = sTI[x]; dD = sTD[x];
                    }
                    pM = uM; pI = uI; pD = uD;
                    if (act && jj >= 0 && jj < C) {
                        const int32_t j = R.c0 + jj;
                        const uint8_t tb = sTg[jj];
                        const int32_t sc = ((qb & tb) && qb != 15 && tb != 15) ? G.match : -G.mis;
                        /* max3(addn(dM, sc), addn(dD, sc), addn(dI, sc)): NEG is below every
                         * value, so the largest predecessor decides and -inf stays -inf */
                        const int32_t mx = max3(dM, dD, dI);
                        int32_t m = mx == NEG ? NEG : mx + sc;
                        if (MODE == 1) m = max(m, 0);
                        /* the gap penalties are non-negative (the host refuses others), so a
                         * step from -inf lands at or below NEG and a clamp at NEG is addn / subn:
                         * nI = max(addn(uM, o), addn(uI, -ge)), nD likewise */
                        const int32_t nI = max3(uM + G.o, uI - G.ge, NEG);
                        const int32_t nD = max3(cM + G.o, cD - G.ge, NEG);
                        /* the backtrace's decisions at (i, j), from the values the recurrences
                         * read (src/Alignment.cpp:188-205, src/Alignment.h:577-597):
                         * subn(subn(uM, go), ge) > subn(uI, ge), and the same for D */
                        const uint32_t dec = (uint32_t)matchMax(dM, dI, dD) |
                                             ((max(uM + G.o, NEG) > max(uI - G.ge, NEG)) ? 0u : 4u) |
                                             ((max(cM + G.o, NEG) > max(cD - G.ge, NEG)) ? 0u : 8u) |
                                             ((MODE == 1 && m > 0) ? 16u : 0u);
                        cM = m; cI = nI; cD = nD;
                        acc |= (DecWord<MODE>)dec << decShiftOf(MODE, t);
                        if (i == R.r1) G.border[R.boff + jj] = make_int4(m, nI, nD, 0);
                        if (jj == C - 1) G.border[R.boff + C + (i - R.r0)] = make_int4(m, nI, nD, 0);
                        if (more && lane == rows - 1) { sTM[jj + 1] = m; sTI[jj + 1] = nI; sTD[jj + 1] = nD; }
                        if (MODE == 1) {
                            if (swBetter(m, i, j, bv, bi, bj)) { bv = m; bi = i; bj = j; }
                        } else if (i == G.L) {
                            if (m > nv) { nv = m; nc = j; nst = OP_M; }
                            if (nI > nv) { nv = nI; nc = j; nst = OP_I; }
                        }
                    }
                    /* every 8 steps (and at the strip's end) the lanes store their words side
                     * by side: one contiguous run of rows words */
                    if ((t & 7) == 7 || t == steps - 1) {
                        if (act) dw[(int64_t)(t >> DEC_SH) * rows + lane] = acc;
                        acc = 0;
                    }
                }
                __syncthreads();
                if (more && lane == rows - 1) { sTM[0] = bM0; sTI[0] = bI0; sTD[0] = bD0; }
                if (!more) {
                    cm = __shfl(cM, rows - 1, 64);
                    ci = __shfl(cI, rows - 1, 64);
                    cd = __shfl(cD, rows - 1, 64);
                }
                __syncthreads();
            }
        };

        Rc R = rectR5(G);
        bool prevOk = R.ok;
        if (R.ok) fill(R);
        for (int k = 0; k < g.ns; ++k) {
            int32_t d0;
            if (prevOk) d0 = max(max3(addn(cm, G.match), addn(cd, G.match), addn(ci, G.match)), 0);
            else d0 = seedD0(G, k);
            __syncthreads();
            if (lane == 0) sSG[k].w = d0;
            __syncthreads();
            if (k + 1 < g.ns) {
                R = rectGap(G, k, sSG[k], sSG[k + 1]);
                prevOk = R.ok;
                if (R.ok) fill(R);
            }
        }
        R = rectR3(G);
        if (R.ok) fill(R);
        for (int k = lane; k < g.ns; k += 64) a.auxSG[g.cs + k].w = sSG[k].w;
        for (int off = 32; off >= 1; off >>= 1) {
            const int32_t ov = __shfl_xor(bv, off, 64), oi = __shfl_xor(bi, off, 64), oj = __shfl_xor(bj, off, 64);
            if (swBetter(ov, oi, oj, bv, bi, bj)) { bv = ov; bi = oi; bj = oj; }
        }
        const int nwLane = R.ok ? (G.L - R.r0) % 64 : 0;
        nv = __shfl(nv, nwLane, 64);
        nc = __shfl(nc, nwLane, 64);
        nst = __shfl(nst, nwLane, 64);
        if (lane == 0) {
            JobSum S;
            S.swV = bv; S.swI = bi; S.swJ = bj; S.nwV = nv; S.nwC = nc; S.nwS = nst; S.pad0 = 0; S.pad1 = 0;
            a.sum[job] = S;
        }
        __syncthreads();
    }
}

template __global__ void dp_fill_kernel<0>(DevIndex, DpArgs);
template __global__ void dp_fill_kernel<1>(DevIndex, DpArgs);

/* ------------------------------------------------------------------ 16-lane wavefront fill */

/* Jobs whose rectangles have at most COLS16 columns and whose chains have at most NS16 seeds
 * are filled 16 lanes per job: each 16-lane DPP row of the wavefront runs its own job, taking
 * the next one from a work counter when it is done, so four jobs of different shapes proceed
 * side by side.  Strips are 16 rows (lane = row), swept along anti-diagonals as in
 * dp_fill_kernel; a 16-row strip leaves fewer lanes idle on its ramp diagonals and on partial
 * strips than a 64-row one.  Per 16-lane job, in LDS: the seeds and the row above the current
 * strip (M, I, D and the target bases).  Lane 0 reads the row's entry for its column from
 * LDS every step and takes it as the out-of-row value of the row_shr shift that brings the
 * other lanes their upper neighbours; the target base reaches lane l one step after lane l - 1
 * the same way.  Each group runs a small state machine (next job, next rectangle, next strip, step) in
 * one loop, so the groups need not be in the same phase.  A group claims its next job when it
 * starts one and fetches that job's number and geometry (into LDS) one step per strip end, so
 * that a job start rarely waits on that chain of dependent reads.  The values, decisions and
 * summaries are those of dp_fill_kernel. */
template <int MODE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) dp_fill16_kernel(DevIndex ix, DpArgs a) {
    constexpr int W = COLS16 + 17;   /* the row is read up to entry C + rows (<= COLS16 + 16) */
    __shared__ int4 sSG[4][NS16];
    __shared__ int32_t sTM[4][W], sTI[4][W], sTD[4][W];
    __shared__ uint8_t sTg[4][W];
    __shared__ JobGeo sNextG[4];   /* the next job's geometry, fetched ahead */
    const int lane = threadIdx.x, grp = lane >> 4, l = lane & 15;
    constexpr int GEO_WORDS = (int)(sizeof(JobGeo) / 4);
    static_assert(GEO_WORDS <= 32, "a job's geometry is read as two words per lane");
    int4* const SG = sSG[grp];
    int32_t* const TM = sTM[grp];
    int32_t* const TI = sTI[grp];
    int32_t* const TD = sTD[grp];
    uint8_t* const TG = sTg[grp];
    const int32_t match = a.match, mis = a.mismatch, o = -(a.gapO + a.gapE), ge = a.gapE;
    const uint32_t mm = (uint32_t)(match + mis);

    enum { FETCH, RECT, STRIP, STEP, DONE };
    int phase = FETCH;
    int64_t job = 0;
    JobGeo g{};
    Geo G{};
    const uint8_t* q = nullptr;
    const uint8_t* tg = nullptr;
    int64_t chromLen = 0;
    int ri = 0;
    bool prevOk = false, act = false;
    Rc R{};
    int32_t Rn = 0, C = 0, s0 = 0, rows = 0, t = 0, steps = 0, i = 0;
    int32_t cM = NEG, cI = NEG, cD = NEG, pM = NEG, pI = NEG, pD = NEG;
    int32_t tbv = 0;
    uint8_t qNext = 0;   /* the read base of the lane's row in the next strip */
    uint32_t mq = 0;   /* bit b: the lane's read base matches target code b (b & base != 0) */
    int32_t bM0 = NEG, bI0 = NEG, bD0 = NEG, cm = NEG, ci = NEG, cd = NEG;
    DecWord<MODE> acc = 0;
    DecWord<MODE>* dw = nullptr;
    int32_t bv = NEG, bi = 0, bj = 0, sbv = NEG, sbj = 0;   /* LOCAL: best cell (lane's rows) */
    int32_t nv = NEG, nc = 0, nst = OP_M;                   /* GLOBAL: scan of R3's last row */
    /* the next job's fetch: stage 0 nothing claimed, 1 queue index claimed (nkRaw, in the
     * group's lane 0), 2 job number read (nJob, -1: queue exhausted), 3 geometry words read
     * (gw0, gw1: words l and l + 16), 4 geometry in sNextG */
    int pst = 0;
    unsigned int nkRaw = 0;
    int64_t nJob = -1;
    uint32_t gw0 = 0, gw1 = 0;
    const uint32_t* const geoWords = reinterpret_cast<const uint32_t*>(a.geo);
    uint32_t* const nextWords = reinterpret_cast<uint32_t*>(&sNextG[grp]);

    while (__any(phase != DONE)) {
        /* ---- the next job's fetch advances one stage at each strip end ---- */
        if (phase == STRIP || phase == RECT) {
            if (pst == 1) {
                const unsigned int k = __shfl(nkRaw, grp * 16, 64);
                nJob = (int64_t)k < a.nWave16Jobs ? a.wave16Jobs[k] : -1;
                pst = 2;
            } else if (pst == 2) {
                if (nJob >= 0) {
                    gw0 = l < GEO_WORDS ? geoWords[nJob * GEO_WORDS + l] : 0u;
                    gw1 = l + 16 < GEO_WORDS ? geoWords[nJob * GEO_WORDS + l + 16] : 0u;
                }
                pst = 3;
            } else if (pst == 3) {
                if (nJob >= 0) {
                    if (l < GEO_WORDS) nextWords[l] = gw0;
                    if (l + 16 < GEO_WORDS) nextWords[l + 16] = gw1;
                }
                pst = 4;
            }
        }
        /* ---- control, per group: until it is stepping a strip or out of jobs ---- */
        while (phase == FETCH || phase == RECT || phase == STRIP) {
            if (phase == FETCH) {
                /* this job: from the fetch ahead, finished here from the stage it reached */
                if (pst == 0) {
                    if (l == 0) nkRaw = atomicAdd(a.fill16Next, 1u);
                    pst = 1;
                }
                if (pst == 1) {
                    const unsigned int k = __shfl(nkRaw, grp * 16, 64);
                    nJob = (int64_t)k < a.nWave16Jobs ? a.wave16Jobs[k] : -1;
                }
                const int64_t jb = nJob;
                if (jb >= 0) g = pst == 4 ? sNextG[grp] : a.geo[jb];
                pst = 0;
                if (jb < 0) {
                    phase = DONE;
                } else {
                    job = jb;
                    /* claim the job after this one; its index is read at the next strip end */
                    if (l == 0) nkRaw = atomicAdd(a.fill16Next, 1u);
                    pst = 1;
                    for (int x = l; x < g.ns; x += 16) SG[x] = a.auxSG[g.cs + x];
                    makeGeo(G, a, g, SG, job);
                    q = ((g.strand == FWD) ? a.seqs : a.rcSeqs) + a.offs[g.rid];
                    tg = ix.tgt + ix.tgtOff[g.tid];
                    chromLen = ix.chromLen[g.tid];
                    ri = 0; prevOk = false;
                    bv = NEG; bi = 0; bj = 0;
                    nv = NEG; nc = 0; nst = OP_M;
                    phase = RECT;
                }
            } else if (phase == RECT) {
                /* the next rectangle in event order, after the first diagonal value of each
                 * seed passed (the corner of the rectangle just filled, or the write history) */
                bool found = false;
                while (ri <= g.ns) {
                    if (ri >= 1) {
                        const int k = ri - 1;
                        const int32_t d0 = prevOk ? max(max3(addn(cm, match), addn(cd, match), addn(ci, match)), 0)
                                                  : seedD0(G, k);
                        if (l == 0) SG[k].w = d0;
                    }
                    R = ri == 0 ? rectR5(G) : (ri < g.ns ? rectGap(G, ri - 1, SG[ri - 1], SG[ri]) : rectR3(G));
                    prevOk = R.ok;
                    if (R.ok) { found = true; break; }
                    ++ri;
                }
                if (!found) {
                    /* job end: the seeds' first diagonal values and the summaries */
                    for (int x = l; x < g.ns; x += 16) a.auxSG[g.cs + x].w = SG[x].w;
                    if (MODE == 1) {
                        for (int off = 8; off >= 1; off >>= 1) {
                            const int32_t ov = __shfl_xor(bv, off, 16), oi = __shfl_xor(bi, off, 16),
                                          oj = __shfl_xor(bj, off, 16);
                            if (swBetter(ov, oi, oj, bv, bi, bj)) { bv = ov; bi = oi; bj = oj; }
                        }
                    }
                    if (l == 0) {
                        JobSum S;
                        S.swV = bv; S.swI = bi; S.swJ = bj; S.nwV = nv; S.nwC = nc; S.nwS = nst; S.pad0 = 0; S.pad1 = 0;
                        a.sum[job] = S;
                    }
                    phase = FETCH;
                } else {
                    /* the row above the rectangle (at event R.ev - 1) and its target bases (N
                     * read as 0, so that one AND decides a match).  The row and the columns
                     * R touches lie in the band of the seed before it (ri - 1, -1 for R5) in
                     * both directions, so no band search is needed */
                    C = R.c1 - R.c0 + 1;
                    Rn = R.r1 - R.r0 + 1;
                    const int e = R.ev, kb = ri - 1;
                    /* the first strip's read bases, fetched with the target bases */
                    qNext = l < min(16, Rn) ? q[R.r0 + l - 1] : (uint8_t)0;
                    for (int x0 = 0; x0 <= C; x0 += 64) {
                        /* four target bases per lane in flight, then the row entries (the row is
                         * in seed kb's rows: no rectangle is read) */
                        uint32_t tb4[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int32_t x = x0 + l + 16 * u;
                            const int64_t tp = g.tStart + (R.c0 + x) - 1;
                            tb4[u] = (x < C && tp >= 0 && tp <= chromLen) ? tg[tp] : 0u;
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int32_t x = x0 + l + 16 * u;
                            if (x <= C) {
                                const int32_t j = R.c0 - 1 + x;
                                TM[x] = vMk<false>(G, R.r0 - 1, j, e - 1, kb);
                                TI[x] = vIk<false>(G, R.r0 - 1, j, e - 1, kb);
                                TD[x] = vDkm<false>(G, R.r0 - 1, j, e - 1, kb, kb);
                                if (x < C) TG[x] = tb4[u] == 15u ? (uint8_t)0 : (uint8_t)tb4[u];
                            }
                        }
                    }
                    s0 = 0;
                    phase = STRIP;
                }
            } else {
                rows = min(16, Rn - s0);
                i = R.r0 + s0 + l;
                act = l < rows;
                const int e = R.ev, kb = ri - 1;
                cM = NEG; cI = NEG; cD = NEG;
                if (act) {
                    /* the strip's left column, at the event before R's (R itself unfilled) */
                    cM = vMk<false>(G, i, R.c0 - 1, e - 1, kb);
                    cI = vIk<false>(G, i, R.c0 - 1, e - 1, kb);
                    cD = vDkm<false>(G, i, R.c0 - 1, e - 1, kb, kb);
                }
                bM0 = cM; bI0 = cI; bD0 = cD;
                /* this strip's read base (fetched a strip ahead), the next strip's in flight */
                const uint32_t qb = act ? qNext : 0u;
                qNext = s0 + 16 + l < Rn ? q[i + 16 - 1] : (uint8_t)0;
                const uint32_t qm = qb == 15 ? 0u : qb;
                mq = ((qm & 1u) ? 0xAAAAu : 0u) | ((qm & 2u) ? 0xCCCCu : 0u) | ((qm & 4u) ? 0xF0F0u : 0u) |
                     ((qm & 8u) ? 0xFF00u : 0u);
                pM = TM[0]; pI = TI[0]; pD = TD[0];
                tbv = 0;
                t = 0;
                steps = C + rows - 1;
                dw = (DecWord<MODE>*)G.bits + R.off / (int64_t)sizeof(DecWord<MODE>) + (int64_t)(s0 >> 4) * stripWords(16, C);
                acc = 0;
                sbv = NEG; sbj = 0;
                phase = STEP;
            }
        }
        /* ---- steps: as many as the stepping group nearest to its strip end has left, so
         * that no group changes phase inside the loop ---- */
        int32_t rem = phase == STEP ? steps - t : 0x7fffffff;
        /* each group's value is uniform in its 16 lanes: read lanes 0, 16, 32, 48 into scalars */
        rem = min(min(__builtin_amdgcn_readlane(rem, 0), __builtin_amdgcn_readlane(rem, 16)),
                  min(__builtin_amdgcn_readlane(rem, 32), __builtin_amdgcn_readlane(rem, 48)));
        if (rem == 0x7fffffff) continue;   /* no group stepping: all are out of jobs */
        if (phase == STEP) {
            /* lane 0's entries of the row above the strip, from LDS one step ahead: entry t + 1
             * (and target base t) for step t.  The bottom lane has written the strip's own row
             * only up to entry t - rows + 1 when step t reads entry t + 2 */
            int32_t aM = TM[t + 1], aI = TI[t + 1], aD = TD[t + 1], aT = TG[t];
            auto step = [&]() {
                /* up = (i - 1, j): row above, one step ago (lane 0: the row above the strip);
                 * diagonal = (i - 1, j - 1): the up value of the step before */
                const int32_t uM = rowUp1(aM, cM), uI = rowUp1(aI, cI), uD = rowUp1(aD, cD);
                tbv = rowUp1(aT, tbv);
                aM = TM[t + 2]; aI = TI[t + 2]; aD = TD[t + 2]; aT = TG[t + 1];
                const int32_t dM = pM, dI = pI, dD = pD;
                pM = uM; pI = uI; pD = uD;
                const int32_t jj = t - l;
                const bool valid = act && (uint32_t)jj < (uint32_t)C;
                /* the recurrences of dp_fill_kernel: -inf (NEG) absorbs, gaps are non-negative,
                 * and every finite value lies far above NEG, so the comparisons of the
                 * backtrace's decisions need no clamping.  Score: -mis, plus match + mis where
                 * the bases match (bit tbv of mq) */
                const int32_t mx = max3(dM, dD, dI);
                const int32_t sv = mx - mis + (int32_t)__umul24(__builtin_amdgcn_ubfe(mq, (uint32_t)tbv, 1), mm);
                int32_t m = selectI(mx == NEG, NEG, sv);
                if (MODE == 1) m = max(m, 0);
                const int32_t a1 = uM + o, b1 = uI - ge, a2 = cM + o, b2 = cD - ge;
                const int32_t nI = max3(a1, b1, NEG), nD = max3(a2, b2, NEG);
                /* matchMax with its quirk: I wins over M, otherwise D if above M */
                if constexpr (MODE == 0) {
                    /* the nibble D-ext, I-ext, M source (2 bits) shifted in bit by bit: after 8
                     * steps the word holds 8 nibbles, the first step's on top */
                    const bool iw = dI > dM;
                    acc = pushBit(acc, a2 <= b2);
                    acc = pushBit(acc, a1 <= b1);
                    acc = pushBit(acc, dD > selectI(iw, INT32_MAX, dM));   /* D above M, I not */
                    acc = pushBit(acc, iw);
                } else {
                    uint32_t dec = dI > dM ? 1u : (dD > dM ? 2u : 0u);
                    dec |= (a1 > b1 ? 0u : 4u) | (a2 > b2 ? 0u : 8u) | (m > 0 ? 16u : 0u);
                    /* bytes enter the word from the top: after 8 steps the first step's is at
                     * the bottom */
                    acc = (acc >> 8) | ((uint64_t)dec << 56);
                }
                cM = valid ? m : cM;
                cI = valid ? nI : cI;
                cD = valid ? nD : cD;
                if (MODE == 1 && valid && m > sbv) { sbv = m; sbj = R.c0 + jj; }
                if (valid && l == rows - 1) { TM[jj + 1] = m; TI[jj + 1] = nI; TD[jj + 1] = nD; }
                /* every 8 steps and at the strip's end the lanes store their words side by side
                 * (a partial word aligned as a full one) */
                const int32_t pos = t & 7;
                if (act && (pos == 7 || t == steps - 1)) {
                    if constexpr (MODE == 0) dw[(int64_t)(t >> DEC_SH) * rows + l] = acc << (4 * (7 - pos));
                    else dw[(int64_t)(t >> DEC_SH) * rows + l] = acc >> (8 * (7 - pos));
                }
                ++t;
            };
            /* two steps per iteration: the registers of the values carried from one step to
             * the next alternate instead of being copied */
            int32_t r = 0;
            for (; r + 2 <= rem; r += 2) { step(); step(); }
            if (r < rem) step();
            if (t == steps) {
                /* strip end: the lanes' last-column cells, the corner of the next strip's top
                 * row, the LOCAL best of each lane */
                if (act) G.border[R.boff + C + (i - R.r0)] = make_int4(cM, cI, cD, 0);
                if (MODE == 1 && act && sbv != NEG && swBetter(sbv, i, sbj, bv, bi, bj)) { bv = sbv; bi = i; bj = sbj; }
                if (l == rows - 1) { TM[0] = bM0; TI[0] = bI0; TD[0] = bD0; }
                if (s0 + rows < Rn) {
                    s0 += 16;
                    phase = STRIP;
                } else {
                    /* rectangle end: its corner, its last row to the border, and in GLOBAL mode
                     * the scan of R3's last row (row L): first maximum in column order, M before
                     * I, strict '>' from -inf */
                    cm = __shfl(cM, grp * 16 + rows - 1, 64);
                    ci = __shfl(cI, grp * 16 + rows - 1, 64);
                    cd = __shfl(cD, grp * 16 + rows - 1, 64);
                    for (int x = l; x < C; x += 16) G.border[R.boff + x] = make_int4(TM[x + 1], TI[x + 1], TD[x + 1], 0);
                    if (MODE == 0 && ri == g.ns) {
                        int32_t v = NEG, vx = 0x7fffffff, vs = OP_M;
                        for (int x = l; x < C; x += 16) {
                            const int32_t mv = TM[x + 1], iv = TI[x + 1];
                            if (mv > v) { v = mv; vx = x; vs = OP_M; }
                            if (iv > v) { v = iv; vx = x; vs = OP_I; }
                        }
                        for (int off = 8; off >= 1; off >>= 1) {
                            const int32_t ov = __shfl_xor(v, off, 16), ox = __shfl_xor(vx, off, 16),
                                          os = __shfl_xor(vs, off, 16);
                            if (ov > v || (ov == v && (ox < vx || (ox == vx && os < vs)))) { v = ov; vx = ox; vs = os; }
                        }
                        if (v != NEG) { nv = v; nc = R.c0 + vx; nst = vs; }
                    }
                    prevOk = true;
                    ++ri;
                    phase = RECT;
                }
            }
        }
    }
}

template __global__ void dp_fill16_kernel<0>(DevIndex, DpArgs);
template __global__ void dp_fill16_kernel<1>(DevIndex, DpArgs);

/* ------------------------------------------------------------------ trace */

/* at least 4 waves per SIMD (at most 128 VGPRs).  The band lookups (vM / vI / vD, cellRect,
 * rowSeg, Rc's methods, ...) are inlined, so the job geometry G stays in registers instead of
 * private memory that every backtrace step would wait on.  At 4 waves 36 VGPRs per lane
 * spill, and the kernel still runs faster than at 3 waves without spills (2.35 vs 2.64 ms
 * per 256k-read batch of the 4 Gbp workload): more waves hide more of the memory waits */
__global__ void __launch_bounds__(256, 4) dp_trace_kernel(DevIndex ix, DpArgs a) {
    /* the quality terms of evaluate, by quality value, in LDS (one read per base) */
    __shared__ double sQM[256], sQX[256], sQG[256];
    for (int t = threadIdx.x; t < 256; t += blockDim.x) {
        sQM[t] = a.qMatch[t];
        sQX[t] = a.qMis[t];
        sQG[t] = a.qGapOpen[t];
    }
    __syncthreads();
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < a.nJobs;
         idx += (int64_t)gridDim.x * blockDim.x) {
        /* jobs by class and area, largest first: serial, one-wavefront, 16-lane, none, so that
         * the lanes of a wavefront trace jobs of similar shape */
        int64_t job = idx;
        if (a.serialJobs) {
            int64_t x = idx;
            if (x < a.nSerialJobs) job = a.serialJobs[x];
            else if ((x -= a.nSerialJobs) < a.nWaveJobs) job = a.waveJobs[x];
            else if ((x -= a.nWaveJobs) < a.nWave16Jobs) job = a.wave16Jobs[x];
            else job = a.noneJobs[x - a.nWave16Jobs];
        }
        const JobGeo g = a.geo[job];
        int4* sgw = a.auxSG + g.cs;
        Geo G;
        makeGeo(G, a, g, sgw, job);
        const bool fwd = g.strand == FWD;
        const uint8_t* q = (fwd ? a.seqs : a.rcSeqs) + a.offs[g.rid];
        const uint8_t* qual = (fwd ? a.quals : a.rcQuals) + a.offs[g.rid];
        const uint8_t* tg = ix.tgt + ix.tgtOff[g.tid];
        const int64_t chromLen = ix.chromLen[g.tid];
        auto tb = [&](int64_t t) -> uint8_t { return (t >= 0 && t <= chromLen) ? tg[t] : (uint8_t)0; };
        const int32_t L = g.L, ns = g.ns;
        JobSum S;
        S.swV = NEG; S.swI = 0; S.swJ = 0; S.nwV = NEG; S.nwC = 0; S.nwS = OP_M;
        if (g.cls == CLS_WAVE || g.cls == CLS_WAVE16) {
            S = a.sum[job];
        } else {
            /* calculateScores(chain) (src/Alignment.cpp:54-70) in event order, one lane */
            auto fillSerial = [&](const Rc& R) {
                if (!R.ok) return;
                const int e = R.ev;
                for (int32_t j = R.c0; j <= R.c1; ++j) {
                    const uint8_t t = tb(g.tStart + j - 1);
                    int32_t uM = vM(G, R.r0 - 1, j, e - 1), uI = vI(G, R.r0 - 1, j, e - 1);
                    for (int32_t i = R.r0; i <= R.r1; ++i) {
                        const uint8_t qb = q[i - 1];
                        const int32_t sc = ((qb & t) && qb != 15 && t != 15) ? G.match : -G.mis;
                        int32_t mD, iD, dD, mL, dL;
                        if (j > R.c0 && i > R.r0) { const int4 v = G.tile[R.idx(i - 1, j - 1)]; mD = v.x; iD = v.y; dD = v.z; }
                        else { mD = vM(G, i - 1, j - 1, e - 1); iD = vI(G, i - 1, j - 1, e - 1); dD = vD(G, i - 1, j - 1, e - 1); }
                        if (j > R.c0) { const int4 v = G.tile[R.idx(i, j - 1)]; mL = v.x; dL = v.z; }
                        else { mL = vM(G, i, j - 1, e - 1); dL = vD(G, i, j - 1, e - 1); }
                        int32_t m = max3(addn(mD, sc), addn(dD, sc), addn(iD, sc));
                        if (G.mode == 1) m = max(m, 0);
                        const int32_t nI = max(addn(uM, G.o), addn(uI, -G.ge));
                        const int32_t nD = max(addn(mL, G.o), addn(dL, -G.ge));
                        G.tile[R.idx(i, j)] = make_int4(m, nI, nD, 0);
                        uM = m; uI = nI;
                        if (G.mode == 1) {
                            if (swBetter(m, i, j, S.swV, S.swI, S.swJ)) { S.swV = m; S.swI = i; S.swJ = j; }
                        } else if (i == L) {
                            if (m > S.nwV) { S.nwV = m; S.nwC = j; S.nwS = OP_M; }
                            if (nI > S.nwV) { S.nwV = nI; S.nwC = j; S.nwS = OP_I; }
                        }
                    }
                }
            };
            if (g.cls == CLS_SERIAL) fillSerial(rectR5(G));
            for (int k = 0; k < ns; ++k) {
                sgw[k].w = seedD0(G, k);
                if (g.cls == CLS_SERIAL && k + 1 < ns) fillSerial(rectGap(G, k, sgw[k], sgw[k + 1]));
            }
            if (g.cls == CLS_SERIAL) fillSerial(rectR3(G));
        }
        /* cells the reference's calculateScores touches */
        int64_t cells = rectR5(G).area() + rectR3(G).area();
        for (int k = 0; k < ns; ++k) {
            const int4 s = G.sg[k];
            cells += s.z;
            if (s.x + s.z <= L && s.y + s.z <= g.tLen) cells += 1;
            if (k + 1 < ns) cells += rectGap(G, k, s, G.sg[k + 1]).area();
        }
        const int T = 2 * ns + 1;
        const int64_t pathCap = (int64_t)L + g.tLen + 3;
        /* ---- backtrace (src/Alignment.cpp:173-276) ----
         * The path is kept as runs of equal ops (BAM CIGAR words), produced end -> start
         * into the tail of the job's CIGAR slot: slot words [0, cgCap) */
        uint32_t* cg = a.cigar + a.pathOff[job] + 2 * job;
        const int64_t cgCap = pathCap + 2;
        int64_t nRun = 0;           /* runs written at cg[cgCap - 1], cg[cgCap - 2], ... */
        uint32_t runOp = 0xffu, runLen = 0;
        auto push = [&](uint32_t op, uint32_t n) {
            if (op == runOp) { runLen += n; return; }
            if (runLen) cg[cgCap - 1 - nRun++] = (runLen << 4) | runOp;
            runOp = op; runLen = n;
        };
        int64_t np = 0;
        int32_t alnTo, alnEnd, i, j;
        int32_t best;
        int32_t s = OP_M;
        const int32_t go = G.go, ge = G.ge;
        if (G.mode == 0) {
            /* scan of the last row, M before I, strict '>' */
            best = NEG; alnEnd = 0; alnTo = L;
            const Rc R3 = rectR3(G);
            const int32_t cEnd = R3.ok ? R3.c0 - 1 : g.tLen;
            const int kL = rowSeg(G, L);
            /* Row L lies in the band of the last seed (every seed starts at or above it).
             * Unless row L is the row just after a seed (whose I row is overwritten from
             * the M row above), I is -inf on row L at every column >= 1, and M is -inf
             * except on the last seed's diagonal and at step-c cells (vMk / vIk with
             * i = L, e = T): -inf never passes the strict '>' from best = -inf, so only
             * those columns are evaluated, in column order. */
            bool general = kL < 0;
            int32_t cand[3] = {-1, -1, -1};
            if (!general) {
                const int4 sk = G.sg[kL];
                const bool inSeedRows = L < sk.x + sk.z;
                const bool stepRow = L == sk.x + sk.z;
                bool prevStepRow = false;
                int4 pk = make_int4(0, 0, 0, 0);
                if (kL >= 1) { pk = G.sg[kL - 1]; prevStepRow = L == pk.x + pk.z; }
                if (stepRow || prevStepRow) general = true;
                else if (inSeedRows) cand[0] = sk.y + (L - sk.x);
            }
            if (general) {
                if (a.dstat) atomicAdd(&a.dstat[3], 1ull);
                for (int32_t c = 1; c <= cEnd; ++c) {
                    const int32_t m = vMk(G, L, c, T, kL);
                    if (m > best) { best = m; alnEnd = c; s = OP_M; }
                    const int32_t x = vIk(G, L, c, T, kL);
                    if (x > best) { best = x; alnEnd = c; s = OP_I; }
                }
            } else {
                for (int t = 0; t < 3; ++t) {
                    const int32_t c = cand[t];
                    if (c < 1 || c > cEnd) continue;
                    const int32_t m = vMk(G, L, c, T, kL);
                    if (m > best) { best = m; alnEnd = c; s = OP_M; }
                    const int32_t x = vIk(G, L, c, T, kL);
                    if (x > best) { best = x; alnEnd = c; s = OP_I; }
                }
            }
            if (R3.ok && S.nwV > best) { best = S.nwV; alnEnd = S.nwC; s = S.nwS; }
        } else {
            /* M.maxCoeff: first maximum in column-major order */
            best = 0; int32_t bi = 0, bj = 0;
            auto visit = [&](int32_t v, int32_t ci, int32_t cj) {
                if (swBetter(v, ci, cj, best, bi, bj)) { best = v; bi = ci; bj = cj; }
            };
            if (S.swV != NEG) visit(S.swV, S.swI, S.swJ);
            for (int k = 0; k < ns; ++k) {
                const int4 sd = G.sg[k];
                const int32_t t = G.match > 0 ? sd.z - 1 : 0;   /* the diagonal is monotone in its index */
                visit(sd.w + t * G.match, sd.x + t, sd.y + t);
                /* the step-c cell; inside a rectangle it is already in the rectangle summary */
                Rc Rs;
                if (sd.x + sd.z <= L && sd.y + sd.z <= g.tLen && !cellRect(G, sd.x + sd.z, sd.y + sd.z, Rs))
                    visit(vM(G, sd.x + sd.z, sd.y + sd.z, T), sd.x + sd.z, sd.y + sd.z);
            }
            alnTo = bi; alnEnd = bj;
        }
        i = alnTo; j = alnEnd;
        Rc Rb;                  /* the rectangle of the previous step (valid when rbOk) */
        bool rbOk = false;
        uint64_t decW = 0;
        int64_t decWi = -1;
        while (true) {
            /* inside a rectangle of a wavefront-filled job, decisions come from its bytes */
            bool inR = false;
            if (!G.serial) {
                if (rbOk && i >= Rb.r0 && i <= Rb.r1 && j >= Rb.c0 && j <= Rb.c1) inR = true;
                else rbOk = inR = cellRect(G, i, j, Rb);
            }
            uint32_t dec = 0u;
            if (inR) {
                int sft;
                const int64_t wi = Rb.decWord(i, j, G.mode, G.hs, sft);
                if (wi != decWi) {
                    decWi = wi;
                    decW = G.mode == 0 ? (uint64_t)((const uint32_t*)G.bits)[wi] : ((const uint64_t*)G.bits)[wi];
                }
                dec = (uint32_t)(decW >> sft) & (G.mode == 0 ? 15u : 255u);
            }
            const bool go_on = G.mode == 0 ? i > 0 : (s != OP_M || (inR ? (dec & 16u) != 0 : vM(G, i, j, T) > 0));
            if (!go_on || np >= pathCap) break;
            if (!inR && s == OP_M && G.match > 0 && i >= 0) {
                /* (a rectangle's rows are never a seed's rows, so only cells outside the
                 * rectangles can lie on a seed diagonal) */
                /* inside a seed diagonal the step is M with -inf I/D neighbours: cross to its second cell */
                const int k = rowSeg(G, i);
                if (k >= 0) {
                    const int4 sd = G.sg[k];
                    const int32_t t = i - sd.x;
                    if (i < sd.x + sd.z && j - sd.y == t && t >= 2) {
                        const int64_t jmp = min((int64_t)(t - 1), pathCap - np);
                        push(OP_M, (uint32_t)jmp);
                        np += jmp; i -= (int32_t)jmp; j -= (int32_t)jmp;
                        continue;
                    }
                }
            }
            push((uint32_t)s, 1u);
            np++;
            if (inR) {
                if (s == OP_M) { s = (int32_t)(dec & 3u); i--; j--; }
                else if (s == OP_I) { s = (dec & 4u) ? OP_I : OP_M; i--; }
                else { s = (dec & 8u) ? OP_D : OP_M; j--; }
            } else if (s == OP_M) { s = matchMax(vM(G, i - 1, j - 1, T), vI(G, i - 1, j - 1, T), vD(G, i - 1, j - 1, T)); i--; j--; }
            else if (s == OP_I) { s = subn(subn(vM(G, i - 1, j, T), go), ge) > subn(vI(G, i - 1, j, T), ge) ? OP_M : OP_I; i--; }
            else { s = subn(subn(vM(G, i, j - 1, T), go), ge) > subn(vD(G, i, j - 1, T), ge) ? OP_M : OP_D; j--; }
            if (G.mode == 0) { if (G.hazard && j < -1) break; }
            else if (G.hazard && (i < -1 || j < -1)) break;
        }
        if (runLen) cg[cgCap - 1 - nRun++] = (runLen << 4) | runOp;
        /* the runs sit at cg[cgCap - nRun, cgCap) in start -> end order */
        const int64_t alnFrom = i, alnStart = g.tStart + j;
        AlnRec rec;
        rec.rid = g.rid;
        rec.tid = g.tid; rec.strand = g.strand; rec.hazard = G.hazard | (G.interior ? 2 : 0);
        rec.tStart = g.tStart; rec.tEnd = g.tEnd;
        rec.alnFrom = alnFrom; rec.alnTo = alnTo; rec.alnStart = alnStart; rec.alnEnd = alnEnd + g.tStart;
        rec.score = best;
        rec.cells = cells;
        /* ---- getAlnMDTag (src/Alignment.cpp:305-353) and evaluate (:446-489), over the runs ---- */
        char* md = a.md + 3 * a.pathOff[job] + 8 * job;
        int64_t ml = 0;
        uint32_t prev = 0xffffffffu, run = 0;
        double x = 0;
        for (int64_t c = 0; c < alnFrom; ++c) x += -a.clip;
        int64_t qi = alnFrom, tj = alnStart;
        /* read, quality and target bytes come from 8-byte words (buffers padded) */
        Bytes8 sQ, sV, sT;
        for (int64_t r = cgCap - nRun; r < cgCap; ++r) {
            const uint32_t rop = cg[r] & 15u, rlen = cg[r] >> 4;
            uint32_t u = 0;
            if (rop == OP_M) {
                /* M runs 8 bases at a time where all 8 read and target bytes are in range: the
                 * words and table terms are loaded first, then summed in base order */
                while (u + 8 <= rlen && qi >= 0 && qi + 8 <= L && tj >= 0 && tj + 8 <= chromLen + 1) {
                    const uint64_t qw = load8u(q + qi), tw = load8u(tg + tj), vw = load8u(qual + qi);
                    const uint64_t both = qw & tw;
                    double tv[8];
                    uint32_t eqm = 0;
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const uint32_t qv = (uint32_t)(vw >> (8 * k)) & 0xffu;
                        const bool eq = ((both >> (8 * k)) & 0xffu) != 0;
                        eqm |= (eq ? 1u : 0u) << k;
                        tv[k] = eq ? sQM[qv] : sQX[qv];
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k) x += tv[k];
                    if (eqm == 0xffu) {
                        run += 8;
                        prev = OP_EQ;
                    } else {
                        for (int k = 0; k < 8; ++k) {
                            if ((eqm >> k) & 1u) { run++; prev = OP_EQ; }
                            else {
                                ml += putNum(md, ml, run);
                                run = 0;
                                md[ml++] = decodeBase((uint8_t)(tw >> (8 * k)));
                                prev = OP_X;
                            }
                        }
                    }
                    qi += 8; tj += 8; u += 8;
                }
            }
            for (; u < rlen; ++u) {
                const bool qin = qi >= 0 && qi < L;
                const uint8_t qb = qin ? (uint8_t)sQ.get(q + qi) : (uint8_t)0;
                const uint8_t tbv = (tj >= 0 && tj <= chromLen) ? (uint8_t)sT.get(tg + tj) : (uint8_t)0;
                const int qv = qin ? (int)sV.get(qual + qi) : 0;
                uint32_t op = rop;
                if (op == OP_M) {
                    x += (qb & tbv) ? sQM[qv] : sQX[qv];
                    op = (qb & tbv) ? OP_EQ : OP_X;
                } else {
                    if (u == 0) x += sQG[qv];   /* the first op of a run opens a gap */
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
        }
        ml += putNum(md, ml, run);
        for (int64_t c = alnTo; c < L; ++c) x += -a.clip;
        /* ---- getAlnCigar (src/Alignment.cpp:278-303): clips around the runs, moved to the
         * head of the slot (the head is written upwards from below the runs' start) ---- */
        int32_t nc = 0;
        const int32_t clip5 = (int32_t)(alnFrom - 0), clip3 = (int32_t)(L - alnTo);
        if (clip5 > 0) cg[nc++] = ((uint32_t)clip5 << 4) | OP_S;
        for (int64_t r = cgCap - nRun; r < cgCap; ++r) cg[nc++] = cg[r];
        if (clip3 > 0) cg[nc++] = ((uint32_t)clip3 << 4) | OP_S;
        rec.nCigar = nc;
        rec.mdLen = (int32_t)ml;
        rec.log10P = x;
        rec.pad = (int32_t)np;
        a.out[job] = rec;
    }
}
