/*
 * msgs_device.h -- HBM layout of the FMD index and the device-side FMD primitives.
 *
 * Occ blocks.  The BWT is cut into 128-position blocks of 64 bytes, one cache
 * sector each:
 *   cnt[4]   number of A, C, G, T in BWT[0, 128*k)            (16 B)
 *   spec[4]  bit j set when BWT[128*k + j] is '\0' or N         (16 B)
 *   lo[4]    low bit of the 2-bit code of position j (A=0 C=1 G=2 T=3;
 *            '\0' and N stored as 0), bit j%32 of word j/32      (16 B)
 *   hi[4]    high bit of the same codes                          (16 B)
 * The four bitmaps share one layout, so a rank query masks each word once and
 * counts C+T, G+T, T and specials with four popcounts per word.  A rank query at
 * any i touches exactly one block.  '\0' counts come from
 * positions minus the four base counts minus N positions; N positions (absent in
 * the synthetic workloads) are kept in a sorted side array.
 *
 * Narrow indexes (BD text shorter than 2^31) hold absolute counts in cnt[] and 32-bit
 * SA / ISA rows.  Wide indexes (up to 2^34) hold 64-bit SA / ISA rows, and cnt[] counts
 * from the start of the block's superblock (2^supShift positions, at most 2^31); the
 * superblock's own counts sit in a small side table (sup, four per superblock), so a
 * block is still one 64-byte sector.
 *
 * The rank semantics are those of WaveletTreeRRR::rank
 * (src/libsds/src/WaveletTreeRRR.cpp:101-125): count in BWT[0..i] inclusive, and
 * size_t(-1) (here -1) for i outside [0, n).
 */
#ifndef MSGS_DEVICE_H
#define MSGS_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

struct __attribute__((aligned(64))) OccBlock {
    uint32_t cnt[4];
    uint32_t spec[4];
    uint32_t lo[4];
    uint32_t hi[4];
};

/* over positions 0..j of a block: ones of the low plane (C, T), of the high plane (G, T),
 * of both (T), and specials */
__device__ __forceinline__ void blockCounts(const uint4& sp, const uint4& pl, const uint4& ph, uint32_t j,
                                            uint32_t& nLo, uint32_t& nHi, uint32_t& nT, uint32_t& nSp) {
    const uint32_t wj = j >> 5, part = (2u << (j & 31)) - 1u;   /* bits 0 .. j%32 */
    const uint32_t S[4] = {sp.x, sp.y, sp.z, sp.w};
    const uint32_t Lw[4] = {pl.x, pl.y, pl.z, pl.w};
    const uint32_t Hw[4] = {ph.x, ph.y, ph.z, ph.w};
    nLo = 0; nHi = 0; nT = 0; nSp = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t m = k < wj ? 0xffffffffu : (k == wj ? part : 0u);
        const uint32_t a = Lw[k] & m, b = Hw[k] & m;
        nLo += __popc(a);
        nHi += __popc(b);
        nT += __popc(a & b);
        nSp += __popc(S[k] & m);
    }
}

/* Everything a kernel needs to read the index; passed by value. */
struct DevIndex {
    const OccBlock* occ;
    const void* sa;            /* full suffix array: uint32_t rows (narrow) or int64_t (wide) */
    const void* isa;           /* inverse suffix array (ISA[SA[i]] = i), same width */
    int32_t wide;              /* 1: 64-bit rows and superblock counts */
    int32_t supShift;          /* wide: log2 of the superblock length in positions */
    const int64_t* sup;        /* wide: A, C, G, T counts before each superblock */
    const uint8_t* bd;         /* BD text */
    const int64_t* nPos;       /* sorted BWT positions holding N */
    int64_t nN;
    int64_t n;                 /* BD text length */
    int64_t C[17];             /* cumulative counts by nt16 code */
    double ll[16];             /* loglik table (src/FMDIndex.h:195-197) */
    const int64_t* chromBD;    /* BD start of each chrom, nChrom+1 entries */
    int32_t nChrom;
    const uint8_t* tgt;        /* chrom sequences, each followed by a 0 byte */
    const int64_t* tgtOff;     /* start of chrom t in tgt */
    const int64_t* chromLen;   /* length of chrom t (without the terminator) */
};

/* SA / ISA row k of either width (a uniform branch on the index kind) */
__device__ __forceinline__ int64_t saAt(const DevIndex& ix, int64_t k) {
    return ix.wide ? ((const int64_t*)ix.sa)[k] : (int64_t)((const uint32_t*)ix.sa)[k];
}
__device__ __forceinline__ int64_t isaAt(const DevIndex& ix, int64_t k) {
    return ix.wide ? ((const int64_t*)ix.isa)[k] : (int64_t)((const uint32_t*)ix.isa)[k];
}
/* the same rows for a search kernel compiled for one index kind: R = int32_t reads a narrow
 * index, R = int64_t a wide one */
template <class R> __device__ __forceinline__ R saRow(const DevIndex& ix, int64_t k) {
    if constexpr (sizeof(R) == 8) return ((const int64_t*)ix.sa)[k];
    else return (R)((const uint32_t*)ix.sa)[k];
}
template <class R> __device__ __forceinline__ R isaRow(const DevIndex& ix, int64_t k) {
    if constexpr (sizeof(R) == 8) return ((const int64_t*)ix.isa)[k];
    else return (R)((const uint32_t*)ix.isa)[k];
}

__device__ __forceinline__ bool amb(uint8_t b) { return !(b == 1 || b == 2 || b == 4 || b == 8); }
__device__ __forceinline__ int toInt(uint8_t b) { return b == 1 ? 1 : b == 2 ? 2 : b == 4 ? 3 : b == 8 ? 4 : 0; }
__device__ __forceinline__ uint8_t comp16(uint8_t b) {
    /* base2comp (src/DNAalphabet.cpp): reverse the 4 bits */
    return (uint8_t)(((b & 1) << 3) | ((b & 2) << 1) | ((b & 4) >> 1) | ((b & 8) >> 3));
}

/* per byte: 0 if the byte is one of the nt16 codes A, C, G, T (1, 2, 4, 8), else nonzero */
__device__ __forceinline__ uint64_t notBasic8(uint64_t v) {
    uint64_t x = v - ((v >> 1) & 0x5555555555555555ull);
    x = (x & 0x3333333333333333ull) + ((x >> 2) & 0x3333333333333333ull);
    x = (x + (x >> 4)) & 0x0f0f0f0f0f0f0f0full;   /* per-byte popcount */
    return (x ^ 0x0101010101010101ull) | (v & 0xf0f0f0f0f0f0f0f0ull);
}

/* the 8 bytes at buf[a .. a+8) from two aligned words (buf padded by 8 bytes past its end) */
__device__ __forceinline__ uint64_t load8(const uint8_t* buf, int64_t a) {
    const uint64_t* w = (const uint64_t*)(buf + (a & ~(int64_t)7));
    const uint32_t sh = 8u * (uint32_t)(a & 7);
    const uint64_t w0 = w[0];
    return sh ? (w0 >> sh) | (w[1] << (64 - sh)) : w0;
}

/* number of leading positions u < len at which read byte rd[u] is A/C/G/T and equals
 * text byte tx[u] (both buffers padded by 8 bytes) */
__device__ __forceinline__ int64_t fwdMatch(const uint8_t* tx, int64_t tp, const uint8_t* rd, int64_t rp, int64_t len) {
    for (int64_t u = 0; u < len; u += 8) {
        const uint64_t v = load8(rd, rp + u);
        const uint64_t bad = notBasic8(v) | (v ^ load8(tx, tp + u));
        if (bad) return min(len, u + (int64_t)(__builtin_ctzll(bad) >> 3));
    }
    return len;
}

/* # of N in BWT[0..i] */
__device__ __forceinline__ int64_t countN(const DevIndex& ix, int64_t i) {
    int64_t lo = 0, hi = ix.nN;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (ix.nPos[mid] <= i) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* rank of '\0', A, C, G, T at i from an already-loaded block */
__device__ __forceinline__ void occFromBlock(const DevIndex& ix, const uint4& h, const uint4& sp, const uint4& pl,
                                             const uint4& ph, int64_t i, int64_t* r) {
    const uint32_t j = (uint32_t)(i & 127);
    uint32_t nLo, nHi, nT, nSp;
    blockCounts(sp, pl, ph, j, nLo, nHi, nT, nSp);
    const uint32_t nC = nLo - nT, nG = nHi - nT;
    const uint32_t nA = (j + 1) - nLo - nHi + nT;
    int64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    if (ix.wide) {
        const int64_t* sb = ix.sup + 4 * (i >> ix.supShift);
        s0 = sb[0]; s1 = sb[1]; s2 = sb[2]; s3 = sb[3];
    }
    r[1] = s0 + h.x + nA - nSp;
    r[2] = s1 + h.y + nC;
    r[3] = s2 + h.z + nG;
    r[4] = s3 + h.w + nT;
    int64_t special = (i & ~(int64_t)127) - (s0 + s1 + s2 + s3 + (int64_t)h.x + h.y + h.z + h.w) + nSp;
    r[0] = ix.nN ? special - countN(ix, i) : special;
}

__device__ __forceinline__ void occ5(const DevIndex& ix, int64_t i, int64_t* r) {
    if (i < 0 || i >= ix.n) { r[0] = r[1] = r[2] = r[3] = r[4] = -1; return; }
    const uint4* b = reinterpret_cast<const uint4*>(ix.occ + (i >> 7));
    uint4 h = b[0], sp = b[1], c0 = b[2], c1 = b[3];
    occFromBlock(ix, h, sp, c0, c1, i, r);
}

/* Both rank rows of one extension; the two block loads are issued together, and
 * only one when both positions share a block. */
__device__ __forceinline__ void occPair(const DevIndex& ix, int64_t a, int64_t b, int64_t* ra, int64_t* rb) {
    const bool va = a >= 0 && a < ix.n, vb = b >= 0 && b < ix.n;
    const int64_t ba = va ? (a >> 7) : 0, bb = vb ? (b >> 7) : 0;
    const uint4* pa = reinterpret_cast<const uint4*>(ix.occ + ba);
    const uint4* pb = reinterpret_cast<const uint4*>(ix.occ + bb);
    uint4 h0 = pa[0], s0 = pa[1], c00 = pa[2], c01 = pa[3];
    uint4 h1, s1, c10, c11;
    if (bb != ba) { h1 = pb[0]; s1 = pb[1]; c10 = pb[2]; c11 = pb[3]; }
    else { h1 = h0; s1 = s0; c10 = c00; c11 = c01; }
    if (va) occFromBlock(ix, h0, s0, c00, c01, a, ra); else { ra[0] = ra[1] = ra[2] = ra[3] = ra[4] = -1; }
    if (vb) occFromBlock(ix, h1, s1, c10, c11, b, rb); else { rb[0] = rb[1] = rb[2] = rb[3] = rb[4] = -1; }
}

/* FMDIndex::backExt (src/FMDIndex.cpp:577-610): bi-interval [p, q, s] extended by b
 * on the left.  The q update adds the '\0' count and the counts of bases above b.
 * All five ranks are counted at both positions in one uniform sequence: lanes of a
 * wavefront extend different intervals, and a conditional cheaper count would run
 * both sequences under divergence. */
__device__ __forceinline__ int64_t dBackExt(const DevIndex& ix, int64_t& p, int64_t& q, int64_t& s, uint8_t b) {
    if (amb(b)) return (s = 0);
    const int c = toInt(b);
    int64_t lo[5], hi[5];
    occPair(ix, p - 1, p + s - 1, lo, hi);
    const int64_t O = lo[c];
    const int64_t sc = hi[c] - O;
    if (sc <= 0) return (s = sc);
    if (sc != s) {
        int64_t add = hi[0] - lo[0];
        for (int j = c + 1; j <= 4; ++j) add += hi[j] - lo[j];
        q = q + add;
    }
    p = ix.C[b] + O;
    s = sc;
    return s;
}

/* Rank rows and backExt for the search kernels, over the row type of one index kind:
 * R = int32_t for narrow indexes (fewer than 2^31 rows, so every rank, row and size fits;
 * the block counts are absolute), R = int64_t for wide ones (block counts plus the
 * superblock's, loaded beside the block).  Same values as the generic 64-bit forms. */
template <class R>
__device__ __forceinline__ void occFromBlockT(const DevIndex& ix, const uint4& h, const uint4& sp, const uint4& pl,
                                              const uint4& ph, R i, R* r) {
    const uint32_t j = (uint32_t)i & 127u;
    uint32_t nLo, nHi, nT, nSp;
    blockCounts(sp, pl, ph, j, nLo, nHi, nT, nSp);
    const uint32_t nC = nLo - nT, nG = nHi - nT;
    const uint32_t nA = (j + 1) - nLo - nHi + nT;
    if constexpr (sizeof(R) == 8) {
        const int64_t* sb = ix.sup + 4 * (i >> ix.supShift);
        const int64_t s0 = sb[0], s1 = sb[1], s2 = sb[2], s3 = sb[3];
        r[1] = s0 + (int64_t)(h.x + nA - nSp);
        r[2] = s1 + (int64_t)(h.y + nC);
        r[3] = s2 + (int64_t)(h.z + nG);
        r[4] = s3 + (int64_t)(h.w + nT);
        const int64_t special = (i & ~(int64_t)127) - (s0 + s1 + s2 + s3) - ((int64_t)h.x + h.y + h.z + h.w) + nSp;
        r[0] = ix.nN ? special - countN(ix, i) : special;
    } else {
        r[1] = (int32_t)(h.x + nA - nSp);
        r[2] = (int32_t)(h.y + nC);
        r[3] = (int32_t)(h.z + nG);
        r[4] = (int32_t)(h.w + nT);
        const int32_t special = (int32_t)(((uint32_t)i & ~127u) - (h.x + h.y + h.z + h.w) + nSp);
        r[0] = ix.nN ? special - (int32_t)countN(ix, i) : special;
    }
}

template <class R>
__device__ __forceinline__ R dBackExtT(const DevIndex& ix, R& p, R& q, R& s, uint8_t b) {
    if (amb(b)) return (s = 0);
    const int c = toInt(b);
    const R a = p - 1, e = p + s - 1;
    const bool va = a >= 0 && (int64_t)a < ix.n, ve = e >= 0 && (int64_t)e < ix.n;
    const R ba = va ? (a >> 7) : 0, be = ve ? (e >> 7) : 0;
    const uint4* pa = reinterpret_cast<const uint4*>(ix.occ + ba);
    const uint4* pe = reinterpret_cast<const uint4*>(ix.occ + be);
    const uint4 h0 = pa[0], s0 = pa[1], c00 = pa[2], c01 = pa[3];
    uint4 h1, s1, c10, c11;
    if (be != ba) { h1 = pe[0]; s1 = pe[1]; c10 = pe[2]; c11 = pe[3]; }
    else { h1 = h0; s1 = s0; c10 = c00; c11 = c01; }
    R lo[5], hi[5];
    if (va) occFromBlockT<R>(ix, h0, s0, c00, c01, a, lo); else { lo[0] = lo[1] = lo[2] = lo[3] = lo[4] = -1; }
    if (ve) occFromBlockT<R>(ix, h1, s1, c10, c11, e, hi); else { hi[0] = hi[1] = hi[2] = hi[3] = hi[4] = -1; }
    /* rank c and the '\0' + above-c counts by selects (no indexed arrays), C[b] from the
     * four uniform base entries */
    const R O = c == 1 ? lo[1] : (c == 2 ? lo[2] : (c == 3 ? lo[3] : lo[4]));
    const R H = c == 1 ? hi[1] : (c == 2 ? hi[2] : (c == 3 ? hi[3] : hi[4]));
    const R sc = H - O;
    if (sc <= 0) return (s = sc);
    if (sc != s) {
        const R add = (hi[0] - lo[0]) + (c < 2 ? hi[2] - lo[2] : 0) + (c < 3 ? hi[3] - lo[3] : 0) +
                      (c < 4 ? hi[4] - lo[4] : 0);
        q = q + add;
    }
    const R Cb = (R)(c == 1 ? ix.C[1] : (c == 2 ? ix.C[2] : (c == 3 ? ix.C[4] : ix.C[8])));
    p = Cb + O;
    s = sc;
    return s;
}

/* FMDIndex::fwdExt (src/FMDIndex.h:226-228) */
__device__ __forceinline__ int64_t dFwdExt(const DevIndex& ix, int64_t& p, int64_t& q, int64_t& s, uint8_t b) {
    return dBackExt(ix, q, p, s, comp16(b));
}

#endif
