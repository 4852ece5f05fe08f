/*
 * msgs_index.hip -- FMD index construction on the device.
 *
 * Suffix array of the BD text by prefix doubling (Manber & Myers) on radix-sorted
 * 64-bit keys: round 0 packs the first 21 symbols (3 bits each; "past the end" is the
 * smallest symbol, then '\0' < A < C < G < T < N), later rounds pack
 * (rank[i], rank[i+h] + 1) with h = 21, 42, 84, ... until every rank is distinct.
 * This is the order libdivsufsort gives (byte order, a suffix that is a prefix of
 * another first), used by the reference's FMDIndex constructor
 * (src/FMDIndex.cpp:29-68).  Then BWT (src/FMDIndex.cpp:385-393) and occ blocks.
 *
 * Narrow texts (n < 2^31) sort all suffixes at once with 32-bit rows.  Wide texts (up to
 * 2^34) cannot: 64-bit rows make a whole-text key/value sort 32 bytes per suffix twice
 * over.  They are cut into ranges of the final suffix array instead: the 4096 buckets of
 * the first four symbols are counted, consecutive buckets are packed into ranges of at
 * most rangeMax suffixes, and each range is sorted on its own.  Round 0 gathers a range's
 * suffixes by their bucket and sorts them by the 21-symbol key; later rounds sort only the
 * suffixes still in groups of two or more (a bitmap marks the settled rows) by
 * ((rank[i] - range start) << 34 | rank[i+h] + 1).  Ranks are the SA row of the first
 * member of a suffix's group and are refined in place, range after range: a rank read in
 * a round may already be finer than that round's, which orders ties by a longer prefix
 * and so stays consistent with the suffix order (the argument of Larsson and Sadakane's
 * qsufsort).  The ranks become the ISA.
 *
 * Block order.  msgseqtk-build indexes a database of blockSize symbols or more block by
 * block (src/msgseqtk-build.cpp:440-467: runs of whole chromosomes taken from the end) and
 * merges the block indexes with FMDIndex::prepend (src/FMDIndex.cpp:267-308, 612-630):
 * within a block the order is the block text's own (a suffix ends at its block's end), and
 * the merge places a suffix of an earlier block after every suffix of a later block that
 * equals it up to and including a '\0' (the interleaving walk starts each earlier-block
 * '\0' row after all later-block '\0' rows).  The wide sort gives that order by coding
 * '\0' per block (later blocks lower, all below A) and cutting every suffix at its block's
 * end; one block is the plain order.
 */
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "msgs_host.h"

namespace {

__device__ __forceinline__ uint64_t sym3(uint8_t b) {
    switch (b) {
    case 0: return 1; case 1: return 2; case 2: return 3; case 4: return 4; case 8: return 5; case 15: return 6;
    default: return 7;
    }
}

__global__ void k_init_keys(const uint8_t* bd, int64_t n, uint64_t* key, uint32_t* val) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = 0;
    for (int t = 0; t < 21; ++t) {
        int64_t j = i + t;
        k = (k << 3) | (j < n ? sym3(bd[j]) : 0);
    }
    key[i] = k;
    val[i] = (uint32_t)i;
}

/* head[k] = 1 where sorted key k starts a new group */
__global__ void k_heads(const uint64_t* key, int64_t n, uint32_t* head) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    head[k] = (k == 0 || key[k] != key[k - 1]) ? (uint32_t)k : 0u;
}

/* rank[val[k]] = start of k's group (max-scan of head) */
__global__ void k_scatter_rank(const uint32_t* val, const uint32_t* gstart, int64_t n, uint32_t* rank) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    rank[val[k]] = gstart[k];
}

__global__ void k_round_keys(const uint32_t* val, const uint32_t* rank, int64_t n, int64_t h, uint64_t* key) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t i = val[k];
    uint64_t r2 = ((int64_t)i + h < n) ? (uint64_t)rank[i + h] + 1 : 0;
    key[k] = ((uint64_t)rank[i] << 32) | r2;
}

/* number of groups (k with gstart[k] == k): a grid-stride count per thread, a block sum in
 * LDS and one atomic per block (one atomic per wavefront serialises ~10^7 atomics on one
 * address for a 10^9-symbol text) */
__global__ void __launch_bounds__(256) k_count_distinct(const uint32_t* gstart, int64_t n, unsigned long long* cnt) {
    __shared__ unsigned long long part[4];
    unsigned long long c = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        c += gstart[k] == (uint32_t)k ? 1 : 0;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = part[0] + part[1] + part[2] + part[3];
        if (t) atomicAdd(cnt, t);
    }
}

/* ---------------------------------------------------------------- wide texts */

constexpr int kBins = 1 << 12;                 /* buckets: the first 12 bits of the key */
constexpr int kMaxBlocks = 9;                  /* '\0' codes of every block fit 4-bit codes */

/* the blocks of the build order: one block, or the reference's block-by-block build */
struct Blocks {
    int32_t k;                  /* blocks (1: the whole text) */
    int32_t bits;               /* bits per symbol code: 3 for one block, else 4 */
    int32_t syms;               /* symbols per round-0 key: 21 or 15 */
    int32_t binSyms;            /* symbols per bucket: 4 or 3 (12 bits) */
    int64_t end[kMaxBlocks];    /* end of each block in the BD text */
};

__device__ __forceinline__ int blockOf(const Blocks& B, int64_t i) {
    int b = 0;
    while (b < B.k - 1 && i >= B.end[b]) ++b;
    return b;
}

/* code of BD byte c in block blk: past the block's end 0, '\0' 1 + (k - 1 - blk) (later
 * blocks lower), A, C, G, T, N from k + 1; with one block the codes of sym3 */
__device__ __forceinline__ uint64_t codeOf(const Blocks& B, uint8_t c, int blk) {
    const uint64_t k = (uint64_t)B.k;
    switch (c) {
    case 0: return 1 + (k - 1 - (uint64_t)blk); case 1: return k + 1; case 2: return k + 2; case 4: return k + 3;
    case 8: return k + 4; case 15: return k + 5;
    default: return k + 6;
    }
}

__device__ __forceinline__ uint32_t binOf(const uint8_t* bd, const Blocks& B, int64_t i) {
    const int blk = blockOf(B, i);
    const int64_t e = B.end[blk];
    uint32_t b = 0;
    for (int t = 0; t < B.binSyms; ++t) b = (b << B.bits) | (uint32_t)(i + t < e ? codeOf(B, bd[i + t], blk) : 0);
    return b;
}

__device__ __forceinline__ uint64_t keyOf(const uint8_t* bd, const Blocks& B, int64_t i) {
    const int blk = blockOf(B, i);
    const int64_t e = B.end[blk];
    uint64_t k = 0;
    for (int t = 0; t < B.syms; ++t) k = (k << B.bits) | (i + t < e ? codeOf(B, bd[i + t], blk) : 0);
    return k;
}

/* suffixes per bucket: LDS counts per block, added to the global table once per block */
__global__ void __launch_bounds__(256) k_bin_hist(const uint8_t* bd, int64_t n, Blocks B, unsigned long long* hist) {
    __shared__ uint32_t h[kBins];
    for (int b = threadIdx.x; b < kBins; b += blockDim.x) h[b] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&h[binOf(bd, B, i)], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < kBins; b += blockDim.x)
        if (h[b]) atomicAdd(&hist[b], (unsigned long long)h[b]);
}

/* the positions whose bucket lies in [b0, b1), in any order: each block takes 4096
 * positions per step and reserves its slots with one global atomic */
__global__ void __launch_bounds__(256) k_bin_select(const uint8_t* bd, int64_t n, Blocks B, uint32_t b0, uint32_t b1,
                                                    int64_t* out, unsigned long long* counter) {
    __shared__ uint32_t sCnt;
    __shared__ unsigned long long sBase;
    constexpr int kPer = 16;
    for (int64_t base = (int64_t)blockIdx.x * 256 * kPer; base < n; base += (int64_t)gridDim.x * 256 * kPer) {
        uint32_t mask = 0;
        for (int r = 0; r < kPer; ++r) {
            const int64_t i = base + r * 256 + threadIdx.x;
            if (i < n) {
                const uint32_t b = binOf(bd, B, i);
                if (b >= b0 && b < b1) mask |= 1u << r;
            }
        }
        if (threadIdx.x == 0) sCnt = 0;
        __syncthreads();
        const uint32_t c = __popc(mask);
        const uint32_t off = c ? atomicAdd(&sCnt, c) : 0u;
        __syncthreads();
        if (threadIdx.x == 0) sBase = atomicAdd(counter, (unsigned long long)sCnt);
        __syncthreads();
        int64_t o = (int64_t)(sBase + off);
        for (int r = 0; r < kPer; ++r)
            if (mask >> r & 1u) out[o++] = base + r * 256 + threadIdx.x;
        __syncthreads();
    }
}

__global__ void k_keys0(const uint8_t* bd, Blocks B, const int64_t* v, int64_t m, uint64_t* key) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) key[j] = keyOf(bd, B, v[j]);
}

/* the rows of a range's unsettled suffixes: flag[j] = 1 when row rs + j is not settled */
__global__ void k_unsettled(const uint32_t* done, int64_t rs, int64_t m, uint8_t* flag) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int64_t p = rs + j;
    flag[j] = (done[p >> 5] >> (p & 31) & 1u) ? 0 : 1;
}

/* round keys of the unsettled rows P[0..m): (rank - rs) << 34 | rank[v + h] + 1 (0 past the
 * end of v's block) */
__global__ void k_round_keys_w(const int64_t* P, int64_t m, const int64_t* sa, const int64_t* rank, Blocks B,
                               int64_t h, int64_t rs, int64_t* v, uint64_t* key) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int64_t s = sa[P[j]];
    v[j] = s;
    const uint64_t r2 = s + h < B.end[blockOf(B, s)] ? (uint64_t)rank[s + h] + 1 : 0;
    key[j] = ((uint64_t)(rank[s] - rs) << 34) | r2;
}

/* sorted suffixes into their rows: P[j], or rs + j when P is NULL */
__global__ void k_place(const int64_t* P, const int64_t* v, int64_t m, int64_t rs, int64_t* sa) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) sa[P ? P[j] : rs + j] = v[j];
}

__global__ void k_heads_w(const uint64_t* key, int64_t m, uint32_t* head) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) head[j] = (j == 0 || key[j] != key[j - 1]) ? (uint32_t)j : 0u;
}

/* rank = the row of the group's first member; a suffix alone in its group is settled */
__global__ void __launch_bounds__(256) k_rank_w(const uint64_t* key, const int64_t* v, const uint32_t* gstart,
                                                const int64_t* P, int64_t m, int64_t rs, int64_t* rank,
                                                uint32_t* done, unsigned long long* settled) {
    __shared__ unsigned long long part[4];
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long one = 0;
    if (j < m) {
        const int64_t g = gstart[j];
        rank[v[j]] = P ? P[g] : rs + g;
        const bool alone = (j == 0 || key[j] != key[j - 1]) && (j == m - 1 || key[j] != key[j + 1]);
        if (alone) {
            const int64_t p = P ? P[j] : rs + j;
            atomicOr(&done[p >> 5], 1u << (p & 31));
            one = 1;
        }
    }
    for (int off = 32; off > 0; off >>= 1) one += __shfl_down(one, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = one;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = part[0] + part[1] + part[2] + part[3];
        if (t) atomicAdd(settled, t);
    }
}

__global__ void k_isa(const uint32_t* sa, int64_t n, uint32_t* isa) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    isa[sa[k]] = (uint32_t)k;
}

/* grid-stride: a launch covers at most 2^32 work-items, fewer than a wide text's rows */
template <class T>
__global__ void k_bwt(const uint8_t* bd, const T* sa, int64_t n, uint8_t* bwt) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = (int64_t)sa[k];
        bwt[k] = s == 0 ? 0 : bd[s - 1];
    }
}

/* one thread per 128-position block: code bit planes, special mask, local base counts */
__global__ void k_occ(const uint8_t* bwt, int64_t n, int64_t nBlk, OccBlock* occ, uint32_t* local) {
    int64_t blk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= nBlk) return;
    OccBlock o;
    uint32_t cnt[4] = {0, 0, 0, 0};
    for (int w = 0; w < 4; ++w) { o.lo[w] = 0; o.hi[w] = 0; o.spec[w] = 0; }
    const int64_t base = blk * 128;
    for (int j = 0; j < 128; ++j) {
        int64_t i = base + j;
        uint32_t c = 0;
        bool special = true;
        if (i < n) {
            uint8_t b = bwt[i];
            if (b == 1) { c = 0; special = false; }
            else if (b == 2) { c = 1; special = false; }
            else if (b == 4) { c = 2; special = false; }
            else if (b == 8) { c = 3; special = false; }
        }
        if (special) { if (i < n) o.spec[j >> 5] |= 1u << (j & 31); }
        else {
            cnt[c]++;
            o.lo[j >> 5] |= (c & 1u) << (j & 31);
            o.hi[j >> 5] |= (c >> 1) << (j & 31);
        }
    }
    for (int t = 0; t < 4; ++t) { o.cnt[t] = 0; local[blk * 4 + t] = cnt[t]; }
    occ[blk] = o;
}

__global__ void k_fill_cnt(OccBlock* occ, const uint32_t* cum, int64_t nBlk) {
    int64_t blk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= nBlk) return;
    for (int t = 0; t < 4; ++t) occ[blk].cnt[t] = cum[blk * 4 + t];
}

inline unsigned grid(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

struct Freer {
    std::vector<void*> ptrs;
    ~Freer() { for (void* p : ptrs) if (p) (void)hipFree(p); }
    template <class T> T* alloc(size_t count, hipError_t& e) {
        void* p = nullptr;
        e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
};

}  // namespace

#define HCHK(x)                                                                 \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                 \
            err = std::string(#x) + ": " + hipGetErrorString(e_);               \
            return MSGS_E_HIP;                                                  \
        }                                                                       \
    } while (0)

namespace {

/* Suffix array of a narrow text (n < 2^31) into db.dSa / db.dIsa (uint32_t rows). */
int sortNarrow(DbState& db, hipStream_t st, std::string& err) {
    const int64_t n = db.n;
    if (n <= 0 || n > (int64_t)INT32_MAX) {
        err = "a narrow index needs a BD text length in [1, 2^31)";
        return MSGS_E_UNSUPPORTED;
    }
    Freer tmp;
    hipError_t e;
    uint64_t* k0 = tmp.alloc<uint64_t>(n, e); HCHK(e);
    uint64_t* k1 = tmp.alloc<uint64_t>(n, e); HCHK(e);
    uint32_t* v0 = tmp.alloc<uint32_t>(n, e); HCHK(e);
    uint32_t* v1 = tmp.alloc<uint32_t>(n, e); HCHK(e);
    uint32_t* rank = tmp.alloc<uint32_t>(n, e); HCHK(e);
    uint32_t* head = tmp.alloc<uint32_t>(n, e); HCHK(e);
    unsigned long long* dcnt = tmp.alloc<unsigned long long>(1, e); HCHK(e);

    size_t sortBytes = 0, scanBytes = 0;
    HCHK(rocprim::radix_sort_pairs(nullptr, sortBytes, k0, k1, v0, v1, (size_t)n, 0u, 64u, st));
    HCHK(rocprim::inclusive_scan(nullptr, scanBytes, head, head, (size_t)n, rocprim::maximum<uint32_t>(), st));
    void* tmpStore = tmp.alloc<uint8_t>(std::max(sortBytes, scanBytes), e); HCHK(e);
    const size_t tmpBytes = std::max(sortBytes, scanBytes);

    int nbits = 1;
    while (nbits < 63 && ((int64_t)1 << nbits) <= n + 1) ++nbits;

    k_init_keys<<<grid(n), 256, 0, st>>>(db.dBd, n, k0, v0);
    HCHK(hipGetLastError());
    int64_t h = 21;
    int endBit = 63;
    for (int round = 0;; ++round) {
        size_t tb = tmpBytes;
        HCHK(rocprim::radix_sort_pairs(tmpStore, tb, k0, k1, v0, v1, (size_t)n, 0u, (unsigned)endBit, st));
        std::swap(k0, k1);
        std::swap(v0, v1);
        k_heads<<<grid(n), 256, 0, st>>>(k0, n, head);
        tb = tmpBytes;
        HCHK(rocprim::inclusive_scan(tmpStore, tb, head, head, (size_t)n, rocprim::maximum<uint32_t>(), st));
        k_scatter_rank<<<grid(n), 256, 0, st>>>(v0, head, n, rank);
        HCHK(hipMemsetAsync(dcnt, 0, sizeof(unsigned long long), st));
        k_count_distinct<<<std::min<unsigned>(grid(n), 8192u), 256, 0, st>>>(head, n, dcnt);
        unsigned long long distinct = 0;
        HCHK(hipMemcpyAsync(&distinct, dcnt, sizeof(distinct), hipMemcpyDeviceToHost, st));
        HCHK(hipStreamSynchronize(st));
        if ((int64_t)distinct == n) break;
        if (round > 40) { err = "suffix sorting did not converge"; return MSGS_E_HIP; }
        k_round_keys<<<grid(n), 256, 0, st>>>(v0, rank, n, h, k0);
        HCHK(hipGetLastError());
        endBit = 32 + nbits;
        h *= 2;
    }
    /* v0 is the suffix array */
    HCHK(hipMalloc(&db.dSa, n * sizeof(uint32_t)));
    HCHK(hipMemcpyAsync(db.dSa, v0, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    HCHK(hipMalloc(&db.dIsa, n * sizeof(uint32_t)));
    k_isa<<<grid(n), 256, 0, st>>>((const uint32_t*)db.dSa, n, (uint32_t*)db.dIsa);
    HCHK(hipGetLastError());
    HCHK(hipStreamSynchronize(st));
    return MSGS_OK;
}

/* Suffix array and ISA of a wide text into db.dSa / db.dIsa (int64_t rows), range by range
 * (the file header describes the rounds). */
int sortWide(DbState& db, hipStream_t st, std::string& err, int64_t rangeMax) {
    const int64_t n = db.n;
    if (n <= 0 || n >= ((int64_t)1 << 34) - 2) {
        err = "BD text length must be in [1, 2^34 - 2)";
        return MSGS_E_UNSUPPORTED;
    }
    Blocks B;
    std::memset(&B, 0, sizeof(B));
    B.k = db.blockEnds.empty() ? 1 : (int32_t)db.blockEnds.size();
    if (B.k > kMaxBlocks) {
        err = "more than 9 build blocks";
        return MSGS_E_UNSUPPORTED;
    }
    for (int b = 0; b < B.k; ++b) B.end[b] = db.blockEnds.empty() ? n : db.blockEnds[b];
    B.bits = B.k == 1 ? 3 : 4;
    B.syms = 63 / B.bits;
    B.binSyms = 12 / B.bits;
    Freer tmp;
    hipError_t e;
    /* bucket sizes and the ranges of the suffix array they pack into */
    unsigned long long* dHist = tmp.alloc<unsigned long long>(kBins + 1, e); HCHK(e);
    HCHK(hipMemsetAsync(dHist, 0, (kBins + 1) * sizeof(unsigned long long), st));
    k_bin_hist<<<2048, 256, 0, st>>>(db.dBd, n, B, dHist);
    HCHK(hipGetLastError());
    std::vector<unsigned long long> hist(kBins);
    HCHK(hipMemcpyAsync(hist.data(), dHist, kBins * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HCHK(hipStreamSynchronize(st));
    int64_t biggest = 0;
    for (unsigned long long c : hist) biggest = std::max<int64_t>(biggest, (int64_t)c);
    const int64_t cap = std::max(rangeMax, biggest);
    if (cap > ((int64_t)1 << 30)) {
        err = "a 4-symbol bucket holds more than 2^30 suffixes";
        return MSGS_E_UNSUPPORTED;
    }
    struct Range { int64_t rs, m; uint32_t b0, b1; int64_t open; };
    std::vector<Range> ranges;
    {
        int64_t rs = 0, m = 0;
        uint32_t b0 = 0;
        for (uint32_t b = 0; b < (uint32_t)kBins; ++b) {
            if (m + (int64_t)hist[b] > cap) {
                ranges.push_back({rs, m, b0, b, m});
                rs += m; m = 0; b0 = b;
            }
            m += (int64_t)hist[b];
        }
        if (m) ranges.push_back({rs, m, b0, (uint32_t)kBins, m});
    }
    int64_t mMax = 0;
    for (const Range& r : ranges) mMax = std::max(mMax, r.m);
    /* whole-text arrays: rows, ranks (the ISA at the end), settled-row bitmap */
    HCHK(hipMalloc(&db.dSa, n * sizeof(int64_t)));
    HCHK(hipMalloc(&db.dIsa, n * sizeof(int64_t)));
    int64_t* sa = (int64_t*)db.dSa;
    int64_t* rank = (int64_t*)db.dIsa;
    uint32_t* done = tmp.alloc<uint32_t>(n / 32 + 1, e); HCHK(e);
    HCHK(hipMemsetAsync(done, 0, (n / 32 + 1) * sizeof(uint32_t), st));
    /* per-range work buffers */
    uint64_t* k0 = tmp.alloc<uint64_t>(mMax, e); HCHK(e);
    uint64_t* k1 = tmp.alloc<uint64_t>(mMax, e); HCHK(e);
    int64_t* v0 = tmp.alloc<int64_t>(mMax, e); HCHK(e);
    int64_t* v1 = tmp.alloc<int64_t>(mMax, e); HCHK(e);
    int64_t* P = tmp.alloc<int64_t>(mMax, e); HCHK(e);
    uint32_t* head = tmp.alloc<uint32_t>(mMax, e); HCHK(e);
    uint8_t* flag = tmp.alloc<uint8_t>(mMax, e); HCHK(e);
    unsigned long long* dCnt = tmp.alloc<unsigned long long>(2, e); HCHK(e);
    size_t sortBytes = 0, scanBytes = 0, selBytes = 0;
    HCHK(rocprim::radix_sort_pairs(nullptr, sortBytes, k0, k1, v0, v1, (size_t)mMax, 0u, 64u, st));
    HCHK(rocprim::inclusive_scan(nullptr, scanBytes, head, head, (size_t)mMax, rocprim::maximum<uint32_t>(), st));
    HCHK(rocprim::select(nullptr, selBytes, rocprim::counting_iterator<int64_t>(0), flag, P, dCnt, (size_t)mMax, st));
    const size_t tmpBytes = std::max(sortBytes, std::max(scanBytes, selBytes));
    void* tmpStore = tmp.alloc<uint8_t>(tmpBytes, e); HCHK(e);
    int mBits = 1;
    while (mBits < 31 && ((int64_t)1 << mBits) <= mMax) ++mBits;

    /* groups of a sorted range: ranks, settled rows; returns the rows settled */
    auto settle = [&](const uint64_t* key, const int64_t* v, const int64_t* rows, int64_t m, int64_t rs,
                      unsigned long long& settled) -> int {
        k_heads_w<<<grid(m), 256, 0, st>>>(key, m, head);
        size_t tb = tmpBytes;
        HCHK(rocprim::inclusive_scan(tmpStore, tb, head, head, (size_t)m, rocprim::maximum<uint32_t>(), st));
        HCHK(hipMemsetAsync(dCnt + 1, 0, sizeof(unsigned long long), st));
        k_rank_w<<<grid(m), 256, 0, st>>>(key, v, head, rows, m, rs, rank, done, dCnt + 1);
        HCHK(hipGetLastError());
        HCHK(hipMemcpyAsync(&settled, dCnt + 1, sizeof(settled), hipMemcpyDeviceToHost, st));
        HCHK(hipStreamSynchronize(st));
        return MSGS_OK;
    };

    /* round 0: each range's suffixes by their first 21 (15) symbols */
    int64_t open = 0;
    for (Range& r : ranges) {
        HCHK(hipMemsetAsync(dCnt, 0, sizeof(unsigned long long), st));
        k_bin_select<<<2048, 256, 0, st>>>(db.dBd, n, B, r.b0, r.b1, v0, dCnt);
        k_keys0<<<grid(r.m), 256, 0, st>>>(db.dBd, B, v0, r.m, k0);
        size_t tb = tmpBytes;
        HCHK(rocprim::radix_sort_pairs(tmpStore, tb, k0, k1, v0, v1, (size_t)r.m, 0u, (unsigned)(B.syms * B.bits), st));
        k_place<<<grid(r.m), 256, 0, st>>>(nullptr, v1, r.m, r.rs, sa);
        unsigned long long got = 0;
        HCHK(hipMemcpyAsync(&got, dCnt, sizeof(got), hipMemcpyDeviceToHost, st));
        unsigned long long settled = 0;
        int rc = settle(k1, v1, nullptr, r.m, r.rs, settled);
        if (rc) return rc;
        if ((int64_t)got != r.m) { err = "bucket gather miscounted"; return MSGS_E_HIP; }
        r.open = r.m - (int64_t)settled;
        open += r.open;
    }
    /* later rounds: the unsettled rows of each range by (rank, rank[i + h]) */
    int64_t h = B.syms;
    for (int round = 1; open > 0; ++round, h *= 2) {
        if (round > 40) { err = "suffix sorting did not converge"; return MSGS_E_HIP; }
        open = 0;
        for (Range& r : ranges) {
            if (r.open == 0) continue;
            k_unsettled<<<grid(r.m), 256, 0, st>>>(done, r.rs, r.m, flag);
            size_t tb = tmpBytes;
            HCHK(rocprim::select(tmpStore, tb, rocprim::counting_iterator<int64_t>(r.rs), flag, P, dCnt, (size_t)r.m, st));
            unsigned long long mu = 0;
            HCHK(hipMemcpyAsync(&mu, dCnt, sizeof(mu), hipMemcpyDeviceToHost, st));
            HCHK(hipStreamSynchronize(st));
            const int64_t m = (int64_t)mu;
            if (m != r.open) { err = "unsettled rows miscounted"; return MSGS_E_HIP; }
            k_round_keys_w<<<grid(m), 256, 0, st>>>(P, m, sa, rank, B, h, r.rs, v0, k0);
            tb = tmpBytes;
            HCHK(rocprim::radix_sort_pairs(tmpStore, tb, k0, k1, v0, v1, (size_t)m, 0u, (unsigned)(34 + mBits), st));
            k_place<<<grid(m), 256, 0, st>>>(P, v1, m, 0, sa);
            unsigned long long settled = 0;
            int rc = settle(k1, v1, P, m, r.rs, settled);
            if (rc) return rc;
            r.open = m - (int64_t)settled;
            open += r.open;
        }
    }
    return MSGS_OK;
}

}  // namespace

/* Suffix array, ISA, BWT and occ blocks of db.dBd (n bytes) into db's device buffers. */
int buildIndexDevice(DbState& db, hipStream_t st, std::string& err, int64_t rangeMax) {
    const int64_t n = db.n;
    int rc = db.wide ? sortWide(db, st, err, rangeMax) : sortNarrow(db, st, err);
    if (rc != MSGS_OK) return rc;
    Freer tmp;
    hipError_t e;
    HCHK(hipMalloc(&db.dBwt, n));
    const unsigned gb = std::min(grid(n), 1u << 18);
    if (db.wide) k_bwt<int64_t><<<gb, 256, 0, st>>>(db.dBd, (const int64_t*)db.dSa, n, db.dBwt);
    else k_bwt<uint32_t><<<gb, 256, 0, st>>>(db.dBd, (const uint32_t*)db.dSa, n, db.dBwt);
    HCHK(hipGetLastError());

    const int64_t nBlk = n / 128 + 1;
    HCHK(hipMalloc(&db.dOcc, nBlk * sizeof(OccBlock)));
    uint32_t* local = tmp.alloc<uint32_t>(nBlk * 4, e); HCHK(e);
    k_occ<<<grid(nBlk), 256, 0, st>>>(db.dBwt, n, nBlk, db.dOcc, local);
    HCHK(hipGetLastError());
    std::vector<uint32_t> h_local(nBlk * 4);
    HCHK(hipMemcpyAsync(h_local.data(), local, nBlk * 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HCHK(hipStreamSynchronize(st));
    /* block counts: absolute (narrow), or from the start of the block's superblock (wide),
     * whose own counts go to db.dSup */
    const int64_t blkPerSup = db.wide ? (int64_t)1 << (db.supShift - 7) : nBlk;
    const int64_t nSup = (nBlk + blkPerSup - 1) / blkPerSup;
    std::vector<int64_t> sup(4 * nSup, 0);
    int64_t run[4] = {0, 0, 0, 0}, base[4] = {0, 0, 0, 0};
    for (int64_t b = 0; b < nBlk; ++b) {
        if (db.wide && b % blkPerSup == 0)
            for (int t = 0; t < 4; ++t) { base[t] = run[t]; sup[4 * (b / blkPerSup) + t] = run[t]; }
        for (int t = 0; t < 4; ++t) {
            const int64_t c = h_local[b * 4 + t];
            h_local[b * 4 + t] = (uint32_t)(run[t] - base[t]);
            run[t] += c;
        }
    }
    HCHK(hipMemcpyAsync(local, h_local.data(), nBlk * 4 * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    if (db.wide) {
        HCHK(hipMalloc(&db.dSup, sup.size() * sizeof(int64_t)));
        HCHK(hipMemcpyAsync(db.dSup, sup.data(), sup.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    }
    k_fill_cnt<<<grid(nBlk), 256, 0, st>>>(db.dOcc, local, nBlk);
    HCHK(hipGetLastError());

    /* N positions of the BWT (none in texts without N) */
    db.nN = db.B[15];
    if (db.nN > 0) {
        std::vector<uint8_t> bwt(n);
        HCHK(hipMemcpyAsync(bwt.data(), db.dBwt, n, hipMemcpyDeviceToHost, st));
        HCHK(hipStreamSynchronize(st));
        std::vector<int64_t> pos;
        for (int64_t i = 0; i < n; ++i) if (bwt[i] == 15) pos.push_back(i);
        HCHK(hipMalloc(&db.dNPos, pos.size() * sizeof(int64_t)));
        HCHK(hipMemcpyAsync(db.dNPos, pos.data(), pos.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    }
    HCHK(hipStreamSynchronize(st));
    return MSGS_OK;
}
