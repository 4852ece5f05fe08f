/*
 * msgs_seed.hip -- from SMEMs to seeds to chains, one lane per read.
 *
 * smem_order_kernel: SMEM_LIST::findSeeds' sort of the SMEMs by loglik
 *   (src/SMEM.cpp:275-276, libstdc++ std::sort restated in msgs_sort.h) and the
 *   maxNSeed cut (:284-300): SMEMs are taken whole until the seed count reaches
 *   maxNSeed.  Each SMEM yields exactly `size` seeds (src/SMEM.cpp:240), so the seed
 *   count is known before any SA lookup and the seed pool is allocated exactly.
 * seed_kernel: SMEM::getSeeds (src/SMEM.cpp:217-269): SA lookups of both intervals,
 *   BD position -> (tid, strand, loc) (src/MetaGenome.h:147-167,474-479; tid by
 *   binary search over the contiguous chrom BD ranges, which finds the same chrom as
 *   the reference's linear scan), forward-strand hits only; then
 *   SeedPair::removeRedundant (src/SeedPair.cpp:71-88).
 * chain_kernel: SeedChain::getChains / dfsSeeds (src/SeedChain.cpp:40-73) as an
 *   explicit-stack DFS; pass 0 counts chains, pass 1 writes them.
 */
#include "msgs_kernels.h"
#include "msgs_sort.h"

namespace {

enum : int32_t { FWD = 1, REV = 2, UNK = 3 };

__device__ __forceinline__ int32_t findTid(const DevIndex& ix, int64_t pos) {
    int32_t lo = 0, hi = ix.nChrom;   /* largest t with chromBD[t] <= pos */
    while (hi - lo > 1) {
        int32_t mid = (lo + hi) >> 1;
        if (ix.chromBD[mid] <= pos) lo = mid; else hi = mid;
    }
    return lo;
}

struct SmemLess {
    const SmemRec* r;
    __device__ bool operator()(int32_t a, int32_t b) const { return r[a].logP < r[b].logP; }
};

/* SeedPair operator< (src/SeedPair.h:225-232) */
struct SeedLess {
    __device__ bool operator()(const SeedRec& l, const SeedRec& r) const {
        return l.from != r.from ? l.from < r.from :
               l.to != r.to ? l.to < r.to :
               l.tid != r.tid ? l.tid < r.tid :
               l.strand != r.strand ? l.strand < r.strand :
               l.start != r.start ? l.start < r.start : l.end < r.end;
    }
};

__device__ __forceinline__ bool containing(const SeedRec& l, const SeedRec& r) {
    return l.tid == r.tid && l.strand == r.strand && l.from <= r.from && l.to >= r.to &&
           l.start <= r.start && l.end >= r.end;
}

__device__ __forceinline__ int64_t locDist(int64_t ls, int64_t le, int64_t rs, int64_t re) {
    bool ov = ls < re && le > rs;
    return ov ? 0 : ls < rs ? rs - le + 1 : ls - re + 1;
}

/* SeedPair::nMismatch <= maxMM && |nIndel| <= maxIndel (src/SeedPair.h:136-158) */
__device__ __forceinline__ bool child(const SeedRec& l, const SeedRec& r, int64_t maxMM, int64_t maxIndel) {
    if (!(l.to <= r.from && l.end <= r.start && l.tid == r.tid && l.strand == r.strand)) return false;
    int64_t mm = min(locDist(l.from, l.to, r.from, r.to), locDist(l.start, l.end, r.start, r.end));
    if (mm > maxMM) return false;
    int64_t ind = (r.from - l.to) - (r.start - l.end);
    return (ind < 0 ? -ind : ind) <= maxIndel;
}

}  // namespace

__global__ void __launch_bounds__(256) smem_order_kernel(DevIndex ix, SeedArgs a) {
    for (int64_t rid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; rid < a.nReads;
         rid += (int64_t)gridDim.x * blockDim.x) {
        const int32_t n = a.smemCnt[rid];
        const SmemRec* r = a.smems + (size_t)rid * a.cap;
        int32_t* ord = a.order + (size_t)rid * a.cap;
        for (int32_t k = 0; k < n; ++k) ord[k] = k;
        msgs_sort::sort(ord, ord + n, SmemLess{r});
        int64_t acc = 0;
        int32_t used = 0;
        for (int32_t u = 0; u < n; ++u) {
            acc += r[ord[u]].size;
            used = u + 1;
            if ((uint64_t)acc >= (uint64_t)a.maxNSeed) break;
        }
        a.nUsed[rid] = used;
        a.nRaw[rid] = acc;
    }
}

__global__ void __launch_bounds__(256) seed_kernel(DevIndex ix, SeedArgs a) {
    unsigned long long nSA = 0;
    for (int64_t rid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; rid < a.nReads;
         rid += (int64_t)gridDim.x * blockDim.x) {
        const int64_t L = a.offs[rid + 1] - a.offs[rid];
        const SmemRec* r = a.smems + (size_t)rid * a.cap;
        const int32_t* ord = a.order + (size_t)rid * a.cap;
        SeedRec* s = a.seeds + a.seedOff[rid];
        int64_t w = 0;
        const int32_t used = a.nUsed[rid];
        for (int32_t u = 0; u < used; ++u) {
            const SmemRec m = r[ord[u]];
            const int64_t len = m.to - m.from;
            for (int64_t i = 0; i < m.size; ++i) {
                for (int side = 0; side < 2; ++side) {
                    const int64_t bd = saAt(ix, side == 0 ? m.fwd + i : m.rev + i);
                    nSA++;
                    const int32_t tid = findTid(ix, bd);
                    const int64_t cs = ix.chromBD[tid], cl = ix.chromBD[tid + 1] - cs;
                    if (cs <= bd && bd < cs + cl / 2) {   /* FWD target strand */
                        const int64_t loc = bd - cs;
                        const int64_t f = side == 0 ? (int64_t)m.from : L - m.to;
                        SeedRec sd;
                        sd.from = f; sd.to = f + len; sd.start = loc; sd.end = loc + len;
                        sd.tid = tid; sd.strand = side == 0 ? FWD : REV; sd.logP = m.logP;
                        s[w++] = sd;
                    }
                }
            }
        }
        /* removeRedundant */
        int64_t n = w;
        if (n > 1) {
            msgs_sort::sort(s, s + n, SeedLess{});
            for (int64_t i = 0; i < n - 1; ++i)
                for (int64_t j = i + 1; j < n; ++j)
                    if (containing(s[i], s[j]) || containing(s[j], s[i])) {
                        s[i].from = 0; s[i].to = 0; s[i].start = 0; s[i].end = 0;
                        s[i].tid = -1; s[i].strand = UNK; s[i].logP = 0;
                        break;
                    }
            int64_t k = 0;
            for (int64_t i = 0; i < n; ++i)
                if (s[i].to - s[i].from != 0) s[k++] = s[i];
            n = k;
        }
        a.seedCnt[rid] = (int32_t)n;
    }
    if (a.stats && nSA) atomicAdd(&a.stats[1], nSA);
}

__global__ void __launch_bounds__(256) chain_kernel(ChainArgs a) {
    for (int64_t rid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; rid < a.nReads;
         rid += (int64_t)gridDim.x * blockDim.x) {
        const int64_t L = a.offs[rid + 1] - a.offs[rid];
        /* maxMismatch = ceil(L * .2), maxIndel = ceil(L * .1) (src/msgseqtk-align.cpp:619-620) */
        const int64_t maxMM = a.maxMM ? a.maxMM[rid] : (int64_t)ceil((double)L * 0.2);
        const int64_t maxIndel = a.maxIndel ? a.maxIndel[rid] : (int64_t)ceil((double)L * 0.1);
        const int32_t n = a.seedCnt[rid];
        const int64_t so = a.seedOff[rid];
        const SeedRec* s = a.seeds + so;
        int32_t* node = a.stackNode + so;
        int32_t* next = a.stackNext + so;
        uint8_t* seen = a.seen + so;
        for (int32_t i = 0; i < n; ++i) seen[i] = 0;
        int32_t nChains = 0;
        int64_t nSeeds = 0;
        int64_t cBase = 0, sBase = 0;
        if (a.pass == 1) { cBase = a.chainOff[rid]; sBase = a.chainSeedOff[rid]; }
        for (int32_t i = 0; i < n; ++i) {
            if (seen[i]) continue;
            int32_t depth = 1;
            node[0] = i;
            next[0] = i + 1;
            seen[i] = 1;
            while (depth > 0) {
                const int32_t d = depth - 1;
                const int32_t cur = node[d];
                uint32_t nx = (uint32_t)next[d];
                const bool leaf = (nx >> 31) == 0;
                int32_t j = (int32_t)(nx & 0x7fffffffu);
                int32_t found = -1;
                while (j < n) {
                    const int32_t jj = j++;
                    if (!seen[jj] && child(s[cur], s[jj], maxMM, maxIndel)) { found = jj; break; }
                }
                if (found >= 0) {
                    next[d] = (int32_t)((uint32_t)j | 0x80000000u);
                    node[depth] = found;
                    next[depth] = found + 1;
                    seen[found] = 1;
                    depth++;
                } else {
                    if (leaf) {
                        if (a.pass == 1) {
                            a.chainStart[cBase + nChains] = (int32_t)(sBase + nSeeds);
                            a.chainLen[cBase + nChains] = depth;
                            for (int32_t k = 0; k < depth; ++k) a.chainSeeds[sBase + nSeeds + k] = node[k];
                        }
                        nChains++;
                        nSeeds += depth;
                    }
                    depth--;
                }
            }
        }
        if (a.pass == 0) { a.nChains[rid] = nChains; a.nChainSeeds[rid] = nSeeds; }
    }
}
