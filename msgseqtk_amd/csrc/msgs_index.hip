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
 */
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

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

__global__ void k_isa(const uint32_t* sa, int64_t n, uint32_t* isa) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    isa[sa[k]] = (uint32_t)k;
}

__global__ void k_bwt(const uint8_t* bd, const uint32_t* sa, int64_t n, uint8_t* bwt) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t s = sa[k];
    bwt[k] = s == 0 ? 0 : bd[s - 1];
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

/* Suffix array, BWT and occ blocks of db.dBd (n bytes) into db's device buffers. */
int buildIndexDevice(DbState& db, hipStream_t st, std::string& err) {
    const int64_t n = db.n;
    if (n <= 0 || n >= (int64_t)0xffffffffLL || n > (int64_t)INT32_MAX) {
        err = "BD text length must be in [1, 2^31)";
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
    k_isa<<<grid(n), 256, 0, st>>>(db.dSa, n, db.dIsa);
    HCHK(hipGetLastError());
    HCHK(hipMalloc(&db.dBwt, n));
    k_bwt<<<grid(n), 256, 0, st>>>(db.dBd, db.dSa, n, db.dBwt);
    HCHK(hipGetLastError());

    const int64_t nBlk = n / 128 + 1;
    HCHK(hipMalloc(&db.dOcc, nBlk * sizeof(OccBlock)));
    uint32_t* local = tmp.alloc<uint32_t>(nBlk * 4, e); HCHK(e);
    k_occ<<<grid(nBlk), 256, 0, st>>>(db.dBwt, n, nBlk, db.dOcc, local);
    HCHK(hipGetLastError());
    std::vector<uint32_t> h_local(nBlk * 4);
    HCHK(hipMemcpyAsync(h_local.data(), local, nBlk * 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HCHK(hipStreamSynchronize(st));
    uint64_t run[4] = {0, 0, 0, 0};
    for (int64_t b = 0; b < nBlk; ++b)
        for (int t = 0; t < 4; ++t) {
            uint64_t c = h_local[b * 4 + t];
            h_local[b * 4 + t] = (uint32_t)run[t];
            run[t] += c;
        }
    HCHK(hipMemcpyAsync(local, h_local.data(), nBlk * 4 * sizeof(uint32_t), hipMemcpyHostToDevice, st));
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
