/*
 * msgs_map.hip -- the whole mapping path over one batch: SMEMs -> seeds -> chains ->
 * DP/backtrace on the device (and, for pairs, the candidate mate pairs: msgs_pair.hip),
 * then the host finishing step of msgseqtk-align (src/msgseqtk-align.cpp:538-758):
 * Alignment::filter, calcMapQ, sort and record output.
 *
 * Host finishing keeps the libm-dependent decisions (exp, log10, round) on the host,
 * where the reference computes them, and reorders with the restated libstdc++ sort.
 */
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_scan.hpp>
#include <x86intrin.h>

#include <algorithm>
#include <array>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "msgs_capi_impl.h"
#include "msgs_samfmt.h"
#include "msgs_sort.h"

namespace {

constexpr int32_t NEG = -1073741824;

/* reverse complement of every read and its qualities: one wavefront per read, lanes
 * on consecutive bytes so that loads and stores coalesce */
__global__ void k_revcom(const uint8_t* seq, const uint8_t* qual, const int64_t* offs, int64_t n, uint8_t* rs,
                         uint8_t* rq) {
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int lane = threadIdx.x & 63;
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < n; r += waves) {
        const int64_t o = offs[r], L = offs[r + 1] - o;
        for (int64_t i = lane; i < L; i += 64) {
            const uint8_t b = seq[o + L - 1 - i];
            rs[o + i] = (uint8_t)(((b & 1) << 3) | ((b & 2) << 1) | ((b & 4) >> 1) | ((b & 8) >> 3));
            rq[o + i] = qual[o + L - 1 - i];
        }
    }
}

/* gather each job's CIGAR ops and MD bytes from their worst-case slots into
 * contiguous arrays (offsets from the host's scan of nCigar / mdLen) */
__global__ void k_compact(const AlnRec* recs, int64_t nJobs, const int64_t* pathOff, const uint32_t* cig,
                          const char* md, const int64_t* cOff, const int64_t* mOff, uint32_t* cOut, char* mOut) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nJobs; j += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t* cs = cig + pathOff[j] + 2 * j;
        const char* ms = md + 3 * pathOff[j] + 8 * j;
        const int32_t nc = recs[j].nCigar, nm = recs[j].mdLen;
        for (int32_t k = 0; k < nc; ++k) cOut[cOff[j] + k] = cs[k];
        for (int32_t k = 0; k < nm; ++k) mOut[mOff[j] + k] = ms[k];
    }
}

/* CIGAR op and MD byte counts of every job, plus a zero at nJobs (scan input) */
__global__ void k_rec_counts(const AlnRec* recs, int64_t nJobs, int64_t* nc, int64_t* nm) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= nJobs; j += (int64_t)gridDim.x * blockDim.x) {
        nc[j] = j < nJobs ? recs[j].nCigar : 0;
        nm[j] = j < nJobs ? recs[j].mdLen : 0;
    }
}

/* wavefront-fill job list, largest tile area first: jobs are bucketed by the bit
 * length of their area (64 buckets, largest bucket first); the order inside a bucket
 * is whatever the atomics give, which changes no result (jobs write disjoint
 * outputs).  hist[64] counts, cur[64] cursors, tot[1] the list length. */
/* two buckets per octave of area (the top bit and the one below it), non-decreasing in area,
 * areas of 2^31 and more sharing the last */
__device__ __forceinline__ int areaBucket(int64_t area) {
    if (area <= 0) return 0;
    const int top = 63 - __clzll((unsigned long long)area);
    const int half = top > 0 ? (int)((area >> (top - 1)) & 1) : 1;
    return min(63, 2 * top + half + 1);
}

__global__ void __launch_bounds__(256) k_wave_hist(const int32_t* cls, const int64_t* area, int64_t nJobs,
                                                    unsigned long long* hist, int32_t want) {
    __shared__ unsigned int h[64];
    if (threadIdx.x < 64) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nJobs && cls[j] == want) atomicAdd(&h[areaBucket(area[j]) & 63], 1u);
    __syncthreads();
    if (threadIdx.x < 64 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

__global__ void k_wave_base(const unsigned long long* hist, unsigned long long* cur, int64_t* tot) {
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int b = 63; b >= 0; --b) { cur[b] = s; s += hist[b]; }
        tot[0] = (int64_t)s;
    }
}

__global__ void __launch_bounds__(256) k_wave_scatter(const int32_t* cls, const int64_t* area, int64_t nJobs,
                                                       unsigned long long* cur, int64_t* wave, int32_t want) {
    __shared__ unsigned int h[64];
    __shared__ unsigned long long base[64];
    if (threadIdx.x < 64) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int b = -1;
    unsigned int r = 0;
    if (j < nJobs && cls[j] == want) { b = areaBucket(area[j]) & 63; r = atomicAdd(&h[b], 1u); }
    __syncthreads();
    if (threadIdx.x < 64 && h[threadIdx.x]) base[threadIdx.x] = atomicAdd(&cur[threadIdx.x], (unsigned long long)h[threadIdx.x]);
    __syncthreads();
    if (b >= 0) wave[base[b] + r] = j;
}

__global__ void k_copy_i64(const int64_t* src, int64_t* dst) { if (threadIdx.x == 0) *dst = *src; }
__global__ void k_zero_u32(unsigned int* p) { if (threadIdx.x == 0) *p = 0; }

__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

inline unsigned gridFor(int64_t n, int t = 256) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + t - 1) / t, 1 << 16));
}

/* device exclusive prefix sum of n int64 values (rocPRIM; a null temp asks for its size) */
inline hipError_t exclusiveSum(void* tmp, size_t& bytes, const int64_t* in, int64_t* out, int64_t n, hipStream_t st) {
    return rocprim::exclusive_scan(tmp, bytes, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), st);
}

template <class T> int64_t exclusiveScan(const std::vector<T>& v, std::vector<int64_t>& off) {
    off.resize(v.size() + 1);
    int64_t s = 0;
    for (size_t i = 0; i < v.size(); ++i) { off[i] = s; s += (int64_t)v[i]; }
    off[v.size()] = s;
    return s;
}

#define CK(x, msg)                                                       \
    do {                                                                 \
        hipError_t e_ = (x);                                             \
        if (e_ != hipSuccess) return fail(c, MSGS_E_HIP, std::string(msg) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

/* ------------------------------------------------------------------ device stages */

struct SeedStage {
    std::vector<int64_t> seedOff;   /* host copy, n+1 */
    std::vector<int32_t> seedCnt;
    int64_t* dSeedOff = nullptr;
    int32_t* dSeedCnt = nullptr;
    SeedRec* dSeeds = nullptr;
};

struct ChainStage {
    int64_t nJobs = 0;
    std::vector<int64_t> chainOff;  /* per read, n+1 */
    int32_t* dJobRead = nullptr;
    int32_t* dChainStart = nullptr;
    int32_t* dChainLen = nullptr;
    int32_t* dChainSeeds = nullptr;
    int64_t nChainSeeds = 0;
};

/* DP results on the host: records plus compacted CIGAR ops / MD bytes of each job
 * at cigOff[j] / mdOff[j] */
struct DpStage {
    PVec<AlnRec> recs;
    PVec<int64_t> cigOff, mdOff;
    PVec<uint32_t> cigar;
    PVec<char> md;
    PVec<int64_t> tot;   /* device totals read back: tiles, paths, wave jobs, CIGAR ops, MD bytes */
    double scoreScale = 1;   /* the DP's integer scores are the reference's alnScore x scoreScale */
    const AlnRec* dRecs = nullptr;   /* the records on the device (the executor's DP pool) */
    double score(int32_t v) const {
        return v == NEG ? -std::numeric_limits<double>::infinity() : (double)v / scoreScale;
    }
};

static int runSeeds(msgs_ctx* c, const DevReads& r, const msgs_params* p, const SmemStage& sm, SeedStage& out,
                    unsigned long long* dStats) {
    const int64_t n = r.n;
    hipStream_t st = c->x().stream;
    hipError_t e = hipSuccess;
    SeedArgs a;
    std::memset(&a, 0, sizeof(a));
    a.offs = r.offs; a.nReads = n; a.smemCnt = sm.cnt; a.smems = sm.recs; a.cap = sm.cap;
    a.maxNSeed = p->maxNSeed;
    a.order = c->x().bufs.get<int32_t>("sd.ord", (size_t)sm.cap * std::max<int64_t>(n, 1), e);
    a.nUsed = c->x().bufs.get<int32_t>("sd.used", n + 1, e);
    a.nRaw = c->x().bufs.get<int64_t>("sd.raw", n + 1, e);
    a.stats = dStats;
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "seed scratch");
    if (n) smem_order_kernel<<<gridFor(n), 256, 0, st>>>(c->db.view(), a);
    std::vector<int64_t> raw(n);
    if (n) CK(hipMemcpyAsync(raw.data(), a.nRaw, n * 8, hipMemcpyDeviceToHost, st), "nRaw");
    CK(hipStreamSynchronize(st), "order kernel");
    const int64_t total = exclusiveScan(raw, out.seedOff);
    out.dSeedOff = c->x().bufs.get<int64_t>("sd.off", n + 1, e);
    out.dSeeds = c->x().bufs.get<SeedRec>("sd.seeds", total + 1, e);
    out.dSeedCnt = c->x().bufs.get<int32_t>("sd.cnt", n + 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "seed pool");
    CK(hipMemcpyAsync(out.dSeedOff, out.seedOff.data(), (n + 1) * 8, hipMemcpyHostToDevice, st), "seedOff");
    a.seedOff = out.dSeedOff; a.seeds = out.dSeeds; a.seedCnt = out.dSeedCnt;
    if (n) seed_kernel<<<gridFor(n), 256, 0, st>>>(c->db.view(), a);
    out.seedCnt.resize(n);
    if (n) CK(hipMemcpyAsync(out.seedCnt.data(), out.dSeedCnt, n * 4, hipMemcpyDeviceToHost, st), "seedCnt");
    CK(hipStreamSynchronize(st), "seed kernel");
    return MSGS_OK;
}

static int runChains(msgs_ctx* c, const DevReads& r, const SeedStage& sd, ChainStage& out,
                     const int64_t* dMaxMM = nullptr, const int64_t* dMaxIndel = nullptr) {
    const int64_t n = r.n;
    hipStream_t st = c->x().stream;
    hipError_t e = hipSuccess;
    const int64_t seedTotal = sd.seedOff[n];
    ChainArgs a;
    std::memset(&a, 0, sizeof(a));
    a.offs = r.offs; a.nReads = n; a.seedOff = sd.dSeedOff; a.seedCnt = sd.dSeedCnt; a.seeds = sd.dSeeds;
    a.maxMM = dMaxMM; a.maxIndel = dMaxIndel;
    a.stackNode = c->x().bufs.get<int32_t>("ch.node", seedTotal + 1, e);
    a.stackNext = c->x().bufs.get<int32_t>("ch.next", seedTotal + 1, e);
    a.seen = c->x().bufs.get<uint8_t>("ch.seen", seedTotal + 1, e);
    a.nChains = c->x().bufs.get<int32_t>("ch.n", n + 1, e);
    a.nChainSeeds = c->x().bufs.get<int64_t>("ch.ns", n + 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "chain scratch");
    a.pass = 0;
    if (n) chain_kernel<<<gridFor(n), 256, 0, st>>>(a);
    std::vector<int32_t> nch(n);
    std::vector<int64_t> ncs(n);
    if (n) {
        CK(hipMemcpyAsync(nch.data(), a.nChains, n * 4, hipMemcpyDeviceToHost, st), "nChains");
        CK(hipMemcpyAsync(ncs.data(), a.nChainSeeds, n * 8, hipMemcpyDeviceToHost, st), "nChainSeeds");
    }
    CK(hipStreamSynchronize(st), "chain pass 0");
    std::vector<int64_t> csOff;
    out.nJobs = exclusiveScan(nch, out.chainOff);
    out.nChainSeeds = exclusiveScan(ncs, csOff);
    int64_t* dChainOff = c->x().bufs.get<int64_t>("ch.coff", n + 1, e);
    int64_t* dCsOff = c->x().bufs.get<int64_t>("ch.soff", n + 1, e);
    out.dChainStart = c->x().bufs.get<int32_t>("ch.start", out.nJobs + 1, e);
    out.dChainLen = c->x().bufs.get<int32_t>("ch.len", out.nJobs + 1, e);
    out.dChainSeeds = c->x().bufs.get<int32_t>("ch.seeds", out.nChainSeeds + 1, e);
    out.dJobRead = c->x().bufs.get<int32_t>("ch.read", out.nJobs + 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "chain pool");
    std::vector<int32_t> jobRead(out.nJobs);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = out.chainOff[i]; j < out.chainOff[i + 1]; ++j) jobRead[j] = (int32_t)i;
    CK(hipMemcpyAsync(dChainOff, out.chainOff.data(), (n + 1) * 8, hipMemcpyHostToDevice, st), "chainOff");
    CK(hipMemcpyAsync(dCsOff, csOff.data(), (n + 1) * 8, hipMemcpyHostToDevice, st), "csOff");
    if (out.nJobs)
        CK(hipMemcpyAsync(out.dJobRead, jobRead.data(), out.nJobs * 4, hipMemcpyHostToDevice, st), "jobRead");
    a.chainOff = dChainOff; a.chainSeedOff = dCsOff;
    a.chainStart = out.dChainStart; a.chainLen = out.dChainLen; a.chainSeeds = out.dChainSeeds;
    a.pass = 1;
    if (n) chain_kernel<<<gridFor(n), 256, 0, st>>>(a);
    CK(hipStreamSynchronize(st), "chain pass 1");
    return MSGS_OK;
}

/* The reference's DP is double arithmetic on the scores (src/Alignment.cpp:72-133).  When every
 * score is a multiple of 2^-k, every value the recurrences form is a multiple of 2^-k far below
 * 2^(53-k) in size, so each double add, subtract, max and compare is exact and equals the same
 * operation on the integers value x 2^k.  The device DP computes those integers (int32, |score
 * x 2^k| <= 10000 as for integer scores); k is the smallest in [0, 8] that makes all four
 * scores integral, and other score sets are refused. */
static bool integral(double x) { return std::isfinite(x) && x == std::floor(x) && std::fabs(x) <= 10000; }
static int dyadicShift(const msgs_params* p) {
    for (int k = 0; k <= 8; ++k) {
        const double f = std::ldexp(1.0, k);
        if (integral(p->match * f) && integral(p->mismatch * f) && integral(p->gapO * f) && integral(p->gapE * f))
            return k;
    }
    return -1;
}

static int runDp(msgs_ctx* c, const DevReads& r, const uint8_t* dQual, const msgs_params* p, const SeedStage& sd,
                 const ChainStage& ch, DpStage& out, double* dpMicros, double* phase = nullptr,
                 unsigned long long* dstat = nullptr) {
    const int64_t n = r.n, nJobs = ch.nJobs;
    const auto t0 = std::chrono::steady_clock::now();
    auto since = [&t0]() { return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count(); };
    hipStream_t st = c->x().stream;
    hipError_t e = hipSuccess;
    const int shift = dyadicShift(p);
    if (shift < 0)
        return fail(c, MSGS_E_UNSUPPORTED,
                    "the device DP needs match/mismatch/gap scores that are multiples of 1/256 (at most 10000 x 256 "
                    "in units of 1/256)");
    if (p->gapO < 0 || p->gapE < 0)
        return fail(c, MSGS_E_UNSUPPORTED, "the device DP needs non-negative gap penalties");
    out.recs.clear();
    if (nJobs == 0) return MSGS_OK;
    DpArgs a;
    std::memset(&a, 0, sizeof(a));
    a.seqs = r.seq; a.quals = dQual; a.offs = r.offs;
    const bool haveRc = r.rcSeq && r.rcQual;
    uint8_t* rcS = haveRc ? r.rcSeq : c->x().bufs.get<uint8_t>("dp.rs", r.total + 64, e);
    uint8_t* rcQ = haveRc ? r.rcQual : c->x().bufs.get<uint8_t>("dp.rq", r.total + 64, e);
    a.rcSeqs = rcS;
    a.rcQuals = rcQ;
    a.seedOff = sd.dSeedOff; a.seeds = sd.dSeeds; a.nJobs = nJobs;
    a.jobRead = ch.dJobRead; a.jobChainStart = ch.dChainStart; a.jobChainLen = ch.dChainLen;
    a.chainSeeds = ch.dChainSeeds;
    a.auxSG = c->x().bufs.get<int4>("dp.sg", ch.nChainSeeds + 1, e);
    a.auxGap = c->x().bufs.get<int64_t>("dp.gap", ch.nChainSeeds + 1, e);
    a.auxGapB = c->x().bufs.get<int64_t>("dp.gapb", ch.nChainSeeds + 1, e);
    a.geo = c->x().bufs.get<JobGeo>("dp.geo", nJobs + 1, e);
    a.sum = c->x().bufs.get<JobSum>("dp.sum", nJobs + 1, e);
    a.jobCls = c->x().bufs.get<int32_t>("dp.cls", nJobs + 1, e);
    a.mode = p->mode;
    const double sf = std::ldexp(1.0, shift);
    out.scoreScale = sf;
    a.match = (int32_t)(p->match * sf); a.mismatch = (int32_t)(p->mismatch * sf);
    a.gapO = (int32_t)(p->gapO * sf); a.gapE = (int32_t)(p->gapE * sf);
    a.clip = p->clip; a.gapE_d = p->gapE;
    a.dstat = dstat;
    a.forceSerial = c->dpSerial;
    /* quality terms of Alignment::evaluate with the host libm (src/Alignment.cpp:460-477,
     * src/QualStr.h:109-117) */
    double hq[3 * 256];
    for (int q = 0; q < 256; ++q) {
        hq[q] = ::log10(1 - ::exp(std::log(10.0) * q / -10.0));
        hq[256 + q] = q / -10.0;
        hq[512 + q] = q / -10.0 - p->gapO;
    }
    double* dq = c->x().bufs.get<double>("dp.qtab", 3 * 256, e);
    a.tileNeed = c->x().bufs.get<int64_t>("dp.tneed", nJobs + 1, e);
    a.pathNeed = c->x().bufs.get<int64_t>("dp.pneed", nJobs + 1, e);
    a.borderNeed = c->x().bufs.get<int64_t>("dp.bneed", nJobs + 1, e);
    a.serialNeed = c->x().bufs.get<int64_t>("dp.sneed", nJobs + 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "dp scratch");
    CK(hipMemcpyAsync(dq, hq, sizeof(hq), hipMemcpyHostToDevice, st), "qtab");
    a.qMatch = dq; a.qMis = dq + 256; a.qGapOpen = dq + 512;
    if (n && !haveRc)
        k_revcom<<<(unsigned)std::min<int64_t>((n + 3) / 4, 65536), 256, 0, st>>>(r.seq, dQual, r.offs, n, rcS, rcQ);
    dp_setup_kernel<<<gridFor(nJobs), 256, 0, st>>>(c->db.view(), a);
    /* tile and path offsets (exclusive scans with the totals at nJobs) and the wavefront
     * job list, all on the device; the host reads back three totals */
    const int nb = (int)((nJobs + 255) / 256);
    size_t scanBytes = 0;
    if (exclusiveSum(nullptr, scanBytes, a.tileNeed, a.tileNeed, nJobs + 1, st) != hipSuccess)
        return fail(c, MSGS_E_HIP, "scan size");
    void* scanTmp = c->x().bufs.get<uint8_t>("dp.scantmp", scanBytes + 1, e);
    int64_t* dTileOff = c->x().bufs.get<int64_t>("dp.toff", nJobs + 1, e);
    int64_t* dPathOff = c->x().bufs.get<int64_t>("dp.poff", nJobs + 1, e);
    int64_t* dBorderOff = c->x().bufs.get<int64_t>("dp.boff", nJobs + 1, e);
    int64_t* dSerialOff = c->x().bufs.get<int64_t>("dp.soff", nJobs + 1, e);
    int64_t* dWave = c->x().bufs.get<int64_t>("dp.wave", nJobs + 1, e);
    int64_t* dWave16 = c->x().bufs.get<int64_t>("dp.wave16", nJobs + 1, e);
    int64_t* dSerialJobs = c->x().bufs.get<int64_t>("dp.serjobs", nJobs + 1, e);
    int64_t* dNoneJobs = c->x().bufs.get<int64_t>("dp.nonejobs", nJobs + 1, e);
    unsigned long long* dHist = c->x().bufs.get<unsigned long long>("dp.whist", 512, e);
    int64_t* dTot = c->x().bufs.get<int64_t>("dp.tot", 8, e);
    unsigned int* dNext = c->x().bufs.get<unsigned int>("dp.next16", 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "dp scans");
    CK(hipMemsetAsync(dHist, 0, 512 * 8, st), "hist");
    CK(exclusiveSum(scanTmp, scanBytes, a.tileNeed, dTileOff, nJobs + 1, st), "tile scan");
    CK(exclusiveSum(scanTmp, scanBytes, a.pathNeed, dPathOff, nJobs + 1, st), "path scan");
    CK(exclusiveSum(scanTmp, scanBytes, a.borderNeed, dBorderOff, nJobs + 1, st), "border scan");
    CK(exclusiveSum(scanTmp, scanBytes, a.serialNeed, dSerialOff, nJobs + 1, st), "serial scan");
    k_wave_hist<<<nb, 256, 0, st>>>(a.jobCls, a.tileNeed, nJobs, dHist, CLS_WAVE);
    k_wave_base<<<1, 64, 0, st>>>(dHist, dHist + 64, dTot + 2);
    k_wave_scatter<<<nb, 256, 0, st>>>(a.jobCls, a.tileNeed, nJobs, dHist + 64, dWave, CLS_WAVE);
    k_wave_hist<<<nb, 256, 0, st>>>(a.jobCls, a.tileNeed, nJobs, dHist + 128, CLS_WAVE16);
    k_wave_base<<<1, 64, 0, st>>>(dHist + 128, dHist + 192, dTot + 5);
    k_wave_scatter<<<nb, 256, 0, st>>>(a.jobCls, a.tileNeed, nJobs, dHist + 192, dWave16, CLS_WAVE16);
    /* the trace's order: the other two classes by area as well */
    k_wave_hist<<<nb, 256, 0, st>>>(a.jobCls, a.serialNeed, nJobs, dHist + 256, CLS_SERIAL);
    k_wave_base<<<1, 64, 0, st>>>(dHist + 256, dHist + 320, dTot + 6);
    k_wave_scatter<<<nb, 256, 0, st>>>(a.jobCls, a.serialNeed, nJobs, dHist + 320, dSerialJobs, CLS_SERIAL);
    k_wave_hist<<<nb, 256, 0, st>>>(a.jobCls, a.tileNeed, nJobs, dHist + 384, CLS_NONE);
    k_wave_base<<<1, 64, 0, st>>>(dHist + 384, dHist + 448, dTot + 7);
    k_wave_scatter<<<nb, 256, 0, st>>>(a.jobCls, a.tileNeed, nJobs, dHist + 448, dNoneJobs, CLS_NONE);
    k_copy_i64<<<1, 64, 0, st>>>(dTileOff + nJobs, dTot + 0);
    k_copy_i64<<<1, 64, 0, st>>>(dPathOff + nJobs, dTot + 1);
    k_copy_i64<<<1, 64, 0, st>>>(dBorderOff + nJobs, dTot + 3);
    k_copy_i64<<<1, 64, 0, st>>>(dSerialOff + nJobs, dTot + 4);
    out.tot.resize(8);
    CK(hipMemcpyAsync(out.tot.data(), dTot, 8 * 8, hipMemcpyDeviceToHost, st), "dp totals");
    CK(hipStreamSynchronize(st), "dp setup");
    const int64_t tiles = out.tot[0], paths = out.tot[1], nWave = out.tot[2], borders = out.tot[3],
                  serialCells = out.tot[4], nWave16 = out.tot[5];
    if (phase) { phase[0] = (double)nJobs; phase[1] = (double)(nWave + nWave16); phase[2] = since(); }
    a.bits = c->x().bufs.get<uint8_t>("dp.bits", tiles + 16, e);   /* read in 16-byte chunks */
    a.border = c->x().bufs.get<int4>("dp.border", borders + 1, e);
    a.tile = c->x().bufs.get<int4>("dp.tile", serialCells + 1, e);
    a.cigar = c->x().bufs.get<uint32_t>("dp.cig", paths + 2 * nJobs + 1, e);
    a.md = c->x().bufs.get<char>("dp.md", 3 * paths + 8 * nJobs + 1, e);
    a.out = c->x().bufs.get<AlnRec>("dp.out", nJobs + 1, e);
    int64_t* dCOff = c->x().bufs.get<int64_t>("dp.coff", nJobs + 1, e);
    int64_t* dMOff = c->x().bufs.get<int64_t>("dp.moff", nJobs + 1, e);
    int64_t* dNc = c->x().bufs.get<int64_t>("dp.nc", nJobs + 1, e);
    int64_t* dNm = c->x().bufs.get<int64_t>("dp.nm", nJobs + 1, e);
    uint32_t* dC = c->x().bufs.get<uint32_t>("dp.cc", paths + 2 * nJobs + 1, e);
    char* dM = c->x().bufs.get<char>("dp.mc", 3 * paths + 8 * nJobs + 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "dp pools");
    a.tileOff = dTileOff; a.pathOff = dPathOff; a.borderOff = dBorderOff; a.serialOff = dSerialOff;
    a.waveJobs = dWave; a.nWaveJobs = nWave;
    a.wave16Jobs = dWave16; a.nWave16Jobs = nWave16; a.fill16Next = dNext;
    if (out.tot[6] + nWave + nWave16 + out.tot[7] == nJobs) {
        a.serialJobs = dSerialJobs; a.nSerialJobs = out.tot[6];
        a.noneJobs = dNoneJobs; a.nNoneJobs = out.tot[7];
    }
    CK(hipEventRecord(c->x().ev2, st), "event");
    if (nWave) {
        /* persistent: as many one-wavefront blocks per CU as are resident at once (blocks
         * started late would take large jobs of the first stride last) */
        static const int fillPerCU[2] = {[]() {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dp_fill_kernel<0>, 64, 0) != hipSuccess || nb < 1) nb = 16;
            return nb;
        }(), []() {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dp_fill_kernel<1>, 64, 0) != hipSuccess || nb < 1) nb = 16;
            return nb;
        }()};
        const int64_t blocks = std::min<int64_t>(nWave, (int64_t)std::max(1, c->nCU) * fillPerCU[p->mode ? 1 : 0]);
        if (p->mode == 0) dp_fill_kernel<0><<<(unsigned)blocks, 64, 0, st>>>(c->db.view(), a);
        else dp_fill_kernel<1><<<(unsigned)blocks, 64, 0, st>>>(c->db.view(), a);
    }
    if (nWave16) {
        /* persistent, four jobs per block; groups take jobs from a counter until none is left */
        static const int fill16PerCU[2] = {[]() {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dp_fill16_kernel<0>, 64, 0) != hipSuccess || nb < 1) nb = 8;
            return nb;
        }(), []() {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dp_fill16_kernel<1>, 64, 0) != hipSuccess || nb < 1) nb = 8;
            return nb;
        }()};
        const int64_t blocks = std::min<int64_t>((nWave16 + 3) / 4,
                                                 (int64_t)std::max(1, c->nCU) * fill16PerCU[p->mode ? 1 : 0]);
        k_zero_u32<<<1, 64, 0, st>>>(dNext);
        if (p->mode == 0) dp_fill16_kernel<0><<<(unsigned)blocks, 64, 0, st>>>(c->db.view(), a);
        else dp_fill16_kernel<1><<<(unsigned)blocks, 64, 0, st>>>(c->db.view(), a);
    }
    dp_trace_kernel<<<gridFor(nJobs), 256, 0, st>>>(c->db.view(), a);
    CK(hipEventRecord(c->x().ev3, st), "event");
    /* compaction of CIGAR ops / MD bytes, offsets scanned on the device */
    k_rec_counts<<<gridFor(nJobs + 1), 256, 0, st>>>(a.out, nJobs, dNc, dNm);
    CK(exclusiveSum(scanTmp, scanBytes, dNc, dCOff, nJobs + 1, st), "cigar scan");
    CK(exclusiveSum(scanTmp, scanBytes, dNm, dMOff, nJobs + 1, st), "md scan");
    k_compact<<<gridFor(nJobs), 256, 0, st>>>(a.out, nJobs, dPathOff, a.cigar, a.md, dCOff, dMOff, dC, dM);
    out.recs.resize(nJobs);
    out.cigOff.resize(nJobs + 1);
    out.mdOff.resize(nJobs + 1);
    CK(hipMemcpyAsync(out.recs.data(), a.out, nJobs * sizeof(AlnRec), hipMemcpyDeviceToHost, st), "recs");
    out.dRecs = a.out;
    CK(hipMemcpyAsync(out.cigOff.data(), dCOff, (nJobs + 1) * 8, hipMemcpyDeviceToHost, st), "coff");
    CK(hipMemcpyAsync(out.mdOff.data(), dMOff, (nJobs + 1) * 8, hipMemcpyDeviceToHost, st), "moff");
    CK(hipStreamSynchronize(st), "dp kernel");
    if (phase) phase[3] = since();
    const int64_t totC = out.cigOff[nJobs], totM = out.mdOff[nJobs];
    out.cigar.resize(totC + 1);
    out.md.resize(totM + 1);
    CK(hipMemcpyAsync(out.cigar.data(), dC, (totC + 1) * 4, hipMemcpyDeviceToHost, st), "cigar");
    CK(hipMemcpyAsync(out.md.data(), dM, totM + 1, hipMemcpyDeviceToHost, st), "md");
    CK(hipStreamSynchronize(st), "dp compaction");
    if (phase) phase[4] = since();
    if (dpMicros) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->x().ev2, c->x().ev3);
        *dpMicros = ms * 1000.0;
    }
    return MSGS_OK;
}

/* ------------------------------------------------------------------ host finishing */

namespace {

enum { FWD = 1, REV = 2 };

struct HostRead {
    const uint8_t* seq;
    const uint8_t* qual;
    int64_t L;
    const char* name;
    int64_t nameLen;
};

struct HAln {
    const AlnRec* r;
    double log10P;
    double postP = 0;
    uint8_t mapQ = 0xff;
};

/* Eigen's VectorXd::sum for SSE2 2-wide packets: two packet accumulators over
 * multiples of 4, one more packet, horizontal add, then the scalar tail. */
double packetSum(const std::vector<double>& v) {
    const size_t n = v.size();
    if (n == 0) return 0;
    const size_t a2 = (n / 4) * 4, a1 = (n / 2) * 2;
    if (a1 == 0) return v[0];
    double p0 = v[0], p1 = v[1];
    if (a1 > 2) {
        double q0 = v[2], q1 = v[3];
        for (size_t k = 4; k < a2; k += 4) { p0 += v[k]; p1 += v[k + 1]; q0 += v[k + 2]; q1 += v[k + 3]; }
        p0 += q0; p1 += q1;
        if (a1 > a2) { p0 += v[a2]; p1 += v[a2 + 1]; }
    }
    double s = p0 + p1;
    for (size_t k = a1; k < n; ++k) s += v[k];
    return s;
}

/* calcMapQ (src/Alignment.cpp:491-512 and :514-539) on log-likelihoods; pr is scratch */
void mapq(const std::vector<double>& ll, std::vector<double>& pr, std::vector<double>& postP, std::vector<uint8_t>& mq,
          bool pe) {
    const size_t n = ll.size();
    pr.assign(ll.begin(), ll.end());
    double maxV = *std::max_element(pr.begin(), pr.end());
    const double inf = std::numeric_limits<double>::infinity();
    double scale = (maxV != -inf && maxV < -256) ? -256 - maxV : 0;
    for (size_t i = 0; i < n; ++i) pr[i] = std::exp(pr[i] + scale);
    double sum = packetSum(pr);
    postP.resize(n);
    mq.resize(n);
    for (size_t i = 0; i < n; ++i) {
        double pp = (!pe || sum > 0) ? pr[i] / sum : 1.0 / (double)n;
        postP[i] = pp;
        double q = -10 * ::log10(1 - pp);
        mq[i] = (uint8_t)std::min(::round(q), 250.0);
    }
}

struct Finisher {
    const DbState& db;
    const msgs_params& P;
    const DpStage& dp;
    bool emit = true;
    bool bam = false;    /* records in the BAM layout instead of SAM text */
    int64_t mapped = 0;
    const double log10e = ::log10(M_E);
    /* time-stamp-counter ticks spent in: [0] SAM text, [1] pairing (0: the candidate pairs come from the
     * device), [2] MAPQ, [3] alignment lists */
    mutable uint64_t tk[4] = {0, 0, 0, 0};

    std::string cigarOf(int64_t job) const {
        static const char kOps[] = "MIDNSHP=XB";
        const AlnRec& r = dp.recs[job];
        const uint32_t* cg = dp.cigar.data() + dp.cigOff[job];
        std::string s;
        for (int k = 0; k < r.nCigar; ++k) { s += std::to_string(cg[k] >> 4); s += kOps[cg[k] & 15]; }
        return s;
    }

    /* one SAM text record (field layout of htslib's sam_format1) */
    void line(msgs_sam::Out& out, const HostRead& rd, bool rc, int flag, int32_t tid, int64_t pos, int mapq,
              const uint32_t* cg, int nc, int32_t isize, const int64_t* tags, const char* md, int mdLen, float xh,
              float xp) const {
        static const char kOps[] = "MIDNSHP=XB";
        if (!emit) return;
        const uint64_t t0 = __rdtsc();
        if (bam) {
            msgs_sam::BamFields f{rd.name, (int)rd.nameLen, flag, tid, pos, mapq, cg, nc, isize, rd.seq, rd.qual,
                                  rd.L, rc, tags != nullptr, tags ? tags[0] : 0, md, mdLen, xh,
                                  tags ? tags[1] : 0, tags ? tags[2] : 0, xp};
            char* p = out.reserve(msgs_sam::bamBound(f));
            out.commit(msgs_sam::putBam(p, f));
            tk[0] += __rdtsc() - t0;
            return;
        }
        const std::string* cn = tid >= 0 ? &db.names[tid] : nullptr;
        const size_t bound = (size_t)rd.nameLen + (cn ? cn->size() : 1) + 2 * (size_t)rd.L + 16 * (size_t)nc +
                             (size_t)mdLen + 256;
        char* p = out.reserve(bound);
        std::memcpy(p, rd.name, rd.nameLen); p += rd.nameLen;
        *p++ = '\t'; p = msgs_sam::putInt(p, flag);
        *p++ = '\t';
        if (cn) { std::memcpy(p, cn->data(), cn->size()); p += cn->size(); } else *p++ = '*';
        *p++ = '\t'; p = msgs_sam::putInt(p, pos + 1);
        *p++ = '\t'; p = msgs_sam::putInt(p, mapq);
        *p++ = '\t';
        if (nc == 0) *p++ = '*';
        for (int k = 0; k < nc; ++k) { p = msgs_sam::putInt(p, cg[k] >> 4); *p++ = kOps[cg[k] & 15]; }
        std::memcpy(p, "\t*\t1\t", 5); p += 5;   /* no mate: mtid -1, mpos 0 (printed + 1) */
        p = msgs_sam::putInt(p, isize);
        *p++ = '\t';
        if (rd.L == 0) *p++ = '*';
        else p = msgs_sam::putSeq(p, rd.seq, rd.L, rc);
        *p++ = '\t';
        if (rd.L == 0 || (rc ? rd.qual[rd.L - 1] : rd.qual[0]) == 0xff) *p++ = '*';
        else p = msgs_sam::putQual(p, rd.qual, rd.L, rc);
        if (tags) {
            std::memcpy(p, "\tAS:i:", 6); p = msgs_sam::putInt(p + 6, tags[0]);
            std::memcpy(p, "\tMD:Z:", 6); p += 6;
            std::memcpy(p, md, mdLen); p += mdLen;
            std::memcpy(p, "\tXH:f:", 6); p = msgs_sam::putG6(p + 6, xh);
            std::memcpy(p, "\tNH:i:", 6); p = msgs_sam::putInt(p + 6, tags[1]);
            std::memcpy(p, "\tXN:i:", 6); p = msgs_sam::putInt(p + 6, tags[2]);
            std::memcpy(p, "\tXP:f:", 6); p = msgs_sam::putG6(p + 6, xp);
        }
        *p++ = '\n';
        out.commit(p);
        tk[0] += __rdtsc() - t0;
    }

    void unmapped(msgs_sam::Out& s, const HostRead& rd, int flag) const {
        line(s, rd, false, flag, -1, 0, 0, nullptr, 0, 0, nullptr, nullptr, 0, 0, 0);
    }

    /* exportBAM + output() (src/Alignment.cpp:429-444, src/msgseqtk-align.cpp:538-557) */
    void alnRecord(msgs_sam::Out& s, const HostRead& rd, const HAln& a, int flagOr, int mapq, int32_t isize,
                   int64_t nh, int64_t xn, double postP) const {
        const AlnRec& r = *a.r;
        const int64_t job = &r - dp.recs.data();
        const bool rev = r.strand == REV;
        int flag = (rev ? 16 : 0) | flagOr;
        int64_t tags[3] = {(int64_t)(int)std::round(dp.score(r.score)), nh, xn};
        line(s, rd, rev, flag, r.tid, r.alnStart, mapq, dp.cigar.data() + dp.cigOff[job], r.nCigar, isize, tags,
             dp.md.data() + dp.mdOff[job], r.mdLen, (float)a.log10P, (float)postP);
    }

    /* per-thread scratch, reused across reads */
    std::vector<HAln> fBuf, rBuf;
    std::vector<double> ll, pr, pp;
    std::vector<uint8_t> mq;

    /* filter (src/Alignment.cpp:548-555): drop alnScore < L * minScoreRate */
    void alnList(const HostRead& rd, int64_t j0, int64_t j1, std::vector<HAln>& v) const {
        const uint64_t t0 = __rdtsc();
        v.clear();
        const double thr = (double)rd.L * P.minScoreRate;
        for (int64_t j = j0; j < j1; ++j) {
            const AlnRec& r = dp.recs[j];
            const double sc = dp.score(r.score);
            if (!(sc < thr)) v.push_back(HAln{&r, r.log10P});
        }
        tk[3] += __rdtsc() - t0;
    }

    void calcMapQ(std::vector<HAln>& v) {
        const uint64_t t0 = __rdtsc();
        ll.resize(v.size());
        for (size_t i = 0; i < v.size(); ++i) ll[i] = v[i].log10P / log10e;
        mapq(ll, pr, pp, mq, false);
        for (size_t i = 0; i < v.size(); ++i) { v[i].postP = pp[i]; v[i].mapQ = mq[i]; }
        tk[2] += __rdtsc() - t0;
    }

    static void sortAlns(std::vector<HAln>& v) {
        msgs_sort::sort(v.data(), v.data() + v.size(), [](const HAln& l, const HAln& r) { return l.log10P > r.log10P; });
    }

    void outputAlns(msgs_sam::Out& s, const HostRead& rd, const std::vector<HAln>& v, int extra) const {
        size_t num = P.maxReport == 0 ? v.size() : std::min((size_t)(uint32_t)P.maxReport, v.size());
        for (size_t i = 0; i < num; ++i)
            alnRecord(s, rd, v[i], extra | (i > 0 ? 256 : 0), v[i].mapQ, 0, (int64_t)num, (int64_t)v.size(), v[i].postP);
    }

    /* main_SE task (src/msgseqtk-align.cpp:607-649) */
    void se(msgs_sam::Out& s, const HostRead& rd, int32_t nSeeds, int64_t j0, int64_t j1) {
        if (nSeeds == 0) { unmapped(s, rd, 4); return; }
        std::vector<HAln>& v = fBuf;
        alnList(rd, j0, j1, v);
        if (v.empty()) { unmapped(s, rd, 4); return; }
        mapped++;
        calcMapQ(v);
        sortAlns(v);
        outputAlns(s, rd, v, 0);
    }

    /* main_PE task (src/msgseqtk-align.cpp:679-754); the candidate pairs q[0, nq) come from the
     * device (getPairs, sort and filter, msgs_pair.hip) */
    void pe(msgs_sam::Out& s, const HostRead& r1, const HostRead& r2, int32_t ns1, int32_t ns2, int64_t a0, int64_t a1,
            int64_t b0, int64_t b1, const PairRec* q, int32_t nq) {
        if (ns1 == 0 && ns2 == 0) { unmapped(s, r1, 4 | 8 | 1 | 64); unmapped(s, r2, 4 | 8 | 1 | 128); return; }
        std::vector<HAln>& f = fBuf;
        std::vector<HAln>& r = rBuf;
        alnList(r1, a0, a1, f);
        alnList(r2, b0, b1, r);
        if (f.empty() && r.empty()) { unmapped(s, r1, 4 | 8 | 1 | 64); unmapped(s, r2, 4 | 8 | 1 | 128); return; }
        const uint64_t tp1 = __rdtsc();
        if (nq > 0) {
            mapped += 2;
            ll.resize(nq);
            for (int32_t i = 0; i < nq; ++i) ll[i] = q[i].ll;
            mapq(ll, pr, pp, mq, true);
            tk[2] += __rdtsc() - tp1;
            size_t num = P.maxReport == 0 ? (size_t)nq : std::min((size_t)(uint32_t)P.maxReport, (size_t)nq);
            for (size_t i = 0; i < num; ++i) {
                const AlnRec &a = dp.recs[q[i].f], &b = dp.recs[q[i].r];
                const HAln hf{&a, a.log10P}, hr{&b, b.log10P};
                int32_t fi = a.alnStart < b.alnStart ? q[i].isize : -q[i].isize;
                int sec = i > 0 ? 256 : 0;
                alnRecord(s, r1, hf, 1 | 2 | 64 | sec, mq[i], fi, (int64_t)num, (int64_t)nq, pp[i]);
                alnRecord(s, r2, hr, 1 | 2 | 128 | sec, mq[i], -fi, (int64_t)num, (int64_t)nq, pp[i]);
            }
        } else if (!P.noMixed) {
            mapped += (f.empty() ? 0 : 1) + (r.empty() ? 0 : 1);
            if (!f.empty()) { calcMapQ(f); sortAlns(f); outputAlns(s, r1, f, 1 | 64); }
            if (!r.empty()) { calcMapQ(r); sortAlns(r); outputAlns(s, r2, r, 1 | 128); }
        }
    }
};

/* [start, end) of each of n newline-separated names */
void splitNames(const char* names, int64_t n, std::vector<int64_t>& off) {
    off.assign(2 * n, 0);
    const char* p = names ? names : "";
    const char* b = p;
    for (int64_t i = 0; i < n; ++i) {
        const char* e = std::strchr(p, '\n');
        if (!e) e = p + std::strlen(p);
        off[2 * i] = p - b;
        off[2 * i + 1] = e - b;
        p = *e ? e + 1 : e;
    }
}

}  // namespace

/* ------------------------------------------------------------------ one batch, all stages */

struct BatchResult {
    SmemStage sm;
    SeedStage sd;
    ChainStage ch;
    DpStage dp;
    PairStage pr;   /* paired batches: the candidate mate pairs */
};

static int uploadQual(msgs_ctx* c, const msgs_reads* r, const DevReads& d, const char* name, uint8_t** out) {
    hipError_t e = hipSuccess;
    uint8_t* q = c->x().bufs.get<uint8_t>(name, d.total + 64, e);   /* padded for word reads */
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "quality buffer");
    if (d.total) {
        if (r->qual) CK(hipMemcpyAsync(q, r->qual + r->offs[0], d.total, hipMemcpyHostToDevice, c->x().stream), "qual");
        else k_fill_u8<<<gridFor(d.total), 256, 0, c->x().stream>>>(q, d.total, 30);
    }
    *out = q;
    return MSGS_OK;
}

static int runAll(msgs_ctx* c, const DevReads& r, const uint8_t* dQual, const msgs_params* p, BatchResult& b,
                  unsigned long long* dStats, double* dpMicros) {
    int rc = runSmem(c, r, p, b.sm, dStats);
    if (rc) return rc;
    rc = runSeeds(c, r, p, b.sm, b.sd, dStats);
    if (rc) return rc;
    rc = runChains(c, r, b.sd, b.ch);
    if (rc) return rc;
    return runDp(c, r, dQual, p, b.sd, b.ch, b.dp, dpMicros);
}

static int checkParams(msgs_ctx* c, const msgs_params* p) {
    if (!p) return fail(c, MSGS_E_ARG, "no parameters");
    if (p->minLen <= 0) return fail(c, MSGS_E_ARG, "minLen must be positive");
    if (p->maxNSeed <= 0) return fail(c, MSGS_E_ARG, "maxNSeed must be positive");
    if (p->mode != 0 && p->mode != 1) return fail(c, MSGS_E_ARG, "mode must be 0 (GLOBAL) or 1 (LOCAL)");
    return MSGS_OK;
}

static int impl_seed_batch(msgs_ctx* c, const msgs_reads* reads, const msgs_params* p, int64_t* counts,
                               int64_t* rows6, double* logP, int64_t cap, int64_t* needed) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    int rc = checkParams(c, p);
    if (rc) return rc;
    DevReads r;
    rc = uploadReads(c, reads, "rd", r);
    if (rc) return rc;
    SmemStage sm;
    SeedStage sd;
    rc = runSmem(c, r, p, sm, nullptr);
    if (rc) return rc;
    rc = runSeeds(c, r, p, sm, sd, nullptr);
    if (rc) return rc;
    const int64_t n = r.n;
    std::vector<SeedRec> seeds(sd.seedOff[n] + 1);
    CK(hipMemcpyAsync(seeds.data(), sd.dSeeds, seeds.size() * sizeof(SeedRec), hipMemcpyDeviceToHost, c->x().stream), "seeds");
    CK(hipStreamSynchronize(c->x().stream), "sync");
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i) total += sd.seedCnt[i];
    if (needed) *needed = total;
    if (total > cap) return fail(c, MSGS_E_CAP, "row capacity");
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (counts) counts[i] = sd.seedCnt[i];
        for (int32_t j = 0; j < sd.seedCnt[i]; ++j, ++k) {
            const SeedRec& s = seeds[sd.seedOff[i] + j];
            int64_t* o = rows6 + 6 * k;
            o[0] = s.from; o[1] = s.to; o[2] = s.start; o[3] = s.end; o[4] = s.tid; o[5] = s.strand;
            if (logP) logP[k] = s.logP;
        }
    }
    return MSGS_OK;
}

extern "C" int msgs_seed_batch(msgs_ctx* c, const msgs_reads* reads, const msgs_params* p, int64_t* counts,
                               int64_t* rows6, double* logP, int64_t cap, int64_t* needed) {
    try {
        return impl_seed_batch(c, reads, p, counts, rows6, logP, cap, needed);
    } catch (const std::bad_alloc&) {
        return fail(c, MSGS_E_NOMEM, "host memory");
    }
}

static int impl_align_batch(msgs_ctx* c, const msgs_reads* reads, const msgs_params* p, int64_t* counts,
                                int64_t* rows10, double* scores, char* cigars, int64_t cigCap, int64_t cap,
                                int64_t* needed) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    int rc = checkParams(c, p);
    if (rc) return rc;
    DevReads r;
    rc = uploadReads(c, reads, "rd", r);
    if (rc) return rc;
    uint8_t* dQual = nullptr;
    rc = uploadQual(c, reads, r, "rd.qual", &dQual);
    if (rc) return rc;
    BatchResult b;
    rc = runAll(c, r, dQual, p, b, nullptr, nullptr);
    if (rc) return rc;
    const int64_t nJobs = b.ch.nJobs;
    if (needed) *needed = nJobs;
    if (nJobs > cap) return fail(c, MSGS_E_CAP, "row capacity");
    Finisher F{c->db, *p, b.dp};
    std::string cig;
    for (int64_t i = 0; i < r.n; ++i)
        if (counts) counts[i] = b.ch.chainOff[i + 1] - b.ch.chainOff[i];
    for (int64_t j = 0; j < nJobs; ++j) {
        const AlnRec& a = b.dp.recs[j];
        int64_t* o = rows10 + 10 * j;
        o[0] = a.tid; o[1] = a.strand; o[2] = a.tStart; o[3] = a.tEnd; o[4] = a.alnFrom; o[5] = a.alnTo;
        o[6] = a.alnStart; o[7] = a.alnEnd; o[8] = a.hazard; o[9] = a.cells;
        if (scores) scores[j] = b.dp.score(a.score);
        cig += F.cigarOf(j);
        cig += '\n';
    }
    if (cigars) {
        if ((int64_t)cig.size() + 1 > cigCap) return fail(c, MSGS_E_CAP, "cigar capacity");
        std::memcpy(cigars, cig.c_str(), cig.size() + 1);
    }
    return MSGS_OK;
}

extern "C" int msgs_align_batch(msgs_ctx* c, const msgs_reads* reads, const msgs_params* p, int64_t* counts,
                                int64_t* rows10, double* scores, char* cigars, int64_t cigCap, int64_t cap,
                                int64_t* needed) {
    try {
        return impl_align_batch(c, reads, p, counts, rows10, scores, cigars, cigCap, cap, needed);
    } catch (const std::bad_alloc&) {
        return fail(c, MSGS_E_NOMEM, "host memory");
    }
}

/* Alignment::calcMapQ / AlignmentPE::calcMapQ on log-likelihoods (natural log) */
extern "C" void msgs_util_mapq(const double* loglik, int64_t n, int32_t pe, double* postP, uint8_t* mapQ) {
    if (n <= 0) return;
    std::vector<double> ll(loglik, loglik + n), pr, pp;
    std::vector<uint8_t> mq;
    mapq(ll, pr, pp, mq, pe != 0);
    for (int64_t i = 0; i < n; ++i) {
        if (postP) postP[i] = pp[i];
        if (mapQ) mapQ[i] = mq[i];
    }
}

/* ------------------------------------------------------------------ caller seeds and chains */

/* a seed row (from, to, start, end, tid, strand) as a SeedRec */
static SeedRec seedOf(const int64_t* r) {
    SeedRec s;
    s.from = r[0]; s.to = r[1]; s.start = r[2]; s.end = r[3]; s.tid = (int32_t)r[4]; s.strand = (int32_t)r[5]; s.logP = 0;
    return s;
}

/* device copies of per-read seed lists */
static int uploadSeeds(msgs_ctx* c, const std::vector<SeedRec>& seeds, const std::vector<int64_t>& off, SeedStage& sd) {
    const int64_t n = (int64_t)off.size() - 1;
    hipError_t e = hipSuccess;
    hipStream_t st = c->x().stream;
    sd.seedOff = off;
    sd.seedCnt.resize(n);
    for (int64_t i = 0; i < n; ++i) sd.seedCnt[i] = (int32_t)(off[i + 1] - off[i]);
    sd.dSeedOff = c->x().bufs.get<int64_t>("cs.off", n + 1, e);
    sd.dSeedCnt = c->x().bufs.get<int32_t>("cs.cnt", n + 1, e);
    sd.dSeeds = c->x().bufs.get<SeedRec>("cs.seeds", seeds.size() + 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "seed buffers");
    CK(hipMemcpyAsync(sd.dSeedOff, off.data(), (n + 1) * 8, hipMemcpyHostToDevice, st), "seed offsets");
    if (n) CK(hipMemcpyAsync(sd.dSeedCnt, sd.seedCnt.data(), n * 4, hipMemcpyHostToDevice, st), "seed counts");
    if (!seeds.empty())
        CK(hipMemcpyAsync(sd.dSeeds, seeds.data(), seeds.size() * sizeof(SeedRec), hipMemcpyHostToDevice, st), "seeds");
    return MSGS_OK;
}

static int impl_chain_batch(msgs_ctx* c, const int64_t* rows6, const int64_t* seedOffs, int64_t nSets,
                            const int64_t* maxMM, const int64_t* maxIndel, int64_t* chainCounts, int64_t* out,
                            int64_t cap, int64_t* needed) {
    if (!c) return MSGS_E_ARG;
    if (nSets < 0 || (nSets > 0 && (!seedOffs || !maxMM || !maxIndel))) return fail(c, MSGS_E_ARG, "bad arguments");
    if (nSets == 0) { if (needed) *needed = 0; return MSGS_OK; }
    std::vector<int64_t> off(nSets + 1);
    for (int64_t i = 0; i <= nSets; ++i) {
        off[i] = seedOffs[i] - seedOffs[0];
        if (i && (off[i] < off[i - 1] || off[i] - off[i - 1] > (1 << 24))) return fail(c, MSGS_E_ARG, "bad seed offsets");
    }
    if (off[nSets] > 0 && !rows6) return fail(c, MSGS_E_ARG, "no seed rows");
    std::vector<SeedRec> seeds(off[nSets]);
    for (int64_t k = 0; k < off[nSets]; ++k) seeds[k] = seedOf(rows6 + 6 * (seedOffs[0] + k));
    SeedStage sd;
    int rc = uploadSeeds(c, seeds, off, sd);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    hipStream_t st = c->x().stream;
    DevReads r;
    r.n = nSets;
    r.offs = c->x().bufs.get<int64_t>("cs.roffs", nSets + 1, e);   /* lengths unused: limits given */
    int64_t* dMM = c->x().bufs.get<int64_t>("cs.mm", nSets, e);
    int64_t* dID = c->x().bufs.get<int64_t>("cs.id", nSets, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "chain limits");
    CK(hipMemsetAsync(r.offs, 0, (nSets + 1) * 8, st), "offsets");
    CK(hipMemcpyAsync(dMM, maxMM, nSets * 8, hipMemcpyHostToDevice, st), "maxMismatch");
    CK(hipMemcpyAsync(dID, maxIndel, nSets * 8, hipMemcpyHostToDevice, st), "maxIndel");
    ChainStage ch;
    rc = runChains(c, r, sd, ch, dMM, dID);
    if (rc) return rc;
    std::vector<int32_t> cStart(ch.nJobs + 1), cLen(ch.nJobs + 1), cSeeds(ch.nChainSeeds + 1);
    if (ch.nJobs) {
        CK(hipMemcpyAsync(cStart.data(), ch.dChainStart, ch.nJobs * 4, hipMemcpyDeviceToHost, st), "chains");
        CK(hipMemcpyAsync(cLen.data(), ch.dChainLen, ch.nJobs * 4, hipMemcpyDeviceToHost, st), "chains");
    }
    if (ch.nChainSeeds) CK(hipMemcpyAsync(cSeeds.data(), ch.dChainSeeds, ch.nChainSeeds * 4, hipMemcpyDeviceToHost, st), "chains");
    CK(hipStreamSynchronize(st), "chain copy");
    const int64_t total = ch.nJobs + ch.nChainSeeds;
    if (needed) *needed = total;
    if (total > cap || !out) return fail(c, MSGS_E_CAP, "chain output capacity");
    int64_t k = 0;
    for (int64_t i = 0; i < nSets; ++i) {
        if (chainCounts) chainCounts[i] = ch.chainOff[i + 1] - ch.chainOff[i];
        for (int64_t j = ch.chainOff[i]; j < ch.chainOff[i + 1]; ++j) {
            out[k++] = cLen[j];
            for (int32_t t = 0; t < cLen[j]; ++t) out[k++] = cSeeds[cStart[j] + t];
        }
    }
    return MSGS_OK;
}

extern "C" int msgs_chain_batch(msgs_ctx* c, const int64_t* rows6, const int64_t* seedOffs, int64_t nSets,
                                const int64_t* maxMismatch, const int64_t* maxIndel, int64_t* chainCounts,
                                int64_t* out, int64_t cap, int64_t* needed) {
    try {
        return impl_chain_batch(c, rows6, seedOffs, nSets, maxMismatch, maxIndel, chainCounts, out, cap, needed);
    } catch (const std::bad_alloc&) {
        return fail(c, MSGS_E_NOMEM, "host memory");
    }
}

static int impl_dp_batch(msgs_ctx* c, const msgs_reads* reads, const int64_t* chainOffs, const int64_t* chainSeedOffs,
                         const int64_t* rows6, const msgs_params* p, int64_t* rows10, double* scores, double* log10P,
                         char* cigars, int64_t cigCap, char* mds, int64_t mdCap, int64_t* needed) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    int rc = checkParams(c, p);
    if (rc) return rc;
    if (!reads || !chainOffs || !chainSeedOffs) return fail(c, MSGS_E_ARG, "bad arguments");
    const int64_t n = reads->n;
    const int64_t nJobs = chainOffs[n] - chainOffs[0];
    if (needed) *needed = nJobs;
    /* every chain is non-empty, its seeds are exact-length pairs inside the read and the
     * chromosome, on one tid and strand, in order on both sequences (SeedChain's invariant,
     * src/SeedPair.h:156-158); the DP's geometry relies on it */
    std::vector<SeedRec> seeds;
    std::vector<int64_t> seedOff(n + 1, 0), chainOff(n + 1, 0);
    std::vector<int32_t> jobRead(nJobs), cStart(nJobs), cLen(nJobs), cSeeds;
    int64_t j = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t L = reads->offs[i + 1] - reads->offs[i];
        seedOff[i] = (int64_t)seeds.size();
        chainOff[i] = j;
        if (chainOffs[i + 1] < chainOffs[i]) return fail(c, MSGS_E_ARG, "chain offsets must not decrease");
        for (int64_t q = chainOffs[i]; q < chainOffs[i + 1]; ++q, ++j) {
            const int64_t s0 = chainSeedOffs[q], s1 = chainSeedOffs[q + 1];
            if (s1 <= s0) return fail(c, MSGS_E_ARG, "empty chain");
            jobRead[j] = (int32_t)i;
            cStart[j] = (int32_t)cSeeds.size();
            cLen[j] = (int32_t)(s1 - s0);
            for (int64_t k = s0; k < s1; ++k) {
                const SeedRec sr = seedOf(rows6 + 6 * k);
                if (sr.tid < 0 || sr.tid >= c->db.nChrom || (sr.strand != 1 && sr.strand != 2) || sr.from < 0 ||
                    sr.to > L || sr.to <= sr.from || sr.end - sr.start != sr.to - sr.from || sr.start < 0 ||
                    sr.end > c->db.chromLen[sr.tid])
                    return fail(c, MSGS_E_ARG, "seed outside the read or the chromosome");
                if (k > s0) {
                    const SeedRec pr = seedOf(rows6 + 6 * (k - 1));
                    if (pr.tid != sr.tid || pr.strand != sr.strand || pr.to > sr.from || pr.end > sr.start)
                        return fail(c, MSGS_E_ARG, "chain seeds not compatible and ordered");
                }
                cSeeds.push_back((int32_t)(seeds.size() - seedOff[i]));
                seeds.push_back(sr);
            }
        }
    }
    seedOff[n] = (int64_t)seeds.size();
    chainOff[n] = j;
    DevReads r;
    rc = uploadReads(c, reads, "dq", r);
    if (rc) return rc;
    uint8_t* dQual = nullptr;
    rc = uploadQual(c, reads, r, "dq.qual", &dQual);
    if (rc) return rc;
    SeedStage sd;
    rc = uploadSeeds(c, seeds, seedOff, sd);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    hipStream_t st = c->x().stream;
    ChainStage ch;
    ch.nJobs = nJobs;
    ch.chainOff = chainOff;
    ch.nChainSeeds = (int64_t)cSeeds.size();
    ch.dJobRead = c->x().bufs.get<int32_t>("dq.read", nJobs + 1, e);
    ch.dChainStart = c->x().bufs.get<int32_t>("dq.start", nJobs + 1, e);
    ch.dChainLen = c->x().bufs.get<int32_t>("dq.len", nJobs + 1, e);
    ch.dChainSeeds = c->x().bufs.get<int32_t>("dq.seeds", cSeeds.size() + 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "chain buffers");
    if (nJobs) {
        CK(hipMemcpyAsync(ch.dJobRead, jobRead.data(), nJobs * 4, hipMemcpyHostToDevice, st), "chains");
        CK(hipMemcpyAsync(ch.dChainStart, cStart.data(), nJobs * 4, hipMemcpyHostToDevice, st), "chains");
        CK(hipMemcpyAsync(ch.dChainLen, cLen.data(), nJobs * 4, hipMemcpyHostToDevice, st), "chains");
        CK(hipMemcpyAsync(ch.dChainSeeds, cSeeds.data(), cSeeds.size() * 4, hipMemcpyHostToDevice, st), "chains");
    }
    DpStage dp;
    rc = runDp(c, r, dQual, p, sd, ch, dp, nullptr);
    if (rc) return rc;
    Finisher F{c->db, *p, dp};
    std::string cig, md;
    for (int64_t k = 0; k < nJobs; ++k) {
        const AlnRec& a = dp.recs[k];
        if (rows10) {
            int64_t* o = rows10 + 10 * k;
            o[0] = a.tid; o[1] = a.strand; o[2] = a.tStart; o[3] = a.tEnd; o[4] = a.alnFrom; o[5] = a.alnTo;
            o[6] = a.alnStart; o[7] = a.alnEnd; o[8] = a.hazard; o[9] = a.cells;
        }
        if (scores) scores[k] = dp.score(a.score);
        if (log10P) log10P[k] = a.log10P;
        cig += F.cigarOf(k);
        cig += '\n';
        md.append(dp.md.data() + dp.mdOff[k], (size_t)a.mdLen);
        md += '\n';
    }
    if (cigars) {
        if ((int64_t)cig.size() + 1 > cigCap) return fail(c, MSGS_E_CAP, "cigar capacity");
        std::memcpy(cigars, cig.c_str(), cig.size() + 1);
    }
    if (mds) {
        if ((int64_t)md.size() + 1 > mdCap) return fail(c, MSGS_E_CAP, "MD capacity");
        std::memcpy(mds, md.c_str(), md.size() + 1);
    }
    return MSGS_OK;
}

extern "C" int msgs_dp_batch(msgs_ctx* c, const msgs_reads* reads, const int64_t* chainOffs,
                             const int64_t* chainSeedOffs, const int64_t* rows6, const msgs_params* p, int64_t* rows10,
                             double* scores, double* log10P, char* cigars, int64_t cigCap, char* mds, int64_t mdCap,
                             int64_t* needed) {
    try {
        return impl_dp_batch(c, reads, chainOffs, chainSeedOffs, rows6, p, rows10, scores, log10P, cigars, cigCap, mds,
                             mdCap, needed);
    } catch (const std::bad_alloc&) {
        return fail(c, MSGS_E_NOMEM, "host memory");
    }
}

/* SAM text of a whole batch (SE when mates is NULL; otherwise b holds the reads followed by
 * their mates, mate i as read n + i), finished on host threads in read-order chunks
 * (part t holds the records of the t-th chunk); lineOffs[i] is where
 * read i's records start in the parts' concatenation.  Returns the number of mapped reads. */
static int64_t samText(msgs_ctx* c, const msgs_params* p, const msgs_reads* reads, const msgs_reads* mates,
                       const BatchResult& b, bool emit, std::vector<std::string>& part, std::vector<int64_t>& lineOffs,
                       const std::vector<int64_t>* names1 = nullptr, const std::vector<int64_t>* names2 = nullptr,
                       int threads = 0, double* phaseUs = nullptr) {
    const int64_t n = reads->n;
    std::vector<int64_t> own1, own2;
    if (!names1) { splitNames(reads->names, n, own1); names1 = &own1; }
    if (mates && !names2) { splitNames(mates->names, n, own2); names2 = &own2; }
    const std::vector<int64_t>& nm1 = *names1;
    const std::vector<int64_t>& nm2 = mates ? *names2 : own2;
    lineOffs.assign(n + 1, 0);
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(threads > 0 ? threads : c->hostThreads, (n + 255) / 256));
    part.resize(nt);
    for (std::string& x : part) x.clear();   /* keeps the capacity of recycled parts */
    std::vector<int64_t> mappedPart(nt, 0);
    std::vector<std::array<uint64_t, 4>> tkPart(nt);
    const uint64_t tsc0 = __rdtsc();
    const auto w0 = std::chrono::steady_clock::now();
    auto work = [&](int t) {
        const int64_t i0 = n * t / nt, i1 = n * (t + 1) / nt;
        Finisher F{c->db, *p, b.dp};
        F.emit = emit;
        F.bam = c->recordFormat == 1;
        std::vector<uint8_t> q30;
        auto host = [&](const msgs_reads* r, const std::vector<int64_t>& nm, int64_t i) {
            HostRead h;
            h.seq = r->seq + r->offs[i];
            h.L = r->offs[i + 1] - r->offs[i];
            if (r->qual) h.qual = r->qual + r->offs[i];
            else { if ((int64_t)q30.size() < h.L) q30.assign(h.L, 30); h.qual = q30.data(); }
            if (nm[2 * i + 1] > nm[2 * i]) { h.name = r->names + nm[2 * i]; h.nameLen = nm[2 * i + 1] - nm[2 * i]; }
            else { h.name = "r"; h.nameLen = 1; }
            return h;
        };
        msgs_sam::Out o;
        o.buf = &part[t];
        for (int64_t i = i0; i < i1; ++i) {
            lineOffs[i] = (int64_t)o.size();
            HostRead h1 = host(reads, nm1, i);
            if (!mates) {
                F.se(o, h1, b.sd.seedCnt[i], b.ch.chainOff[i], b.ch.chainOff[i + 1]);
            } else {
                HostRead h2 = host(mates, nm2, i);
                F.pe(o, h1, h2, b.sd.seedCnt[i], b.sd.seedCnt[n + i], b.ch.chainOff[i], b.ch.chainOff[i + 1],
                     b.ch.chainOff[n + i], b.ch.chainOff[n + i + 1], b.pr.pairs.data() + b.pr.off[i], b.pr.kept[i]);
            }
        }
        o.finish();
        mappedPart[t] = F.mapped;
        for (int k = 0; k < 4; ++k) tkPart[t][k] = F.tk[k];
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    int64_t base = 0, mapped = 0;
    for (int t = 0; t < nt; ++t) {
        const int64_t i0 = n * t / nt, i1 = n * (t + 1) / nt;
        for (int64_t i = i0; i < i1; ++i) lineOffs[i] += base;
        base += (int64_t)part[t].size();
        mapped += mappedPart[t];
    }
    lineOffs[n] = base;
    if (phaseUs) {
        /* ticks to microseconds with the counter's rate over this call */
        const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - w0).count();
        const double perUs = (double)(__rdtsc() - tsc0) / std::max(1.0, us);
        for (int k = 0; k < 4; ++k) {
            uint64_t sum = 0;
            for (int t = 0; t < nt; ++t) sum += tkPart[t][k];
            phaseUs[k] = perUs > 0 ? (double)sum / perUs : 0;
        }
    }
    return mapped;
}

static void copyParts(const std::vector<std::string>& part, char* out) {
    int64_t k = 0;
    for (const std::string& s : part) { std::memcpy(out + k, s.data(), s.size()); k += (int64_t)s.size(); }
}

/* reads followed by their mates as one batch (mate i becomes read n + i) */
static void joinMates(const msgs_reads* a, const msgs_reads* b, HostBatch& out) {
    const int64_t n = a->n;
    const int64_t ta = n ? a->offs[n] - a->offs[0] : 0, tb = n ? b->offs[n] - b->offs[0] : 0;
    out.seq.resize(ta + tb);
    out.qual.resize(ta + tb);
    if (ta) std::memcpy(out.seq.data(), a->seq + a->offs[0], ta);
    if (tb) std::memcpy(out.seq.data() + ta, b->seq + b->offs[0], tb);
    if (ta) { if (a->qual) std::memcpy(out.qual.data(), a->qual + a->offs[0], ta); else std::memset(out.qual.data(), 30, ta); }
    if (tb) { if (b->qual) std::memcpy(out.qual.data() + ta, b->qual + b->offs[0], tb); else std::memset(out.qual.data() + ta, 30, tb); }
    out.offs.assign(2 * n + 1, 0);
    for (int64_t i = 0; i <= n; ++i) out.offs[i] = a->offs[i] - a->offs[0];
    for (int64_t i = 1; i <= n; ++i) out.offs[n + i] = ta + b->offs[i] - b->offs[0];
    out.view.seq = out.seq.data();
    out.view.qual = out.qual.data();
    out.view.offs = out.offs.data();
    out.view.names = nullptr;
    out.view.n = 2 * n;
}

static int impl_map_batch(msgs_ctx* c, const msgs_reads* reads, const msgs_reads* mates, const msgs_params* p,
                              char* out, int64_t cap, int64_t* lineOffs, int64_t* needed) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    int rc = checkParams(c, p);
    if (rc) return rc;
    if (!reads) return fail(c, MSGS_E_ARG, "no reads");
    if (mates && mates->n != reads->n) return fail(c, MSGS_E_ARG, "mate batches differ in size");
    HostBatch joined;
    const msgs_reads* all = reads;
    if (mates) { joinMates(reads, mates, joined); all = &joined.view; }
    BatchResult b;
    {
        DevReads r;
        rc = uploadReads(c, all, "rd", r);
        if (rc) return rc;
        uint8_t* dQual = nullptr;
        rc = uploadQual(c, all, r, "rd.qual", &dQual);
        if (rc) return rc;
        rc = runAll(c, r, dQual, p, b, nullptr, nullptr);
        if (rc) return rc;
        if (mates) {
            rc = runPairs(c, reads->n, b.ch.chainOff, b.dp.dRecs, r.offs, true, b.dp.scoreScale, p, b.pr);
            if (rc) return rc;
        }
    }
    std::vector<std::string> parts;
    std::vector<int64_t> lo;
    samText(c, p, reads, mates, b, true, parts, lo);
    const int64_t total = lo[reads->n];
    if (needed) *needed = total;
    if (total > cap || !out) return fail(c, MSGS_E_CAP, "output capacity");
    copyParts(parts, out);
    if (lineOffs) std::memcpy(lineOffs, lo.data(), lo.size() * 8);
    return MSGS_OK;
}

extern "C" int msgs_map_batch(msgs_ctx* c, const msgs_reads* reads, const msgs_reads* mates, const msgs_params* p,
                              char* out, int64_t cap, int64_t* lineOffs, int64_t* needed) {
    try {
        return impl_map_batch(c, reads, mates, p, out, cap, lineOffs, needed);
    } catch (const std::bad_alloc&) {
        return fail(c, MSGS_E_NOMEM, "host memory");
    }
}

/* ------------------------------------------------------------------ resident batches */

static std::vector<uint8_t> hostCopy(const uint8_t* p, int64_t n) { return p ? std::vector<uint8_t>(p, p + n) : std::vector<uint8_t>(); }

extern "C" int msgs_batch_upload(msgs_ctx* c, int32_t slot, const msgs_reads* reads, const msgs_reads* mates) {
    if (!c) return MSGS_E_ARG;
    if (slot < 0 || slot >= 16 || !reads) return fail(c, MSGS_E_ARG, "bad slot or reads");
    if (mates && mates->n != reads->n) return fail(c, MSGS_E_ARG, "mate batches differ in size");
    Slot& S = c->slots[slot];
    S.release();
    const msgs_reads* rs[2] = {reads, mates};
    for (int m = 0; m < (mates ? 2 : 1); ++m) {
        const msgs_reads* r = rs[m];
        HostBatch& hb = S.host[m];
        const int64_t n = r->n, total = n ? r->offs[n] - r->offs[0] : 0;
        hb.seq = hostCopy(r->seq + (n ? r->offs[0] : 0), total);
        hb.qual = r->qual ? hostCopy(r->qual + (n ? r->offs[0] : 0), total) : std::vector<uint8_t>(total, 30);
        hb.offs.assign(n + 1, 0);
        for (int64_t i = 0; i <= n; ++i) hb.offs[i] = n ? r->offs[i] - r->offs[0] : 0;
        hb.names = r->names ? std::string(r->names) : std::string();
        hb.view.seq = hb.seq.empty() ? nullptr : hb.seq.data();
        hb.view.qual = hb.qual.empty() ? nullptr : hb.qual.data();
        hb.view.offs = hb.offs.data();
        hb.view.names = hb.names.empty() ? nullptr : hb.names.c_str();
        hb.view.n = n;
        splitNames(hb.view.names, n, hb.nameOff);
    }
    /* on the device the mates follow the reads as one batch */
    HostBatch joined;
    const msgs_reads* all = &S.host[0].view;
    if (mates) { joinMates(&S.host[0].view, &S.host[1].view, joined); all = &joined.view; }
    char pfx[32];
    std::snprintf(pfx, sizeof pfx, "slot%d", slot);
    int rc = uploadReads(c, all, pfx, S.dev);
    if (rc) return rc;
    std::snprintf(pfx, sizeof pfx, "slot%d.q", slot);
    rc = uploadQual(c, all, S.dev, pfx, &S.dQual);
    if (rc) return rc;
    /* reverse complements for the DP, once per resident batch */
    {
        hipError_t e = hipSuccess;
        std::snprintf(pfx, sizeof pfx, "slot%d.rs", slot);
        S.dev.rcSeq = c->x().bufs.get<uint8_t>(pfx, S.dev.total + 64, e);
        std::snprintf(pfx, sizeof pfx, "slot%d.rq", slot);
        S.dev.rcQual = c->x().bufs.get<uint8_t>(pfx, S.dev.total + 64, e);
        if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "reverse complements");
        const int64_t n = S.dev.n;
        if (n)
            k_revcom<<<(unsigned)std::min<int64_t>((n + 3) / 4, 65536), 256, 0, c->x().stream>>>(
                S.dev.seq, S.dQual, S.dev.offs, n, S.dev.rcSeq, S.dev.rcQual);
    }
    if (hipStreamSynchronize(c->x().stream) != hipSuccess) return fail(c, MSGS_E_HIP, "upload");
    /* pinned copies for the per-run ingest */
    S.inSeq.resize(S.dev.total + 1);
    S.inQual.resize(S.dev.total + 1);
    if (S.dev.total) {
        std::memcpy(S.inSeq.data(), all->seq + all->offs[0], S.dev.total);
        if (all->qual) std::memcpy(S.inQual.data(), all->qual + all->offs[0], S.dev.total);
        else std::memset(S.inQual.data(), 30, S.dev.total);
    }
    S.paired = mates != nullptr;
    S.used = true;
    return MSGS_OK;
}

/* GPU stages of a resident batch, then its host finishing on the slot's worker thread */
static int batchSubmit(msgs_ctx* c, int32_t slot, const msgs_params* p, int32_t emitSam);

extern "C" int msgs_batch_submit(msgs_ctx* c, int32_t slot, const msgs_params* p, int32_t emitSam) {
    try {
        return batchSubmit(c, slot, p, emitSam);
    } catch (const std::bad_alloc&) {
        return fail(c, MSGS_E_NOMEM, "host memory");
    }
}

/* the device stages of a submitted batch, on the calling worker's executor; fills S.stats */
static int batchGpu(msgs_ctx* c, Slot& S, std::chrono::steady_clock::time_point t0) {
    const msgs_params* p = &S.params;
    hipError_t e = hipSuccess;
    unsigned long long* dStats = c->x().bufs.get<unsigned long long>("stats", 24, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "stats");
    CK(hipMemsetAsync(dStats, 0, 24 * 8, c->x().stream), "stats");
    std::shared_ptr<BatchResult> b = S.res;
    double seedUs = 0, smemUs = 0, dpUs = 0, filterUs = 0, ingestUs = 0;
    if (c->ingest && S.dev.total) {
        /* the batch's bases and qualities from pinned host memory, then the reverse
         * complements, on this worker's stream (other batches' stages run meanwhile) */
        hipStream_t st = c->x().stream;
        const int64_t n = S.dev.n;
        CK(hipEventRecord(c->x().ev6, st), "event");
        CK(hipMemcpyAsync(S.dev.seq, S.inSeq.data(), S.dev.total, hipMemcpyHostToDevice, st), "ingest bases");
        CK(hipMemcpyAsync(S.dQual, S.inQual.data(), S.dev.total, hipMemcpyHostToDevice, st), "ingest qualities");
        k_revcom<<<(unsigned)std::min<int64_t>((n + 3) / 4, 65536), 256, 0, st>>>(S.dev.seq, S.dQual, S.dev.offs, n,
                                                                                  S.dev.rcSeq, S.dev.rcQual);
        CK(hipEventRecord(c->x().ev7, st), "event");
    }
    double wall[4] = {0, 0, 0, 0};   /* host wall time of the SMEM, seed, chain and DP stages */
    double dpPhase[5] = {0, 0, 0, 0, 0};
    int rc;
    {
        auto us = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
            return std::chrono::duration<double, std::micro>(b - a).count();
        };
        auto s0 = std::chrono::steady_clock::now();
        rc = runSmem(c, S.dev, p, b->sm, dStats);
        if (rc) return rc;
        auto s1 = std::chrono::steady_clock::now();
        smemUs += c->x().lastSmemMicros;
        filterUs += c->x().lastFilterMicros;
        rc = runSeeds(c, S.dev, p, b->sm, b->sd, dStats);
        if (rc) return rc;
        auto s2 = std::chrono::steady_clock::now();
        rc = runChains(c, S.dev, b->sd, b->ch);
        if (rc) return rc;
        auto s3 = std::chrono::steady_clock::now();
        seedUs += us(s0, s3);
        double k = 0;
        rc = runDp(c, S.dev, S.dQual, p, b->sd, b->ch, b->dp, &k, dpPhase, dStats + 10);
        if (rc) return rc;
        if (S.paired) {
            rc = runPairs(c, S.host[0].view.n, b->ch.chainOff, b->dp.dRecs, S.dev.offs, true, b->dp.scoreScale, p,
                          b->pr);
            if (rc) return rc;
        }
        auto s4 = std::chrono::steady_clock::now();
        dpUs += k;
        wall[0] = us(s0, s1); wall[1] = us(s1, s2); wall[2] = us(s2, s3); wall[3] = us(s3, s4);
    }
    unsigned long long hs[24];
    CK(hipMemcpyAsync(hs, dStats, sizeof(hs), hipMemcpyDeviceToHost, c->x().stream), "stats");
    CK(hipStreamSynchronize(c->x().stream), "stats");
    if (c->ingest && S.dev.total) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->x().ev6, c->x().ev7);
        ingestUs = ms * 1000.0;
    }
    double* st = S.stats;
    std::memset(st, 0, sizeof(S.stats));
    st[0] = (double)(S.host[0].view.n * (S.paired ? 2 : 1));
    st[2] = (double)hs[0];
    st[3] = (double)hs[1];
    st[5] = smemUs;
    st[6] = dpUs;
    st[8] = seedUs;
    st[10] = (double)hs[2];
    st[11] = (double)hs[3];
    st[12] = (double)hs[5];
    st[13] = (double)hs[6];
    st[14] = (double)hs[7];
    st[15] = filterUs;
    st[16] = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    for (int k = 0; k < 4; ++k) st[17 + k] = wall[k];
    for (int k = 0; k < 5; ++k) st[21 + k] = dpPhase[k];
    st[26] = (double)hs[8];
    st[27] = (double)hs[9];
    for (int k = 0; k < 4; ++k) st[32 + k] = (double)hs[10 + k];
    st[36] = c->ingest ? 2.0 * (double)S.dev.total : 0.0;
    st[37] = ingestUs;
    for (int k = 0; k < 4; ++k) st[38 + k] = (double)hs[14 + k];
    return MSGS_OK;
}

static int batchSubmit(msgs_ctx* c, int32_t slot, const msgs_params* p, int32_t emitSam) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (slot < 0 || slot >= 16 || !c->slots[slot].used) return fail(c, MSGS_E_STATE, "slot not uploaded");
    int rc = checkParams(c, p);
    if (rc) return rc;
    Slot& S = c->slots[slot];
    S.join();
    if (!S.res) {
        if (!c->spareResults.empty()) { S.res = c->spareResults.back(); c->spareResults.pop_back(); }
        else S.res = std::make_shared<BatchResult>();
    }
    S.params = *p;
    S.pending = true;
    S.rc = MSGS_OK;
    S.err.clear();
    const bool emit = emitSam != 0;
    const auto t0 = std::chrono::steady_clock::now();
    S.worker = std::thread([c, &S, emit, t0]() {
        (void)hipSetDevice(c->device);
        /* device stages on a free worker executor (several batches can be in them at once) */
        Exec* X = c->acquireWork();
        int rc;
        try {
            rc = batchGpu(c, S, t0);
        } catch (const std::bad_alloc&) {
            rc = fail(c, MSGS_E_NOMEM, "host memory");
        }
        if (rc) S.err = X->err;
        c->releaseWork(X);
        if (rc) { S.rc = rc; return; }
        /* host finishing */
        const auto h0 = std::chrono::steady_clock::now();
        int64_t mapped = 0;
        try {
            mapped = samText(c, &S.params, &S.host[0].view, S.paired ? &S.host[1].view : nullptr, *S.res, emit,
                             S.parts, S.lineOffs, &S.host[0].nameOff, S.paired ? &S.host[1].nameOff : nullptr,
                             std::max(1, c->hostThreads - 1),   /* one core stays with the device stages */
                             S.stats + 28);
        } catch (const std::bad_alloc&) {
            S.rc = MSGS_E_NOMEM;
            S.err = "host memory while finishing records";
        }
        int64_t cells = 0;
        for (const AlnRec& a : S.res->dp.recs) cells += a.cells;
        S.stats[4] = (double)cells;
        const auto t1 = std::chrono::steady_clock::now();
        S.stats[1] = (double)mapped;
        S.stats[7] = std::chrono::duration<double, std::micro>(t1 - t0).count();
        S.stats[9] = std::chrono::duration<double, std::micro>(t1 - h0).count();
    });
    return MSGS_OK;
}

/* wait for a submitted batch; its counters and records become the context's last run */
extern "C" int msgs_batch_wait(msgs_ctx* c, int32_t slot) {
    if (!c) return MSGS_E_ARG;
    if (slot < 0 || slot >= 16) return fail(c, MSGS_E_ARG, "bad slot");
    Slot& S = c->slots[slot];
    if (!S.pending) return fail(c, MSGS_E_STATE, "slot has no submitted run");
    S.join();
    if (S.res) { c->spareResults.push_back(S.res); S.res.reset(); }
    if (S.rc) return fail(c, S.rc, S.err);
    std::memcpy(c->stats, S.stats, sizeof(c->stats));
    std::swap(c->lastParts, S.parts);
    std::swap(c->lastLineOffs, S.lineOffs);
    return MSGS_OK;
}

extern "C" int msgs_batch_run(msgs_ctx* c, int32_t slot, const msgs_params* p, int32_t emitSam) {
    int rc = msgs_batch_submit(c, slot, p, emitSam);
    if (rc) return rc;
    return msgs_batch_wait(c, slot);
}

extern "C" int msgs_batch_stats(msgs_ctx* c, double* stats, int32_t n) {
    if (!c || !stats) return MSGS_E_ARG;
    for (int32_t i = 0; i < n && i < 48; ++i) stats[i] = c->stats[i];
    return MSGS_OK;
}

extern "C" int msgs_batch_sam(msgs_ctx* c, char* out, int64_t cap, int64_t* needed) {
    if (!c) return MSGS_E_ARG;
    const int64_t total = c->lastLineOffs.empty() ? 0 : c->lastLineOffs.back();
    if (needed) *needed = total;
    if (!out || total + 1 > cap) return fail(c, MSGS_E_CAP, "output capacity");
    copyParts(c->lastParts, out);
    out[total] = 0;
    return MSGS_OK;
}

/* ------------------------------------------------------------------ autotuner */

/* Time the SMEM search of resident batch `slot` for each wavefronts-per-CU setting in
 * `cand` (lanes per CU, multiples of 64) and keep the fastest in the context.  The
 * search's results do not depend on the setting (reads are independent; the work queue
 * only changes which lane runs which read). */
extern "C" int msgs_tune(msgs_ctx* c, int32_t slot, const msgs_params* p, const int32_t* cand, int32_t nCand,
                         int32_t* best, double* bestMicros) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (slot < 0 || slot >= 16 || !c->slots[slot].used) return fail(c, MSGS_E_STATE, "slot not uploaded");
    int rc = checkParams(c, p);
    if (rc) return rc;
    static const int32_t kDefault[] = {128, 256, 384, 512, 768, 1024};
    if (!cand || nCand <= 0) { cand = kDefault; nCand = 6; }
    Slot& S = c->slots[slot];
    for (int32_t i = 0; i < nCand; ++i)
        if (cand[i] < 64 || cand[i] > 4096 || cand[i] % 64) return fail(c, MSGS_E_ARG, "lanes per CU");
    const int keep = c->lanesPerCU, keepG = c->groupLanes, keepT = c->textPath, keepF = c->readFilter;
    const int keepS = c->splitLanes, keepP = c->pairChains;
    /* form 2 search front ends and kernels: (lanes per read, text path, read filter, lanes
     * per read of the split launch for the reads marked to start early, paired chains) */
    static const int kKernels[][5] = {{8, 1, 0, 0, 0}, {16, 1, 0, 0, 0}, {0, 1, 0, 0, 0}, {0, 0, 0, 0, 0},
                                      {8, 1, 1, 0, 0}, {0, 0, 1, 0, 0}, {0, 0, 0, 4, 0}, {0, 0, 0, 8, 0},
                                      {0, 0, 0, 16, 0}, {0, 0, 0, 0, 1}, {0, 1, 0, 0, 1}};
    /* the other forms: one lane per read, unpaired or paired */
    const int nK = c->form == 2 ? 11 : 2;
    int32_t bestL = keep;
    int bestG = keepG, bestT = keepT, bestF = keepF, bestS = keepS, bestP = keepP;
    double bestUs = 1e300;
    auto restore = [&]() {
        c->lanesPerCU = keep; c->groupLanes = keepG; c->textPath = keepT; c->readFilter = keepF; c->splitLanes = keepS;
        c->pairChains = keepP;
    };
    for (int ki = 0; ki < nK; ++ki) {
        if (c->form == 2) {
            c->groupLanes = kKernels[ki][0]; c->textPath = kKernels[ki][1]; c->readFilter = kKernels[ki][2];
            c->splitLanes = kKernels[ki][3]; c->pairChains = kKernels[ki][4];
        } else {
            c->pairChains = ki;
        }
        for (int32_t i = 0; i < nCand; ++i) {
            c->lanesPerCU = cand[i];
            double t = 1e300;
            for (int rep = 0; rep < 2; ++rep) {
                SmemStage sm;
                rc = runSmem(c, S.dev, p, sm, nullptr);
                if (rc) { restore(); return rc; }
                t = std::min(t, c->x().lastStageMicros + c->x().lastFilterMicros);
            }
            if (t < bestUs) {
                bestUs = t; bestL = cand[i]; bestG = c->groupLanes; bestT = c->textPath; bestF = c->readFilter;
                bestS = c->splitLanes; bestP = c->pairChains;
            }
        }
    }
    c->lanesPerCU = bestL;
    c->groupLanes = bestG;
    c->textPath = bestT;
    c->readFilter = bestF;
    c->splitLanes = bestS;
    c->pairChains = bestP;
    if (best) *best = bestL;
    if (bestMicros) *bestMicros = bestUs;
    return MSGS_OK;
}
