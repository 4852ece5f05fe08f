/*
 * msgs_noalloc.hip -- SMEM search over reads already in device memory, with no
 * allocation: msgs_smem_batch_noalloc (include/msgs_hip.h).  It has no counterpart
 * in the reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).
 *
 * The search's scratch (the work-queue counter, a padded copy of the reads -- the search
 * loads read bytes as aligned 8-byte words, up to 24 bytes past a read's end --, the
 * per-read loglik prefix sums and the per-lane candidate lists) lives in `scratch`, a fixed device array of
 * kNoAllocScratch bytes that every GPU the library runs on holds until the process
 * ends.  Its device address is looked up once per device at context creation
 * (msgs_ctx_create -> noallocInit) and kept per device; host code uses that address,
 * never the array's name.  Calls on one GPU share the array: calls that overlap on one
 * GPU race.  Each call clears what must start at zero (the queue counter) with a fill
 * kernel.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>

#include "msgs_capi_impl.h"
#include "msgs_kernels.h"

namespace {

constexpr size_t kNoAllocScratch = size_t(128) << 20;
constexpr int kMaxDevices = 64;
/* header: [0] queue counter (2 x u64), [16] longest read, [24] offs[n], [32] largest count */
constexpr size_t kHeader = 256;

void* g_scratch[kMaxDevices] = {};
std::mutex g_scratchMu;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

static_assert(sizeof(msgs_smem_rec) == sizeof(SmemRec), "msgs_smem_rec mirrors SmemRec");
static_assert(offsetof(msgs_smem_rec, fwdStart) == offsetof(SmemRec, fwd), "msgs_smem_rec mirrors SmemRec");
static_assert(offsetof(msgs_smem_rec, logP) == offsetof(SmemRec, logP), "msgs_smem_rec mirrors SmemRec");

/* the scratch array's device address on `device` (the current device), looked up once */
int noallocInit(int device) {
    if (device < 0 || device >= kMaxDevices) return MSGS_E_ARG;
    std::lock_guard<std::mutex> g(g_scratchMu);
    if (g_scratch[device]) return MSGS_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MSGS_E_HIP;
    g_scratch[device] = p;
    return MSGS_OK;
}

namespace {

/* zero bytes [a, b) */
__global__ void k_na_zero(unsigned char* p, int64_t a, int64_t b) {
    const int64_t i = a + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < b) p[i] = 0;
}

/* clears the queue counters and the header's reductions */
__global__ void k_na_clear(unsigned long long* hdr) {
    if (threadIdx.x < 6) hdr[threadIdx.x] = 0;
}

/* longest read and offs[n] into the header (hdr[2], hdr[3]) */
__global__ void k_na_lengths(const int64_t* offs, int64_t n, unsigned long long* hdr) {
    unsigned long long mx = 0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t L = offs[r + 1] - offs[r];
        mx = L > (int64_t)mx ? (unsigned long long)L : mx;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long y = __shfl_xor(mx, o);
        mx = y > mx ? y : mx;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(&hdr[2], mx);
    if (blockIdx.x == 0 && threadIdx.x == 0) hdr[3] = (unsigned long long)offs[n];
}

/* largest SMEM count into hdr[4] */
__global__ void k_na_maxcount(const int32_t* cnt, int64_t n, unsigned long long* hdr) {
    unsigned long long mx = 0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        mx = (unsigned long long)cnt[r] > mx ? (unsigned long long)cnt[r] : mx;
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long y = __shfl_xor(mx, o);
        mx = y > mx ? y : mx;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(&hdr[4], mx);
}

}  // namespace

extern "C" int msgs_smem_batch_noalloc(msgs_ctx* c, const uint8_t* dSeq, const int64_t* dOffs, int64_t n,
                                       const msgs_params* p, int32_t cap, int32_t* dCounts, msgs_smem_rec* dRecs,
                                       int32_t* maxCount) {
    if (!c) return MSGS_E_ARG;
    if (!c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (!p || n < 0 || cap < 1 || (n > 0 && (!dSeq || !dOffs || !dCounts || !dRecs)))
        return fail(c, MSGS_E_ARG, "bad arguments");
    if (p->minLen <= 0) return fail(c, MSGS_E_ARG, "minLen must be positive");
    if (c->device < 0 || c->device >= kMaxDevices || !g_scratch[c->device])
        return fail(c, MSGS_E_STATE, "allocation-free scratch not set up on this device");
    if (maxCount) *maxCount = 0;
    if (n == 0) return MSGS_OK;
    double T;
    if (!evalueThreshold((double)c->db.n, p->maxEvalue, &T))
        return fail(c, MSGS_E_UNSUPPORTED, "e-value predicate is not monotone near its threshold");
    hipStream_t st = c->own.stream;
    unsigned char* base = static_cast<unsigned char*>(g_scratch[c->device]);
    unsigned long long* hdr = reinterpret_cast<unsigned long long*>(base);
    /* read lengths first: the prefix sums and candidate lists are sized by them */
    k_na_clear<<<1, 64, 0, st>>>(hdr);
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
    k_na_lengths<<<g, 256, 0, st>>>(dOffs, n, hdr);
    unsigned long long lens[2] = {0, 0};
    if (hipMemcpyAsync(lens, hdr + 2, sizeof lens, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(c, MSGS_E_HIP, "read lengths");
    const int64_t maxL = (int64_t)lens[0], endOff = (int64_t)lens[1];
    if (endOff < 0) return fail(c, MSGS_E_ARG, "bad read offsets");
    const int gate1 = std::isfinite(p->maxEvalue) ? 1 : 0;
    /* the reads' bytes [0, offs[n]) followed by 64 zero bytes */
    const size_t seqBytes = ((size_t)endOff + 64 + 255) / 256 * 256;
    /* prefix sums: indexes offs[r] + r .. offs[r] + r + L of each read */
    const size_t prefBytes = gate1 ? (((size_t)(endOff + n + 1) * sizeof(double) + 255) / 256) * 256 : 0;
    if (kHeader + seqBytes + prefBytes >= kNoAllocScratch)
        return fail(c, MSGS_E_CAP, "reads exceed the allocation-free scratch (read copy and loglik prefix sums)");
    const int64_t candCap = maxL + 2;
    if (c->db.wide && maxL >= SMEM_WIDE_MAX_READ - 2)
        return fail(c, MSGS_E_UNSUPPORTED, "reads of 2^26 bases or more on a 64-bit-row index");
    const size_t laneBytes = (size_t)candCap * SMEM_CAND_BYTES;
    const int64_t lanesFit = (int64_t)((kNoAllocScratch - kHeader - seqBytes - prefBytes) / laneBytes) / 256 * 256;
    const int64_t lanesMax = (int64_t)std::max(1, c->nCU) * c->lanesPerCU;
    const int64_t lanes = std::min<int64_t>({std::max<int64_t>(256, (n + 255) / 256 * 256), lanesMax / 256 * 256, lanesFit});
    if (lanes < 256) return fail(c, MSGS_E_CAP, "reads exceed the allocation-free scratch (candidate lists)");
    unsigned char* seq = base + kHeader;
    if (endOff > 0 && hipMemcpyAsync(seq, dSeq, (size_t)endOff, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return fail(c, MSGS_E_HIP, "read copy");
    k_na_zero<<<1, 64, 0, st>>>(seq, endOff, endOff + 64);
    SmemArgs a;
    std::memset(&a, 0, sizeof(a));
    a.seqs = seq; a.offs = dOffs; a.nReads = n;
    a.minLen = p->minLen;
    a.gate1 = gate1;
    a.T1 = T; a.T2 = T;
    a.cap = cap;
    a.cnt = dCounts;
    a.out = reinterpret_cast<SmemRec*>(dRecs);
    a.queue = hdr;
    a.cand = base + kHeader + seqBytes + prefBytes;
    a.candCap = candCap;
    a.nLanes = lanes;
    if (gate1) {
        double* pref = reinterpret_cast<double*>(base + kHeader + seqBytes);
        smem_prefix_kernel<<<(unsigned)std::min<int64_t>((n + 255) / 256, 4096), 256, 0, st>>>(c->db.view(), a, pref);
        a.prefLL = pref;
    }
    if (c->db.wide) smem_kernel_t<true, int64_t><<<(unsigned)(lanes / 256), 256, 0, st>>>(c->db.view(), a);
    else smem_kernel_t<true, int32_t><<<(unsigned)(lanes / 256), 256, 0, st>>>(c->db.view(), a);
    k_na_maxcount<<<g, 256, 0, st>>>(dCounts, n, hdr);
    unsigned long long mx = 0;
    if (hipGetLastError() != hipSuccess) return fail(c, MSGS_E_HIP, "smem launch");
    if (hipMemcpyAsync(&mx, hdr + 4, sizeof mx, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(c, MSGS_E_HIP, "smem kernel");
    if (maxCount) *maxCount = (int32_t)mx;
    if ((int64_t)mx > cap) return fail(c, MSGS_E_CAP, "SMEM slots per read");
    return MSGS_OK;
}
