/*
 * msgs_capi.hip -- the extern "C" face of libmsgs_hip.so: context, resident
 * database, FMD primitives and the seeding batch.  See include/msgs_hip.h.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "msgs_capi_impl.h"

/* ------------------------------------------------------------------ state */

void DbState::release() {
    void* ptrs[] = {dBd, dSa, dIsa, dBwt, dOcc, dNPos, dChromBD, dTgtOff, dChromLen, dTgt, dSup};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    dBd = nullptr; dSa = nullptr; dIsa = nullptr; dBwt = nullptr; dOcc = nullptr; dNPos = nullptr; dSup = nullptr;
    wide = false;
    dChromBD = nullptr; dTgtOff = nullptr; dChromLen = nullptr; dTgt = nullptr;
    n = 0;
}

DevIndex DbState::view() const {
    DevIndex v;
    v.occ = dOcc; v.sa = dSa; v.isa = dIsa; v.bd = dBd; v.nPos = dNPos; v.nN = nN; v.n = n;
    v.wide = wide ? 1 : 0; v.supShift = supShift; v.sup = dSup;
    for (int i = 0; i < 17; ++i) v.C[i] = C[i];
    for (int i = 0; i < 16; ++i) v.ll[i] = ll[i];
    v.chromBD = dChromBD; v.nChrom = nChrom; v.tgt = dTgt; v.tgtOff = dTgtOff; v.chromLen = dChromLen;
    return v;
}

int fail(msgs_ctx* c, int code, const std::string& msg) {
    if (c) c->x().err = msg;
    return code;
}

/* ------------------------------------------------------------------ executors */

namespace {
thread_local Exec* tlsExec = nullptr;   /* set by a batch worker for its run */
}

Exec& msgs_ctx::x() { return tlsExec ? *tlsExec : own; }

bool Exec::create(unsigned streamFlags) {
    return hipStreamCreateWithFlags(&stream, streamFlags) == hipSuccess && hipEventCreate(&ev0) == hipSuccess &&
           hipEventCreate(&ev1) == hipSuccess && hipEventCreate(&ev2) == hipSuccess &&
           hipEventCreate(&ev3) == hipSuccess && hipEventCreate(&ev4) == hipSuccess &&
           hipEventCreate(&ev5) == hipSuccess && hipEventCreate(&ev6) == hipSuccess &&
           hipEventCreate(&ev7) == hipSuccess;
}

void Exec::destroy() {
    if (stream) (void)hipStreamSynchronize(stream);
    bufs.releaseAll();
    for (hipEvent_t* e : {&ev0, &ev1, &ev2, &ev3, &ev4, &ev5, &ev6, &ev7})
        if (*e) { (void)hipEventDestroy(*e); *e = nullptr; }
    if (stream) { (void)hipStreamDestroy(stream); stream = nullptr; }
}

/* a free worker executor, bound to the calling thread until releaseWork */
Exec* msgs_ctx::acquireWork() {
    std::unique_lock<std::mutex> lk(workMu);
    int k = -1;
    workCv.wait(lk, [&]() {
        for (int i = 0; i < std::min(workStreams, kWorkExecs); ++i)
            if (!workBusy[i]) { k = i; return true; }
        return false;
    });
    workBusy[k] = true;
    tlsExec = &work[k];
    return &work[k];
}

void msgs_ctx::releaseWork(Exec* e) {
    tlsExec = nullptr;
    {
        std::lock_guard<std::mutex> lk(workMu);
        workBusy[e - work] = false;
    }
    workCv.notify_one();
}

/* ------------------------------------------------------------------ e-value thresholds */

static uint64_t okey(double x) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}
static double fromKey(uint64_t k) {
    uint64_t u = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
    double x;
    std::memcpy(&x, &u, 8);
    return x;
}

/* Largest double T with (double)n * exp(T) <= maxE, using the host libm exp that
 * the reference calls (SMEM::evalue, src/SMEM.h:104-110).  NaN when no logP passes,
 * +inf when every logP passes.  Returns false if the predicate is not monotone in
 * a 1000-ulp neighbourhood of T. */
bool evalueThreshold(double n, double maxE, double* T) {
    auto ok = [&](double x) { return n * std::exp(x) <= maxE; };
    const double inf = std::numeric_limits<double>::infinity();
    if (std::isnan(maxE) || !ok(-inf)) { *T = std::numeric_limits<double>::quiet_NaN(); return true; }
    if (ok(inf)) { *T = inf; return true; }
    uint64_t lo = okey(-DBL_MAX), hi = okey(DBL_MAX);
    if (!ok(-DBL_MAX)) { *T = -inf; return true; }
    while (hi - lo > 1) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (ok(fromKey(mid))) lo = mid; else hi = mid;
    }
    for (int d = 1; d <= 1000; ++d) {
        if (!ok(fromKey(lo - d))) return false;
        if (ok(fromKey(lo + d))) return false;
    }
    *T = fromKey(lo);
    return true;
}

/* ------------------------------------------------------------------ small kernels */

__global__ void k_rank(DevIndex ix, const uint8_t* sym, const int64_t* pos, int64_t* out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t r[5];
    occ5(ix, pos[i], r);
    uint8_t s = sym[i];
    /* N from the sorted side array; the other IUPAC codes never occur in a BD text */
    const bool in = pos[i] >= 0 && pos[i] < ix.n;
    out[i] = s == 0 ? r[0] : s == 1 ? r[1] : s == 2 ? r[2] : s == 4 ? r[3] : s == 8 ? r[4]
           : !in ? -1 : s == 15 ? countN(ix, pos[i]) : 0;
}

__global__ void k_ext(DevIndex ix, int64_t* pqs, const uint8_t* bases, int32_t fwd, int64_t* ret, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t p = pqs[3 * i], q = pqs[3 * i + 1], s = pqs[3 * i + 2];
    ret[i] = fwd ? dFwdExt(ix, p, q, s, bases[i]) : dBackExt(ix, p, q, s, bases[i]);
    pqs[3 * i] = p; pqs[3 * i + 1] = q; pqs[3 * i + 2] = s;
}

/* the bi-interval after a pattern search: mode 0 = FMDIndex::count's backward search from the
 * last base (src/FMDIndex.cpp:163-176), 1 = locateAllFwd's forward search from the first base
 * (:529-551), 2 = locateAllRev's backward search (:553-575); modes 1 and 2 give s = 0 for an
 * empty pattern or an ambiguous start base, as the reference returns no locations then */
__global__ void k_search(DevIndex ix, const uint8_t* seqs, const int64_t* offs, int64_t n, int32_t mode,
                         int64_t* pqs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t* pat = seqs + offs[i];
    const int64_t L = offs[i + 1] - offs[i];
    int64_t p = 0, q = 0, s = 0;
    if (L > 0) {
        const uint8_t b = mode == 1 ? pat[0] : pat[L - 1];
        if (mode == 0 || !amb(b)) {
            p = ix.C[b]; q = ix.C[comp16(b)]; s = ix.C[b + 1] - ix.C[b];
            if (mode == 1) {
                for (int64_t k = 1; k < L && dFwdExt(ix, p, q, s, pat[k]) > 0; ++k) continue;
            } else {
                for (int64_t k = L - 1; k > 0 && dBackExt(ix, p, q, s, pat[k - 1]) > 0; --k) continue;
            }
        }
    }
    pqs[3 * i] = p; pqs[3 * i + 1] = q; pqs[3 * i + 2] = s;
}

/* SA values of the given rows (FMDIndex::accessSA, src/FMDIndex.cpp:520-527) */
__global__ void k_sa(DevIndex ix, const int64_t* rows, int64_t* out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = saAt(ix, rows[i]);
}

static inline unsigned gridFor(int64_t n, int t = 256) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

/* the search kernels compiled for the database's index kind: 32-bit rows for narrow indexes,
 * 64-bit rows for wide ones */
template <class Row>
static void launchGroupSearchT(int G, unsigned grid, const DevIndex& ix, const SmemArgs& a, hipStream_t st) {
    if (G == 4) smem_group_kernel<4, Row><<<grid, 256, 0, st>>>(ix, a);
    else if (G == 8) smem_group_kernel<8, Row><<<grid, 256, 0, st>>>(ix, a);
    else smem_group_kernel<16, Row><<<grid, 256, 0, st>>>(ix, a);
}
static void launchGroupSearch(bool wide, int G, unsigned grid, const DevIndex& ix, const SmemArgs& a, hipStream_t st) {
    if (wide) launchGroupSearchT<int64_t>(G, grid, ix, a, st);
    else launchGroupSearchT<int32_t>(G, grid, ix, a, st);
}
template <class Row>
static void launchLaneSearchT(bool fast, bool pair, unsigned grid, const DevIndex& ix, const SmemArgs& a,
                              hipStream_t st) {
    if (pair) {
        if (fast) smem_pair_kernel<true, Row><<<grid, 256, 0, st>>>(ix, a);
        else smem_pair_kernel<false, Row><<<grid, 256, 0, st>>>(ix, a);
    } else {
        if (fast) smem_kernel_t<true, Row><<<grid, 256, 0, st>>>(ix, a);
        else smem_kernel_t<false, Row><<<grid, 256, 0, st>>>(ix, a);
    }
}
static void launchLaneSearch(bool wide, bool fast, bool pair, unsigned grid, const DevIndex& ix, const SmemArgs& a,
                             hipStream_t st) {
    if (wide) launchLaneSearchT<int64_t>(fast, pair, grid, ix, a, st);
    else launchLaneSearchT<int32_t>(fast, pair, grid, ix, a, st);
}

/* ------------------------------------------------------------------ context */

extern "C" {

void msgs_params_default(msgs_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->minLen = 17; p->maxEvalue = 0.01; p->maxNSeed = 500;
    p->mode = 0; p->maxReport = 1;
    p->match = 1; p->mismatch = 3; p->gapO = 5; p->gapE = 2; p->clip = 4;
    p->minScoreRate = 0.65;
    p->insMean = 200; p->insSD = 20; p->insMin = 0; p->insMax = 400;
    p->maxNPair = 0;
}

int msgs_ctx_create(int device, msgs_ctx** out) {
    if (!out) return MSGS_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return MSGS_E_HIP;
    if (hipSetDevice(device) != hipSuccess) return MSGS_E_HIP;
    msgs_ctx* c = new msgs_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->nCU = prop.multiProcessorCount;
    c->hostThreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    /* the worker streams do not synchronize with the null stream */
    bool ok = c->own.create(hipStreamDefault);
    for (Exec& e : c->work) ok = ok && e.create(hipStreamNonBlocking);
    ok = ok && noallocInit(device) == MSGS_OK;
    if (!ok) {
        c->own.destroy();
        for (Exec& e : c->work) e.destroy();
        delete c;
        return MSGS_E_HIP;
    }
    *out = c;
    return MSGS_OK;
}

void msgs_ctx_destroy(msgs_ctx* c) {
    if (!c) return;
    for (auto& s : c->slots) s.join();
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->own.stream);
    c->db.release();
    c->lix.release();
    for (auto& s : c->slots) s.release();
    for (Exec& e : c->work) e.destroy();
    c->own.destroy();
    delete c;
}

const char* msgs_last_error(msgs_ctx* c) { return c ? c->own.err.c_str() : "no context"; }

/* ------------------------------------------------------------------ database */

static const uint8_t kBasic[16] = {0, 1, 2, 2, 4, 4, 4, 4, 8, 8, 8, 8, 8, 8, 8, 15};
static inline uint8_t compBase(uint8_t b) {
    return (uint8_t)(((b & 1) << 3) | ((b & 2) << 1) | ((b & 4) >> 1) | ((b & 8) >> 3));
}

int msgs_db_build(msgs_ctx* c, const uint8_t* seqs, const int64_t* lens, int32_t nChrom, const char* names) {
    if (!c || !seqs || !lens || nChrom <= 0) return fail(c, MSGS_E_ARG, "bad arguments");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MSGS_E_HIP, "hipSetDevice");
    c->db.release();
    c->lix.release();
    DbState& db = c->db;
    db.names.clear(); db.chromLen.clear(); db.chromBD.clear(); db.tgtOff.clear();
    db.nChrom = nChrom;
    int64_t total = 0, seqTotal = 0;
    for (int32_t t = 0; t < nChrom; ++t) {
        if (lens[t] < 0) return fail(c, MSGS_E_ARG, "negative chrom length");
        total += 2 * (lens[t] + 1);
        seqTotal += lens[t];
    }
    /* narrow indexes (32-bit rows) below 2^31 symbols, wide ones (64-bit rows, superblock
     * counts) up to 2^34 */
    if (total >= ((int64_t)1 << 34) - 2) return fail(c, MSGS_E_UNSUPPORTED, "BD text of 2^34 or more bases");
    /* names */
    const char* p = names ? names : "";
    for (int32_t t = 0; t < nChrom; ++t) {
        const char* e = p;
        while (*e && *e != '\n') ++e;
        std::string nm(p, e);
        if (nm.empty()) { char buf[32]; std::snprintf(buf, sizeof buf, "g%04d", t); nm = buf; }
        db.names.push_back(nm);
        p = *e ? e + 1 : e;
    }
    /* BD text (MetaGenome::getBDSeq, src/MetaGenome.cpp:104-113) and targets */
    std::vector<uint8_t> bd(total);
    db.tgtHost.assign(seqTotal + nChrom, 0);
    int64_t off = 0, bdOff = 0, tOff = 0;
    for (int32_t t = 0; t < nChrom; ++t) {
        const int64_t L = lens[t];
        db.chromLen.push_back(L);
        db.chromBD.push_back(bdOff);
        db.tgtOff.push_back(tOff);
        for (int64_t i = 0; i < L; ++i) {
            uint8_t b = seqs[off + i];
            if (b > 15) return fail(c, MSGS_E_ARG, "chrom bases must be nt16 codes");
            bd[bdOff + i] = kBasic[b];
            bd[bdOff + L + 1 + (L - 1 - i)] = compBase(kBasic[b]);
            db.tgtHost[tOff + i] = b;
        }
        bd[bdOff + L] = 0;
        bd[bdOff + 2 * L + 1] = 0;
        off += L;
        bdOff += 2 * (L + 1);
        tOff += L + 1;
    }
    db.chromBD.push_back(bdOff);
    db.n = total;
    /* counts (FMDIndex::buildCounts, src/FMDIndex.cpp:118-134) */
    std::fill(db.B, db.B + 256, 0);
    std::fill(db.C, db.C + 256, 0);
    for (int64_t i = 0; i < total; ++i) db.B[bd[i]]++;
    int64_t S = 0;
    for (int i = 0; i <= 16; ++i) { db.C[i] = S; S += db.B[i]; }
    for (int i = 17; i < 256; ++i) db.C[i] = 0;
    for (int b = 0; b < 16; ++b)
        if (db.B[b] != db.B[compBase((uint8_t)b)]) return fail(c, MSGS_E_DB, "input seq is not bi-directional");
    if (db.C[15] - (db.B[1] + db.B[2] + db.B[4] + db.B[8]) - db.B[0] > 0)
        return fail(c, MSGS_E_DB, "input seq does not allow IUPAC-extended bases");
    for (int b = 0; b < 16; ++b) db.ll[b] = std::log((double)db.B[kBasic[b]]) - std::log((double)total);
    for (int b = 0; b < 16; ++b)
        if (!(db.ll[b] <= 0)) return fail(c, MSGS_E_DB, "base log-frequency above zero");

    hipStream_t st = c->x().stream;
    /* padded: the SMEM kernel reads the text in aligned 8-byte words */
    if (hipMalloc(&db.dBd, total + 64) != hipSuccess) return fail(c, MSGS_E_NOMEM, "BD text");
    if (hipMemcpyAsync(db.dBd, bd.data(), total, hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(c, MSGS_E_HIP, "BD upload");
    std::string err;
    /* msgseqtk-build's block-by-block build (src/msgseqtk-build.cpp:403-415, 440-467): a BD
     * text of blockSize symbols or more is indexed in runs of whole chromosomes taken from the
     * last one, each run closed when one more chromosome would reach blockSize, and the run
     * indexes merged; the suffix order follows that build (msgs_index.hip) */
    db.blockEnds.clear();
    const int64_t blockSize = c->buildBlockMb * ((int64_t)1 << 20);
    if (blockSize > 0 && total >= blockSize) {
        std::vector<int64_t> bdLen(nChrom);
        for (int32_t t = 0; t < nChrom; ++t) bdLen[t] = 2 * (lens[t] + 1);
        auto span = [&](int64_t a, int64_t b) { return db.chromBD[b] - db.chromBD[a]; };   /* chroms [a, b) */
        std::vector<int64_t> starts;   /* first chrom of each block, last block first */
        int64_t blockEnd = nChrom;
        for (int64_t i = nChrom; i > 0; --i) {
            const int64_t blockStart = i - 1;
            if (blockStart == 0 || (span(blockStart, blockEnd) < blockSize && span(blockStart - 1, blockEnd) >= blockSize)) {
                starts.push_back(blockStart);
                blockEnd = blockStart;
            }
        }
        std::reverse(starts.begin(), starts.end());
        for (size_t b = 0; b < starts.size(); ++b)
            db.blockEnds.push_back(b + 1 < starts.size() ? db.chromBD[starts[b + 1]] : total);
        if (db.blockEnds.size() == 1) db.blockEnds.clear();
    }
    db.wide = c->wideIndex != 0 || total > (int64_t)INT32_MAX || !db.blockEnds.empty();
    /* wide_index = 1 puts superblock and suffix-sorting range boundaries into small texts */
    db.supShift = c->wideIndex ? 14 : 31;
    const int64_t rangeMax = c->wideIndex ? (int64_t)1 << 16 : (int64_t)1 << 29;
    int rc = buildIndexDevice(db, st, err, rangeMax);
    if (rc != MSGS_OK) return fail(c, rc, err);
    /* targets and coordinates */
    if (hipMalloc(&db.dTgt, db.tgtHost.size() + 64) != hipSuccess || hipMemset(db.dTgt, 0, db.tgtHost.size() + 64) != hipSuccess ||
        hipMalloc(&db.dTgtOff, nChrom * sizeof(int64_t)) != hipSuccess ||
        hipMalloc(&db.dChromLen, nChrom * sizeof(int64_t)) != hipSuccess ||
        hipMalloc(&db.dChromBD, (nChrom + 1) * sizeof(int64_t)) != hipSuccess)
        return fail(c, MSGS_E_NOMEM, "targets");
    if (hipMemcpyAsync(db.dTgt, db.tgtHost.data(), db.tgtHost.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(db.dTgtOff, db.tgtOff.data(), nChrom * sizeof(int64_t), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(db.dChromLen, db.chromLen.data(), nChrom * sizeof(int64_t), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(db.dChromBD, db.chromBD.data(), (nChrom + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(c, MSGS_E_HIP, "target upload");
    if (hipStreamSynchronize(st) != hipSuccess) return fail(c, MSGS_E_HIP, "sync");
    return MSGS_OK;
}

int64_t msgs_db_length(msgs_ctx* c) { return c ? c->db.n : 0; }

int msgs_db_counts(msgs_ctx* c, int64_t* B, int64_t* C) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (B) std::memcpy(B, c->db.B, sizeof(c->db.B));
    if (C) std::memcpy(C, c->db.C, sizeof(c->db.C));
    return MSGS_OK;
}

double msgs_db_loglik(msgs_ctx* c, int32_t b) {
    if (!c || !c->db.n || b < 0 || b > 15) return std::numeric_limits<double>::quiet_NaN();
    return c->db.ll[b];
}

int msgs_db_copy(msgs_ctx* c, uint8_t* bd, uint8_t* bwt, int64_t* sa) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    DbState& db = c->db;
    hipStream_t st = c->x().stream;
    if (bd && hipMemcpyAsync(bd, db.dBd, db.n, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(c, MSGS_E_HIP, "copy");
    if (bwt && hipMemcpyAsync(bwt, db.dBwt, db.n, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(c, MSGS_E_HIP, "copy");
    if (sa && db.wide) {
        if (hipMemcpyAsync(sa, db.dSa, db.n * 8, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(c, MSGS_E_HIP, "copy");
    } else if (sa) {
        std::vector<uint32_t> s(db.n);
        if (hipMemcpyAsync(s.data(), db.dSa, db.n * 4, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(c, MSGS_E_HIP, "copy");
        if (hipStreamSynchronize(st) != hipSuccess) return fail(c, MSGS_E_HIP, "sync");
        for (int64_t i = 0; i < db.n; ++i) sa[i] = s[i];
    }
    if (hipStreamSynchronize(st) != hipSuccess) return fail(c, MSGS_E_HIP, "sync");
    return MSGS_OK;
}

/* ------------------------------------------------------------------ FMD primitives */

int msgs_rank_batch(msgs_ctx* c, const uint8_t* sym, const int64_t* pos, int64_t* out, int64_t n) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (n <= 0) return MSGS_OK;
    hipError_t e = hipSuccess;
    uint8_t* dS = c->x().bufs.get<uint8_t>("rk.s", n, e);
    int64_t* dP = c->x().bufs.get<int64_t>("rk.p", n, e);
    int64_t* dO = c->x().bufs.get<int64_t>("rk.o", n, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "rank buffers");
    hipStream_t st = c->x().stream;
    (void)hipMemcpyAsync(dS, sym, n, hipMemcpyHostToDevice, st);
    (void)hipMemcpyAsync(dP, pos, n * 8, hipMemcpyHostToDevice, st);
    k_rank<<<gridFor(n), 256, 0, st>>>(c->db.view(), dS, dP, dO, n);
    (void)hipMemcpyAsync(out, dO, n * 8, hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) return fail(c, MSGS_E_HIP, "rank kernel");
    return MSGS_OK;
}

int msgs_ext_batch(msgs_ctx* c, int64_t* pqs, const uint8_t* bases, int32_t fwd, int64_t* ret, int64_t n) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (n <= 0) return MSGS_OK;
    hipError_t e = hipSuccess;
    int64_t* dQ = c->x().bufs.get<int64_t>("ex.q", 3 * n, e);
    uint8_t* dB = c->x().bufs.get<uint8_t>("ex.b", n, e);
    int64_t* dR = c->x().bufs.get<int64_t>("ex.r", n, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "ext buffers");
    hipStream_t st = c->x().stream;
    (void)hipMemcpyAsync(dQ, pqs, 3 * n * 8, hipMemcpyHostToDevice, st);
    (void)hipMemcpyAsync(dB, bases, n, hipMemcpyHostToDevice, st);
    k_ext<<<gridFor(n), 256, 0, st>>>(c->db.view(), dQ, dB, fwd, dR, n);
    (void)hipMemcpyAsync(pqs, dQ, 3 * n * 8, hipMemcpyDeviceToHost, st);
    if (ret) (void)hipMemcpyAsync(ret, dR, n * 8, hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) return fail(c, MSGS_E_HIP, "ext kernel");
    return MSGS_OK;
}

int msgs_search_batch(msgs_ctx* c, const uint8_t* seqs, const int64_t* offs, int64_t n, int32_t mode, int64_t* pqs) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (mode < 0 || mode > 2 || n < 0 || (n > 0 && (!seqs || !offs || !pqs))) return fail(c, MSGS_E_ARG, "bad arguments");
    if (n == 0) return MSGS_OK;
    const int64_t total = offs[n] - offs[0];
    for (int64_t i = 0; i < n; ++i)
        if (offs[i + 1] < offs[i]) return fail(c, MSGS_E_ARG, "pattern offsets must not decrease");
    for (int64_t k = 0; k < total; ++k)
        if (seqs[offs[0] + k] > 15) return fail(c, MSGS_E_ARG, "pattern bases must be nt16 codes");
    hipError_t e = hipSuccess;
    uint8_t* dS = c->x().bufs.get<uint8_t>("se.s", total + 1, e);
    int64_t* dO = c->x().bufs.get<int64_t>("se.o", n + 1, e);
    int64_t* dQ = c->x().bufs.get<int64_t>("se.q", 3 * n, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "search buffers");
    std::vector<int64_t> o(n + 1);
    for (int64_t i = 0; i <= n; ++i) o[i] = offs[i] - offs[0];
    hipStream_t st = c->x().stream;
    if (total) (void)hipMemcpyAsync(dS, seqs + offs[0], total, hipMemcpyHostToDevice, st);
    (void)hipMemcpyAsync(dO, o.data(), (n + 1) * 8, hipMemcpyHostToDevice, st);
    k_search<<<gridFor(n), 256, 0, st>>>(c->db.view(), dS, dO, n, mode, dQ);
    (void)hipMemcpyAsync(pqs, dQ, 3 * n * 8, hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) return fail(c, MSGS_E_HIP, "search kernel");
    return MSGS_OK;
}

int msgs_sa_batch(msgs_ctx* c, const int64_t* rows, int64_t n, int64_t* out) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (n < 0 || (n > 0 && (!rows || !out))) return fail(c, MSGS_E_ARG, "bad arguments");
    for (int64_t i = 0; i < n; ++i)
        if (rows[i] < 0 || rows[i] >= c->db.n) return fail(c, MSGS_E_ARG, "suffix-array row out of range");
    if (n == 0) return MSGS_OK;
    hipError_t e = hipSuccess;
    int64_t* dR = c->x().bufs.get<int64_t>("sa.r", n, e);
    int64_t* dO = c->x().bufs.get<int64_t>("sa.o", n, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "SA buffers");
    hipStream_t st = c->x().stream;
    (void)hipMemcpyAsync(dR, rows, n * 8, hipMemcpyHostToDevice, st);
    k_sa<<<gridFor(n), 256, 0, st>>>(c->db.view(), dR, dO, n);
    (void)hipMemcpyAsync(out, dO, n * 8, hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) return fail(c, MSGS_E_HIP, "SA kernel");
    return MSGS_OK;
}

int msgs_fmd_write(msgs_ctx* c, const char* path, int32_t header) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (!path) return fail(c, MSGS_E_ARG, "no path");
    try {
        std::vector<uint8_t> bwt(c->db.n);
        std::vector<int64_t> sa(c->db.n);
        int rc = msgs_db_copy(c, nullptr, bwt.data(), sa.data());
        if (rc) return rc;
        rc = fmdWrite(path, header != 0, bwt.data(), sa.data(), c->db.n);
        return rc ? fail(c, rc, std::string("cannot write ") + path) : MSGS_OK;
    } catch (const std::bad_alloc&) {
        return fail(c, MSGS_E_NOMEM, "host memory");
    }
}

int msgs_db_chrom_seq(msgs_ctx* c, int32_t tid, uint8_t* out, int64_t cap, int64_t* len) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (tid < 0 || tid >= c->db.nChrom) return fail(c, MSGS_E_ARG, "bad chrom id");
    const int64_t L = c->db.chromLen[tid];
    if (len) *len = L;
    if (!out) return MSGS_OK;
    if (cap < L) return fail(c, MSGS_E_CAP, "chrom sequence capacity");
    std::memcpy(out, c->db.tgtHost.data() + c->db.tgtOff[tid], L);
    return MSGS_OK;
}

}  // extern "C"

/* ------------------------------------------------------------------ seeding stage */

/* Upload a read batch into device buffers of the given prefix; returns max length. */
int uploadReads(msgs_ctx* c, const msgs_reads* r, const char* pfx, DevReads& out) {
    if (!r || r->n < 0 || (r->n > 0 && (!r->seq || !r->offs))) return fail(c, MSGS_E_ARG, "bad reads");
    const int64_t n = r->n;
    const int64_t total = n ? r->offs[n] - r->offs[0] : 0;
    int64_t maxL = 0;
    for (int64_t i = 0; i < n; ++i) {
        int64_t L = r->offs[i + 1] - r->offs[i];
        if (L < 0) return fail(c, MSGS_E_ARG, "read offsets must not decrease");
        maxL = std::max(maxL, L);
    }
    if (maxL > 100000) return fail(c, MSGS_E_UNSUPPORTED, "reads longer than 100000 bases");
    for (int64_t i = 0; i < total; ++i)
        if (r->seq[r->offs[0] + i] > 15) return fail(c, MSGS_E_ARG, "read bases must be nt16 codes");
    std::string p(pfx);
    hipError_t e = hipSuccess;
    out.n = n;
    out.maxL = maxL;
    out.total = total;
    /* padded: the SMEM kernel reads bases in aligned 8-byte words */
    out.seq = c->x().bufs.get<uint8_t>(p + ".seq", total + 64, e);
    out.offs = c->x().bufs.get<int64_t>(p + ".offs", n + 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "read buffers");
    std::vector<int64_t> offs(n + 1);
    for (int64_t i = 0; i <= n; ++i) offs[i] = r->offs[i] - r->offs[0];
    hipStream_t st = c->x().stream;
    if (total) (void)hipMemcpyAsync(out.seq, r->seq + r->offs[0], total, hipMemcpyHostToDevice, st);
    (void)hipMemcpyAsync(out.offs, offs.data(), (n + 1) * 8, hipMemcpyHostToDevice, st);
    out.offsHost = offs;
    return MSGS_OK;
}

/* Run the SMEM kernel over resident reads; grows the per-read slot count until no
 * read overflows.  Leaves per-read counts and records in c->x().bufs ("sm.cnt", "sm.out"). */
int runSmem(msgs_ctx* c, const DevReads& r, const msgs_params* p, SmemStage& out, unsigned long long* dStats) {
    const int64_t n = r.n;
    if (p->minLen <= 0) return fail(c, MSGS_E_ARG, "minLen must be positive");
    double T;
    if (!evalueThreshold((double)c->db.n, p->maxEvalue, &T))
        return fail(c, MSGS_E_UNSUPPORTED, "e-value predicate is not monotone near its threshold");
    SmemArgs a;
    std::memset(&a, 0, sizeof(a));
    a.seqs = r.seq; a.offs = r.offs; a.nReads = n;
    a.minLen = p->minLen;
    a.gate1 = std::isfinite(p->maxEvalue) ? 1 : 0;
    a.T1 = T; a.T2 = T;
    /* form 2: skip reads in which no window that an emitted SMEM must contain occurs, and
     * take the SMEM list of exactly matching reads from the learned index */
    a.need = nullptr;
    c->x().lastFilterMicros = 0;
    if (c->form == 2 && c->lix.trained && p->minLen >= 16 && n > 0) {
        uint8_t* need = nullptr;
        SmemRec* exact = nullptr;
        /* timed with the second event pair, read after the search has finished (no host
         * wait between the two) */
        (void)hipEventRecord(c->x().ev2, c->x().stream);
        int rc = runLixFilter(c, r, (int32_t)p->minLen, T, a.gate1, &need, &exact, dStats);
        if (rc) return rc;
        a.need = need;
        a.exact = exact;
        int32_t* order = nullptr;
        rc = runLixOrder(c, n, need, &order, &a.qSplit);
        if (rc) return rc;
        a.order = order;
        (void)hipEventRecord(c->x().ev3, c->x().stream);
    }
    const int G = c->form == 2 ? c->groupLanes : 0;
    /* split: the reads marked to start early (sorted first by the order) go to a launch of
     * the lane-group kernel, the others to the one-lane kernel after it */
    const int split = (G == 0 && a.order && a.qSplit) ? c->splitLanes : 0;
    /* paired chains: the one-lane kernel with two lanes per read */
    const bool pair = G == 0 && !split && c->pairChains;
    const int64_t lanesMax = (int64_t)std::max(1, c->nCU) * c->lanesPerCU;
    const int64_t want = n * std::max(G, 1) * (pair ? 2 : 1);
    const int64_t lanes = std::max<int64_t>(256, (std::min<int64_t>(((want + 255) / 256) * 256, lanesMax) / 256) * 256);
    const int64_t lanesSplit = std::max<int64_t>(256, (lanesMax / 256) * 256);
    const int64_t candCap = r.maxL + 2;
    hipError_t e = hipSuccess;
    const int64_t nCand = candCap * std::max(lanes, split ? lanesSplit / split : 0);
    const bool wide = c->db.wide;
    if (wide && r.maxL >= SMEM_WIDE_MAX_READ - 2)
        return fail(c, MSGS_E_UNSUPPORTED, "reads of 2^26 bases or more on a 64-bit-row index");
    a.cand = c->x().bufs.get<uint8_t>("sm.cand", nCand * SMEM_CAND_BYTES, e);
    a.candCap = candCap;
    a.nLanes = lanes;
    a.queue = c->x().bufs.get<unsigned long long>("sm.q", 2, e);
    if (!split) { a.qSplit = nullptr; a.qPart = 0; }
    a.cnt = c->x().bufs.get<int32_t>("sm.cnt", n + 1, e);
    a.stats = dStats;
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "smem scratch");
    int cap = c->smemCap;
    hipStream_t st = c->x().stream;
    for (int attempt = 0; attempt < 3; ++attempt) {
        a.cap = cap;
        a.out = c->x().bufs.get<SmemRec>("sm.out", (size_t)cap * std::max<int64_t>(n, 1), e);
        if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "smem output");
        (void)hipMemsetAsync(a.queue, 0, 16, st);
        (void)hipEventRecord(c->x().ev0, st);
        if (n > 0 && a.gate1) {
            double* pref = c->x().bufs.get<double>("sm.pref", r.total + n + 1, e);
            if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "loglik prefix sums");
            smem_prefix_kernel<<<(unsigned)std::min<int64_t>((n + 255) / 256, 4096), 256, 0, st>>>(c->db.view(), a, pref);
            a.prefLL = pref;
        }
        /* ev4 .. ev5: the search kernel(s) alone */
        (void)hipEventRecord(c->x().ev4, st);
        if (n > 0 && split) {
            SmemArgs a1 = a, a2 = a;
            a1.qPart = 1;
            a2.qPart = 2;
            a2.queue = a.queue + 1;
            const unsigned g1 = (unsigned)(lanesSplit / 256), g2 = (unsigned)(lanes / 256);
            launchGroupSearch(wide, split, g1, c->db.view(), a1, st);
            launchLaneSearch(wide, c->textPath != 0, false, g2, c->db.view(), a2, st);
        } else if (n > 0 && pair) {
            /* each lane of a pair writes its own list; the merge puts them in order */
            SmemArgs k = a;
            k.out = c->x().bufs.get<SmemRec>("sm.outA", (size_t)cap * n, e);
            k.outB = c->x().bufs.get<SmemRec>("sm.outB", (size_t)cap * n, e);
            k.tagA = c->x().bufs.get<int32_t>("sm.tagA", (size_t)cap * n, e);
            k.tagB = c->x().bufs.get<int32_t>("sm.tagB", (size_t)cap * n, e);
            k.cnt = c->x().bufs.get<int32_t>("sm.cntA", n + 1, e);
            k.cntB = c->x().bufs.get<int32_t>("sm.cntB", n + 1, e);
            if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "smem pair lists");
            (void)hipMemsetAsync(k.cntB, 0, (n + 1) * sizeof(int32_t), st);
            const unsigned grid = (unsigned)(lanes / 256);
            const bool fast = !(c->form == 0 || (c->form == 2 && !c->textPath));
            launchLaneSearch(wide, fast, true, grid, c->db.view(), k, st);
            (void)hipEventRecord(c->x().ev5, st);
            smem_merge_kernel<<<(unsigned)std::min<int64_t>((n + 255) / 256, 65535), 256, 0, st>>>(k, a.out, a.cnt);
        } else if (n > 0) {
            const unsigned grid = (unsigned)(lanes / 256);
            if (G == 8 || G == 16) launchGroupSearch(wide, G, grid, c->db.view(), a, st);
            else launchLaneSearch(wide, !(c->form == 0 || (c->form == 2 && !c->textPath)), false, grid, c->db.view(), a, st);
        }
        if (!(n > 0 && pair)) (void)hipEventRecord(c->x().ev5, st);
        (void)hipEventRecord(c->x().ev1, st);
        if (hipGetLastError() != hipSuccess) return fail(c, MSGS_E_HIP, "smem launch");
        /* largest count */
        std::vector<int32_t> cnt(n);
        if (n) (void)hipMemcpyAsync(cnt.data(), a.cnt, n * 4, hipMemcpyDeviceToHost, st);
        if (hipStreamSynchronize(st) != hipSuccess) return fail(c, MSGS_E_HIP, "smem kernel");
        int mx = 0;
        for (int64_t i = 0; i < n; ++i) mx = std::max(mx, cnt[i]);
        if (mx <= cap) {
            float ms = 0, ks = 0;
            (void)hipEventElapsedTime(&ms, c->x().ev0, c->x().ev1);
            (void)hipEventElapsedTime(&ks, c->x().ev4, c->x().ev5);
            c->x().lastStageMicros = ms * 1000.0;
            c->x().lastSmemMicros = ks * 1000.0;
            if (a.need) {
                float fm = 0;
                (void)hipEventElapsedTime(&fm, c->x().ev2, c->x().ev3);
                c->x().lastFilterMicros = fm * 1000.0;
            }
            out.cap = cap;
            out.cnt = a.cnt;
            out.recs = a.out;
            out.cntHost.swap(cnt);
            return MSGS_OK;
        }
        cap = mx;
    }
    return fail(c, MSGS_E_HIP, "SMEM slot growth did not settle");
}

/* SMEM_LIST::findMEMS of every read (mem_kernel) and the loglik of each read's list; the
 * records stay on the device at *recs ([n][cap]), counts and logliks come to the host */
static int runMems(msgs_ctx* c, const DevReads& r, const msgs_params* p, int& capOut, SmemRec** recs,
                   std::vector<int32_t>& cnt, std::vector<double>& ll) {
    const int64_t n = r.n;
    double T;
    if (!evalueThreshold((double)c->db.n, p->maxEvalue, &T))
        return fail(c, MSGS_E_UNSUPPORTED, "e-value predicate is not monotone near its threshold");
    SmemArgs a;
    std::memset(&a, 0, sizeof(a));
    a.seqs = r.seq; a.offs = r.offs; a.nReads = n;
    a.minLen = p->minLen;
    a.T1 = T; a.T2 = T;
    hipError_t e = hipSuccess;
    a.cnt = c->x().bufs.get<int32_t>("mm.cnt", n + 1, e);
    double* dLL = c->x().bufs.get<double>("mm.ll", n + 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "mem scratch");
    hipStream_t st = c->x().stream;
    cnt.assign(n, 0);
    ll.assign(n, 0.0);
    int cap = c->smemCap;
    for (int attempt = 0; attempt < 3; ++attempt) {
        a.cap = cap;
        a.out = c->x().bufs.get<SmemRec>("mm.out", (size_t)cap * std::max<int64_t>(n, 1), e);
        if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "mem output");
        if (n > 0) {
            const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 65535);
            if (c->db.wide) mem_kernel<int64_t><<<grid, 256, 0, st>>>(c->db.view(), a, dLL);
            else mem_kernel<int32_t><<<grid, 256, 0, st>>>(c->db.view(), a, dLL);
            if (hipGetLastError() != hipSuccess) return fail(c, MSGS_E_HIP, "mem launch");
            (void)hipMemcpyAsync(cnt.data(), a.cnt, n * 4, hipMemcpyDeviceToHost, st);
            (void)hipMemcpyAsync(ll.data(), dLL, n * 8, hipMemcpyDeviceToHost, st);
        }
        if (hipStreamSynchronize(st) != hipSuccess) return fail(c, MSGS_E_HIP, "mem kernel");
        int mx = 0;
        for (int64_t i = 0; i < n; ++i) mx = std::max(mx, cnt[i]);
        if (mx <= cap) {
            capOut = cap;
            *recs = a.out;
            return MSGS_OK;
        }
        cap = mx;
    }
    return fail(c, MSGS_E_HIP, "MEM slot growth did not settle");
}

extern "C" int msgs_mems_batch(msgs_ctx* c, const msgs_reads* reads, const msgs_params* p, int64_t* counts,
                               int64_t* rows5, double* logP, double* readLoglik, int64_t cap, int64_t* needed) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (!reads || !p) return fail(c, MSGS_E_ARG, "bad arguments");
    DevReads r;
    int rc = uploadReads(c, reads, "rd", r);
    if (rc) return rc;
    int slots = 0;
    SmemRec* dRecs = nullptr;
    std::vector<int32_t> cnt;
    std::vector<double> ll;
    rc = runMems(c, r, p, slots, &dRecs, cnt, ll);
    if (rc) return rc;
    const int64_t n = r.n;
    std::vector<SmemRec> recs((size_t)slots * n);
    if (n) (void)hipMemcpyAsync(recs.data(), dRecs, recs.size() * sizeof(SmemRec), hipMemcpyDeviceToHost, c->x().stream);
    if (hipStreamSynchronize(c->x().stream) != hipSuccess) return fail(c, MSGS_E_HIP, "copy");
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i) total += cnt[i];
    if (needed) *needed = total;
    if (readLoglik)
        for (int64_t i = 0; i < n; ++i) readLoglik[i] = ll[i];
    if (total > cap) return fail(c, MSGS_E_CAP, "row capacity");
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (counts) counts[i] = cnt[i];
        for (int j = 0; j < cnt[i]; ++j, ++k) {
            const SmemRec& m = recs[(size_t)i * slots + j];
            if (rows5) {
                int64_t* o = rows5 + 5 * k;
                o[0] = m.from; o[1] = m.to; o[2] = m.fwd; o[3] = m.rev; o[4] = m.size;
            }
            if (logP) logP[k] = m.logP;
        }
    }
    return MSGS_OK;
}

/* the loglik of each read's MEM list (of each pair's two lists, the first read's plus the
 * mate's) in one database */
static int memLoglik(msgs_ctx* c, const msgs_reads* reads, const msgs_reads* mates, const msgs_params* p,
                     std::vector<double>& out) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    DevReads r;
    int rc = uploadReads(c, reads, "rd", r);
    if (rc) return rc;
    int slots = 0;
    SmemRec* dRecs = nullptr;
    std::vector<int32_t> cnt;
    rc = runMems(c, r, p, slots, &dRecs, cnt, out);
    if (rc || !mates) return rc;
    DevReads r2;
    rc = uploadReads(c, mates, "rd2", r2);
    if (rc) return rc;
    std::vector<double> ll2;
    rc = runMems(c, r2, p, slots, &dRecs, cnt, ll2);
    if (rc) return rc;
    for (size_t i = 0; i < out.size(); ++i) out[i] = out[i] + ll2[i];   /* SMEM_LIST::loglik(MEM_LIST_PE) */
    return MSGS_OK;
}

extern "C" int msgs_clean_batch(msgs_ctx* ref, msgs_ctx* bg, const msgs_reads* reads, const msgs_reads* mates,
                                const msgs_params* p, double minLod, double* refLoglik, double* bgLoglik,
                                uint8_t* keep) {
    if (!ref || !bg) return MSGS_E_ARG;
    if (!reads || !p || (mates && mates->n != reads->n)) return fail(ref, MSGS_E_ARG, "bad arguments");
    try {
        std::vector<double> rl, bl;
        int rc = memLoglik(ref, reads, mates, p, rl);
        if (rc) return rc;
        rc = memLoglik(bg, reads, mates, p, bl);
        if (rc) return rc;
        for (int64_t i = 0; i < reads->n; ++i) {
            const double lod = -rl[i] + bl[i];   /* msgseqtk-clean main_SE / main_PE */
            if (refLoglik) refLoglik[i] = rl[i];
            if (bgLoglik) bgLoglik[i] = bl[i];
            if (keep) keep[i] = lod > minLod ? 1 : 0;
        }
        return MSGS_OK;
    } catch (const std::bad_alloc&) {
        return fail(ref, MSGS_E_NOMEM, "host memory");
    }
}

extern "C" int msgs_smem_batch(msgs_ctx* c, const msgs_reads* reads, const msgs_params* p, int64_t* counts,
                               int64_t* rows5, double* logP, int64_t cap, int64_t* needed) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (!reads || !p) return fail(c, MSGS_E_ARG, "bad arguments");
    DevReads r;
    int rc = uploadReads(c, reads, "rd", r);
    if (rc) return rc;
    SmemStage sm;
    rc = runSmem(c, r, p, sm, nullptr);
    if (rc) return rc;
    const int64_t n = r.n;
    std::vector<SmemRec> recs((size_t)sm.cap * n);
    if (n) (void)hipMemcpyAsync(recs.data(), sm.recs, recs.size() * sizeof(SmemRec), hipMemcpyDeviceToHost, c->x().stream);
    if (hipStreamSynchronize(c->x().stream) != hipSuccess) return fail(c, MSGS_E_HIP, "copy");
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i) total += sm.cntHost[i];
    if (needed) *needed = total;
    if (total > cap) return fail(c, MSGS_E_CAP, "row capacity");
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (counts) counts[i] = sm.cntHost[i];
        for (int j = 0; j < sm.cntHost[i]; ++j, ++k) {
            const SmemRec& m = recs[(size_t)i * sm.cap + j];
            int64_t* o = rows5 + 5 * k;
            o[0] = m.from; o[1] = m.to; o[2] = m.fwd; o[3] = m.rev; o[4] = m.size;
            if (logP) logP[k] = m.logP;
        }
    }
    return MSGS_OK;
}

/* ------------------------------------------------------------------ utilities */

#include "msgs_sort.h"

/* The permutation std::sort gives for keys compared with '<' (host run of the
 * restated libstdc++ sort, exposed so the restatement can be checked). */
extern "C" void msgs_util_sort_by_key(const double* keys, int32_t* perm, int64_t n) {
    struct Less {
        const double* k;
        bool operator()(int32_t a, int32_t b) const { return k[a] < k[b]; }
    };
    msgs_sort::sort(perm, perm + n, Less{keys});
}

/* ------------------------------------------------------------------ reference database files */

#include <fstream>

/* MetaGenome + its sequences from the reference's database files: DB.mtg
 * ("MSGseqTK" + 3 int32 version numbers, then size_t #genomes and per genome the
 * strings id and name, size_t #chroms and per chrom the string name and int64 length;
 * strings are size_t length + bytes: src/util/ProgEnv.cpp:24-64,
 * src/MetaGenome.cpp:42-73, src/Genome.cpp:66-97) and DB.mgs (per chrom len+1 nt16
 * bytes incl. the 0 terminator: src/Genome.cpp:29-44).  The FMD index is rebuilt on
 * the device from the sequences (the .fmd file is not needed). */
extern "C" int msgs_db_load(msgs_ctx* c, const char* prefix) {
    if (!c || !prefix) return MSGS_E_ARG;
    std::string pre(prefix);
    std::ifstream mtg(pre + ".mtg", std::ios::binary), mgs(pre + ".mgs", std::ios::binary);
    if (!mtg) return fail(c, MSGS_E_ARG, "cannot open " + pre + ".mtg");
    if (!mgs) return fail(c, MSGS_E_ARG, "cannot open " + pre + ".mgs");
    char magic[8];
    int32_t ver[3];
    mtg.read(magic, 8);
    mtg.read((char*)ver, sizeof(ver));
    if (!mtg || std::string(magic, 8) != "MSGseqTK") return fail(c, MSGS_E_DB, "not a MSGseqTK database file");
    auto readStr = [&](std::string& s) -> bool {
        uint64_t len = 0;
        mtg.read((char*)&len, 8);
        if (!mtg || len > (1u << 30)) return false;
        s.resize(len);
        mtg.read(&s[0], len);
        return (bool)mtg;
    };
    uint64_t NG = 0;
    mtg.read((char*)&NG, 8);
    if (!mtg || NG > (1u << 24)) return fail(c, MSGS_E_DB, "bad genome count");
    std::vector<std::string> names;
    std::vector<int64_t> lens;
    for (uint64_t g = 0; g < NG; ++g) {
        std::string id, nm;
        uint64_t NC = 0;
        if (!readStr(id) || !readStr(nm)) return fail(c, MSGS_E_DB, "truncated genome record");
        mtg.read((char*)&NC, 8);
        if (!mtg || NC > (1u << 24)) return fail(c, MSGS_E_DB, "bad chrom count");
        for (uint64_t k = 0; k < NC; ++k) {
            std::string cn;
            int64_t len = 0;
            if (!readStr(cn)) return fail(c, MSGS_E_DB, "truncated chrom record");
            mtg.read((char*)&len, 8);
            if (!mtg || len < 0) return fail(c, MSGS_E_DB, "bad chrom length");
            names.push_back(cn);
            lens.push_back(len);
        }
    }
    int64_t total = 0;
    for (int64_t L : lens) total += L;
    std::vector<uint8_t> seqs(total);
    int64_t off = 0;
    for (size_t t = 0; t < lens.size(); ++t) {
        mgs.read((char*)seqs.data() + off, lens[t]);
        char term = 1;
        mgs.read(&term, 1);
        if (!mgs || term != 0) return fail(c, MSGS_E_DB, "truncated or malformed sequence file");
        off += lens[t];
    }
    std::string joined;
    for (size_t t = 0; t < names.size(); ++t) { joined += names[t]; joined += '\n'; }
    int rc = msgs_db_build(c, seqs.data(), lens.data(), (int32_t)lens.size(), joined.c_str());
    if (rc) return rc;
    /* DB.fmd (FMDIndex::load, src/FMDIndex.cpp:153-161), when present: its counts and BWT must
     * be those of the index just built from the sequences */
    const std::string fmd = pre + ".fmd";
    if (std::ifstream(fmd, std::ios::binary).good()) {
        int64_t n = 0, B[256], C[256];
        rc = msgs_fmd_read(fmd.c_str(), 1, &n, nullptr, B, C, nullptr, nullptr, nullptr, nullptr);
        if (rc) return fail(c, MSGS_E_DB, "unreadable FMD-index file " + fmd);
        if (n != c->db.n || std::memcmp(B, c->db.B, sizeof B) != 0 || std::memcmp(C, c->db.C, sizeof C) != 0)
            return fail(c, MSGS_E_DB, "FMD-index file " + fmd + " does not match the database sequences");
        std::vector<uint8_t> fb(n), db(n);
        rc = msgs_fmd_read(fmd.c_str(), 1, &n, fb.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        if (rc == MSGS_OK) rc = msgs_db_copy(c, nullptr, db.data(), nullptr);
        if (rc || fb != db) return fail(c, MSGS_E_DB, "FMD-index file " + fmd + " holds another BWT");
    }
    return MSGS_OK;
}

/* chrom names and lengths of the resident database (for record headers) */
extern "C" int msgs_db_chrom(msgs_ctx* c, int32_t tid, char* name, int64_t cap, int64_t* len) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (tid < 0 || tid >= c->db.nChrom) return fail(c, MSGS_E_ARG, "bad tid");
    if (len) *len = c->db.chromLen[tid];
    if (name) {
        const std::string& s = c->db.names[tid];
        if ((int64_t)s.size() + 1 > cap) return fail(c, MSGS_E_CAP, "name capacity");
        std::memcpy(name, s.c_str(), s.size() + 1);
    }
    return MSGS_OK;
}
extern "C" int32_t msgs_db_nchrom(msgs_ctx* c) { return c ? c->db.nChrom : 0; }

/* context options: "form" (0 plain FMD, 1 unique-interval text path, 2 the batched form:
 * lane groups per read, text path and the learned-index read filter), "group_lanes",
 * "lanes_per_cu", "smem_cap", "host_threads" */
extern "C" int msgs_ctx_set(msgs_ctx* c, const char* key, int64_t value) {
    if (!c || !key) return MSGS_E_ARG;
    std::string k(key);
    if (k == "form") { if (value < 0 || value > 2) return fail(c, MSGS_E_ARG, "form 0, 1 or 2"); c->form = (int)value; }
    else if (k == "group_lanes") { if (value != 0 && value != 8 && value != 16) return fail(c, MSGS_E_ARG, "group_lanes: 0, 8 or 16"); c->groupLanes = (int)value; }
    else if (k == "lanes_per_cu") { if (value < 64 || value > 8192 || value % 64) return fail(c, MSGS_E_ARG, "lanes_per_cu: multiple of 64 in [64, 8192]"); c->lanesPerCU = (int)value; }
    else if (k == "smem_cap") { if (value < 1 || value > 4096) return fail(c, MSGS_E_ARG, "smem_cap in [1, 4096]"); c->smemCap = (int)value; }
    else if (k == "host_threads") { if (value < 1 || value > 256) return fail(c, MSGS_E_ARG, "host_threads in [1, 256]"); c->hostThreads = (int)value; }
    else if (k == "read_filter") { if (value != 0 && value != 1) return fail(c, MSGS_E_ARG, "read_filter: 0 or 1"); c->readFilter = (int)value; }
    else if (k == "text_path") { if (value != 0 && value != 1) return fail(c, MSGS_E_ARG, "text_path: 0 or 1"); c->textPath = (int)value; }
    else if (k == "split_lanes") {
        if (value != 0 && value != 4 && value != 8 && value != 16) return fail(c, MSGS_E_ARG, "split_lanes: 0, 4, 8 or 16");
        c->splitLanes = (int)value;
    }
    else if (k == "pair_chains") { if (value != 0 && value != 1) return fail(c, MSGS_E_ARG, "pair_chains: 0 or 1"); c->pairChains = (int)value; }
    else if (k == "streams") { if (value < 1 || value > kWorkExecs) return fail(c, MSGS_E_ARG, "streams in [1, 4]"); c->workStreams = (int)value; }
    else if (k == "ingest") { if (value != 0 && value != 1) return fail(c, MSGS_E_ARG, "ingest: 0 or 1"); c->ingest = (int)value; }
    else if (k == "record_format") { if (value != 0 && value != 1) return fail(c, MSGS_E_ARG, "record_format: 0 or 1"); c->recordFormat = (int)value; }
    else if (k == "dp_serial") { if (value != 0 && value != 1) return fail(c, MSGS_E_ARG, "dp_serial: 0 or 1"); c->dpSerial = (int)value; }
    else if (k == "wide_index") { if (value != 0 && value != 1) return fail(c, MSGS_E_ARG, "wide_index: 0 or 1"); c->wideIndex = (int)value; }
    else if (k == "build_block_mb") { if (value < 0 || value > (1 << 20)) return fail(c, MSGS_E_ARG, "build_block_mb in [0, 2^20]"); c->buildBlockMb = value; }
    else return fail(c, MSGS_E_ARG, "unknown option " + k);
    return MSGS_OK;
}

extern "C" int msgs_ctx_get(msgs_ctx* c, const char* key, int64_t* value) {
    if (!c || !key || !value) return MSGS_E_ARG;
    std::string k(key);
    if (k == "form") *value = c->form;
    else if (k == "group_lanes") *value = c->groupLanes;
    else if (k == "lanes_per_cu") *value = c->lanesPerCU;
    else if (k == "smem_cap") *value = c->smemCap;
    else if (k == "host_threads") *value = c->hostThreads;
    else if (k == "read_filter") *value = c->readFilter;
    else if (k == "text_path") *value = c->textPath;
    else if (k == "split_lanes") *value = c->splitLanes;
    else if (k == "pair_chains") *value = c->pairChains;
    else if (k == "streams") *value = c->workStreams;
    else if (k == "ingest") *value = c->ingest;
    else if (k == "dp_serial") *value = c->dpSerial;
    else if (k == "record_format") *value = c->recordFormat;
    else if (k == "wide_index") *value = c->wideIndex;
    else if (k == "build_block_mb") *value = c->buildBlockMb;
    else if (k == "index_blocks") *value = c->db.blockEnds.empty() ? 1 : (int64_t)c->db.blockEnds.size();
    else if (k == "index_wide") *value = c->db.wide ? 1 : 0;
    else return fail(c, MSGS_E_ARG, "unknown option " + k);
    return MSGS_OK;
}

/* ------------------------------------------------------------------ counter calibration */

namespace {
__global__ void __launch_bounds__(256) k_calib_copy(const uint4* src, uint4* dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

/* nLoads gathers of one 64-byte line each at hashed line indexes of a buffer of nLines
 * lines (the access pattern of occ-block / SA lookups with no locality) */
__global__ void __launch_bounds__(256) k_gather(const uint4* buf, int64_t nLines, int64_t nLoads, uint64_t seed,
                                                unsigned long long* sink) {
    uint32_t acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nLoads; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + (uint64_t)i * 0x9e3779b97f4a7c15ull;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        const uint4* l = buf + 4 * (int64_t)(z % (uint64_t)nLines);
        const uint4 a = l[0], b = l[1], c = l[2], d = l[3];
        acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^ d.z ^ d.w;
    }
    if (acc == 0x9e3779b9u) atomicAdd(sink, 1ull);
}
}  // namespace

/* One device-to-device copy of `bytes` (a multiple of 16) by a plain kernel, so that
 * hardware byte counters can be scaled against a transfer of known size. */
extern "C" int msgs_util_calib_copy(msgs_ctx* c, int64_t bytes, double* micros) {
    if (!c || bytes <= 0 || bytes % 16) return fail(c, MSGS_E_ARG, "bytes must be a positive multiple of 16");
    hipError_t e = hipSuccess;
    uint4* a = c->x().bufs.get<uint4>("calib.a", bytes / 16, e);
    uint4* b = c->x().bufs.get<uint4>("calib.b", bytes / 16, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "calibration buffers");
    const int64_t n = bytes / 16;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, (int64_t)std::max(1, c->nCU) * 64);
    (void)hipEventRecord(c->x().ev0, c->x().stream);
    k_calib_copy<<<blocks, 256, 0, c->x().stream>>>(a, b, n);
    (void)hipEventRecord(c->x().ev1, c->x().stream);
    if (hipStreamSynchronize(c->x().stream) != hipSuccess) return fail(c, MSGS_E_HIP, "calibration copy");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->x().ev0, c->x().ev1);
    if (micros) *micros = ms * 1000.0;
    return MSGS_OK;
}

/* Random-gather ceiling: nLoads 64-byte line reads at uniformly hashed positions of a
 * device buffer of `bytes` (a multiple of 64), timed with HIP events (second of two
 * launches). */
extern "C" int msgs_util_random_gather(msgs_ctx* c, int64_t bytes, int64_t nLoads, double* micros) {
    if (!c || bytes <= 0 || bytes % 64 || nLoads <= 0) return fail(c, MSGS_E_ARG, "bytes: positive multiple of 64; nLoads > 0");
    hipError_t e = hipSuccess;
    uint4* buf = c->x().bufs.get<uint4>("gather.buf", bytes / 16, e);
    unsigned long long* sink = c->x().bufs.get<unsigned long long>("gather.sink", 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "gather buffer");
    if (hipMemsetAsync(buf, 0x5a, bytes, c->x().stream) != hipSuccess) return fail(c, MSGS_E_HIP, "gather fill");
    const unsigned blocks = (unsigned)std::max(1, c->nCU) * 32;
    float ms = 0;
    for (int k = 0; k < 2; ++k) {
        (void)hipEventRecord(c->x().ev0, c->x().stream);
        k_gather<<<blocks, 256, 0, c->x().stream>>>(buf, bytes / 64, nLoads, 0x1234567ull + k, sink);
        (void)hipEventRecord(c->x().ev1, c->x().stream);
        if (hipStreamSynchronize(c->x().stream) != hipSuccess) return fail(c, MSGS_E_HIP, "gather");
        (void)hipEventElapsedTime(&ms, c->x().ev0, c->x().ev1);
    }
    if (micros) *micros = ms * 1000.0;
    return MSGS_OK;
}
