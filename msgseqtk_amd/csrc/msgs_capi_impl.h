/*
 * msgs_capi_impl.h -- context internals shared by the C-ABI translation units.
 */
#ifndef MSGS_CAPI_IMPL_H
#define MSGS_CAPI_IMPL_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <new>
#include <utility>
#include <string>
#include <memory>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "msgs_host.h"
#include "msgs_lix.h"

/* grow-only named device buffers */
struct BufPool {
    struct B { void* p = nullptr; size_t cap = 0; };
    std::map<std::string, B> m;
    template <class T> T* get(const std::string& name, size_t count, hipError_t& e) {
        B& b = m[name];
        size_t need = std::max<size_t>(count, 1) * sizeof(T);
        if (need > b.cap) {
            if (b.p) (void)hipFree(b.p);
            b.p = nullptr;
            b.cap = 0;
            size_t want = need + need / 4;
            hipError_t r = hipMalloc(&b.p, want);
            if (r != hipSuccess) { e = r; b.p = nullptr; return nullptr; }
            b.cap = want;
        }
        return (T*)b.p;
    }
    void releaseAll() {
        for (auto& kv : m) if (kv.second.p) (void)hipFree(kv.second.p);
        m.clear();
    }
};

struct DevReads {
    int64_t n = 0, maxL = 0, total = 0;
    uint8_t* seq = nullptr;
    int64_t* offs = nullptr;
    std::vector<int64_t> offsHost;
    /* resident batches: reverse complements of the reads and their qualities, made once at
     * upload (NULL: the DP stage makes them per run) */
    uint8_t* rcSeq = nullptr;
    uint8_t* rcQual = nullptr;
};

struct SmemStage {
    int cap = 0;
    int32_t* cnt = nullptr;
    SmemRec* recs = nullptr;
    std::vector<int32_t> cntHost;
};

/* host copy of a batch (names, bases and qualities are needed for the records) */
struct HostBatch {
    std::vector<uint8_t> seq, qual;
    std::vector<int64_t> offs;
    std::string names;
    std::vector<int64_t> nameOff;   /* [start, end) of each name in `names` */
    msgs_reads view{};
};

/* page-locked host memory for device-to-host results (fast DMA), non-coherent so that
 * the host reads it through its caches (results are only moved by explicit copies);
 * resize leaves new elements uninitialised (they are overwritten by the copy) */
template <class T> struct PinnedAlloc {
    using value_type = T;
    PinnedAlloc() = default;
    template <class U> PinnedAlloc(const PinnedAlloc<U>&) {}
    T* allocate(size_t n) {
        void* p = nullptr;
        if (hipHostMalloc(&p, std::max<size_t>(1, n) * sizeof(T), hipHostMallocNonCoherent) != hipSuccess)
            throw std::bad_alloc();
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_t) { (void)hipHostFree(p); }
    template <class U> void construct(U*) noexcept {}
    template <class U, class... A> void construct(U* p, A&&... a) { ::new ((void*)p) U(std::forward<A>(a)...); }
    template <class U> bool operator==(const PinnedAlloc<U>&) const { return true; }
    template <class U> bool operator!=(const PinnedAlloc<U>&) const { return false; }
};
template <class T> using PVec = std::vector<T, PinnedAlloc<T>>;

struct BatchResult;

/* candidate mate pairs of a paired batch after AlignmentPE's sort and filter: read pair i's
 * are pairs[off[i] .. off[i] + kept[i]) in the reference's order */
struct PairStage {
    std::vector<int64_t> off;
    PVec<int32_t> kept;
    PVec<PairRec> pairs;
};

/* getPairs + sort + filter on the device for n read pairs whose alignment records are
 * dRecs[jobOff[i] .. jobOff[i+1]) (read i) and dRecs[jobOff[n+i] .. jobOff[n+i+1]) (its
 * mate).  scoreFilter: first drop records with alnScore < L * minScoreRate (Alignment::filter;
 * read lengths from dReadOffs[2n+1], scores are integers x scoreScale); otherwise the lists
 * are taken as given. */
int runPairs(msgs_ctx* c, int64_t n, const std::vector<int64_t>& jobOff, const AlnRec* dRecs,
             const int64_t* dReadOffs, bool scoreFilter, double scoreScale, const msgs_params* p, PairStage& out);

/* a resident batch (reads kept in HBM between runs); paired batches hold the reads
 * followed by their mates on the device.  A submitted run finishes its records on a
 * host worker thread while the caller goes on with other batches. */
struct Slot {
    bool used = false;
    bool paired = false;
    HostBatch host[2];
    DevReads dev;
    uint8_t* dQual = nullptr;
    /* the device batch's bases and qualities in pinned host memory (reads then mates), copied
     * to the device again at the start of every run when the context option "ingest" is 1 */
    PVec<uint8_t> inSeq, inQual;
    std::thread worker;
    bool pending = false;
    std::shared_ptr<BatchResult> res;
    msgs_params params{};
    std::vector<std::string> parts;
    std::vector<int64_t> lineOffs;
    double stats[48] = {};
    int rc = 0;
    std::string err;
    void join() {
        if (worker.joinable()) worker.join();
        pending = false;   /* res is kept: its pinned buffers serve the next run */
    }
    void release() { join(); used = false; }
};

/* what one host thread needs to drive the device path: a stream, timing events, device
 * pools and the last stage timings.  The caller's thread uses the context's own; each
 * submitted batch runs its device stages on one of the worker executors, so that two
 * batches' kernels can overlap on the GPU. */
struct Exec {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr, ev4 = nullptr, ev5 = nullptr;
    hipEvent_t ev6 = nullptr, ev7 = nullptr;   /* per-run ingest of a resident batch */
    BufPool bufs;
    double lastSmemMicros = 0;    /* the SMEM search kernel(s) of the last SMEM stage */
    double lastStageMicros = 0;   /* the whole SMEM stage (prefix sums, search, merge) */
    double lastFilterMicros = 0;
    /* the insert model's table in bufs "pr.tab": its parameters and range when filled */
    double pairTabKey[6] = {};
    bool pairTabValid = false;
    std::string err;
    bool create(unsigned streamFlags);
    void destroy();
};
constexpr int kWorkExecs = 4;   /* worker executors created; "streams" of them are used */

struct msgs_ctx {
    int device = 0;
    int nCU = 256;
    int lanesPerCU = 1024;
    int smemCap = 32;
    int readFilter = 1;        /* form 2: also skip reads that cannot emit an SMEM */
    int textPath = 1;          /* form 2 with one lane per read: follow unique intervals through the text */
    int form = 1;
    int groupLanes = 8;   /* form 2: lanes per read of the SMEM kernel (0 = one lane per read) */
    int splitLanes = 0;   /* form 2 with one lane per read: lanes per read for the reads marked
                           * to start early, searched by a lane-group launch first (0 = off) */
    int pairChains = 0;   /* one-lane search kernel: 1 = two lanes per read (forward and backward
                           * searches of findAllSMEMS on separate lanes, smem_pair_kernel) */
    int ingest = 0;       /* resident batches: 1 = every run first copies the batch's reads and
                           * qualities from pinned host memory and makes the reverse complements */
    int recordFormat = 0; /* records of msgs_map_batch / msgs_batch_sam: 0 SAM text, 1 BAM */
    int dpSerial = 0;     /* 1 = every DP job takes the serial fill (M, I, D kept per cell) instead
                           * of the wavefront fill */
    int64_t buildBlockMb = 2048; /* msgseqtk-build's -b (MB): databases of this many MB or more take
                                  * the suffix order of its block-by-block build (0: one block) */
    int wideIndex = 0;    /* databases built next: 0 = the wide index layout (64-bit rows) only from
                           * 2^31 symbols on; 1 = always, with 2^14-position superblocks and
                           * 2^16-suffix sorting ranges (the wide code paths on small texts) */
    Exec own;                      /* the caller's thread */
    Exec work[kWorkExecs];         /* submitted batches' device stages */
    int workStreams = 4;           /* how many run at once */
    bool workBusy[kWorkExecs] = {};
    std::mutex workMu;
    std::condition_variable workCv;
    DbState db;
    LixState lix;
    Slot slots[16];
    double stats[48] = {};
    int hostThreads = 16;
    /* the executor of the calling thread */
    Exec& x();
    Exec* acquireWork();
    void releaseWork(Exec* e);
    /* SAM text of the last msgs_batch_run: per-thread parts in read order and the
     * offset of each read's records in their concatenation */
    std::vector<std::string> lastParts;
    std::vector<int64_t> lastLineOffs;
    /* finished batch results, recycled so that their pinned buffers are reused */
    std::vector<std::shared_ptr<BatchResult>> spareResults;
};

int fail(msgs_ctx* c, int code, const std::string& msg);
/* writes DB.fmd of a BD text with this BWT and suffix array (msgs_fmd.cpp) */
int fmdWrite(const std::string& path, bool header, const uint8_t* bwt, const int64_t* sa, int64_t n);
/* looks up the allocation-free scratch array's address on `device` (the current device) */
int noallocInit(int device);
bool evalueThreshold(double n, double maxE, double* T);
int uploadReads(msgs_ctx* c, const msgs_reads* r, const char* pfx, DevReads& out);
int runSmem(msgs_ctx* c, const DevReads& r, const msgs_params* p, SmemStage& out, unsigned long long* dStats);
int runLixFilter(msgs_ctx* c, const DevReads& r, int32_t minLen, double T, int32_t gate1, uint8_t** needOut,
                 SmemRec** exactOut, unsigned long long* dStats);
/* the search order (reads marked to start early first); *splitOut: device count of those */
int runLixOrder(msgs_ctx* c, int64_t n, const uint8_t* need, int32_t** orderOut,
                const unsigned long long** splitOut);

#endif
