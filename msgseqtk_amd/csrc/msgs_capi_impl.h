/*
 * msgs_capi_impl.h -- context internals shared by the C-ABI translation units.
 */
#ifndef MSGS_CAPI_IMPL_H
#define MSGS_CAPI_IMPL_H

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <memory>
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
    msgs_reads view{};
};

struct BatchResult;

/* a resident batch (reads kept in HBM between runs); paired batches hold the reads
 * followed by their mates on the device.  A submitted run finishes its records on a
 * host worker thread while the caller goes on with other batches. */
struct Slot {
    bool used = false;
    bool paired = false;
    HostBatch host[2];
    DevReads dev;
    uint8_t* dQual = nullptr;
    std::thread worker;
    bool pending = false;
    std::shared_ptr<BatchResult> res;
    msgs_params params{};
    std::vector<std::string> parts;
    std::vector<int64_t> lineOffs;
    double stats[16] = {};
    void join() {
        if (worker.joinable()) worker.join();
        pending = false;
        res.reset();
    }
    void release() { join(); used = false; }
};

struct msgs_ctx {
    int device = 0;
    int nCU = 256;
    int lanesPerCU = 1024;
    int smemCap = 32;
    int form = 1;
    int groupLanes = 8;   /* form 2: lanes per read of the SMEM kernel (0 = one lane per read) */
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    std::string err;
    DbState db;
    LixState lix;
    BufPool bufs;
    Slot slots[16];
    double stats[16] = {};
    double lastSmemMicros = 0;
    double lastFilterMicros = 0;
    int hostThreads = 16;
    /* SAM text of the last msgs_batch_run: per-thread parts in read order and the
     * offset of each read's records in their concatenation */
    std::vector<std::string> lastParts;
    std::vector<int64_t> lastLineOffs;
};

int fail(msgs_ctx* c, int code, const std::string& msg);
bool evalueThreshold(double n, double maxE, double* T);
int uploadReads(msgs_ctx* c, const msgs_reads* r, const char* pfx, DevReads& out);
int runSmem(msgs_ctx* c, const DevReads& r, const msgs_params* p, SmemStage& out, unsigned long long* dStats);
int runLixFilter(msgs_ctx* c, const DevReads& r, int32_t minLen, double T, uint8_t** needOut,
                 unsigned long long* dStats);

#endif
