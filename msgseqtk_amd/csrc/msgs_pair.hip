/*
 * msgs_pair.hip -- paired-end candidate pairs on the device: AlignmentPE::getPairs, the insert
 * model's likelihood terms, AlignmentPE::sort and AlignmentPE::filter
 * (src/Alignment.cpp:547-581, src/Alignment.h:446-575,635-639, src/PairingScheme.cpp:18-25).
 *
 * One lane per read pair.  A first pass counts each read pair's candidates (mates on opposite
 * strands) and the largest insert size; a host scan places every read pair's candidates in one
 * pool; the second pass writes them in getPairs order (the read's alignments outer, the mate's
 * inner), sorts them by log10lik with the restated libstdc++ sort (equal keys end where the
 * reference's std::sort leaves them), removes the filtered ones in order (remove_if keeps the
 * order of the rest) and cuts the list at maxNPair.
 *
 * The insert model enters only through log10(pr(isize)) and log(pr(isize)), functions of the
 * integer insert size.  The host evaluates them with its own libm, as the reference does
 * (boost's normal pdf expression, then glibc exp, log10, log), for every insert size inside
 * [insMin, insMax] up to the largest one the batch holds, and the device looks them up; the
 * likelihood sums are IEEE additions and divisions in the reference's order (no contraction),
 * so the keys and their sort order are the reference's.  MAPQ (calcMapQ: exp, log10, round over
 * arbitrary doubles) stays with the host's libm.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "msgs_capi_impl.h"
#include "msgs_sort.h"

namespace {

constexpr int32_t NEG = -1073741824;   /* the DP's -inf score */

struct PairArgs {
    const AlnRec* recs;
    const int64_t* jobOff;      /* 2n+1 */
    const int64_t* readOffs;    /* 2n+1 (read lengths for the score filter) */
    int64_t n;
    int32_t scoreFilter, uniform;   /* uniform: insMean == 0, pr = 1 for every insert size */
    double minScoreRate, scoreScale;
    double insMin, insMax, log10e;
    const double* tab;          /* [2k] = log10(pr(tlo + k)), [2k+1] = log(pr(tlo + k)), k < tn */
    int64_t tlo, tn;
    int32_t noDiscordant, noTailOver, noContain, noOverlap;
    int64_t maxNPair;
    int64_t* cnt;               /* n: candidates per read pair */
    unsigned int* maxIs;        /* largest insert size + 1 */
    const int64_t* off;         /* n: pool offsets */
    PairRec* pool;
    int32_t* kept;              /* n */
    int32_t* bad;               /* set when an insert size in range has no table entry */
};

/* Alignment::filter (src/Alignment.cpp:547-554): alnScore < readLen * minScoreRate drops */
__device__ __forceinline__ bool keepRec(const PairArgs& a, int64_t read, const AlnRec& x) {
    if (!a.scoreFilter) return true;
    const double thr = (double)(a.readOffs[read + 1] - a.readOffs[read]) * a.minScoreRate;
    const double sc = x.score == NEG ? -INFINITY : (double)x.score / a.scoreScale;
    return !(sc < thr);
}

/* AlignmentPE::getInsertSize (src/Alignment.cpp:540-545) */
__device__ __forceinline__ int32_t insertSize(const AlnRec& x, const AlnRec& y) {
    if (x.tid != y.tid) return -1;
    return (int32_t)(max(x.alnEnd, y.alnEnd) - min(x.alnStart, y.alnStart));
}

__global__ void __launch_bounds__(256) pair_count_kernel(PairArgs a) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a0 = a.jobOff[i], a1 = a.jobOff[i + 1], b0 = a.jobOff[a.n + i], b1 = a.jobOff[a.n + i + 1];
        int64_t c = 0;
        int32_t mx = -1;
        for (int64_t x = a0; x < a1; ++x) {
            const AlnRec X = a.recs[x];
            if (!keepRec(a, i, X)) continue;
            for (int64_t y = b0; y < b1; ++y) {
                const AlnRec& Y = a.recs[y];
                if ((X.strand & Y.strand) != 0 || !keepRec(a, a.n + i, Y)) continue;
                ++c;
                mx = max(mx, insertSize(X, Y));
            }
        }
        a.cnt[i] = c;
        if (c) atomicMax(a.maxIs, (unsigned int)(mx + 1));
    }
}

struct L10Greater {
    __device__ bool operator()(const PairRec& l, const PairRec& r) const { return l.l10 > r.l10; }
};

__global__ void __launch_bounds__(256) pair_kernel(PairArgs a) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a0 = a.jobOff[i], a1 = a.jobOff[i + 1], b0 = a.jobOff[a.n + i], b1 = a.jobOff[a.n + i + 1];
        PairRec* P = a.pool + a.off[i];
        int64_t m = 0;
        /* getPairs with the pair's likelihoods (AlignmentPE::log10lik / loglik,
         * src/Alignment.h:491-499) */
        for (int64_t x = a0; x < a1; ++x) {
            const AlnRec X = a.recs[x];
            if (!keepRec(a, i, X)) continue;
            for (int64_t y = b0; y < b1; ++y) {
                const AlnRec& Y = a.recs[y];
                if ((X.strand & Y.strand) != 0 || !keepRec(a, a.n + i, Y)) continue;
                PairRec q;
                q.f = (int32_t)x;
                q.r = (int32_t)y;
                q.isize = insertSize(X, Y);
                q.pad = 0;
                double t10, te;
                const double xs = (double)q.isize;
                if (a.uniform) {
                    t10 = 0.0;   /* log10(1) and log(1) */
                    te = 0.0;
                } else if (a.insMin <= xs && xs <= a.insMax) {
                    const int64_t k = (int64_t)q.isize - a.tlo;
                    if (k >= 0 && k < a.tn) {
                        t10 = a.tab[2 * k];
                        te = a.tab[2 * k + 1];
                    } else {
                        atomicOr(a.bad, 1);
                        t10 = te = 0.0;
                    }
                } else {
                    t10 = -INFINITY;   /* pr = 0 outside the range */
                    te = -INFINITY;
                }
                q.l10 = X.log10P + Y.log10P + t10;
                q.ll = X.log10P / a.log10e + Y.log10P / a.log10e + te;
                P[m++] = q;
            }
        }
        /* AlignmentPE::sort (src/Alignment.h:635-639) */
        msgs_sort::sort(P, P + m, L10Greater{});
        /* AlignmentPE::filter (src/Alignment.cpp:556-571): the removals in sequence are one
         * order-keeping pass over the union of their predicates */
        int64_t w = 0;
        for (int64_t k = 0; k < m; ++k) {
            const PairRec q = P[k];
            const AlnRec& A = a.recs[q.f];
            const AlnRec& B = a.recs[q.r];
            const double xs = (double)q.isize;
            bool rm = !(a.insMin <= xs && xs <= a.insMax);
            if (a.noDiscordant) rm = rm || !(A.tid == B.tid && (A.strand & B.strand) == 0);
            if (a.noTailOver)
                rm = rm || (A.alnStart <= B.alnStart && A.alnEnd > B.alnStart) ||
                     (B.alnStart <= A.alnStart && B.alnEnd > A.alnStart);
            if (a.noContain)
                rm = rm || (A.alnStart <= B.alnStart && A.alnEnd >= B.alnEnd) ||
                     (B.alnStart <= A.alnStart && B.alnEnd >= A.alnEnd);
            if (a.noOverlap) rm = rm || (A.alnStart < B.alnEnd && A.alnEnd > B.alnStart);
            if (!rm) P[w++] = q;
        }
        if (a.maxNPair > 0 && w > a.maxNPair) w = a.maxNPair;
        a.kept[i] = (int32_t)w;
    }
}

__global__ void k_pair_zero(unsigned int* p, int32_t* q) {
    if (threadIdx.x == 0) { *p = 0; *q = 0; }
}

inline unsigned gridOf(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
}

/* PairingScheme::pr (src/PairingScheme.cpp:18-25) with boost::math::pdf(normal(m, s), x)
 * evaluated in boost's expression order */
double pairPr(const msgs_params* p, double x) {
    if (p->insMean == 0) return 1;
    if (p->insMin <= x && x <= p->insMax) {
        double ex = x - p->insMean;
        ex *= -ex;
        ex /= 2 * p->insSD * p->insSD;
        double r = std::exp(ex);
        r /= p->insSD * 2.506628274631000502415765284811045253e+00;
        return r;
    }
    return 0;
}

#define PCK(x, msg)                                                          \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) return fail(c, MSGS_E_HIP, std::string(msg) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

int runPairs(msgs_ctx* c, int64_t n, const std::vector<int64_t>& jobOff, const AlnRec* dRecs,
             const int64_t* dReadOffs, bool scoreFilter, double scoreScale, const msgs_params* p, PairStage& out) {
    hipStream_t st = c->x().stream;
    hipError_t e = hipSuccess;
    out.off.assign(n + 1, 0);
    out.kept.resize(n);
    out.pairs.resize(0);
    if (n == 0) return MSGS_OK;
    if ((int64_t)jobOff.size() < 2 * n + 1) return fail(c, MSGS_E_ARG, "pairing: job offsets");
    if (jobOff[2 * n] >= (int64_t)INT32_MAX) return fail(c, MSGS_E_UNSUPPORTED, "pairing: more than 2^31 alignments");
    PairArgs a;
    std::memset(&a, 0, sizeof(a));
    a.recs = dRecs;
    a.readOffs = dReadOffs;
    a.n = n;
    a.scoreFilter = scoreFilter ? 1 : 0;
    a.uniform = p->insMean == 0 ? 1 : 0;
    a.minScoreRate = p->minScoreRate;
    a.scoreScale = scoreScale;
    a.insMin = p->insMin;
    a.insMax = p->insMax;
    a.log10e = ::log10(M_E);
    a.noDiscordant = p->noDiscordant;
    a.noTailOver = p->noTailOver;
    a.noContain = p->noContain;
    a.noOverlap = p->noOverlap;
    a.maxNPair = p->maxNPair;
    int64_t* dJobOff = c->x().bufs.get<int64_t>("pr.joff", 2 * n + 1, e);
    a.cnt = c->x().bufs.get<int64_t>("pr.cnt", n, e);
    int64_t* dOff = c->x().bufs.get<int64_t>("pr.off", n + 1, e);
    a.kept = c->x().bufs.get<int32_t>("pr.kept", n, e);
    unsigned int* flags = c->x().bufs.get<unsigned int>("pr.flags", 2, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "pairing buffers");
    a.jobOff = dJobOff;
    a.off = dOff;
    a.maxIs = flags;
    a.bad = (int32_t*)(flags + 1);
    PCK(hipMemcpyAsync(dJobOff, jobOff.data(), (2 * n + 1) * 8, hipMemcpyHostToDevice, st), "pairing: job offsets");
    k_pair_zero<<<1, 64, 0, st>>>(a.maxIs, a.bad);
    pair_count_kernel<<<gridOf(n), 256, 0, st>>>(a);
    std::vector<int64_t> cnt(n);
    unsigned int maxIs = 0;
    PCK(hipMemcpyAsync(cnt.data(), a.cnt, n * 8, hipMemcpyDeviceToHost, st), "pairing: counts");
    PCK(hipMemcpyAsync(&maxIs, a.maxIs, 4, hipMemcpyDeviceToHost, st), "pairing: insert sizes");
    PCK(hipStreamSynchronize(st), "pairing: count kernel");
    for (int64_t i = 0; i < n; ++i) out.off[i + 1] = out.off[i] + cnt[i];
    const int64_t total = out.off[n];
    /* the insert model's terms for the integer insert sizes in [insMin, insMax] that occur
     * at most (-1 for mates on different chromosomes) */
    if (!a.uniform && total > 0) {
        const double lo = std::max(std::ceil(p->insMin), -1.0);
        const double hi = std::min(std::floor(p->insMax), (double)maxIs - 1.0);
        const int64_t tlo = lo <= hi ? (int64_t)lo : 0;
        const int64_t tn = lo <= hi ? (int64_t)(hi - lo) + 1 : 0;
        Exec& X = c->x();
        const double key[6] = {p->insMean, p->insSD, p->insMin, p->insMax, (double)tlo, (double)tn};
        double* dTab = X.bufs.get<double>("pr.tab", 2 * std::max<int64_t>(tn, 1), e);
        if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "pairing table");
        if (!X.pairTabValid || std::memcmp(key, X.pairTabKey, sizeof(key)) != 0) {
            std::vector<double> tab(2 * std::max<int64_t>(tn, 1));
            for (int64_t k = 0; k < tn; ++k) {
                const double pr = pairPr(p, (double)(tlo + k));
                tab[2 * k] = std::log10(pr);
                tab[2 * k + 1] = std::log(pr);
            }
            PCK(hipMemcpyAsync(dTab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, st), "pairing table");
            PCK(hipStreamSynchronize(st), "pairing table");
            std::memcpy(X.pairTabKey, key, sizeof(key));
            X.pairTabValid = true;
        }
        a.tab = dTab;
        a.tlo = tlo;
        a.tn = tn;
    }
    a.pool = c->x().bufs.get<PairRec>("pr.pool", total, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "pair pool");
    PCK(hipMemcpyAsync(dOff, out.off.data(), (n + 1) * 8, hipMemcpyHostToDevice, st), "pairing: offsets");
    pair_kernel<<<gridOf(n), 256, 0, st>>>(a);
    out.pairs.resize(total);
    int32_t bad = 0;
    PCK(hipMemcpyAsync(out.kept.data(), a.kept, n * 4, hipMemcpyDeviceToHost, st), "pairing: kept");
    if (total) PCK(hipMemcpyAsync(out.pairs.data(), a.pool, total * sizeof(PairRec), hipMemcpyDeviceToHost, st), "pairs");
    PCK(hipMemcpyAsync(&bad, a.bad, 4, hipMemcpyDeviceToHost, st), "pairing: flags");
    PCK(hipStreamSynchronize(st), "pairing kernel");
    if (bad) return fail(c, MSGS_E_HIP, "pairing: an insert size in range had no table entry");
    return MSGS_OK;
}

extern "C" double msgs_util_insert_pr(const msgs_params* p, double x) { return p ? pairPr(p, x) : 0.0; }

/* AlignmentPE::getPairs + sort + filter over caller alignment lists (see include/msgs_hip.h) */
static int impl_pair_batch(msgs_ctx* c, int64_t n, const int64_t* alnOffs, const int64_t* rows4, const double* log10P,
                           const msgs_params* p, int64_t* pairCounts, int64_t* pairRows3, double* pairLik2,
                           int64_t cap, int64_t* needed) {
    if (!c) return MSGS_E_ARG;
    if (!p || n < 0 || (n > 0 && (!alnOffs || !pairCounts))) return fail(c, MSGS_E_ARG, "pairing arguments");
    const int64_t nAln = n > 0 ? alnOffs[2 * n] : 0;
    if (n > 0 && (alnOffs[0] != 0 || nAln < 0 || (nAln > 0 && (!rows4 || !log10P))))
        return fail(c, MSGS_E_ARG, "pairing: alignment offsets");
    for (int64_t i = 0; i < 2 * n; ++i)
        if (alnOffs[i + 1] < alnOffs[i]) return fail(c, MSGS_E_ARG, "pairing: alignment offsets decrease");
    if (nAln >= (int64_t)INT32_MAX) return fail(c, MSGS_E_UNSUPPORTED, "pairing: more than 2^31 alignments");
    hipStream_t st = c->x().stream;
    hipError_t e = hipSuccess;
    std::vector<AlnRec> h(std::max<int64_t>(nAln, 1));
    for (int64_t k = 0; k < nAln; ++k) {
        AlnRec& r = h[k];
        std::memset(&r, 0, sizeof(r));
        r.tid = (int32_t)rows4[4 * k];
        r.strand = (int32_t)rows4[4 * k + 1];
        r.alnStart = rows4[4 * k + 2];
        r.alnEnd = rows4[4 * k + 3];
        r.log10P = log10P[k];
    }
    AlnRec* dRecs = c->x().bufs.get<AlnRec>("pr.recs", h.size(), e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "pairing: alignments");
    if (nAln) PCK(hipMemcpyAsync(dRecs, h.data(), nAln * sizeof(AlnRec), hipMemcpyHostToDevice, st), "pairing: alignments");
    std::vector<int64_t> jobOff(alnOffs, alnOffs + (n > 0 ? 2 * n + 1 : 1));
    PairStage ps;
    int rc = runPairs(c, n, jobOff, dRecs, nullptr, false, 1.0, p, ps);
    if (rc) return rc;
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i) total += ps.kept[i];
    if (needed) *needed = total;
    for (int64_t i = 0; i < n; ++i) pairCounts[i] = ps.kept[i];
    if (total > cap) return fail(c, MSGS_E_CAP, "pair capacity");
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int32_t j = 0; j < ps.kept[i]; ++j, ++k) {
            const PairRec& q = ps.pairs[ps.off[i] + j];
            if (pairRows3) { pairRows3[3 * k] = q.f; pairRows3[3 * k + 1] = q.r; pairRows3[3 * k + 2] = q.isize; }
            if (pairLik2) { pairLik2[2 * k] = q.l10; pairLik2[2 * k + 1] = q.ll; }
        }
    return MSGS_OK;
}

extern "C" int msgs_pair_batch(msgs_ctx* c, int64_t n, const int64_t* alnOffs, const int64_t* rows4,
                               const double* log10P, const msgs_params* p, int64_t* pairCounts, int64_t* pairRows3,
                               double* pairLik2, int64_t cap, int64_t* needed) {
    try {
        return impl_pair_batch(c, n, alnOffs, rows4, log10P, p, pairCounts, pairRows3, pairLik2, cap, needed);
    } catch (const std::bad_alloc&) {
        return fail(c, MSGS_E_NOMEM, "host memory");
    }
}
