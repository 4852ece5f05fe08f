/*
 * msgs_kernels.h -- kernel argument blocks and launch declarations shared by the
 * HIP translation units of libmsgs_hip.so.
 */
#ifndef MSGS_KERNELS_H
#define MSGS_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "msgs_device.h"

/* one SMEM: read interval [from, to), bi-interval (fwd, rev, size), loglik */
struct SmemRec {
    int32_t from, to;
    int64_t fwd, rev, size;
    double logP;
};

/* a candidate of a search's lock-step list: bi-interval, loglik, its moving side as a
 * text position (X, -1 when the row values are exact) and its fixed read end (pos); rows
 * of the index's width.  In the kernels' global scratch a candidate is packed into 24 or
 * 32 bytes (SMEM_CAND_BYTES at most per entry). */
template <class Row>
struct SmemCandT {
    Row F, R, S, X;
    double LL;
    int32_t pos, pad;
};

constexpr int64_t SMEM_CAND_BYTES = 32;
/* reads of at least this many bases are refused on 64-bit-row indexes (the packed
 * candidate keeps 26 bits of read position) */
constexpr int64_t SMEM_WIDE_MAX_READ = (int64_t)1 << 26;

struct SmemArgs {
    const uint8_t* seqs;
    const int64_t* offs;
    int64_t nReads;
    int64_t minLen;
    double T1, T2;             /* phase-1 / phase-2 e-value thresholds on logP */
    int32_t gate1;             /* 0 when maxEvalue is not finite (phase-1 gate skipped) */
    int32_t cap;               /* SMEM slots per read */
    int32_t* cnt;              /* SMEMs emitted per read (may exceed cap) */
    SmemRec* out;              /* [nReads][cap] */
    void* cand;                /* candidate scratch, nLanes x candCap entries of SMEM_CAND_BYTES */
    int64_t candCap;
    int64_t nLanes;
    const uint8_t* need;       /* per read (may be NULL), bits 0-1: 0 = the read cannot emit an SMEM,
                                * skip it; 2 = its whole SMEM list is exact[r] alone; 1 or 3 =
                                * search; bit 2: start it early */
    const SmemRec* exact;      /* per read, read where need[r] == 2 */
    const int32_t* order;      /* the order in which the queue hands out reads (may be NULL) */
    const double* prefLL;      /* per read at offs[r] + r: L+1 left-to-right loglik prefix sums */
    unsigned long long* queue; /* next queue entry (counted from this launch's first entry) */
    const unsigned long long* qSplit; /* queue entries before *qSplit are the reads marked to
                                       * start early (NULL: no split) */
    int32_t qPart;             /* 0 = the whole queue, 1 = entries before the split, 2 = from it */
    unsigned long long* stats; /* [0] occ-block extensions, [2] text steps, [3] SA/ISA lookups,
                                * [5] lane iterations, [6] wavefront iterations, [7] max of [6] */
    /* paired chains (smem_pair_kernel): the even lane of a pair writes out / cnt / tagA, the odd
     * lane outB / cntB / tagB; tags number the search of each SMEM (0 = the first search,
     * k = the k-th forward or backward search); smem_merge_kernel interleaves them */
    SmemRec* outB;
    int32_t* cntB;
    int32_t* tagA;
    int32_t* tagB;
};

/* Row: int32_t for narrow indexes, int64_t for wide ones (DevIndex::wide) */
template <bool FAST, class Row> __global__ void smem_kernel_t(DevIndex ix, SmemArgs a);
/* the same search with two lanes per read: after the first search from the midpoint, the
 * forward searches run on one lane and the backward searches on the other */
template <bool FAST, class Row> __global__ void smem_pair_kernel(DevIndex ix, SmemArgs a);
/* the paired lists of each read in the reference's order, into out / cnt */
__global__ void smem_merge_kernel(SmemArgs a, SmemRec* out, int32_t* cnt);
/* SMEM_LIST::findMEMS per read (out / cnt, a.minLen, a.T2) and the list's loglik */
template <class Row> __global__ void mem_kernel(DevIndex ix, SmemArgs a, double* listLL);
/* the same search with one read per group of G lanes (same output) */
template <int G, class Row> __global__ void smem_group_kernel(DevIndex ix, SmemArgs a);
/* per-read loglik prefix sums (SmemArgs::prefLL) */
__global__ void smem_prefix_kernel(DevIndex ix, SmemArgs a, double* pref);

/* one seed (SeedPair, src/SeedPair.h:120-128) */
struct SeedRec {
    int64_t from, to, start, end;
    int32_t tid, strand;
    double logP;
};

struct SeedArgs {
    const int64_t* offs;      /* read offsets (for L) */
    int64_t nReads;
    const int32_t* smemCnt;
    const SmemRec* smems;      /* [nReads][cap] */
    int32_t cap;
    int64_t maxNSeed;
    int32_t* order;            /* [nReads][cap] SMEM permutation after the loglik sort */
    int32_t* nUsed;            /* SMEMs whose seeds are taken */
    int64_t* nRaw;             /* seeds before removeRedundant */
    const int64_t* seedOff;    /* exclusive scan of nRaw */
    SeedRec* seeds;
    int32_t* seedCnt;          /* seeds after removeRedundant */
    unsigned long long* stats; /* [1] SA lookups */
};

__global__ void smem_order_kernel(DevIndex ix, SeedArgs a);
__global__ void seed_kernel(DevIndex ix, SeedArgs a);

/* chains: per read, a list of chains, each a run of seed indices */
struct ChainArgs {
    const int64_t* offs;
    int64_t nReads;
    const int64_t* seedOff;
    const int32_t* seedCnt;
    const SeedRec* seeds;
    int32_t* stackNode;        /* per read scratch [seedOff .. + seedCnt) */
    int32_t* stackNext;
    uint8_t* seen;
    int32_t* nChains;          /* pass 1 outputs */
    int64_t* nChainSeeds;
    const int64_t* chainOff;   /* pass 2 inputs: exclusive scans */
    const int64_t* chainSeedOff;
    int32_t* chainStart;       /* [chain] start index into chainSeeds */
    int32_t* chainLen;
    int32_t* chainSeeds;       /* seed index (relative to the read's seeds) */
    int32_t pass;
    const int64_t* maxMM;      /* per read maxMismatch / maxIndel (NULL: ceil(L * .2) / ceil(L * .1)) */
    const int64_t* maxIndel;
};
__global__ void chain_kernel(ChainArgs a);

/* DP over one chain (one alignment job) */
struct AlnRec {
    int32_t rid, tid, strand, hazard;
    int64_t tStart, tEnd, alnFrom, alnTo, alnStart, alnEnd;
    int32_t score;             /* integer-valued alnScore */
    int32_t nCigar, mdLen;
    int32_t pad;
    double log10P;
    int64_t cells;
};

/* one candidate mate pair (AlignmentPE, src/Alignment.h:446-575): the read's and the mate's
 * alignments (job or row indexes), the insert size, and the pair's log10 and natural-log
 * likelihoods (AlignmentPE::log10lik / loglik) */
struct PairRec {
    int32_t f, r, isize, pad;
    double l10, ll;
};

/* DP job classes: no rectangle, wavefront fill (one wave per job), serial fill */
enum : int32_t { CLS_NONE = 0, CLS_WAVE = 1, CLS_SERIAL = 2, CLS_WAVE16 = 3 };

/* per-job geometry in matrix coordinates (row i = query base i-1 + 1, column j =
 * target base tStart + j - 1); R5 = rows 1..r5r1 x cols 1..r5c1, R3 = rows r3r0..L x
 * cols r3c0..tLen (empty when a side is empty) */
struct JobGeo {
    int64_t tStart, tEnd, r3off, r3boff;   /* R3's cell and border offsets (job-relative) */
    int32_t L, tLen, ns, cs;        /* cs: first chain-seed slot of the job */
    int32_t tid, strand, rid, cls;  /* cls: 0 no rectangle, 1 wavefront fill (one job per wavefront,
                                     * 64-row strips), 2 serial fill, 3 wavefront fill of 16 lanes
                                     * (four jobs per wavefront, 16-row strips) */
    int32_t r5r1, r5c1, r3r0, r3c0;
};
/* rectangle summaries left by the fill: the column-major first maximum of M over all
 * rectangle cells (LOCAL), and the in-order scan of M/I over R3's last row (GLOBAL) */
struct JobSum {
    int32_t swV, swI, swJ, nwV, nwC, nwS, pad0, pad1;
};

struct DpArgs {
    const uint8_t* seqs;       /* reads */
    const uint8_t* quals;
    const uint8_t* rcSeqs;     /* reverse complements, same offsets */
    const uint8_t* rcQuals;
    const int64_t* offs;
    const int64_t* seedOff;
    const SeedRec* seeds;
    int64_t nJobs;
    const int32_t* jobRead;    /* read of each job */
    const int32_t* jobChainStart;
    const int32_t* jobChainLen;
    const int32_t* chainSeeds;
    int4* auxSG;               /* per chain seed: bi, bj, len, d0 (first diagonal value) */
    int64_t* auxGap;           /* per chain seed: cell offset of the following gap rect (job-relative) */
    int64_t* auxGapB;          /* per chain seed: border offset of the following gap rect (job-relative) */
    JobGeo* geo;
    JobSum* sum;
    int32_t* jobCls;
    const int64_t* waveJobs;   /* jobs filled by the wavefront kernel, largest first */
    int64_t nWaveJobs;
    const int64_t* wave16Jobs; /* jobs filled 16 lanes each, largest first */
    int64_t nWave16Jobs;
    const int64_t* serialJobs; /* serially filled jobs, largest first (NULL: the trace takes jobs
                                * in job order) */
    int64_t nSerialJobs;
    const int64_t* noneJobs;   /* jobs without a rectangle */
    int64_t nNoneJobs;
    unsigned int* fill16Next;  /* work counter of the 16-lane fill (zero before its launch) */
    const int64_t* tileOff;    /* per job byte offset into the decision pool */
    uint8_t* bits;             /* decisions of wavefront-filled jobs, per strip (64 or 16 rows) of a
                                * rectangle: one word per row and 8 sweep steps (GLOBAL: 32-bit words,
                                * 4 bits per cell; LOCAL: 64-bit words, 8 bits); per cell the backtrace's next state
                                * from M (bits 0-1), I (bit 2: 1 = stay), D (bit 3: 1 = stay) and, in
                                * LOCAL mode, M > 0 (bit 4) */
    const int64_t* borderOff;  /* per job offset into the border pool */
    int4* border;              /* M, I, D of each rectangle's last row, then its last column */
    const int64_t* serialOff;  /* per job offset into the tile pool (serially filled jobs only) */
    int4* tile;                /* M, I, D per rectangle cell of serially filled jobs */
    const int64_t* pathOff;    /* per job: scan of path bounds (L + tLen + 3); the CIGAR slot is at
                                * pathOff + 2*job, the MD slot at 3*pathOff + 8*job */
    uint32_t* cigar;
    char* md;
    AlnRec* out;
    int32_t mode;              /* 0 GLOBAL, 1 LOCAL */
    int32_t match, mismatch, gapO, gapE;   /* integer-valued scores */
    double clip;
    const double* qMatch;      /* log10(1 - phredQ2P(q)) by q */
    const double* qMis;        /* q / PHRED_SCALE by q */
    const double* qGapOpen;    /* q / PHRED_SCALE - gapO by q */
    double gapE_d;
    int32_t forceSerial;       /* every job filled serially (context option dp_serial) */
    unsigned long long* dstat; /* optional counters (NULL: none): serially filled jobs, their cells,
                                * the largest one's cells, full last-row scans; of the jobs filled 16
                                * lanes each: rectangles, their cells, lane-steps of their strips
                                * (16 x (C + rows - 1) per strip), strips */
    int64_t* tileNeed;         /* rectangle cells */
    int64_t* pathNeed;
    int64_t* borderNeed;       /* border cells */
    int64_t* serialNeed;       /* rectangle cells of serially filled jobs */
};
/* seeds, rectangles and sizes of each job */
__global__ void dp_setup_kernel(DevIndex ix, DpArgs a);
/* rectangle fills, one wavefront per listed job (MODE 0 GLOBAL, 1 LOCAL) */
template <int MODE> __global__ void dp_fill_kernel(DevIndex ix, DpArgs a);
/* rectangle fills of small jobs, 16 lanes per job, four jobs per wavefront */
template <int MODE> __global__ void dp_fill16_kernel(DevIndex ix, DpArgs a);
/* remaining fills, backtrace, CIGAR, MD and evaluate, one lane per job */
__global__ void dp_trace_kernel(DevIndex ix, DpArgs a);

#endif
