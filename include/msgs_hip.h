/*
 * msgs_hip.h -- C-ABI of libmsgs_hip.so, the MI355X (gfx950) read-mapping hot path
 * of MSGseqTK.  Plain pointers and sizes only; every call returns 0 or a negative
 * MSGS_E* code and never throws.  A context is used by one host thread at a time.
 *
 * The reference has no FFI: its boundary is the C++ class API in namespace
 * EGriceLab::MSGseqTK.  Each entry point below names the reference interface
 * it replaces (paths relative to the reference source tree).
 */
#ifndef MSGS_HIP_H
#define MSGS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MSGS_OK = 0,
    MSGS_E_ARG = -1,        /* bad argument */
    MSGS_E_HIP = -2,        /* HIP runtime error */
    MSGS_E_NOMEM = -3,      /* device or host allocation failed */
    MSGS_E_STATE = -4,      /* call out of order (e.g. no database) */
    MSGS_E_DB = -5,         /* database refused (not bi-directional / IUPAC-extended, like
                               FMDIndex::buildCounts, src/FMDIndex.cpp:118-134) */
    MSGS_E_CAP = -6,        /* caller buffer too small; *needed reports the size */
    MSGS_E_UNSUPPORTED = -7 /* parameters outside what the device path implements */
};

typedef struct msgs_ctx msgs_ctx;

/* Mapping parameters; defaults are the reference's (src/SMEM.h:204-207,
 * src/SMEM.cpp:17, src/ScoreScheme.cpp:14-18, src/PairingScheme.cpp:14-16,
 * src/msgseqtk-align.cpp:59-62).  mode 0 = GLOBAL (Needleman-Wunsch), 1 = LOCAL. */
typedef struct {
    int64_t minLen; double maxEvalue; int64_t maxNSeed;
    int32_t mode; int32_t maxReport;
    double match, mismatch, gapO, gapE, clip;
    double minScoreRate;
    double insMean, insSD, insMin, insMax;
    int32_t noMixed, noDiscordant, noTailOver, noContain, noOverlap, pad0;
    int64_t maxNPair;
} msgs_params;

void msgs_params_default(msgs_params* p);

/* A batch of reads: nt16 codes (A=1 C=2 G=4 T=8 ...) concatenated, Phred scores
 * (already minus 33, >= 2) in the same layout (may be NULL: all 30), offsets[n+1],
 * and names ('\n'-separated, may be NULL). */
typedef struct {
    const uint8_t* seq;
    const uint8_t* qual;
    const int64_t* offs;
    const char* names;
    int64_t n;
} msgs_reads;

/* ---- context ---- */
int msgs_ctx_create(int device, msgs_ctx** out);
void msgs_ctx_destroy(msgs_ctx* ctx);
const char* msgs_last_error(msgs_ctx* ctx);
/* Context options (no reference counterpart; tuning knobs of the device path):
 * "form" 0 = plain FMD extensions, 1 = unique intervals followed through the text (default),
 *   2 = form 1 plus the learned index in front of the search (needs msgs_lix_train): reads
 *   that occur exactly once with unique halves get their SMEM list from two learned-index
 *   lower bounds, and (with "read_filter" 1, the default) reads in which no SMEM can occur
 *   are skipped;
 * "group_lanes" lanes per read of the form-2 search kernel (0 = one lane per read, 8, 16);
 * "text_path" form 2 with one lane per read: 1 follows unique intervals through the text
 *   (as form 1), 0 extends with occ blocks only (as form 0);
 * "split_lanes" form 2 with one lane per read: 0 = off, or 4 / 8 / 16 lanes per read for the
 *   reads the learned index marks as likely long searches (no 16-mer at the read's middle),
 *   searched by a lane-group launch before the one-lane launch takes the others;
 * "pair_chains" one lane per read (any form): 1 = two lanes per read, the first search and
 *   then the forward searches of SMEM_LIST::findAllSMEMS on one, its backward searches on
 *   the other, the two lists merged in the reference's order (default 0);
 * "lanes_per_cu" persistent lanes of the SMEM kernel per CU; "smem_cap" SMEM slots per
 * read before a regrow; "host_threads" threads of the host finishing step; "streams" how
 * many submitted batches run their device stages at once (1..4, default 4); "ingest" 1 = a
 * submitted run of a resident batch first copies the batch's bases and qualities to the device
 * again from pinned host memory (streaming input; default 0); "dp_serial" 1 = every DP job is
 * filled by the serial path (all of M, I, D kept per cell) that otherwise takes only jobs with
 * more than 64 seeds or a rectangle wider than 512 columns (default 0); "wide_index" (read when a
 * database is built or loaded) 0 = the wide index layout (64-bit SA / ISA rows and
 * bi-intervals, occ-block counts relative to superblocks) only for BD texts of 2^31 symbols or
 * more (default), 1 = for every database, with superblocks of 2^14 positions and suffix-sorting
 * ranges of 2^16 suffixes, so that small texts run the wide code paths; msgs_ctx_get
 * "index_wide" tells which layout the resident database has; "build_block_mb" (read when a
 * database is built or loaded) msgseqtk-build's -b: a database whose BD text has this many
 * MB (2^20 symbols) or more gets the suffix-array order of msgseqtk-build's block-by-block
 * build, whole chromosomes in blocks below that size merged with FMDIndex::prepend
 * (src/msgseqtk-build.cpp:440-467, src/FMDIndex.cpp:267-308, 612-630), so that row numbers
 * and DB.fmd match a database the reference built with the same -b (default 2048, the
 * reference's; 0 = one block whatever the size); msgs_ctx_get "index_blocks" tells how many
 * blocks the resident database was ordered by (row numbers and bi-intervals follow it; the
 * mapping output matches the reference's over a database it built with the same -b).  Results
 * never depend on the other settings, except "record_format": the records msgs_map_batch and msgs_batch_sam
 * return are SAM text lines (0, default) or the same records in the BAM layout (1): for each,
 * what htslib's bam_write1 puts in the BGZF stream (int32 block_size, the 32-byte core, read name,
 * CIGAR, 4-bit SEQ, QUAL and tags as the reference's BAM::setAux stores them), uncompressed. */
int msgs_ctx_set(msgs_ctx* ctx, const char* key, int64_t value);
int msgs_ctx_get(msgs_ctx* ctx, const char* key, int64_t* value);

/* ---- database: MetaGenome + FMDIndex resident in HBM ----
 * Replaces MetaGenome::load/loadSeq + FMDIndex::load (src/MetaGenome.cpp:42-73,
 * src/FMDIndex.cpp:153-161) and, for building, FMDIndex(const DNAseq&, ...)
 * (src/FMDIndex.cpp:29-116) over MetaGenome::getBDSeq (src/MetaGenome.cpp:104-113).
 * The suffix array is built on the device.  names: '\n'-separated chrom names.  BD texts of
 * up to 2^34 - 3 symbols (chrom lengths summing to about 8.5 Gbp); longer ones return
 * MSGS_E_UNSUPPORTED. */
int msgs_db_build(msgs_ctx* ctx, const uint8_t* chromSeqs, const int64_t* chromLens, int32_t nChrom,
                  const char* names);
/* The reference's database files DB.mtg + DB.mgs (MetaGenome::load / loadSeq,
 * src/MetaGenome.cpp:42-73, src/Genome.cpp:29-97, header src/util/ProgEnv.cpp:24-64);
 * the FMD index is rebuilt on the device from the sequences. */
int msgs_db_load(msgs_ctx* ctx, const char* prefix);
/* FMDIndex::load of a DB.fmd file (src/FMDIndex.cpp:153-161; with the program-info header of
 * src/util/ProgEnv.cpp:24-64 when header is 1): the BWT symbol count *n and, into the buffers
 * given (any may be NULL), the BWT decoded from its RRR wavelet tree, the counts B[256] / C[256],
 * gapSA (*nGap values) and the sampled SA (*nSampled values).  Host only (no context); MSGS_E_DB
 * for a malformed file.  msgs_db_load also reads DB.fmd when it exists and checks it against the
 * index it builds from the sequences. */
int msgs_fmd_read(const char* path, int32_t header, int64_t* n, uint8_t* bwt, int64_t* B256, int64_t* C256,
                  int64_t* gapSA, int64_t* nGap, int64_t* saSampled, int64_t* nSampled);
/* FMDIndex(bdSeq, true, 32) + save (src/FMDIndex.cpp:29-116,143-151) for the resident
 * database: the reference's DB.fmd format (RRR wavelet tree of the BWT, gapSA, SA sampled at
 * every 32nd row and every '\0' row), with the program-info header when header is 1. */
int msgs_fmd_write(msgs_ctx* ctx, const char* path, int32_t header);
/* the same file from host arrays: the BWT (n nt16 codes) and suffix array of a BD text */
int msgs_fmd_write_arrays(const char* path, int32_t header, const uint8_t* bwt, const int64_t* sa, int64_t n);
/* MetaGenome::numChroms / getChromName / getChrom(i).size() (src/MetaGenome.h:64-112) */
int32_t msgs_db_nchrom(msgs_ctx* ctx);
int msgs_db_chrom(msgs_ctx* ctx, int32_t tid, char* name, int64_t cap, int64_t* len);
/* FMDIndex::length / getBaseCount / getCumCount (src/FMDIndex.h:152-179) */
int64_t msgs_db_length(msgs_ctx* ctx);
int msgs_db_counts(msgs_ctx* ctx, int64_t* B256, int64_t* C256);
/* Copies of the BD text, BWT and suffix array (any may be NULL), for parity checks. */
int msgs_db_copy(msgs_ctx* ctx, uint8_t* bd, uint8_t* bwt, int64_t* sa);
/* FMDIndex::loglik (src/FMDIndex.h:195-197) */
double msgs_db_loglik(msgs_ctx* ctx, int32_t base);

/* ---- FMD primitives, batched (device) ----
 * rank(s, i) of WaveletTreeRRR (src/libsds/src/WaveletTreeRRR.cpp:101-125): #s in BWT[0..i]. */
int msgs_rank_batch(msgs_ctx* ctx, const uint8_t* sym, const int64_t* pos, int64_t* out, int64_t n);
/* FMDIndex::backExt / fwdExt (src/FMDIndex.cpp:577-610, src/FMDIndex.h:226-228):
 * pqs is n rows of (p, q, s), updated in place; ret[i] = new s. */
int msgs_ext_batch(msgs_ctx* ctx, int64_t* pqs, const uint8_t* bases, int32_t forward, int64_t* ret, int64_t n);

/* Pattern searches, one lane per pattern (patterns as reads: nt16 codes + offsets[n+1]);
 * pqs[3i..3i+2] = the bi-interval (p, q, s) after the search.  mode 0 = FMDIndex::count's
 * backward search from the last base (src/FMDIndex.cpp:163-176; count = max(s, 0)), mode 1 =
 * FMDIndex::locateAllFwd's forward search from the first base (src/FMDIndex.cpp:529-551; the
 * locations are SA[p .. p+s)), mode 2 = locateAllRev's backward search (src/FMDIndex.cpp:553-575;
 * SA[q .. q+s)); modes 1 and 2 give s = 0 for an empty pattern or an ambiguous start base. */
int msgs_search_batch(msgs_ctx* ctx, const uint8_t* seqs, const int64_t* offs, int64_t n, int32_t mode, int64_t* pqs);
/* FMDIndex::accessSA (src/FMDIndex.cpp:520-527) of n suffix-array rows (0 <= row < length). */
int msgs_sa_batch(msgs_ctx* ctx, const int64_t* rows, int64_t n, int64_t* out);
/* MetaGenome::getSeq(tid) (src/MetaGenome.h:319-321): the chromosome's nt16 codes (*len bases,
 * without the terminator); out may be NULL to ask the length. */
int msgs_db_chrom_seq(msgs_ctx* ctx, int32_t tid, uint8_t* out, int64_t cap, int64_t* len);

/* ---- seeding ----
 * SMEM_LIST::findAllSMEMS (src/SMEM.cpp:186-215), one read per lane.  Per read i,
 * counts[i] SMEMs in emission order; rows5 = (from, to, fwdStart, revStart, size).
 * Returns MSGS_E_CAP with *needed set when cap rows are not enough.  On a database with the
 * wide index layout (index_wide 1), a read of 2^26 - 2 bases or more makes this and every
 * entry that searches it (map, align, batch runs) return MSGS_E_UNSUPPORTED. */
int msgs_smem_batch(msgs_ctx* ctx, const msgs_reads* reads, const msgs_params* p, int64_t* counts,
                    int64_t* rows5, double* logP, int64_t cap, int64_t* needed);
/* SMEM_LIST::findMEMS (src/SMEM.cpp:174-184, SMEM::findMEM :42-62) with p->minLen and
 * p->maxEvalue, one read per lane; rows5 / logP / counts / cap / needed as msgs_smem_batch,
 * readLoglik[i] = SMEM_LIST::loglik (src/SMEM.cpp:307-312) of read i's list (may be NULL). */
int msgs_mems_batch(msgs_ctx* ctx, const msgs_reads* reads, const msgs_params* p, int64_t* counts,
                    int64_t* rows5, double* logP, double* readLoglik, int64_t cap, int64_t* needed);
/* msgseqtk-clean's per-read decision (src/msgseqtk-clean.cpp:413-488): the loglik of the
 * read's MEM list in the reference database `ref` and in the background database `bg`
 * (SMEM_LIST::findMEMS / findMEMS_PE and SMEM_LIST::loglik, src/SMEM.h:373-394; with mates,
 * the first read's list plus the mate's), and keep[i] = (-refLoglik[i] + bgLoglik[i] >
 * minLod).  mates may be NULL (single-end); any output may be NULL.  A failure inside one
 * database's search is reported by msgs_last_error of that database's context. */
int msgs_clean_batch(msgs_ctx* ref, msgs_ctx* bg, const msgs_reads* reads, const msgs_reads* mates,
                     const msgs_params* p, double minLod, double* refLoglik, double* bgLoglik, uint8_t* keep);
/* SMEM_LIST::findSeeds (src/SMEM.cpp:271-305) incl. SeedPair::removeRedundant
 * (src/SeedPair.cpp:71-88).  rows6 = (from, to, start, end, tid, strand). */
int msgs_seed_batch(msgs_ctx* ctx, const msgs_reads* reads, const msgs_params* p, int64_t* counts,
                    int64_t* rows6, double* logP, int64_t cap, int64_t* needed);

/* One SMEM in device memory (the layout msgs_smem_batch_noalloc writes): read interval
 * [from, to), bi-interval (fwdStart, revStart, size), loglik. */
typedef struct {
    int32_t from, to;
    int64_t fwdStart, revStart, size;
    double logP;
} msgs_smem_rec;

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the SMEM
 * search of msgs_smem_batch (SMEM_LIST::findAllSMEMS per read) over n reads already in device
 * memory (dSeq nt16 codes, dOffs[n+1] offsets into dSeq), written to caller device memory:
 * dCounts[i] SMEMs of read i (it may exceed cap: only the first cap are stored) at
 * dRecs[i * cap ...].  It allocates nothing: its scratch (work queue, loglik prefix sums,
 * candidate lists) is a fixed 128 MiB device array per GPU, held until the process ends;
 * inputs whose prefix sums (8 bytes per base + 8 per read) or candidate lists (32 bytes per
 * base of the longest read + 64, times 256 lanes at least) do not fit return MSGS_E_CAP.
 * *maxCount (may be NULL) is the largest count; MSGS_E_CAP when it exceeds cap.  Calls that
 * overlap on one GPU share the scratch and race.  Runs on the context's own stream and waits
 * for it. */
int msgs_smem_batch_noalloc(msgs_ctx* ctx, const uint8_t* dSeq, const int64_t* dOffs, int64_t n,
                            const msgs_params* p, int32_t cap, int32_t* dCounts, msgs_smem_rec* dRecs,
                            int32_t* maxCount);

/* ---- whole path ----
 * msgseqtk-align main_SE / main_PE (src/msgseqtk-align.cpp:595-758): seeds, chains
 * (SeedChain::getChains, src/SeedChain.cpp:40-73), Alignment::buildAlignments
 * (src/Alignment.cpp:420-427), filter/evaluate/calcMapQ/sort and record output.
 * mates == NULL selects single-end.  SAM text of read (pair) i is
 * out[lineOffs[i] .. lineOffs[i+1]).  *needed reports the bytes required. */
int msgs_map_batch(msgs_ctx* ctx, const msgs_reads* reads, const msgs_reads* mates, const msgs_params* p,
                   char* out, int64_t cap, int64_t* lineOffs, int64_t* needed);

/* Per-chain alignments (for parity checks of the DP): per read counts[i] rows of
 * (tid, strand, tStart, tEnd, alnFrom, alnTo, alnStart, alnEnd, hazard, cells), score,
 * and CIGAR strings joined by '\n' into cigars (cigCap bytes). */
int msgs_align_batch(msgs_ctx* ctx, const msgs_reads* reads, const msgs_params* p, int64_t* counts,
                     int64_t* rows10, double* scores, char* cigars, int64_t cigCap, int64_t cap,
                     int64_t* needed);

/* SeedChain::getChains (src/SeedChain.cpp:40-73) over caller seed lists: set i is
 * rows6[seedOffs[i] .. seedOffs[i+1]) of (from, to, start, end, tid, strand), in SeedPair order
 * (as SMEM_LIST::findSeeds returns them), chained with maxMismatch[i] / maxIndel[i].  Output:
 * chainCounts[i] chains of set i; out holds, set after set and chain after chain, the chain's
 * seed count followed by its seeds' indexes into the set; *needed = entries of out. */
int msgs_chain_batch(msgs_ctx* ctx, const int64_t* rows6, const int64_t* seedOffs, int64_t nSets,
                     const int64_t* maxMismatch, const int64_t* maxIndel, int64_t* chainCounts, int64_t* out,
                     int64_t cap, int64_t* needed);
/* Alignment::buildAlignments + evaluate (src/Alignment.cpp:420-427,446-489) over caller chains:
 * read i's chains are chainOffs[i] .. chainOffs[i+1]; chain k's seeds are
 * rows6[chainSeedOffs[k] .. chainSeedOffs[k+1]) (from, to, start, end, tid, strand; compatible and
 * in order, inside the read and the chromosome; others return MSGS_E_ARG).  Per chain: rows10 as
 * msgs_align_batch, the score, log10P (Alignment::evaluate), and the CIGAR and MD strings each
 * ended by '\n' in cigars / mds (any output may be NULL); *needed = chains. */
int msgs_dp_batch(msgs_ctx* ctx, const msgs_reads* reads, const int64_t* chainOffs, const int64_t* chainSeedOffs,
                  const int64_t* rows6, const msgs_params* p, int64_t* rows10, double* scores, double* log10P,
                  char* cigars, int64_t cigCap, char* mds, int64_t mdCap, int64_t* needed);
/* AlignmentPE::getPairs + sort + filter (src/Alignment.cpp:556-581, src/Alignment.h:635-639) on
 * the device over caller alignment lists, as the PE driver calls them (src/msgseqtk-align.cpp:
 * 712-717): for read pair i (i < n) the read's alignments are rows alnOffs[i] .. alnOffs[i+1] and
 * its mate's rows alnOffs[n+i] .. alnOffs[n+i+1] of rows4 (tid, strand 1 = FWD / 2 = REV,
 * alnStart, alnEnd) and log10P (Alignment::log10lik), the lists as Alignment::filter left them
 * (alnOffs[0] = 0, 2n+1 entries).  Uses p->insMean, insSD, insMin, insMax (PairingScheme's
 * fields after updateParam / updateRange), noDiscordant, noTailOver, noContain, noOverlap,
 * maxNPair.  Output: pairCounts[i] pairs of read pair i, and read pair after read pair, each
 * pair's rows3 (row of the read's alignment, row of the mate's, insert size) and lik2
 * (AlignmentPE::log10lik, loglik) in the reference's order; either may be NULL.  *needed =
 * pairs; MSGS_E_CAP when more than cap (pairCounts is filled). */
int msgs_pair_batch(msgs_ctx* ctx, int64_t n, const int64_t* alnOffs, const int64_t* rows4, const double* log10P,
                    const msgs_params* p, int64_t* pairCounts, int64_t* pairRows3, double* pairLik2, int64_t cap,
                    int64_t* needed);

/* ---- resident batches (bench): reads copied to HBM once, mapped many times ----
 * slot 0..15.  msgs_batch_run maps the resident batch and keeps the records on the
 * host side of the context; msgs_batch_stats reports counters of the last run:
 * stats[0] reads, [1] mapped reads, [2] extension calls, [3] SA lookups, [4] DP cells,
 * [5] seeding kernel microseconds (HIP events), [6] DP kernel microseconds,
 * [7] whole-run microseconds, [8] seed-stage microseconds, [9] host finishing microseconds,
 * [10] unique-interval text steps, [11] SA/ISA lookups of the SMEM kernel,
 * [12] SMEM-kernel lane iterations, [13] its wavefront iterations (busiest lane of each
 * wavefront, summed), [14] the largest wavefront iteration count, [15] learned-index read
 * filter microseconds (form 2), [16] microseconds until the GPU stages were done,
 * [17]..[20] host wall microseconds of the SMEM, seed, chain and DP stages, [21] DP jobs,
 * [22] wavefront-fill DP jobs, [23]..[25] microseconds into the DP stage when its pool
 * sizes, its records and its CIGAR/MD bytes were back on the host, [26] reads whose SMEM
 * list came from the learned index's exact-read path (form 2), [27] reads the form-2
 * filter skipped, [28]..[31] host finishing thread-microseconds (summed over threads) in
 * SAM text, PE pairing, MAPQ and alignment lists, [32] DP jobs filled serially by one
 * lane, [33] their rectangle cells, [34] the largest one's cells, [35] DP jobs whose
 * GLOBAL last-row scan visits every column (n <= 48), [36] bytes copied host-to-device by the
 * run's ingest (context option "ingest" 1: the batch's bases and qualities from pinned host
 * memory at the start of every run, then its reverse complements), [37] their microseconds,
 * [38]..[41] of the DP jobs filled 16 lanes each: rectangles, their cells, the lane-steps of
 * their strips (16 x (columns + rows - 1) per strip of at most 16 rows) and the strips. */
int msgs_batch_upload(msgs_ctx* ctx, int32_t slot, const msgs_reads* reads, const msgs_reads* mates);
int msgs_batch_run(msgs_ctx* ctx, int32_t slot, const msgs_params* p, int32_t emitSam);
/* msgs_batch_run in two halves: submit returns at once; a worker thread of the slot runs
 * the device stages (on one of the context's worker streams, see "streams") and then the
 * host finishing (records, MAPQ, SAM text), so the caller can submit further batches
 * meanwhile; wait joins it and makes its counters and SAM the context's last run
 * (msgs_batch_stats / msgs_batch_sam). */
int msgs_batch_submit(msgs_ctx* ctx, int32_t slot, const msgs_params* p, int32_t emitSam);
int msgs_batch_wait(msgs_ctx* ctx, int32_t slot);
int msgs_batch_stats(msgs_ctx* ctx, double* stats, int32_t n);
int msgs_batch_sam(msgs_ctx* ctx, char* out, int64_t cap, int64_t* needed);
/* autotuner (no reference counterpart): times the SMEM search of resident batch `slot`
 * with each lanes-per-CU value in cand[0..nCand) (NULL: 128..1024), in form 2 for each
 * search kernel too ("group_lanes" 8 / 16, or one lane per read with "text_path" 1 / 0)
 * with and without the "read_filter", and one lane per read with "split_lanes" 4 / 8 / 16,
 * with "pair_chains" 0 / 1 in every form, keeps the fastest in the context and reports the
 * lanes per CU; results never depend on the choice. */
int msgs_tune(msgs_ctx* ctx, int32_t slot, const msgs_params* p, const int32_t* cand, int32_t nCand, int32_t* best,
              double* bestMicros);

/* ---- learned index (no reference counterpart) ----
 * A model of the SA rank of 16-mers: piecewise-linear root over `segments` rank
 * quantiles + a 2-layer MLP (16 -> 64 -> 1), trained on the device (fp32, Adam) on
 * `samples` (key, rank) pairs drawn from the index; inference in bf16 MFMA; exact
 * lower bounds by a search bounded by the measured max error with a full-search
 * fallback.  Answers FMDIndex::count for 16-mers (src/FMDIndex.cpp:163-176) and, with
 * the context option "form" = 2, lets the SMEM stage skip reads that have no 16-mer
 * in the index (exact when minLen >= 16: every SMEM that long contains one).
 * info: [0] max error (MLP), [1] max error (root alone), [2] first loss, [3] last loss,
 * [4] training milliseconds. */
int msgs_lix_train(msgs_ctx* ctx, int32_t segments, int64_t samples, int32_t steps, int32_t batch, double lr,
                   uint64_t seed, double* info);
/* exact lower bounds lb[i] of keys (16 symbols, 3 bits each: '\0'=1 A=2 C=3 G=4 T=5 N=6,
 * first symbol most significant); useMlp 1 = model + bound, 0 = root alone, -1 = plain
 * binary search; probes = key-table reads, fell = 1 where the bound missed. */
int msgs_lix_query(msgs_ctx* ctx, const uint64_t* keys, int64_t n, int32_t useMlp, int64_t* lb, int32_t* probes,
                   int32_t* fell);
/* raw predictions: mode 0 bf16 MFMA, mode 1 the same arithmetic on VALU */
int msgs_lix_predict(msgs_ctx* ctx, const uint64_t* keys, int64_t n, int32_t mode, double* out);
/* gradient of the mean squared error at weights W (1153 floats: W1[16][64], b1, W2, b2)
 * over B rows of features X (B x 16) and targets T, by the training kernels */
int msgs_lix_grad(msgs_ctx* ctx, const float* W, const float* X, const float* T, int32_t B, float* grad,
                  float* loss);
/* read (get) and/or replace (set) the 1153 MLP weights */
int msgs_lix_weights(msgs_ctx* ctx, float* get, const float* set);
/* knot table: S+1 keys and lower bounds */
int msgs_lix_knots(msgs_ctx* ctx, uint64_t* keys, int64_t* ranks, int32_t* S);

/* counter calibration (no reference counterpart): one device copy of `bytes` (multiple
 * of 16) by a plain kernel, timed; hardware byte counters are scaled against it */
int msgs_util_calib_copy(msgs_ctx* ctx, int64_t bytes, double* micros);
/* random-access ceiling (no reference counterpart): nLoads reads of one 64-byte line each
 * at hashed positions of a device buffer of `bytes` (multiple of 64); *micros = kernel time */
int msgs_util_random_gather(msgs_ctx* ctx, int64_t bytes, int64_t nLoads, double* micros);
/* test hook of the SAM writer (no reference counterpart): the "%g" text of each float as
 * printed in XH:f / XP:f tags, NUL-terminated in out[16 * i .. 16 * i + 16) */
void msgs_util_format_g6(const float* v, int64_t n, char* out);

/* ---- utilities ----
 * The permutation std::sort (libstdc++) gives for n keys compared with '<' -- the
 * reference's unstable sorts (src/SMEM.cpp:275-276, src/SeedPair.cpp:75,
 * src/Alignment.h:629-633) as the library restates them for the device; runs on
 * the host.  perm must hold 0..n-1 on entry. */
void msgs_util_sort_by_key(const double* keys, int32_t* perm, int64_t n);
/* Alignment::calcMapQ (src/Alignment.cpp:491-512) and, pe = 1, AlignmentPE::calcMapQ (:514-539)
 * over n log-likelihoods (natural log): posterior probabilities and MAPQ; runs on the host. */
void msgs_util_mapq(const double* loglik, int64_t n, int32_t pe, double* postP, uint8_t* mapQ);
/* PairingScheme::pr (src/PairingScheme.cpp:18-25) of the insert model p->insMean, insSD, insMin,
 * insMax at x: boost's normal pdf inside [insMin, insMax], 0 outside, 1 when insMean is 0 -- the
 * values the device pairing's likelihood terms are made from; runs on the host. */
double msgs_util_insert_pr(const msgs_params* p, double x);

#ifdef __cplusplus
}
#endif
#endif /* MSGS_HIP_H */
