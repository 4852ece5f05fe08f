/*
 * orc_capi.cpp -- extern "C" face of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 * Loaded through ctypes by tests/ and by bench.py's cpu_baseline leg.
 */
#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "msgs_oracle.hpp"

using namespace orc;

extern "C" {

/* Parameter block; field order and types match msgs_params in include/msgs_hip.h. */
typedef struct {
    int64_t minLen; double maxEvalue; int64_t maxNSeed;
    int32_t mode; int32_t maxReport;
    double match, mismatch, gapO, gapE, clip;
    double minScoreRate;
    double insMean, insSD, insMin, insMax;
    int32_t noMixed, noDiscordant, noTailOver, noContain, noOverlap, pad0;
    int64_t maxNPair;
} orc_params;

static Params toParams(const orc_params* p) {
    Params P;
    P.minLen = p->minLen; P.maxEvalue = p->maxEvalue; P.maxNSeed = p->maxNSeed;
    P.mode = p->mode; P.maxReport = p->maxReport;
    P.match = p->match; P.mismatch = p->mismatch; P.gapO = p->gapO; P.gapE = p->gapE; P.clip = p->clip;
    P.minScoreRate = p->minScoreRate;
    P.insMean = p->insMean; P.insSD = p->insSD; P.insMin = p->insMin; P.insMax = p->insMax;
    P.noMixed = p->noMixed; P.noDiscordant = p->noDiscordant; P.noTailOver = p->noTailOver;
    P.noContain = p->noContain; P.noOverlap = p->noOverlap; P.maxNPair = p->maxNPair;
    return P;
}

static std::vector<std::string> splitNames(const char* names, int64_t n) {
    std::vector<std::string> v;
    const char* p = names;
    for (int64_t i = 0; i < n; ++i) {
        const char* e = p;
        while (*e && *e != '\n') ++e;
        v.emplace_back(p, e);
        p = *e ? e + 1 : e;
    }
    return v;
}

void* orc_db_create(const uint8_t* seqs, const int64_t* lens, int32_t nChrom, const char* names,
                    const int64_t* saOrNull) {
    try {
        Db* db = new Db();
        buildMeta(*db, seqs, lens, nChrom, splitNames(names, nChrom));
        buildIndex(*db, saOrNull);
        return db;
    } catch (...) {
        return nullptr;
    }
}
void orc_db_free(void* h) { delete (Db*)h; }

/* msgseqtk-build's blocks for chromosomes of these lengths: BD ends into ends (at most cap),
 * returns their number (0: one step) */
int64_t orc_block_ends(const int64_t* lens, int32_t nChrom, int64_t blockSize, int64_t* ends, int64_t cap) {
    std::vector<int64_t> bdLen(nChrom);
    for (int32_t t = 0; t < nChrom; ++t) bdLen[t] = 2 * (lens[t] + 1);
    std::vector<int64_t> e = buildBlockEnds(bdLen, blockSize);
    for (size_t i = 0; i < e.size() && (int64_t)i < cap; ++i) ends[i] = e[i];
    return (int64_t)e.size();
}

/* suffix array of a BD text in the block-by-block build order (k = 0: one block) */
void orc_suffix_array(const uint8_t* bd, int64_t n, const int64_t* ends, int64_t k, int64_t* sa) {
    std::vector<int64_t> e(ends, ends + k);
    std::vector<int64_t> s = suffixArrayBlocks(bd, n, e);
    std::memcpy(sa, s.data(), n * sizeof(int64_t));
}
int64_t orc_db_len(void* h) { return ((Db*)h)->n; }
void orc_db_counts(void* h, int64_t* B, int64_t* C) {
    Db* db = (Db*)h;
    std::memcpy(B, db->B, sizeof(db->B));
    std::memcpy(C, db->Cc, sizeof(db->Cc));
}
void orc_db_copy(void* h, uint8_t* bd, uint8_t* bwt, int64_t* sa) {
    Db* db = (Db*)h;
    if (bd) std::memcpy(bd, db->bd.data(), db->n);
    if (bwt) std::memcpy(bwt, db->bwt.data(), db->n);
    if (sa) std::memcpy(sa, db->sa.data(), db->n * sizeof(int64_t));
}
int64_t orc_rank(void* h, int32_t s, int64_t i) { return rank(*(Db*)h, (nt16_t)s, i); }
int64_t orc_backext(void* h, int64_t* pqs, int32_t b) { return backExt(*(Db*)h, pqs[0], pqs[1], pqs[2], (nt16_t)b); }
int64_t orc_fwdext(void* h, int64_t* pqs, int32_t b) { return fwdExt(*(Db*)h, pqs[0], pqs[1], pqs[2], (nt16_t)b); }
double orc_loglik(void* h, int32_t b) { return loglik(*(Db*)h, (nt16_t)b); }
int64_t orc_tid(void* h, int64_t pos) { return getTid(*(Db*)h, pos); }

static std::vector<Read> makeReads(const uint8_t* seq, const uint8_t* qual, const int64_t* offs,
                                   const char* names, int64_t n) {
    std::vector<std::string> nm = names ? splitNames(names, n) : std::vector<std::string>(n, "r");
    std::vector<Read> v;
    v.reserve(n);
    std::vector<uint8_t> q30;
    for (int64_t i = 0; i < n; ++i) {
        int64_t L = offs[i + 1] - offs[i];
        if (!qual) { q30.assign(L, 30); }
        v.push_back(makeRead(nm[i], seq + offs[i], qual ? qual + offs[i] : q30.data(), L));
    }
    return v;
}

/* SMEM lists of each read; rows of out5 are (from, to, fwdStart, revStart, size). */
int64_t orc_smems(void* h, const uint8_t* seq, const int64_t* offs, int64_t n, int64_t minLen, double maxE,
                  int32_t mems, int64_t* counts, int64_t* out5, double* logP, int64_t cap) {
    Db* db = (Db*)h;
    std::vector<Read> reads = makeReads(seq, nullptr, offs, nullptr, n);
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        std::vector<Smem> v = mems ? findMEMS(*db, reads[i], minLen, maxE) : findAllSMEMS(*db, reads[i], minLen, maxE);
        counts[i] = (int64_t)v.size();
        for (const Smem& m : v) {
            if (k < cap) {
                int64_t* o = out5 + 5 * k;
                o[0] = m.from; o[1] = m.to; o[2] = m.fwdStart; o[3] = m.revStart; o[4] = m.size;
                logP[k] = m.logP;
            }
            ++k;
        }
    }
    return k;
}

/* seeds of each read (after removeRedundant); rows (from, to, start, end, tid, strand) */
int64_t orc_seeds(void* h, const uint8_t* seq, const int64_t* offs, int64_t n, int64_t minLen, double maxE,
                  int64_t maxNSeed, int64_t* counts, int64_t* out6, double* logP, int64_t cap) {
    Db* db = (Db*)h;
    std::vector<Read> reads = makeReads(seq, nullptr, offs, nullptr, n);
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        std::vector<Seed> v = findSeeds(*db, reads[i], minLen, maxE, maxNSeed);
        counts[i] = (int64_t)v.size();
        for (const Seed& s : v) {
            if (k < cap) {
                int64_t* o = out6 + 6 * k;
                o[0] = s.from; o[1] = s.to; o[2] = s.start; o[3] = s.end; o[4] = s.tid; o[5] = s.strand;
                logP[k] = s.logP;
            }
            ++k;
        }
    }
    return k;
}

/* removeRedundant / getChains on caller-supplied seeds (rows as in orc_seeds). */
int64_t orc_remove_redundant(int64_t* rows6, double* logP, int64_t n) {
    std::vector<Seed> v(n);
    for (int64_t i = 0; i < n; ++i) {
        int64_t* r = rows6 + 6 * i;
        v[i] = Seed{r[0], r[1], r[2], r[3], r[4], (int32_t)r[5], logP[i]};
    }
    removeRedundant(v);
    for (size_t i = 0; i < v.size(); ++i) {
        int64_t* r = rows6 + 6 * i;
        r[0] = v[i].from; r[1] = v[i].to; r[2] = v[i].start; r[3] = v[i].end; r[4] = v[i].tid; r[5] = v[i].strand;
        logP[i] = v[i].logP;
    }
    return (int64_t)v.size();
}
/* chains: out holds, per chain, the count followed by the seed indices into the input rows */
int64_t orc_get_chains(const int64_t* rows6, int64_t n, int64_t maxMM, int64_t maxIndel, int64_t* out, int64_t cap) {
    std::vector<Seed> v(n);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t* r = rows6 + 6 * i;
        v[i] = Seed{r[0], r[1], r[2], r[3], i /* index carried in tid */, (int32_t)r[5], 0};
    }
    /* chaining compares tids: carry the real tid separately */
    std::vector<Seed> w(n);
    for (int64_t i = 0; i < n; ++i) w[i] = Seed{v[i].from, v[i].to, v[i].start, v[i].end, rows6[6 * i + 4], v[i].strand, (double)i};
    auto chains = getChains(w, maxMM, maxIndel);
    int64_t k = 0;
    for (auto& ch : chains) {
        if (k < cap) out[k] = (int64_t)ch.size();
        ++k;
        for (auto& s : ch) { if (k < cap) out[k] = (int64_t)s.logP; ++k; }
    }
    return k;
}

/* Full-rectangle alignment (Alignment_test form).  res: alnFrom, alnTo, alnStart, alnEnd, hazard; score. */
int64_t orc_align_full(const orc_params* p, const uint8_t* q, int64_t L, const uint8_t* t, int64_t tLen,
                       int32_t mode, int64_t* res, double* scoreOut, char* cigar, char* md, int64_t strCap) {
    Params P = toParams(p);
    std::vector<uint8_t> qual(L, 30);
    Read r = makeRead("q", q, qual.data(), L);
    Read rc = revcomRead(r);
    std::vector<nt16_t> tg(t, t + tLen);
    tg.push_back(0);
    Aln a = alignFull(P, r, rc, tg, mode);
    res[0] = a.alnFrom; res[1] = a.alnTo; res[2] = a.alnStart; res[3] = a.alnEnd; res[4] = a.hazard;
    *scoreOut = a.alnScore;
    std::string c = cigarString(a), m = mdTag(a);
    std::snprintf(cigar, strCap, "%s", c.c_str());
    std::snprintf(md, strCap, "%s", m.c_str());
    return a.dpCells;
}

/* Per-chain alignments of each read: rows (tid, strand, tStart, tEnd, alnFrom, alnTo, alnStart, alnEnd,
 * hazard, cells), score, and '\n'-joined CIGAR strings. */
int64_t orc_alignments(void* h, const orc_params* p, const uint8_t* seq, const uint8_t* qual, const int64_t* offs,
                       int64_t n, int64_t* counts, int64_t* rows10, double* scores, char* cigars, int64_t cigCap,
                       int64_t cap) {
    Db* db = (Db*)h;
    Params P = toParams(p);
    std::vector<Read> reads = makeReads(seq, qual, offs, nullptr, n);
    int64_t k = 0;
    std::string cig;
    for (int64_t i = 0; i < n; ++i) {
        const Read& r = reads[i];
        Read rc = revcomRead(r);
        std::vector<Seed> seeds = findSeeds(*db, r, P.minLen, P.maxEvalue, P.maxNSeed);
        auto chains = getChains(seeds, (int64_t)std::ceil(r.L * 0.2), (int64_t)std::ceil(r.L * 0.1));
        counts[i] = (int64_t)chains.size();
        for (auto& ch : chains) {
            Aln a = alignChain(*db, P, r, rc, ch);
            if (k < cap) {
                int64_t* o = rows10 + 10 * k;
                o[0] = a.tid; o[1] = a.qStrand; o[2] = a.tStart; o[3] = a.tEnd; o[4] = a.alnFrom; o[5] = a.alnTo;
                o[6] = a.alnStart; o[7] = a.alnEnd; o[8] = a.hazard; o[9] = a.dpCells;
                scores[k] = a.alnScore;
            }
            cig += cigarString(a);
            cig += '\n';
            ++k;
        }
    }
    if (cigars) std::snprintf(cigars, cigCap, "%s", cig.c_str());
    return k;
}

/* Whole-path SAM text, one read (SE) or one pair (PE) per task, OpenMP over tasks when nThreads > 1.
 * offs2 != NULL selects paired-end mode (mates are reads i of batch 1 and 2).  Output is the
 * concatenation in input order; lineOffs[i] gives where read/pair i starts.  Returns bytes
 * needed, or -1 if cap is too small. */
int64_t orc_map(void* h, const orc_params* p, const uint8_t* seq, const uint8_t* qual, const int64_t* offs,
                const char* names, const uint8_t* seq2, const uint8_t* qual2, const int64_t* offs2,
                const char* names2, int64_t n, int32_t nThreads, char* out, int64_t cap, int64_t* lineOffs,
                double* seconds) {
    return mapBatch(*(Db*)h, toParams(p), seq, qual, offs, names, seq2, qual2, offs2, names2, n, nThreads, out, cap,
                    lineOffs, seconds, nullptr, nullptr);
}

/* the caller's parameter block as oracle parameters (same layout as msgs_params) */
void orc_params_convert(const orc_params* p, void* out) { *(Params*)out = toParams(p); }

}  // extern "C"

namespace orc {
int64_t mapBatch(const Db& dbr, const Params& P, const uint8_t* seq, const uint8_t* qual, const int64_t* offs,
                 const char* names, const uint8_t* seq2, const uint8_t* qual2, const int64_t* offs2,
                 const char* names2, int64_t n, int32_t nThreads, char* out, int64_t cap, int64_t* lineOffs,
                 double* seconds, const SeedFn* seedFn, const ChainFn* chainFn) {
    const Db* db = &dbr;
    std::vector<Read> r1 = makeReads(seq, qual, offs, names, n);
    std::vector<Read> r2;
    if (offs2) r2 = makeReads(seq2, qual2, offs2, names2, n);
    std::vector<std::string> res(n);
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nThreads > 0 ? nThreads : 1)
    for (int64_t i = 0; i < n; ++i) {
        if (seedFn)
            res[i] = offs2 ? mapPE(*db, P, r1[i], r2[i], *seedFn, *chainFn) : mapSE(*db, P, r1[i], *seedFn, *chainFn);
        else
            res[i] = offs2 ? mapPE(*db, P, r1[i], r2[i]) : mapSE(*db, P, r1[i]);
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    int64_t total = 0;
    for (auto& s : res) total += (int64_t)s.size();
    if (!out) return total;
    if (total > cap) return -1;
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (lineOffs) lineOffs[i] = k;
        std::memcpy(out + k, res[i].data(), res[i].size());
        k += (int64_t)res[i].size();
    }
    if (lineOffs) lineOffs[n] = k;
    return total;
}
}  // namespace orc

extern "C" {

int32_t orc_max_threads() { return omp_get_max_threads(); }

/* The oracle's alphabet tables in the layout of ref_alphabet (oracle/ref_harness.cpp). */
void orc_alphabet(uint8_t* enc256, char* dec16, uint8_t* comp16, uint8_t* basic16, int32_t* int16, uint8_t* amb16,
                  uint8_t* qenc256) {
    for (int c = 0; c < 256; ++c) { enc256[c] = encodeSym((char)c); qenc256[c] = qualEncode((char)c); }
    for (int b = 0; b < 16; ++b) {
        dec16[b] = decodeBase((nt16_t)b);
        comp16[b] = complement((nt16_t)b);
        basic16[b] = toBasic((nt16_t)b);
        int16[b] = toInt((nt16_t)b);
        amb16[b] = isAmbiguous((nt16_t)b) ? 1 : 0;
    }
}

/* revcomRead of one read: codes and scores */
void orc_revcom_read(const uint8_t* seq, const uint8_t* qual, int64_t L, uint8_t* outSeq, uint8_t* outQual) {
    Read r = makeRead("r", seq, qual, L);
    Read rc = revcomRead(r);
    std::memcpy(outSeq, rc.seq.data(), L);
    std::memcpy(outQual, rc.qual.data(), L);
}

/* Seeds of every SMEM of each read, before the loglik sort, the maxNSeed cut and
 * removeRedundant: per SMEM its size and the number of seeds getSeeds made (the reference
 * asserts they are equal, src/SMEM.cpp:240,267), and the seed rows (from, to, start, end,
 * tid, strand).  Returns the number of seed rows (rows beyond cap are counted, not written). */
int64_t orc_raw_seeds(void* h, const uint8_t* seq, const int64_t* offs, int64_t n, int64_t minLen, double maxE,
                      int64_t* smemCount, int64_t* sizeAndSeeds, int64_t smemCap, int64_t* out6, int64_t cap) {
    Db* db = (Db*)h;
    std::vector<Read> reads = makeReads(seq, nullptr, offs, nullptr, n);
    int64_t k = 0, m = 0;
    for (int64_t i = 0; i < n; ++i) {
        std::vector<Smem> v = findAllSMEMS(*db, reads[i], minLen, maxE);
        smemCount[i] = (int64_t)v.size();
        for (const Smem& s : v) {
            std::vector<Seed> seeds;
            getSeeds(*db, reads[i], s, seeds);
            if (m < smemCap) { sizeAndSeeds[2 * m] = s.size; sizeAndSeeds[2 * m + 1] = (int64_t)seeds.size(); }
            ++m;
            for (const Seed& x : seeds) {
                if (k < cap) {
                    int64_t* o = out6 + 6 * k;
                    o[0] = x.from; o[1] = x.to; o[2] = x.start; o[3] = x.end; o[4] = x.tid; o[5] = x.strand;
                }
                ++k;
            }
        }
    }
    return k;
}

/* One chain of caller-given seeds (rows from, to, start, end, tid, strand) aligned against
 * chromosome tid of the database: res10 as in orc_alignments; the score, Alignment::evaluate's
 * log10P, the CIGAR and the MD tag. */
int64_t orc_align_chain(void* h, const orc_params* p, const uint8_t* seq, const uint8_t* qual, int64_t L,
                        const int64_t* rows6, int64_t nSeeds, int64_t* res10, double* scoreOut, double* log10POut,
                        char* cigar, char* md, int64_t strCap) {
    Db* db = (Db*)h;
    Params P = toParams(p);
    std::vector<uint8_t> q30(L, 30);
    Read r = makeRead("q", seq, qual ? qual : q30.data(), L);
    Read rc = revcomRead(r);
    std::vector<Seed> chain(nSeeds);
    for (int64_t i = 0; i < nSeeds; ++i) {
        const int64_t* x = rows6 + 6 * i;
        chain[i] = Seed{x[0], x[1], x[2], x[3], x[4], (int32_t)x[5], 0};
    }
    Aln a = alignChain(*db, P, r, rc, chain);
    evaluate(P, a);
    res10[0] = a.tid; res10[1] = a.qStrand; res10[2] = a.tStart; res10[3] = a.tEnd; res10[4] = a.alnFrom;
    res10[5] = a.alnTo; res10[6] = a.alnStart; res10[7] = a.alnEnd; res10[8] = a.hazard; res10[9] = a.dpCells;
    *scoreOut = a.alnScore;
    *log10POut = a.log10P;
    std::snprintf(cigar, strCap, "%s", cigarString(a).c_str());
    std::snprintf(md, strCap, "%s", mdTag(a).c_str());
    return 0;
}

/* backExt / fwdExt calls with an unambiguous base (the FMD extensions, two occ-block reads
 * each) made by findAllSMEMS over the reads, single-threaded */
int64_t orc_count_ext(void* h, const uint8_t* seq, const int64_t* offs, int64_t n, int64_t minLen, double maxE) {
    Db* db = (Db*)h;
    std::vector<Read> reads = makeReads(seq, nullptr, offs, nullptr, n);
    g_nExt = 0;
    g_countExt = true;
    for (int64_t i = 0; i < n; ++i) (void)findAllSMEMS(*db, reads[i], minLen, maxE);
    g_countExt = false;
    return g_nExt;
}

/* AlignmentPE::getPairs + sort + filter over caller alignment lists, laid out as msgs_pair_batch
 * takes them (read pair i: rows alnOffs[i] .. alnOffs[i+1] and alnOffs[n+i] .. alnOffs[n+i+1] of
 * rows4 = tid, strand, alnStart, alnEnd, with log10P); returns the number of pairs written (rows3:
 * absolute rows of both alignments and the insert size, lik2: log10lik and loglik), or -1 when
 * more than cap */
int64_t orc_pairs(const orc_params* p, int64_t n, const int64_t* alnOffs, const int64_t* rows4, const double* log10P,
                  int64_t* pairCounts, int64_t* rows3, double* lik2, int64_t cap) {
    const Params P = toParams(p);
    int64_t k = 0;
    std::vector<PairOut> out;
    for (int64_t i = 0; i < n; ++i) {
        std::vector<Aln> a1, a2;
        auto take = [&](int64_t j0, int64_t j1, std::vector<Aln>& v) {
            for (int64_t j = j0; j < j1; ++j) {
                Aln a;
                a.tid = (int32_t)rows4[4 * j];
                a.qStrand = (int32_t)rows4[4 * j + 1];
                a.alnStart = rows4[4 * j + 2];
                a.alnEnd = rows4[4 * j + 3];
                a.log10P = log10P[j];
                v.push_back(a);
            }
        };
        take(alnOffs[i], alnOffs[i + 1], a1);
        take(alnOffs[n + i], alnOffs[n + i + 1], a2);
        pairAlignments(P, a1, a2, out);
        pairCounts[i] = (int64_t)out.size();
        for (const PairOut& q : out) {
            if (k >= cap) return -1;
            rows3[3 * k] = alnOffs[i] + q.f;
            rows3[3 * k + 1] = alnOffs[n + i] + q.r;
            rows3[3 * k + 2] = q.isize;
            lik2[2 * k] = q.l10;
            lik2[2 * k + 1] = q.ll;
            ++k;
        }
    }
    return k;
}

/* Alignment::calcMapQ over alignments with the given log10P values */
void orc_calc_mapq(const double* log10P, int64_t n, double* postP, uint8_t* mapQ) {
    std::vector<Aln> v(n);
    for (int64_t i = 0; i < n; ++i) v[i].log10P = log10P[i];
    calcMapQ(v);
    for (int64_t i = 0; i < n; ++i) { postP[i] = v[i].postP; mapQ[i] = v[i].mapQ; }
}

}  // extern "C"
