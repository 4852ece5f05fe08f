/*
 * msgs_oracle.hpp -- CPU oracle for the MSGseqTK read-mapping hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a literal, single-threaded (optionally
 * OpenMP-per-read) restatement of the reference algorithm, written to be read
 * side by side with the reference sources it cites.  It is the checker for the
 * HIP product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (msgseqtk_amd/) never links it.
 *
 * Every function names the reference file:line it follows.  Pinning status of
 * each part (what independent check fixes it) is listed in oracle/README.md.
 */
#ifndef MSGS_ORACLE_HPP
#define MSGS_ORACLE_HPP

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace orc {

typedef uint8_t nt16_t;

/* ---- A1: nt16 alphabet (src/DNAalphabet.h:37-60, src/DNAalphabet.cpp:14-152) ---- */
enum Base { A = 1, C = 2, G = 4, T = 8, N = 15 };
nt16_t encodeSym(char s);          /* sym2base */
char decodeBase(nt16_t b);         /* base2sym ('$' for 0) */
nt16_t complement(nt16_t b);       /* base2comp */
nt16_t toBasic(nt16_t b);          /* base2basic */
int toInt(nt16_t b);               /* base2int */
bool isAmbiguous(nt16_t b);        /* src/DNAalphabet.h:135-142 */
nt16_t encodeInt(int i);           /* DNAalphabet::encode(int) */
uint8_t qualEncode(char c);        /* quality::encode with the default shift 33 */

/* GLoc::STRAND (src/GLoc.h:20) */
enum Strand { FWD = 1, REV = 2, UNK = 3 };

/* ---- parameters (defaults: src/SMEM.h:204-207, src/SMEM.cpp:17, src/ScoreScheme.cpp:14-18,
 *      src/PairingScheme.cpp:14-16, src/msgseqtk-align.cpp:59-62, src/Alignment.h:385) ---- */
struct Params {
    int64_t minLen = 17;
    double maxEvalue = 0.01;
    int64_t maxNSeed = 500;
    int32_t mode = 0;            /* 0 GLOBAL, 1 LOCAL */
    double match = 1, mismatch = 3, gapO = 5, gapE = 2, clip = 4;
    double minScoreRate = 0.65;
    int32_t maxReport = 1;
    /* paired-end */
    double insMean = 200, insSD = 20, insMin = 0, insMax = 400;
    int32_t noMixed = 0, noDiscordant = 0, noTailOver = 0, noContain = 0, noOverlap = 0;
    int64_t maxNPair = 0;
};

/* ---- A3/A4: metagenome + FMD index ---- */
struct Db {
    /* metagenome (src/MetaGenome.cpp:170-224) */
    std::vector<std::string> names;
    std::vector<std::vector<nt16_t>> chroms;  /* each with a trailing 0 (basic_string terminator) */
    std::vector<int64_t> chromLen;            /* len (without terminator) */
    std::vector<int64_t> bdStart, bdEnd;      /* chromIdx2BDLoc */
    /* FMD index (src/FMDIndex.h:337-343) */
    int64_t n = 0;
    std::vector<nt16_t> bd;                   /* BD text */
    std::vector<int64_t> sa;                  /* full suffix array */
    std::vector<nt16_t> bwt;
    int64_t B[256] = {}, Cc[256] = {};
    std::vector<int64_t> gapSA;
    std::vector<int64_t> occ;                 /* 16 counts per 64-position checkpoint */
    double loglikTab[16] = {};
};

/* Build the BD text of a metagenome (src/MetaGenome.cpp:104-113, :170-224). */
void buildMeta(Db& db, const nt16_t* seqs, const int64_t* lens, int nChrom, const std::vector<std::string>& names);
/* Suffix array by prefix doubling; byte order, shorter suffix first. */
std::vector<int64_t> suffixArray(const nt16_t* t, int64_t n);
/* The suffix order of msgseqtk-build's block-by-block build, blocks ending at ends[] (text
 * order; empty = one block) */
std::vector<int64_t> suffixArrayBlocks(const nt16_t* t, int64_t n, const std::vector<int64_t>& ends);
/* BD ends of the blocks msgseqtk-build forms from chromosome BD lengths (src/msgseqtk-build.cpp:
 * 403-415, 440-467); empty when the text is built in one step */
std::vector<int64_t> buildBlockEnds(const std::vector<int64_t>& chromBDLen, int64_t blockSize);
/* counts / BWT / gapSA / occ (src/FMDIndex.cpp:29-141,385-415); sa may be given. */
void buildIndex(Db& db, const int64_t* saIn);

int64_t rank(const Db& db, nt16_t s, int64_t i);                                 /* A5 */
int64_t backExt(const Db& db, int64_t& p, int64_t& q, int64_t& s, nt16_t b);      /* A6 */
/* counting of backExt/fwdExt calls with an unambiguous base (single-threaded runs only) */
extern bool g_countExt;
extern int64_t g_nExt;
int64_t fwdExt(const Db& db, int64_t& p, int64_t& q, int64_t& s, nt16_t b);       /* A7 */
int64_t accessSA(const Db& db, int64_t i);                                       /* A8 */
double loglik(const Db& db, nt16_t b);                                           /* A9 */
int64_t getTid(const Db& db, int64_t pos);                                       /* A3 */
int getStrand(const Db& db, int64_t tid, int64_t pos);                            /* A3 */
int64_t getLoc(const Db& db, int64_t tid, int64_t pos);                         /* A3 */

/* ---- reads (A2) ---- */
struct Read {
    std::string name;
    std::vector<nt16_t> seq;   /* + trailing 0 kept at seq[L] */
    std::vector<uint8_t> qual;
    int64_t L = 0;
};
Read makeRead(const std::string& name, const nt16_t* s, const uint8_t* q, int64_t L);
Read revcomRead(const Read& r);

/* ---- A10-A15 ---- */
struct Smem { int64_t from, to, fwdStart, revStart, size; double logP; };
std::vector<Smem> findAllSMEMS(const Db& db, const Read& r, int64_t minLen, double maxE);
std::vector<Smem> findMEMS(const Db& db, const Read& r, int64_t minLen, double maxE);

/* ---- A16 ---- */
struct Seed { int64_t from, to, start, end, tid; int32_t strand; double logP; };
std::vector<Seed> findSeeds(const Db& db, const Read& r, int64_t minLen, double maxE, int64_t maxNSeed);
/* SMEM::getSeeds of one SMEM, appended to out (src/SMEM.cpp:217-269) */
void getSeeds(const Db& db, const Read& r, const Smem& m, std::vector<Seed>& out);
void removeRedundant(std::vector<Seed>& seeds);                                  /* A17 */
std::vector<std::vector<Seed>> getChains(const std::vector<Seed>& seeds, int64_t maxMM, int64_t maxIndel); /* A18 */

/* ---- A19-A29 ---- */
struct Aln {
    const Read* read = nullptr; const Read* rcRead = nullptr;
    int32_t tid = -1; int32_t qStrand = UNK;
    int64_t qFrom = 0, qTo = 0, tStart = 0, tEnd = 0;
    int64_t alnFrom = 0, alnTo = 0, alnStart = 0, alnEnd = 0;
    double alnScore = 0;
    std::vector<uint32_t> path;   /* BAM ops 0=M 1=I 2=D */
    double log10P = 0, postP = 0;
    uint8_t mapQ = 0xff;
    int32_t hazard = 0;           /* 1 = backtrace read column -1 (undefined in the reference) */
    int64_t dpCells = 0;          /* recurrence evaluations (algorithmic cells) */
    const std::vector<nt16_t>* target = nullptr;
};
Aln alignChain(const Db& db, const Params& P, const Read& r, const Read& rc, const std::vector<Seed>& chain);
/* full-rectangle DP (Alignment_test constructor form, src/Alignment.h:68-81 + calculateScores()) */
Aln alignFull(const Params& P, const Read& r, const Read& rc, const std::vector<nt16_t>& target, int mode);
std::string cigarString(const Aln& a);
std::string mdTag(const Aln& a);
void evaluate(const Params& P, Aln& a);
void calcMapQ(std::vector<Aln>& alns);

/* ---- A31/A32 ---- */
/* A30: the candidate pairs of two filtered, evaluated alignment lists after AlignmentPE's
 * getPairs, sort and filter: indexes into a1 / a2, insert size, log10lik, loglik */
struct PairOut { int64_t f, r; int32_t isize; double l10, ll; };
void pairAlignments(const Params& P, const std::vector<Aln>& a1, const std::vector<Aln>& a2,
                    std::vector<PairOut>& out);
std::string mapSE(const Db& db, const Params& P, const Read& r);
std::string mapPE(const Db& db, const Params& P, const Read& r1, const Read& r2);
/* the same tasks with the seeding (SMEM_LIST::findSeeds) and chaining (SeedChain::getChains)
 * supplied by the caller */
using SeedFn = std::function<std::vector<Seed>(const Read&)>;
using ChainFn = std::function<std::vector<std::vector<Seed>>(const std::vector<Seed>&, int64_t, int64_t)>;
std::string mapSE(const Db& db, const Params& P, const Read& r, const SeedFn& seeds, const ChainFn& chains);
std::string mapPE(const Db& db, const Params& P, const Read& r1, const Read& r2, const SeedFn& seeds,
                  const ChainFn& chains);
/* a batch of reads (SE: seq2 == nullptr) mapped per read / pair, OpenMP over tasks; returns the
 * bytes of SAM text (out may be null), -1 if cap is too small; see orc_map (orc_capi.cpp) */
int64_t mapBatch(const Db& db, const Params& P, const uint8_t* seq, const uint8_t* qual, const int64_t* offs,
                 const char* names, const uint8_t* seq2, const uint8_t* qual2, const int64_t* offs2,
                 const char* names2, int64_t n, int32_t nThreads, char* out, int64_t cap, int64_t* lineOffs,
                 double* seconds, const SeedFn* seeds, const ChainFn* chains);

} // namespace orc
#endif
