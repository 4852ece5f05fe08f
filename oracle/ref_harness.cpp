/*
 * ref_harness.cpp -- extern "C" entry points over the REFERENCE's own seeding
 * stack (FMDIndex, SMEM_LIST, SeedPair, SeedChain, libsds, libdivsufsort),
 * compiled from /root/reference/src by oracle/Makefile into
 * oracle/_ref/libmsgs_ref.so.  TEST INFRASTRUCTURE ONLY: used to pin the oracle
 * restatement against the real reference code.  MetaGenome.cpp, Alignment.cpp
 * and the CLI need Boost/Eigen/htslib (absent) and are not built, so SMEM
 * searches run with mtg = nullptr (findAllSMEMS/findMEMS only store it).
 */
#include <cstdint>
#include <cstdlib>
#include <iomanip>
#include <sstream>
#include <string>

#include "DNAalphabet.h"
#include "DNAseq.h"
#include "FMDIndex.h"
#include "PrimarySeq.h"
#include "QualStr.h"
#include "SMEM.h"
#include "SeedChain.h"
#include "SeedPair.h"

using namespace EGriceLab::MSGseqTK;

extern "C" {

/* The reference's alphabet tables (src/DNAalphabet.h:45-190): DNAalphabet::encode of every
 * char, and per nt16 code decode / complement / toBasic / toInt / isAmbiguous; plus
 * quality::encode of every char (src/QualStr.h:40-42). */
void ref_alphabet(uint8_t* enc256, char* dec16, uint8_t* comp16, uint8_t* basic16, int32_t* int16, uint8_t* amb16,
                  uint8_t* qenc256) {
    for (int c = 0; c < 256; ++c) {
        enc256[c] = DNAalphabet::encode((char)c);
        qenc256[c] = quality::encode((char)c);
    }
    for (int b = 0; b < 16; ++b) {
        dec16[b] = DNAalphabet::decode((nt16_t)b);
        comp16[b] = DNAalphabet::complement((nt16_t)b);
        basic16[b] = DNAalphabet::toBasic((nt16_t)b);
        int16[b] = DNAalphabet::toInt((nt16_t)b);
        amb16[b] = DNAalphabet::isAmbiguous((nt16_t)b) ? 1 : 0;
    }
}
/* dna::encode of a string (src/DNAseq.cpp) */
int64_t ref_encode(const char* s, uint8_t* out) {
    DNAseq d = dna::encode(std::string(s));
    std::copy(d.begin(), d.end(), out);
    return (int64_t)d.size();
}
/* dna::revcom (src/DNAseq.h:134-152) */
void ref_revcom(const uint8_t* s, int64_t n, uint8_t* out) {
    DNAseq d = dna::revcom(DNAseq(s, s + n));
    std::copy(d.begin(), d.end(), out);
}
/* PrimarySeq(seqStr, name, "", qStr).revcom() (src/PrimarySeq.h:54-57,144-152): codes and scores */
int64_t ref_primary_revcom(const char* seq, const char* qual, uint8_t* outSeq, uint8_t* outQual) {
    PrimarySeq p(std::string(seq), "r", "", std::string(qual));
    PrimarySeq r = p.revcom();
    std::copy(r.getSeq().begin(), r.getSeq().end(), outSeq);
    std::copy(r.getQual().begin(), r.getQual().end(), outQual);
    return (int64_t)r.length();
}

void* ref_index_build(const uint8_t* bd, int64_t n) {
    try {
        return new FMDIndex(DNAseq(bd, bd + n), true, FMDIndex::SA_SAMPLE_RATE);
    } catch (...) {
        return nullptr;
    }
}
void ref_index_free(void* h) { delete (FMDIndex*)h; }

/* msgseqtk-build's block-by-block build (src/msgseqtk-build.cpp:403-415, 440-467, restated:
 * the build program is not part of this library): chromosomes whose BD texts start at
 * chromBD[0..nChrom] (chromBD[nChrom] = n) are taken from the last one into blocks, a block
 * closed when one more chromosome would reach blockSize; each block is indexed by the
 * reference's FMDIndex constructor and prepended with the reference's FMDIndex::prepend;
 * then the reference's buildSA.  A text shorter than blockSize is built in one step. */
void* ref_index_build_blocks(const uint8_t* bd, const int64_t* chromBD, int32_t nChrom, int64_t blockSize) {
    try {
        const int64_t n = chromBD[nChrom];
        if (n < blockSize) return new FMDIndex(DNAseq(bd, bd + n), true, FMDIndex::SA_SAMPLE_RATE);
        auto len = [&](int64_t a, int64_t b) { return chromBD[b] - chromBD[a]; };
        FMDIndex* f = new FMDIndex();
        int64_t blockEnd = nChrom;
        for (int64_t i = nChrom; i > 0; --i) {
            const int64_t blockStart = i - 1;
            if (blockStart == 0 || (len(blockStart, blockEnd) < blockSize && len(blockStart - 1, blockEnd) >= blockSize)) {
                f->prepend(FMDIndex(DNAseq(bd + chromBD[blockStart], bd + chromBD[blockEnd])));
                blockEnd = blockStart;
            }
        }
        f->buildSA(FMDIndex::SA_SAMPLE_RATE);
        return f;
    } catch (...) {
        return nullptr;
    }
}
int64_t ref_length(void* h) { return ((FMDIndex*)h)->length(); }
void ref_counts(void* h, int64_t* B, int64_t* C) {
    const FMDIndex* f = (const FMDIndex*)h;
    for (int b = 0; b < 256; ++b) { B[b] = f->getBaseCount((nt16_t)b); C[b] = f->getCumCount((nt16_t)b); }
}
int32_t ref_access_bwt(void* h, int64_t i) { return ((FMDIndex*)h)->accessBWT(i); }
/* rank(b, i) through the public LF(b, i) = C[b] + rank(b, i) */
int64_t ref_rank(void* h, int32_t b, int64_t i) {
    const FMDIndex* f = (const FMDIndex*)h;
    return f->LF((nt16_t)b, i) - f->getCumCount((nt16_t)b);
}
int64_t ref_access_sa(void* h, int64_t i) { return ((FMDIndex*)h)->accessSA(i); }
int64_t ref_backext(void* h, int64_t* pqs, int32_t b) {
    return ((FMDIndex*)h)->backExt(pqs[0], pqs[1], pqs[2], (nt16_t)b);
}
int64_t ref_fwdext(void* h, int64_t* pqs, int32_t b) {
    return ((FMDIndex*)h)->fwdExt(pqs[0], pqs[1], pqs[2], (nt16_t)b);
}
double ref_loglik(void* h, int32_t b) { return ((FMDIndex*)h)->loglik((nt16_t)b); }

/* SMEM list of one read via the reference's public formatted output
 * "from-to:size:fwdStart:revStart:logP" printed with 17 significant digits. */
int64_t ref_find_smems(void* h, const uint8_t* read, int64_t L, int64_t minLen, double maxE, int32_t mems,
                       int64_t* out5, double* logP, int64_t cap) {
    const FMDIndex* f = (const FMDIndex*)h;
    PrimarySeq seq(DNAseq(read, read + L), "r");
    SMEM_LIST l = mems ? SMEM_LIST::findMEMS(&seq, nullptr, f, minLen, maxE)
                       : SMEM_LIST::findAllSMEMS(&seq, nullptr, f, minLen, maxE);
    int64_t k = 0;
    for (const SMEM& m : l) {
        std::ostringstream os;
        os << std::setprecision(17) << m;
        std::string s = os.str();
        long long from, to, size, fs, rs;
        char rest[128];
        if (std::sscanf(s.c_str(), "%lld-%lld:%lld:%lld:%lld:%127s", &from, &to, &size, &fs, &rs, rest) != 6) return -1;
        if (k < cap) {
            int64_t* o = out5 + 5 * k;
            o[0] = from; o[1] = to; o[2] = fs; o[3] = rs; o[4] = size;
            logP[k] = std::strtod(rest, nullptr);
        }
        ++k;
    }
    return k;
}

/* SeedPair::removeRedundant on rows (from, to, start, end, tid, strand) */
int64_t ref_remove_redundant(int64_t* rows6, double* logP, int64_t n) {
    SeedList seeds;
    for (int64_t i = 0; i < n; ++i) {
        int64_t* r = rows6 + 6 * i;
        seeds.push_back(SeedPair(r[0], r[2], r[1] - r[0], r[4], (GLoc::STRAND)r[5], logP[i]));
    }
    SeedPair::removeRedundant(seeds);
    for (size_t i = 0; i < seeds.size(); ++i) {
        int64_t* r = rows6 + 6 * i;
        const SeedPair& s = seeds[i];
        r[0] = s.getFrom(); r[1] = s.getTo(); r[2] = s.getStart(); r[3] = s.getEnd();
        r[4] = s.getTid(); r[5] = s.getStrand();
        logP[i] = s.loglik();
    }
    return (int64_t)seeds.size();
}

/* SeedChain::getChains; out holds per chain its size then the input row indices
 * (the index travels in the seed's logP field, which chaining does not read). */
int64_t ref_get_chains(const int64_t* rows6, int64_t n, int64_t maxMM, int64_t maxIndel, int64_t* out, int64_t cap) {
    SeedList seeds;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t* r = rows6 + 6 * i;
        seeds.push_back(SeedPair(r[0], r[2], r[1] - r[0], r[4], (GLoc::STRAND)r[5], (double)i));
    }
    ChainList chains = SeedChain::getChains(seeds, maxMM, maxIndel);
    int64_t k = 0;
    for (const SeedChain& ch : chains) {
        if (k < cap) out[k] = (int64_t)ch.size();
        ++k;
        for (const SeedPair& s : ch) { if (k < cap) out[k] = (int64_t)s.loglik(); ++k; }
    }
    return k;
}

}  // extern "C"

/* ------------------------------------------------------------------ DB.fmd files */

#include <fstream>

extern "C" {

/* FMDIndex::save of a reference index to `path` (the body of DB.fmd, no program header) */
int32_t ref_index_save(void* h, const char* path) {
    std::ofstream out(path, std::ios::binary);
    ((FMDIndex*)h)->save(out);
    return out ? 0 : -1;
}
/* FMDIndex::load of the body of a DB.fmd file written elsewhere */
void* ref_index_load(const char* path, int64_t skip) {
    try {
        std::ifstream in(path, std::ios::binary);
        in.seekg(skip);
        FMDIndex* f = new FMDIndex();
        f->load(in);
        if (!in) { delete f; return nullptr; }
        return f;
    } catch (...) {
        return nullptr;
    }
}

}  // extern "C"

/* ------------------------------------------------------------------ CPU baseline */

#include <chrono>
#include <cmath>
#include <sstream>
#include <unordered_map>

#include "StringUtils.h"
#include "msgs_oracle.hpp"

extern "C" {

/* The reference's FMDIndex over the BD text bd[0..n) with the given suffix array, made of
 * the reference's own components the way its constructor makes them (src/FMDIndex.cpp:96-116:
 * counts, BWT, gapSA = SA[0 .. numGaps), SA sampled at i % 32 == 0 or BWT[i] == 0 into an RRR
 * bit set, WaveletTreeRRR over the BWT) and read back through FMDIndex::load from their own
 * save() output in FMDIndex::save's order (src/FMDIndex.cpp:143-161).  Replaces the suffix
 * sort of the constructor with the caller's (equal) suffix array. */
void* ref_index_from_sa(const uint8_t* bd, const int64_t* sa, int64_t n) {
    try {
        std::array<int64_t, 256> B = {}, C = {};
        for (int64_t i = 0; i < n; ++i) B[bd[i]]++;
        int64_t S = 0;
        for (int i = 0; i <= DNAalphabet::SIZE; ++i) { C[i] = S; S += B[i]; }
        DNAseq bwt(n, 0);
        for (int64_t i = 0; i < n; ++i) bwt[i] = sa[i] == 0 ? 0 : bd[sa[i] - 1];
        /* sampled SA first: the wavelet tree's constructor consumes the BWT (as in the
         * reference's constructor, which samples before it builds the tree) */
        BitStr32 bstr(n);
        FMDIndex::SAarr_t sampled;
        sampled.reserve(n / FMDIndex::SA_SAMPLE_RATE + B[0] + 1);
        for (int64_t i = 0; i < n; ++i)
            if (i % FMDIndex::SA_SAMPLE_RATE == 0 || bwt[i] == 0) { bstr.set(i); sampled.push_back(sa[i]); }
        std::stringstream ss;
        ss.write((const char*)B.data(), sizeof(B));
        ss.write((const char*)C.data(), sizeof(C));
        {
            WaveletTreeRRR wt(bwt, 0, DNAalphabet::NT16_MAX, FMDIndex::RRR_SAMPLE_RATE);
            wt.save(ss);
        }
        FMDIndex::GAParr_t gap(sa, sa + B[0]);
        EGriceLab::StringUtils::saveString(gap, ss);
        BitSeqRRR(bstr, FMDIndex::RRR_SAMPLE_RATE).save(ss);
        EGriceLab::StringUtils::saveString(sampled, ss);
        FMDIndex* f = new FMDIndex();
        f->load(ss);
        if (!ss || f->length() != n) { delete f; return nullptr; }
        return f;
    } catch (...) {
        return nullptr;
    }
}

}  // extern "C"

namespace {

/* SMEM_LIST::findSeeds (src/SMEM.cpp:271-305) with the reference's own SMEM search, FMDIndex
 * (RRR wavelet-tree ranks, LF-walk accessSA with the per-read SA cache) and
 * SeedPair::removeRedundant; only the BD-coordinate lookups (MetaGenome::getTid / getLoc /
 * getStrand, whose source needs Boost) come from the oracle's restatement.  The SMEM's
 * SA start rows are read from its formatted output, the reference's public face for them. */
std::vector<orc::Seed> refSeeds(const FMDIndex* f, const orc::Db& db, const orc::Params& P, const orc::Read& r) {
    PrimarySeq seq(DNAseq(r.seq.data(), r.seq.data() + r.L), r.name);
    SMEM_LIST smems = SMEM_LIST::findAllSMEMS(&seq, nullptr, f, P.minLen, P.maxEvalue);
    std::sort(smems.begin(), smems.end(), [](const SMEM& l, const SMEM& rr) { return l.loglik() < rr.loglik(); });
    SeedList all;
    SMEM::SAmap_t cache;
    const bool cached = smems.size() > 1;
    char buf[256];
    for (const SMEM& m : smems) {
        std::ostringstream os;
        os << std::setprecision(17) << m;
        long long from, to, size, fs, rs;
        std::snprintf(buf, sizeof buf, "%s", os.str().c_str());
        if (std::sscanf(buf, "%lld-%lld:%lld:%lld:%lld:", &from, &to, &size, &fs, &rs) != 5) continue;
        const int64_t len = to - from;
        for (int64_t i = 0; i < size; ++i) {
            for (int side = 0; side < 2; ++side) {
                const int64_t row = (side == 0 ? fs : rs) + i;
                int64_t bd;
                if (cached) {
                    auto it = cache.find(row);
                    bd = it != cache.end() ? it->second : (cache[row] = f->accessSA(row));
                } else {
                    bd = f->accessSA(row);
                }
                const int64_t tid = orc::getTid(db, bd);
                if (orc::getStrand(db, tid, bd) != orc::FWD) continue;
                const int64_t start = orc::getLoc(db, tid, bd);
                all.push_back(SeedPair(side == 0 ? from : r.L - to, start, len, tid,
                                       side == 0 ? GLoc::FWD : GLoc::REV, m.loglik()));
            }
        }
        if ((int64_t)all.size() >= P.maxNSeed) break;
    }
    SeedPair::removeRedundant(all);
    std::vector<orc::Seed> out;
    out.reserve(all.size());
    for (const SeedPair& s : all)
        out.push_back(orc::Seed{s.getFrom(), s.getTo(), s.getStart(), s.getEnd(), s.getTid(), (int32_t)s.getStrand(),
                                s.loglik()});
    return out;
}

/* SeedChain::getChains (src/SeedChain.cpp:40-73), the reference's own */
std::vector<std::vector<orc::Seed>> refChains(const std::vector<orc::Seed>& seeds, int64_t maxMM, int64_t maxIndel) {
    SeedList sl;
    sl.reserve(seeds.size());
    for (const orc::Seed& s : seeds)
        sl.push_back(SeedPair(s.from, s.start, s.to - s.from, s.tid, (GLoc::STRAND)s.strand, s.logP));
    ChainList chains = SeedChain::getChains(sl, maxMM, maxIndel);
    std::vector<std::vector<orc::Seed>> out;
    out.reserve(chains.size());
    for (const SeedChain& ch : chains) {
        std::vector<orc::Seed> v;
        for (const SeedPair& s : ch)
            v.push_back(orc::Seed{s.getFrom(), s.getTo(), s.getStart(), s.getEnd(), s.getTid(), (int32_t)s.getStrand(),
                                  s.loglik()});
        out.push_back(std::move(v));
    }
    return out;
}

}  // namespace

extern "C" {

/* The CPU baseline: msgseqtk-align's per-read (per-pair) tasks with the reference's own
 * seeding and chaining over `fmd` (ref_index_from_sa) and the oracle's DP, evaluation, MAPQ,
 * pairing and records over the oracle database `orcDb` (orc_db_create of liborc.so); OpenMP
 * over tasks.  Arguments and result as orc_map (oracle/orc_capi.cpp); `params` is the
 * msgs_params block converted by orc_params_convert. */
int64_t ref_baseline_map(void* fmd, void* orcDb, const void* params, const uint8_t* seq, const uint8_t* qual,
                         const int64_t* offs, const char* names, const uint8_t* seq2, const uint8_t* qual2,
                         const int64_t* offs2, const char* names2, int64_t n, int32_t nThreads, char* out,
                         int64_t cap, int64_t* lineOffs, double* seconds) {
    const FMDIndex* f = (const FMDIndex*)fmd;
    const orc::Db& db = *(const orc::Db*)orcDb;
    const orc::Params& P = *(const orc::Params*)params;
    orc::SeedFn sf = [f, &db, &P](const orc::Read& r) { return refSeeds(f, db, P, r); };
    orc::ChainFn cf = refChains;
    return orc::mapBatch(db, P, seq, qual, offs, names, seq2, qual2, offs2, names2, n, nThreads, out, cap, lineOffs,
                         seconds, &sf, &cf);
}

}  // extern "C"
