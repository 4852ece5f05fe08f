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
