/*
 * msgseqtk_api.h -- C++ drop-in face of the MI355X mapping path in namespace
 * EGriceLab::MSGseqTK: the classes msgseqtk-align's tasks call (src/msgseqtk-align.cpp:595-758),
 * with the reference's names and signatures, backed by libmsgs_hip.so through its C-ABI
 * (include/msgs_hip.h).  Every per-read call is a device batch of one; results are those of the
 * reference (the device path is bit-exact against it).  Only the types and members on the
 * mapping path are provided; index construction, merging and BAM export are not.
 *
 * Reference interfaces: FMDIndex (src/FMDIndex.h:122-284), SMEM / SMEM_LIST (src/SMEM.h:40-135,
 * 335-389), SeedPair (src/SeedPair.h:32-100), SeedChain (src/SeedChain.h:31-109), Alignment
 * (src/Alignment.h:60-130,395-436), MetaGenome (src/MetaGenome.h:64-167,319-321,474-479),
 * PrimarySeq (src/PrimarySeq.h:37-152), DNAalphabet / dna:: (src/DNAalphabet.h:37-190,
 * src/DNAseq.h:134-212), GLoc (src/GLoc.h:20-80).
 *
 * Errors: constructors and loads throw std::invalid_argument / std::runtime_error as the
 * reference does; a device failure throws std::runtime_error with the library's message.
 */
#ifndef MSGSEQTK_API_H
#define MSGSEQTK_API_H

#include <cmath>
#include <cstdint>
#include <istream>
#include <memory>
#include <ostream>
#include <string>
#include <utility>
#include <vector>

struct msgs_ctx;

namespace EGriceLab {
namespace MSGseqTK {

using std::string;
using std::vector;

typedef uint8_t nt16_t;
typedef std::basic_string<nt16_t> DNAseq;
typedef std::basic_string<uint8_t> QualStr;

/* nt16 alphabet (src/DNAalphabet.h:37-190) */
class DNAalphabet {
public:
    enum Base { A = 1, C = 2, G = 4, T = 8, U = T, R = A | G, Y = C | T, S = G | C, W = A | T, K = G | T, M = A | C,
                B = C | G | T, D = A | G | T, H = A | C | T, V = A | C | G, N = A | C | G | T };
    static const nt16_t GAP_BASE = 0;
    static const nt16_t NT16_MAX = 0xF;
    static nt16_t encode(char s);
    static char decode(nt16_t b);
    static nt16_t complement(nt16_t b);
    static nt16_t toBasic(nt16_t b);
    static int toInt(nt16_t b);
    static bool isAmbiguous(nt16_t b) { return !(b == A || b == C || b == G || b == T); }
};

namespace dna {
DNAseq encode(const string& seqStr);
string decode(const DNAseq& seq);
DNAseq revcom(const DNAseq& seq);
}  // namespace dna

namespace quality {
const uint8_t DEFAULT_Q_SHIFT = 33;
const uint8_t MIN_Q_SCORE = 2;
/* quality::encode (src/QualStr.h:40-42) */
inline uint8_t encode(char c, uint8_t qShift = DEFAULT_Q_SHIFT) {
    const uint8_t v = (uint8_t)(c - qShift);
    return v > MIN_Q_SCORE ? v : MIN_Q_SCORE;
}
QualStr encode(const string& qStr, uint8_t qShift = DEFAULT_Q_SHIFT);
}  // namespace quality

/* a read (src/PrimarySeq.h:37-152): bases, Phred scores (qualities missing in the input
 * take the shift value, as fixQual does), name */
class PrimarySeq {
public:
    PrimarySeq() = default;
    PrimarySeq(const DNAseq& seq, const string& name, const string& desc = "", const string& qStr = "",
               uint8_t qShift = quality::DEFAULT_Q_SHIFT);
    PrimarySeq(const string& seqStr, const string& name, const string& desc = "", const string& qStr = "",
               uint8_t qShift = quality::DEFAULT_Q_SHIFT);
    const DNAseq& getSeq() const { return seq; }
    const QualStr& getQual() const { return qual; }
    const string& getName() const { return name; }
    const string& getDesc() const { return desc; }
    nt16_t getBase(size_t i) const { return seq[i]; }
    size_t length() const { return seq.length(); }
    PrimarySeq revcom() const;
private:
    DNAseq seq;
    string name, desc;
    QualStr qual;
};

/* genomic location with strand (src/GLoc.h:20-80) */
class GLoc {
public:
    enum STRAND { FWD = 1, REV = 2, UNK = 3 };
    GLoc() = default;
    GLoc(int64_t start, int64_t end, int32_t tid, STRAND strand) : start(start), end(end), tid(tid), strand(strand) {}
    int64_t getStart() const { return start; }
    int64_t getEnd() const { return end; }
    int64_t length() const { return end - start; }
    int32_t getTid() const { return tid; }
    STRAND getStrand() const { return strand; }
private:
    int64_t start = 0, end = 0;
    int32_t tid = -1;
    STRAND strand = UNK;
};

class Device;   /* a device database (implementation) */

/* the genome/chromosome catalogue and sequences (src/MetaGenome.h): chromosomes in order,
 * the BD text built from them (src/MetaGenome.cpp:104-113) and the BD coordinate lookups */
class MetaGenome {
public:
    MetaGenome();
    ~MetaGenome();
    /* one chromosome (Genome::addChrom + MetaGenome::addGenome, one genome per chromosome) */
    MetaGenome& addChrom(const string& name, const DNAseq& seq);
    /* DB.mtg / DB.mgs (src/MetaGenome.cpp:42-73, src/Genome.cpp:29-97); the stream is set bad
     * on a format error, as the reference's loaders do */
    std::istream& load(std::istream& in);
    std::istream& loadSeq(std::istream& in);
    size_t numChroms() const { return names.size(); }
    const string& getChromName(size_t tid) const { return names[tid]; }
    const DNAseq& getSeq(size_t tid) const { return seqs[tid]; }
    int64_t getChromLength(size_t tid) const { return (int64_t)seqs[tid].length() + 1; }
    int64_t BDSize() const;
    DNAseq getBDSeq() const;
    /* BD position -> chromosome / strand / chromosome-relative location
     * (src/MetaGenome.h:147-167,474-479) */
    int64_t getTid(int64_t pos) const;
    GLoc::STRAND getStrand(int64_t tid, int64_t pos) const;
    int64_t getLoc(int64_t tid, int64_t pos) const;
    /* the device database of these chromosomes (built on first use) */
    std::shared_ptr<Device> device() const;
private:
    vector<string> names;
    vector<DNAseq> seqs;
    vector<int64_t> bdStart;
    mutable std::shared_ptr<Device> dev;
    friend class FMDIndex;
};

/* the FMD index of a BD text, resident in HBM (src/FMDIndex.h:122-284) */
class FMDIndex {
public:
    static const int RRR_SAMPLE_RATE = 32;
    static const int SA_SAMPLE_RATE = 32;
    FMDIndex();
    ~FMDIndex();
    /* the index of a BD text (the text of MetaGenome::getBDSeq); throws std::invalid_argument
     * when the text is not a BD text (FMDIndex::buildCounts, src/FMDIndex.cpp:118-134) */
    explicit FMDIndex(const DNAseq& bdSeq, bool buildSAsampled = true, int saSampleRate = SA_SAMPLE_RATE);
    int64_t length() const;
    int64_t getBaseCount(nt16_t b) const;
    int64_t getCumCount(nt16_t b) const;
    double loglik(nt16_t b) const;
    int64_t backExt(int64_t& p, int64_t& q, int64_t& s, nt16_t b) const;
    int64_t fwdExt(int64_t& p, int64_t& q, int64_t& s, nt16_t b) const;
    int64_t LF(nt16_t b, int64_t i) const;
    int64_t accessSA(int64_t i) const;
    nt16_t accessBWT(int64_t i) const;
    int64_t count(const DNAseq& pattern) const;
    vector<GLoc> locateAll(const DNAseq& pattern, GLoc::STRAND strand = GLoc::FWD) const;
    vector<GLoc> locateAllFwd(const DNAseq& pattern) const { return locateAll(pattern, GLoc::FWD); }
    vector<GLoc> locateAllRev(const DNAseq& pattern) const { return locateAll(pattern, GLoc::REV); }
    /* the BD text (the reference inverts the BWT; here the device keeps it) */
    DNAseq getSeq() const;
    /* DB.fmd (src/FMDIndex.cpp:143-161): the counts, the BWT (RRR wavelet tree), gapSA and the
     * sampled SA are read and checked for consistency; the device index is bound to the BD text
     * of the MetaGenome passed to the first search (it is rebuilt on the device from the
     * sequences, and its BWT must equal the file's) */
    std::istream& load(std::istream& in);
    /* the device database, bound on first use (to mtg's chromosomes when unbound) */
    std::shared_ptr<Device> device(const MetaGenome* mtg = nullptr) const;
private:
    mutable std::shared_ptr<Device> dev;
    std::shared_ptr<DNAseq> loadedBwt;   /* BWT read from a .fmd file */
    int64_t loadedB[256] = {}, loadedC[256] = {};
};

/* a seed (src/SeedPair.h:32-100): read interval [from, to), target interval [start, end) on
 * chromosome tid, strand, loglik */
class SeedPair {
public:
    SeedPair() = default;
    SeedPair(int64_t from, int64_t start, int64_t len, int64_t tid = -1, GLoc::STRAND strand = GLoc::UNK,
             double logP = 0)
        : from(from), start(start), len(len), tid(tid), strand(strand), logP(logP) {}
    int64_t getFrom() const { return from; }
    int64_t getTo() const { return from + len; }
    int64_t getStart() const { return start; }
    int64_t getEnd() const { return start + len; }
    int64_t getTid() const { return tid; }
    GLoc::STRAND getStrand() const { return strand; }
    int64_t length() const { return len; }
    double loglik() const { return logP; }
    /* src/SeedPair.h:156-158 */
    static bool isCompatitable(const SeedPair& lhs, const SeedPair& rhs) {
        return lhs.getTo() <= rhs.getFrom() && lhs.getEnd() <= rhs.getStart() && lhs.tid == rhs.tid &&
               lhs.strand == rhs.strand;
    }
private:
    int64_t from = 0, start = 0, len = 0, tid = -1;
    GLoc::STRAND strand = GLoc::UNK;
    double logP = 0;
};
typedef vector<SeedPair> SeedList;
typedef std::pair<SeedList, SeedList> SeedListPE;

/* one SMEM (src/SMEM.h:40-135): read interval [from, to), bi-interval (fwdStart, revStart, size)
 * and loglik, with the read, MetaGenome and index it was found in */
class SMEM {
public:
    SMEM() = default;
    /* evaluates the loglik from the index's base frequencies (SMEM::evaluate, src/SMEM.cpp:19-24) */
    SMEM(const PrimarySeq* seq, const MetaGenome* mtg, const FMDIndex* fmdidx, int64_t from, int64_t to,
         int64_t fwdStart, int64_t revStart, int64_t size);
    /* with the loglik already known (the device's) */
    SMEM(const PrimarySeq* seq, const MetaGenome* mtg, const FMDIndex* fmdidx, int64_t from, int64_t to,
         int64_t fwdStart, int64_t revStart, int64_t size, double logP)
        : seq(seq), mtg(mtg), fmdidx(fmdidx), from(from), to(to), fwdStart(fwdStart), revStart(revStart), size(size),
          logP(logP) {}
    const PrimarySeq* getSeq() const { return seq; }
    const MetaGenome* getMtg() const { return mtg; }
    const FMDIndex* getFmdidx() const { return fmdidx; }
    int64_t getFrom() const { return from; }
    int64_t getTo() const { return to; }
    int64_t getSize() const { return size; }
    int64_t getFwdStart() const { return fwdStart; }
    int64_t getRevStart() const { return revStart; }
    double loglik() const { return logP; }
    int64_t length() const { return to - from; }
    bool empty() const { return length() == 0; }
    bool isValid() const { return size > 0; }
    /* SeedPairs of this SMEM's occurrences on forward chromosome strands (SMEM::getSeeds,
     * src/SMEM.cpp:217-242); the suffix-array rows are looked up on the device in one batch */
    SeedList getSeeds() const;
    SMEM& evaluate();
    static const int64_t MIN_LENGTH = 17;
    static const int64_t MAX_NSEED = 500;
    static const double MAX_EVALUE;
private:
    const PrimarySeq* seq = nullptr;
    const MetaGenome* mtg = nullptr;
    const FMDIndex* fmdidx = nullptr;
    int64_t from = 0, to = 0, fwdStart = 0, revStart = 0, size = 0;
    double logP = 0;
};
typedef SMEM MEM;

/* a list of SMEMs and the seeding entry points (src/SMEM.h:335-389) */
class SMEM_LIST : public vector<SMEM> {
public:
    double loglik() const;
    static SMEM_LIST findAllSMEMS(const PrimarySeq* seq, const MetaGenome* mtg, const FMDIndex* fmdidx,
                                  int64_t minLen = SMEM::MIN_LENGTH, double maxEvalue = SMEM::MAX_EVALUE);
    static SMEM_LIST findMEMS(const PrimarySeq* seq, const MetaGenome* mtg, const FMDIndex* fmdidx,
                              int64_t minLen = SMEM::MIN_LENGTH, double maxEvalue = SMEM::MAX_EVALUE);
    static SeedList findSeeds(const PrimarySeq* seq, const MetaGenome* mtg, const FMDIndex* fmdidx,
                              int64_t minLen = SMEM::MIN_LENGTH, double maxEvalue = SMEM::MAX_EVALUE,
                              int64_t maxNSeed = SMEM::MAX_NSEED);
    static SeedListPE findSeedsPE(const PrimarySeq* fwdSeq, const PrimarySeq* revSeq, const MetaGenome* mtg,
                                  const FMDIndex* fmdidx, int64_t minLen = SMEM::MIN_LENGTH,
                                  double maxEvalue = SMEM::MAX_EVALUE, int64_t maxNSeed = SMEM::MAX_NSEED) {
        return SeedListPE(findSeeds(fwdSeq, mtg, fmdidx, minLen, maxEvalue, maxNSeed),
                          findSeeds(revSeq, mtg, fmdidx, minLen, maxEvalue, maxNSeed));
    }
};
typedef SMEM_LIST MEM_LIST;

/* a chain of compatible seeds (src/SeedChain.h:31-109) */
class SeedChain : public SeedList {
public:
    int64_t getFrom() const { return front().getFrom(); }
    int64_t getTo() const { return back().getTo(); }
    int64_t getStart() const { return front().getStart(); }
    int64_t getEnd() const { return back().getEnd(); }
    int64_t getTid() const { return front().getTid(); }
    GLoc::STRAND getStrand() const { return front().getStrand(); }
    static vector<SeedChain> getChains(const SeedList& inputSeeds, int64_t maxMismatch, int64_t maxIndel);
};
typedef vector<SeedChain> ChainList;

/* alignment scores (src/ScoreScheme.h:76-117, defaults src/ScoreScheme.cpp:14-18) */
struct ScoreScheme {
    double match = 1, mismatch = 3, gapO = 5, gapE = 2, clip = 4;
};

class Alignment;
typedef vector<Alignment> ALIGN_LIST;

/* one alignment of a read along a chain (src/Alignment.h:60-130,395-436) */
class Alignment {
public:
    enum MODE { GLOBAL, LOCAL };
    static ScoreScheme ss;
    Alignment() = default;
    int64_t getAlnFrom() const { return alnFrom; }
    int64_t getAlnTo() const { return alnTo; }
    int64_t getAlnStart() const { return alnStart; }
    int64_t getAlnEnd() const { return alnEnd; }
    int32_t getAlnQLen() const { return (int32_t)(alnTo - alnFrom); }
    int32_t getAlnTLen() const { return (int32_t)(alnEnd - alnStart); }
    double getAlnScore() const { return alnScore; }
    double getLog10P() const { return log10P; }
    double log10lik() const { return log10P; }
    double loglik() const;
    double getPostP() const { return postP; }
    uint8_t getMapQ() const { return mapQ; }
    int32_t getTid() const { return tid; }
    GLoc::STRAND getStrand() const { return qStrand; }
    int64_t getTStart() const { return tStart; }
    int64_t getTEnd() const { return tEnd; }
    /* getAlnCigar as SAM text, getAlnMDTag (src/Alignment.cpp:278-353) */
    const string& getAlnCigarStr() const { return cigar; }
    const string& getAlnMDTag() const { return md; }
    static int32_t mdTag2alnQLen(const string& mdTag);
    static ALIGN_LIST buildAlignments(const PrimarySeq* read, const PrimarySeq* rcRead, const MetaGenome& mtg,
                                      const ChainList& chains, MODE alnMode = GLOBAL);
    static ALIGN_LIST& filter(ALIGN_LIST& alnList, double minScoreRate);
    static ALIGN_LIST& evaluate(ALIGN_LIST& alnList) { return alnList; }   /* done by buildAlignments */
    static ALIGN_LIST& calcMapQ(ALIGN_LIST& alnList);
    static ALIGN_LIST& sort(ALIGN_LIST& alnList);
private:
    friend class AlignmentPE;
    const PrimarySeq* read = nullptr;
    int32_t tid = -1;
    GLoc::STRAND qStrand = GLoc::UNK;
    int64_t tStart = 0, tEnd = 0, alnFrom = 0, alnTo = 0, alnStart = 0, alnEnd = 0;
    double alnScore = 0, log10P = 0, postP = 0;
    uint8_t mapQ = 0xff;
    string cigar, md;
};

/* the insert model (src/PairingScheme.h:26-122, src/PairingScheme.cpp:14-25): a normal density
 * truncated to [min, max], or 1 everywhere when the mean is 0 */
class PairingScheme {
public:
    PairingScheme() : m(DEFAULT_MEAN_INSERT) { updateParam(); updateRange(); }
    double getMean() const { return m; }
    double getSD() const { return s; }
    double getMin() const { return min; }
    double getMax() const { return max; }
    void setMean(double v) { m = v; }
    void setSD(double v) { s = v; }
    void setMin(double v) { min = v; }
    void setMax(double v) { max = v; }
    void setRange(double lo, double hi) { min = lo; max = hi; }
    void updateParam() {
        if (m == 0) s = 1;
        if (s == 0) s = m * DEFAULT_CV_INSERT;
    }
    void updateRange() {
        if (m > 0 && s > 0) {
            min = m - s * DEFAULT_OUTLIER_DIST;
            max = m + s * DEFAULT_OUTLIER_DIST;
        }
    }
    double pr(double x) const;
    double loglik(double x) const { return std::log(pr(x)); }
    double log10lik(double x) const { return std::log10(pr(x)); }
    static const double DEFAULT_MEAN_INSERT, DEFAULT_CV_INSERT, DEFAULT_OUTLIER_DIST;
private:
    double m = 0, s = 0, min = 0, max = 0;
};

class AlignmentPE;
typedef vector<AlignmentPE> PAIR_LIST;

/* a read's alignment with its mate's (src/Alignment.h:446-575) */
class AlignmentPE {
public:
    AlignmentPE() = default;
    AlignmentPE(const Alignment* fwd, const Alignment* rev) : fwdAln(fwd), revAln(rev) {
        setInsertSize(getInsertSize(*fwd, *rev));
    }
    bool isConcordant() const { return fwdAln->tid == revAln->tid && (fwdAln->qStrand & revAln->qStrand) == 0; }
    bool isTailOver() const {
        return (fwdAln->alnStart <= revAln->alnStart && fwdAln->alnEnd > revAln->alnStart) ||
               (revAln->alnStart <= fwdAln->alnStart && revAln->alnEnd > fwdAln->alnStart);
    }
    bool isContained() const {
        return (fwdAln->alnStart <= revAln->alnStart && fwdAln->alnEnd >= revAln->alnEnd) ||
               (revAln->alnStart <= fwdAln->alnStart && revAln->alnEnd >= fwdAln->alnEnd);
    }
    bool isOverlap() const { return fwdAln->alnStart < revAln->alnEnd && fwdAln->alnEnd > revAln->alnStart; }
    double getAlnScore() const { return fwdAln->alnScore + revAln->alnScore; }
    double log10lik() const { return fwdAln->log10lik() + revAln->log10lik() + ps.log10lik(getInsertSize()); }
    double loglik() const { return fwdAln->loglik() + revAln->loglik() + ps.loglik(getInsertSize()); }
    int32_t getInsertSize() const { return isize; }
    void setInsertSize(int32_t v) { isize = v; }
    uint8_t getMapQ() const { return mapQ; }
    double getPostP() const { return postP; }

    const Alignment* fwdAln = nullptr;
    const Alignment* revAln = nullptr;
    int32_t isize = -1;
    uint8_t mapQ = 0xff;
    double postP = 0;

    static PAIR_LIST getPairs(const ALIGN_LIST& fwdAlnList, const ALIGN_LIST& revAlnList);
    static PAIR_LIST& filter(PAIR_LIST& pairList, bool noDiscordant, bool noTailOver, bool noContain, bool noOverlap,
                             int64_t maxNPair = MAX_NPAIR);
    static PAIR_LIST& calcMapQ(PAIR_LIST& pairList);
    static PAIR_LIST& sort(PAIR_LIST& pairList);
    static int32_t getInsertSize(const Alignment& lhs, const Alignment& rhs);
    static PairingScheme ps;
    static const int64_t MAX_NPAIR = 0;
};

}  // namespace MSGseqTK
}  // namespace EGriceLab

#endif
