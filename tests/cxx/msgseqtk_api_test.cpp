/*
 * msgseqtk_api_test.cpp -- the C++ drop-in face on the GPU, through the reference's own test
 * cases and invariants:
 *   SMEM_test (test/SMEM_test.cpp): four copies of a chromosome with an N, read ACGTAGTA,
 *     findMEMS / findSeeds with minLen 1 and no e-value cut; every seed's read segment equals its
 *     target segment (isValidSeed), and every MEM has as many forward-strand seeds as occurrences;
 *   SeedChain_test (test/SeedChain_test.cpp): two chromosomes, seeds of length >= 4, chains with
 *     maxMismatch / maxIndel of 0.2 / 0.1 read lengths; seeds in a chain are compatible and match
 *     (isValidChain);
 *   Alignment_test (test/Alignment_test.cpp): its five read/target pairs (perfect match, match in
 *     the middle, mismatch, insertion, deletion; LOCAL for the first three), aligned along the
 *     chains of their seeds by Alignment::buildAlignments; the CIGAR covers the whole read and
 *     mdTag2alnQLen(MD) equals the aligned query length.
 *   the PE driver's pairing (src/msgseqtk-align.cpp:700-717): mates whose target segments occur
 *     twice give two in-range candidate pairs with equal log10lik and two out of range; getPairs,
 *     sort, filter and calcMapQ through AlignmentPE, and the same lists through the device pairing
 *     entry msgs_pair_batch, which must give the same pairs, order, insert sizes and log10lik.
 * Plus known answers that pin the device results: FMD counts of the BD text, count / locateAll
 * against a naive scan, accessSA against a suffix sort, the alignments' scores and strings of the
 * Alignment_test cases, calcMapQ / sort on the reference's formulas.  Prints PASS and exits 0.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <limits>
#include <map>
#include <numeric>
#include <sstream>

#include "msgseqtk/Alignment.h"
#include "msgseqtk/FMDIndex.h"
#include "msgseqtk/MetaGenome.h"
#include "msgseqtk/SMEM.h"
#include "msgseqtk/SeedChain.h"
#include "msgs_hip.h"

using namespace std;
using namespace EGriceLab::MSGseqTK;

static int failures = 0;
#define CHECK(cond)                                                                            \
    do {                                                                                       \
        if (!(cond)) {                                                                         \
            cerr << __FILE__ << ":" << __LINE__ << ": check failed: " #cond << endl;           \
            ++failures;                                                                        \
        }                                                                                      \
    } while (0)

static const double inf = std::numeric_limits<double>::infinity();

static bool isValidSeed(const DNAseq& target, const DNAseq& query, const SeedPair& seed) {
    const DNAseq tSeg = target.substr(seed.getStart(), seed.length());
    const DNAseq qSeg = seed.getStrand() == GLoc::FWD ? query.substr(seed.getFrom(), seed.length())
                                                      : dna::revcom(query).substr(seed.getFrom(), seed.length());
    if (tSeg != qSeg) cerr << "unmatched seed at " << seed.getFrom() << " -> " << seed.getStart() << endl;
    return tSeg == qSeg;
}

/* query length of a CIGAR string (M, I, S, =, X consume the query) */
static int64_t cigar2QLen(const string& cigar) {
    int64_t len = 0, num = 0;
    for (char c : cigar) {
        if (c >= '0' && c <= '9') { num = num * 10 + (c - '0'); continue; }
        if (c == 'M' || c == 'I' || c == 'S' || c == '=' || c == 'X') len += num;
        num = 0;
    }
    return len;
}

static vector<int64_t> naiveFind(const DNAseq& text, const DNAseq& pat) {
    vector<int64_t> out;
    for (size_t k = text.find(pat); k != DNAseq::npos; k = text.find(pat, k + 1)) out.push_back((int64_t)k);
    return out;
}

static void smemTest() {
    const DNAseq chr = dna::encode("ACGTCGTAGTACTACGNACGTCGTAGTACTACG");
    MetaGenome mtg;
    mtg.addChrom("chr1_1", chr).addChrom("chr1_2", chr).addChrom("chr2_1", chr).addChrom("chr2_2", chr);
    FMDIndex fmdidx(mtg.getBDSeq());
    CHECK(fmdidx.getSeq() == mtg.getBDSeq());
    CHECK(fmdidx.length() == mtg.BDSize());

    const PrimarySeq read("ACGTAGTA", "seq1");
    const MEM_LIST mems = SMEM_LIST::findMEMS(&read, &mtg, &fmdidx, 1, inf);
    CHECK(!mems.empty());
    for (const MEM& mem : mems) {
        const SeedList seeds = mem.getSeeds();
        CHECK((int64_t)seeds.size() == mem.getSize());   /* every occurrence once on a forward strand */
        for (const SeedPair& seed : seeds) CHECK(isValidSeed(mtg.getSeq(seed.getTid()), read.getSeq(), seed));
    }
    const SeedList seeds = SMEM_LIST::findSeeds(&read, &mtg, &fmdidx, 1, inf);
    CHECK(!seeds.empty());
    for (const SeedPair& seed : seeds) CHECK(isValidSeed(mtg.getSeq(seed.getTid()), read.getSeq(), seed));

    /* SMEMs: maximal, and each occurs exactly size times in the BD text */
    const SMEM_LIST smems = SMEM_LIST::findAllSMEMS(&read, &mtg, &fmdidx, 1, inf);
    CHECK(!smems.empty());
    const DNAseq bd = mtg.getBDSeq();
    for (const SMEM& s : smems) {
        const DNAseq pat = read.getSeq().substr(s.getFrom(), s.length());
        CHECK((int64_t)naiveFind(bd, pat).size() == s.getSize());
        if (s.getFrom() > 0) CHECK(naiveFind(bd, read.getSeq().substr(s.getFrom() - 1, s.length() + 1)).empty());
        if (s.getTo() < (int64_t)read.length()) CHECK(naiveFind(bd, read.getSeq().substr(s.getFrom(), s.length() + 1)).empty());
        double ll = 0;
        for (int64_t i = s.getFrom(); i < s.getTo(); ++i) ll += fmdidx.loglik(read.getBase(i));
        CHECK(std::fabs(ll - s.loglik()) < 1e-9);
    }

    /* counts, count / locateAll, accessSA */
    int64_t B[16] = {0};
    for (nt16_t b : bd) ++B[b];
    for (nt16_t b = 0; b < 16; ++b) CHECK(fmdidx.getBaseCount(b) == B[b]);
    for (const string p : {"ACG", "TACTACG", "GTA", "CGNA", "A", "TTTT"}) {
        const DNAseq pat = dna::encode(p);
        const vector<int64_t> occ = naiveFind(bd, pat);
        if (p.find('N') == string::npos) CHECK(fmdidx.count(pat) == (int64_t)occ.size());
        vector<int64_t> got;
        for (const GLoc& l : fmdidx.locateAllFwd(pat)) got.push_back(l.getStart());
        std::sort(got.begin(), got.end());
        if (p.find('N') == string::npos) CHECK(got == occ);
    }
    vector<int64_t> sa(bd.size());
    std::iota(sa.begin(), sa.end(), 0);
    /* the BD text's suffix order: '\0' ends compare as the text's own trailing 0 does */
    std::sort(sa.begin(), sa.end(), [&](int64_t a, int64_t b) { return bd.compare(a, DNAseq::npos, bd, b, DNAseq::npos) < 0; });
    int64_t saOk = 0;
    for (int64_t i = 0; i < (int64_t)bd.size(); ++i) saOk += fmdidx.accessSA(i) == sa[i];
    CHECK(saOk == (int64_t)bd.size());
    for (int64_t i = 0; i < (int64_t)bd.size(); ++i) {
        const nt16_t w = sa[i] == 0 ? 0 : bd[sa[i] - 1];
        CHECK(fmdidx.accessBWT(i) == w);
    }
    cout << "SMEM_test: " << mems.size() << " MEMs, " << seeds.size() << " seeds, " << smems.size() << " SMEMs" << endl;
}

static void seedChainTest() {
    const DNAseq chr1 = dna::encode("ACGTCGTAACGTGGTAACGTTGTA");
    const DNAseq chr2 = dna::encode("ACGTCGTAACGTGGTAACGTTGTA");
    MetaGenome mtg;
    mtg.addChrom("chr1", chr1).addChrom("chr2", chr2);
    FMDIndex fmdidx(mtg.getBDSeq());
    CHECK(mtg.BDSize() == fmdidx.length());

    const PrimarySeq read("ACGTAGTA", "seq1");
    const SeedList seeds = SMEM_LIST::findSeeds(&read, &mtg, &fmdidx, 4, inf);
    const ChainList chains = SeedChain::getChains(seeds, read.length() * 0.2, read.length() * 0.1);
    CHECK(!seeds.empty());
    CHECK(!chains.empty());
    size_t inChains = 0;
    for (const SeedChain& chain : chains) {
        inChains += chain.size();
        for (size_t k = 0; k < chain.size(); ++k) {
            if (k + 1 < chain.size()) CHECK(SeedPair::isCompatitable(chain[k], chain[k + 1]));
            CHECK(isValidSeed(mtg.getSeq(chain[k].getTid()), read.getSeq(), chain[k]));
        }
    }
    CHECK(inChains >= seeds.size());   /* every seed starts or joins a chain */
    cout << "SeedChain_test: " << seeds.size() << " seeds, " << chains.size() << " chains" << endl;
}

struct AlnCase {
    const char* read;
    const char* target;
    Alignment::MODE mode;
};

static void alignmentTest() {
    const AlnCase cases[] = {
        {"ATCG", "AAAATCGCCC", Alignment::LOCAL},                                  /* perfect match */
        {"TTTATCGAAA", "AAAATCGCCC", Alignment::LOCAL},                            /* match in the middle */
        {"AAAATCGATCGATCGATCGTTT", "GGGATCGATCAATCGATCGCCC", Alignment::LOCAL},    /* mismatch */
        {"AAAATCGATCAATCGATCGTTT", "AAAATCGATCATCGATCGTTT", Alignment::GLOBAL},    /* insertion */
        {"AAAATCGATCGTCGATCGTTT", "AAAATCGATCGATCGATCGTTT", Alignment::GLOBAL},    /* deletion */
    };
    int k = 0;
    for (const AlnCase& c : cases) {
        ++k;
        MetaGenome mtg;
        mtg.addChrom("target" + std::to_string(k), dna::encode(c.target));
        FMDIndex fmdidx(mtg.getBDSeq());
        const PrimarySeq read(c.read, "r" + std::to_string(k));
        const PrimarySeq rcRead = read.revcom();
        const SeedList seeds = SMEM_LIST::findSeeds(&read, &mtg, &fmdidx, 3, inf);
        const ChainList chains = SeedChain::getChains(seeds, read.length() * 0.2, read.length() * 0.1);
        ALIGN_LIST alns = Alignment::buildAlignments(&read, &rcRead, mtg, chains, c.mode);
        CHECK(!alns.empty());
        CHECK(alns.size() == chains.size());
        for (const Alignment& a : alns) {
            CHECK(cigar2QLen(a.getAlnCigarStr()) == (int64_t)read.length());
            CHECK(a.getAlnQLen() == Alignment::mdTag2alnQLen(a.getAlnMDTag()));
            if (c.mode == Alignment::GLOBAL) CHECK(a.getAlnFrom() == 0 && a.getAlnTo() == (int64_t)read.length());
        }
        /* the best alignment of each case, known by hand (scores 1 / 3 / 5 / 2; soft clips do not
         * enter the score; the two placements of case 4's insertion tie) */
        Alignment::calcMapQ(alns);
        Alignment::sort(alns);
        const Alignment& best = alns.front();
        for (size_t i = 1; i < alns.size(); ++i) CHECK(alns[i - 1].getLog10P() >= alns[i].getLog10P());
        double sumP = 0;
        for (const Alignment& a : alns) sumP += a.getPostP();
        CHECK(std::fabs(sumP - 1) < 1e-9);
        cout << "Alignment_test " << k << ": " << alns.size() << " alignments, best " << best.getAlnCigarStr()
             << " MD:Z:" << best.getAlnMDTag() << " score " << best.getAlnScore() << " mapQ " << (int)best.getMapQ()
             << endl;
        switch (k) {
        case 1: CHECK(best.getAlnCigarStr() == "4M" && best.getAlnMDTag() == "4" && best.getAlnScore() == 4); break;
        case 2:
            CHECK(best.getAlnCigarStr() == "3S4M3S" && best.getAlnMDTag() == "4" && best.getAlnScore() == 4);
            break;
        case 3:
            CHECK(best.getAlnCigarStr() == "3S16M3S" && best.getAlnMDTag() == "7A8" && best.getAlnScore() == 15 - 3);
            break;
        case 4:
            CHECK(((best.getAlnCigarStr() == "10M1I11M" && best.getAlnMDTag() == "10^A11") ||
                   (best.getAlnCigarStr() == "11M1I10M" && best.getAlnMDTag() == "11^A10")) &&
                  best.getAlnScore() == 21 - 7);
            break;
        case 5:
            CHECK(best.getAlnCigarStr() == "11M1D10M" && best.getAlnMDTag() == "21" && best.getAlnScore() == 21 - 7);
            break;
        }
    }
    /* mdTag2alnQLen on its own (src/Alignment.cpp:402-418) */
    CHECK(Alignment::mdTag2alnQLen("10^A11") == 22);
    CHECK(Alignment::mdTag2alnQLen("7A8") == 16);
    CHECK(Alignment::mdTag2alnQLen("0T0^GG3") == 6);
    CHECK(Alignment::mdTag2alnQLen("") == 0);
}

static void pairTest() {
    /* a 2 kb chromosome (fixed LCG) whose [100, 160) reappears at [1000, 1060) and whose
     * [300, 360) reappears at [1200, 1260); read 1 = [100, 160) forward, read 2 = the reverse
     * complement of [300, 360) */
    uint64_t x = 12345;
    string chrom(2000, 'A');
    for (char& c : chrom) { x = x * 6364136223846793005ULL + 1442695040888963407ULL; c = "ACGT"[(x >> 33) & 3]; }
    chrom.replace(1000, 60, chrom.substr(100, 60));
    chrom.replace(1200, 60, chrom.substr(300, 60));
    MetaGenome mtg;
    mtg.addChrom("pe", dna::encode(chrom));
    FMDIndex fmdidx(mtg.getBDSeq());
    const PrimarySeq r1(chrom.substr(100, 60), "p1");
    const PrimarySeq r2 = PrimarySeq(chrom.substr(300, 60), "p2").revcom();
    const PrimarySeq rc1 = r1.revcom(), rc2 = r2.revcom();
    auto align = [&](const PrimarySeq& r, const PrimarySeq& rc) {
        const SeedList seeds = SMEM_LIST::findSeeds(&r, &mtg, &fmdidx, 17, 0.01);
        const ChainList chains = SeedChain::getChains(seeds, r.length() * 0.2, r.length() * 0.1);
        ALIGN_LIST alns = Alignment::buildAlignments(&r, &rc, mtg, chains, Alignment::GLOBAL);
        Alignment::filter(alns, 0.65);
        return alns;
    };
    const ALIGN_LIST f = align(r1, rc1), g = align(r2, rc2);
    CHECK(f.size() == 2 && g.size() == 2);
    AlignmentPE::ps.setMean(250);
    AlignmentPE::ps.setSD(30);
    AlignmentPE::ps.updateParam();
    AlignmentPE::ps.updateRange();
    CHECK(AlignmentPE::ps.getMin() == -50 && AlignmentPE::ps.getMax() == 550);
    PAIR_LIST pairs = AlignmentPE::getPairs(f, g);
    CHECK(pairs.size() == 4);
    AlignmentPE::sort(pairs);
    for (size_t i = 1; i < pairs.size(); ++i) CHECK(pairs[i - 1].log10lik() >= pairs[i].log10lik());
    PAIR_LIST kept = pairs;
    AlignmentPE::filter(kept, false, false, false, false);
    CHECK(kept.size() == 2);
    for (const AlignmentPE& q : kept) {
        CHECK(q.getInsertSize() == 260 && q.isConcordant() && !q.isOverlap());
        CHECK(q.log10lik() == kept.front().log10lik());
    }
    AlignmentPE::calcMapQ(kept);
    CHECK(kept[0].getPostP() == 0.5 && kept[1].getPostP() == 0.5 && kept[0].getMapQ() == 3);
    /* the same lists through the device pairing */
    vector<int64_t> offs = {0, (int64_t)f.size(), (int64_t)(f.size() + g.size())}, rows;
    vector<double> lp;
    for (const ALIGN_LIST* l : {&f, &g})
        for (const Alignment& a : *l) {
            rows.insert(rows.end(), {a.getTid(), (int64_t)a.getStrand(), a.getAlnStart(), a.getAlnEnd()});
            lp.push_back(a.log10lik());
        }
    msgs_params p;
    msgs_params_default(&p);
    p.insMean = 250; p.insSD = 30; p.insMin = -50; p.insMax = 550;
    msgs_ctx* ctx = nullptr;
    CHECK(msgs_ctx_create(0, &ctx) == MSGS_OK);
    int64_t cnt = 0, need = 0, r3[3 * 4];
    double lik[2 * 4];
    CHECK(msgs_pair_batch(ctx, 1, offs.data(), rows.data(), lp.data(), &p, &cnt, r3, lik, 4, &need) == MSGS_OK);
    CHECK(cnt == (int64_t)kept.size());
    for (int64_t k = 0; k < cnt && k < (int64_t)kept.size(); ++k) {
        CHECK(&f[r3[3 * k]] == kept[k].fwdAln && &g[r3[3 * k + 1] - (int64_t)f.size()] == kept[k].revAln);
        CHECK(r3[3 * k + 2] == kept[k].getInsertSize());
        CHECK(lik[2 * k] == kept[k].log10lik() && lik[2 * k + 1] == kept[k].loglik());
    }
    msgs_ctx_destroy(ctx);
    cout << "pairing: " << pairs.size() << " candidates, " << kept.size() << " in range, insert size "
         << kept.front().getInsertSize() << ", mapQ " << (int)kept.front().getMapQ() << endl;
}

int main() {
    try {
        smemTest();
        seedChainTest();
        alignmentTest();
        pairTest();
    } catch (const std::exception& e) {
        cerr << "exception: " << e.what() << endl;
        return 2;
    }
    if (failures) {
        cerr << failures << " checks failed" << endl;
        return 1;
    }
    cout << "PASS" << endl;
    return 0;
}
