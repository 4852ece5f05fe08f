/*
 * msgs_oracle.cpp -- CPU oracle (TEST INFRASTRUCTURE ONLY, see msgs_oracle.hpp).
 *
 * A literal restatement of the reference MSGseqTK v1.2.3 read-mapping path.
 * Data structures are deliberately simple (full SA, 64-position occ
 * checkpoints, dense double DP matrices) so that each function can be read
 * against the reference line it cites.
 */
#include "msgs_oracle.hpp"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>

namespace orc {

static const double kInf = std::numeric_limits<double>::infinity();

/* ------------------------------------------------------------------ A1 */
/* src/DNAalphabet.cpp:14-152 (tables restated from their definitions) */
static const nt16_t kComp[16] = {0, 8, 4, 12, 2, 10, 6, 14, 1, 9, 5, 13, 3, 11, 7, 15};
static const nt16_t kBasic[16] = {0, 1, 2, 2, 4, 4, 4, 4, 8, 8, 8, 8, 8, 8, 8, 15};
static const char kSym[17] = "$ACMGRSVTWYHKDBN";

nt16_t encodeSym(char s) {
    /* sym2base: '0'..'4' -> A,C,G,T,N; IUPAC letters (either case); U -> T; else 0 */
    switch (s) {
    case '0': return A; case '1': return C; case '2': return G; case '3': return T; case '4': return N;
    default: break;
    }
    char u = (s >= 'a' && s <= 'z') ? char(s - 32) : s;
    switch (u) {
    case 'A': return A; case 'B': return C | G | T; case 'C': return C; case 'D': return A | G | T;
    case 'G': return G; case 'H': return A | C | T; case 'K': return G | T; case 'M': return A | C;
    case 'N': return N; case 'R': return A | G; case 'S': return G | C; case 'T': return T;
    case 'U': return T; case 'V': return A | C | G; case 'W': return A | T; case 'Y': return C | T;
    default: return 0;
    }
}
char decodeBase(nt16_t b) { return b < 16 ? kSym[b] : '='; }
nt16_t complement(nt16_t b) { return b < 16 ? kComp[b] : 0; }
nt16_t toBasic(nt16_t b) { return b < 16 ? kBasic[b] : 0; }
int toInt(nt16_t b) { return b == A ? 1 : b == C ? 2 : b == G ? 3 : b == T ? 4 : 0; }
bool isAmbiguous(nt16_t b) { return !(b == A || b == C || b == G || b == T); }
nt16_t encodeInt(int i) { return i == 1 ? A : i == 2 ? C : i == 3 ? G : i == 4 ? T : 0; }
/* quality::encode (src/QualStr.h:40-42): the difference is taken in uint8_t, then clamped at 2 */
uint8_t qualEncode(char c) { return std::max<uint8_t>((uint8_t)(c - 33), 2); }

/* ------------------------------------------------------------------ A3 */
/* BD text: seq + 0 + revcom(seq) + 0 after toBasic (src/MetaGenome.cpp:104-113);
 * BD locations per chrom: cumulative 2*(len+1) (src/MetaGenome.cpp:170-224). */
void buildMeta(Db& db, const nt16_t* seqs, const int64_t* lens, int nChrom,
               const std::vector<std::string>& names) {
    db.names = names;
    db.chroms.clear(); db.chromLen.clear(); db.bdStart.clear(); db.bdEnd.clear();
    int64_t off = 0, bdOff = 0, total = 0;
    for (int c = 0; c < nChrom; ++c) total += 2 * (lens[c] + 1);
    db.bd.assign(total, 0);
    for (int c = 0; c < nChrom; ++c) {
        const int64_t L = lens[c];
        std::vector<nt16_t> s(seqs + off, seqs + off + L);
        s.push_back(0);
        db.chroms.push_back(s);
        db.chromLen.push_back(L);
        db.bdStart.push_back(bdOff);
        db.bdEnd.push_back(bdOff + 2 * (L + 1));
        nt16_t* bd = db.bd.data() + bdOff;
        for (int64_t i = 0; i < L; ++i) bd[i] = toBasic(seqs[off + i]);
        bd[L] = 0;
        for (int64_t i = 0; i < L; ++i) bd[L + 1 + i] = complement(toBasic(seqs[off + L - 1 - i]));
        bd[2 * L + 1] = 0;
        off += L;
        bdOff += 2 * (L + 1);
    }
    db.n = total;
}

/* getTid: first chrom whose BD range holds pos (linear find_if, src/MetaGenome.h:474-479) */
int64_t getTid(const Db& db, int64_t pos) {
    for (size_t t = 0; t < db.bdStart.size(); ++t)
        if (db.bdStart[t] <= pos && pos < db.bdEnd[t]) return (int64_t)t;
    return -1;
}
/* getStrand (src/MetaGenome.h:161-167) */
int getStrand(const Db& db, int64_t tid, int64_t pos) {
    int64_t s = db.bdStart[tid], len = db.bdEnd[tid] - db.bdStart[tid];
    return (s <= pos && pos < s + len / 2) ? FWD : REV;
}
/* getLoc (src/MetaGenome.h:147-153) */
int64_t getLoc(const Db& db, int64_t tid, int64_t pos) {
    if (getStrand(db, tid, pos) == FWD) return pos - db.bdStart[tid];
    return db.bdEnd[tid] - pos - 1;
}

/* ------------------------------------------------------------------ A4 */
/* Suffix array of t[0..n) in byte order where a suffix that is a proper
 * prefix of another sorts first -- the order divsufsort produces
 * (src/FMDIndex.cpp:29-68).  Prefix doubling (Manber & Myers). */
std::vector<int64_t> suffixArray(const nt16_t* t, int64_t n) { return suffixArrayBlocks(t, n, {}); }

/* The order of msgseqtk-build's block-by-block build (src/msgseqtk-build.cpp:440-467): each
 * block indexed by divsufsort over its own text (src/FMDIndex.cpp:29-68), the block indexes
 * merged by FMDIndex::prepend (src/FMDIndex.cpp:267-308) with buildInterleavingBS
 * (:612-630), whose walk starts the rows of an earlier block's '\0' suffixes after all rows
 * of the later block's '\0' suffixes.  So: a suffix ends at its block's end (ends[] in text
 * order; empty = one block), and at a '\0' shared by suffixes of two blocks the later
 * block's suffix sorts first.  Ranks of round 0: '\0' of block b = k - 1 - b, base codes
 * above every '\0' rank; past a block's end -1. */
std::vector<int64_t> suffixArrayBlocks(const nt16_t* t, int64_t n, const std::vector<int64_t>& ends) {
    const int64_t k = ends.empty() ? 1 : (int64_t)ends.size();
    std::vector<int64_t> sa(n), rk(n), tmp(n), bend(n);
    for (int64_t i = 0, b = 0; i < n; ++i) {
        while (b + 1 < k && i >= ends[b]) ++b;
        bend[i] = ends.empty() ? n : ends[b];
        sa[i] = i;
        rk[i] = t[i] == 0 ? k - 1 - b : k + t[i];
    }
    if (n <= 1) return sa;
    for (int64_t h = 1;; h <<= 1) {
        auto key2 = [&](int64_t i) { return i + h < bend[i] ? rk[i + h] : (int64_t)-1; };
        std::sort(sa.begin(), sa.end(), [&](int64_t a, int64_t b) {
            if (rk[a] != rk[b]) return rk[a] < rk[b];
            return key2(a) < key2(b);
        });
        tmp[sa[0]] = 0;
        for (int64_t k = 1; k < n; ++k) {
            int64_t a = sa[k - 1], b = sa[k];
            tmp[b] = tmp[a] + ((rk[a] != rk[b] || key2(a) != key2(b)) ? 1 : 0);
        }
        rk.swap(tmp);
        if (rk[sa[n - 1]] == n - 1) break;
    }
    return sa;
}

std::vector<int64_t> buildBlockEnds(const std::vector<int64_t>& chromBDLen, int64_t blockSize) {
    const int64_t NC = (int64_t)chromBDLen.size();
    std::vector<int64_t> start(NC + 1, 0);
    for (int64_t i = 0; i < NC; ++i) start[i + 1] = start[i] + chromBDLen[i];
    std::vector<int64_t> ends;
    if (blockSize <= 0 || start[NC] < blockSize) return ends;   /* one step */
    auto len = [&](int64_t a, int64_t b) { return start[b] - start[a]; };
    int64_t blockEnd = NC;
    std::vector<int64_t> firsts;
    for (int64_t i = NC; i > 0; --i) {
        const int64_t blockStart = i - 1;
        if (blockStart == 0 || (len(blockStart, blockEnd) < blockSize && len(blockStart - 1, blockEnd) >= blockSize)) {
            firsts.push_back(blockStart);
            blockEnd = blockStart;
        }
    }
    for (size_t b = firsts.size(); b-- > 0;) ends.push_back(b == 0 ? start[NC] : start[firsts[b - 1]]);
    if (ends.size() == 1) ends.clear();
    return ends;
}

/* buildCounts (src/FMDIndex.cpp:118-134), buildBWT (:385-393), buildGap (:406-410). */
void buildIndex(Db& db, const int64_t* saIn) {
    const int64_t n = db.n;
    if (saIn) db.sa.assign(saIn, saIn + n);
    else db.sa = suffixArray(db.bd.data(), n);
    std::fill(db.B, db.B + 256, 0);
    std::fill(db.Cc, db.Cc + 256, 0);
    for (int64_t i = 0; i < n; ++i) db.B[db.bd[i]]++;
    int64_t S = 0;
    for (int i = 0; i <= 16; ++i) { db.Cc[i] = S; S += db.B[i]; }
    for (int b = 0; b < 16; ++b)
        if (db.B[b] != db.B[complement((nt16_t)b)])
            throw std::invalid_argument("input seq is not bi-directional");
    /* getExtBaseCount = C[15] - (A+C+G+T) - B[0] */
    if (db.Cc[15] - (db.B[A] + db.B[C] + db.B[G] + db.B[T]) - db.B[0] > 0)
        throw std::invalid_argument("input seq does not allow IUPAC-extended bases");
    db.bwt.assign(n, 0);
    for (int64_t i = 0; i < n; ++i) db.bwt[i] = db.sa[i] == 0 ? 0 : db.bd[db.sa[i] - 1];
    db.gapSA.assign(db.sa.begin(), db.sa.begin() + db.B[0]);
    /* occ checkpoints: counts of each symbol in bwt[0 .. 64k) */
    const int64_t nb = n / 64 + 1;
    db.occ.assign(nb * 16, 0);
    int64_t cnt[16] = {};
    for (int64_t i = 0; i < n; ++i) {
        if (i % 64 == 0) std::copy(cnt, cnt + 16, &db.occ[(i / 64) * 16]);
        cnt[db.bwt[i]]++;
    }
    if (n % 64 == 0) std::copy(cnt, cnt + 16, &db.occ[(n / 64) * 16]);
    /* loglik(b) = log(B[toBasic b]) - log(length()) (src/FMDIndex.h:195-197) */
    for (int b = 0; b < 16; ++b)
        db.loglikTab[b] = std::log((double)db.B[toBasic((nt16_t)b)]) - std::log((double)n);
}

/* ------------------------------------------------------------------ A5 */
/* rank(s, i) = #s in bwt[0..i] inclusive; size_t(-1) when i >= n
 * (src/libsds/src/WaveletTreeRRR.cpp:101-125). */
int64_t rank(const Db& db, nt16_t s, int64_t i) {
    if ((uint64_t)i >= (uint64_t)db.n) return -1;
    int64_t blk = i / 64;
    int64_t r = db.occ[blk * 16 + s];
    for (int64_t k = blk * 64; k <= i; ++k) r += db.bwt[k] == s;
    return r;
}

/* ------------------------------------------------------------------ A6 */
/* FMDIndex::backExt (src/FMDIndex.cpp:577-610) */
bool g_countExt = false;
int64_t g_nExt = 0;
int64_t backExt(const Db& db, int64_t& p, int64_t& q, int64_t& s, nt16_t b) {
    if (isAmbiguous(b)) return (s = 0);
    if (g_countExt) ++g_nExt;   /* single-threaded counting runs only (orc_count_ext) */
    int64_t sB[5], qB[5];
    int c = toInt(b);
    int64_t O = rank(db, b, p - 1);
    sB[c] = rank(db, b, p + s - 1) - O;
    if (sB[c] <= 0) return (s = sB[c]);
    if (sB[c] != s) {
        sB[0] = rank(db, 0, p + s - 1) - rank(db, 0, p - 1);
        for (int j = c + 1; j <= 4; ++j)
            sB[j] = rank(db, encodeInt(j), p + s - 1) - rank(db, encodeInt(j), p - 1);
        qB[0] = q;
        qB[4] = qB[0] + sB[0];
        for (int j = 4; j > c; --j) qB[j - 1] = qB[j] + sB[j];
        q = qB[c];
    }
    p = db.Cc[b] + O;
    s = sB[c];
    return s;
}
/* FMDIndex::fwdExt (src/FMDIndex.h:226-228) */
int64_t fwdExt(const Db& db, int64_t& p, int64_t& q, int64_t& s, nt16_t b) {
    return backExt(db, q, p, s, complement(b));
}
/* accessSA (src/FMDIndex.cpp:520-527): any SA sampling returns SA[i]. */
int64_t accessSA(const Db& db, int64_t i) { return db.sa[i]; }
double loglik(const Db& db, nt16_t b) { return db.loglikTab[b & 15]; }

/* ------------------------------------------------------------------ A2 */
Read makeRead(const std::string& name, const nt16_t* s, const uint8_t* q, int64_t L) {
    Read r;
    r.name = name;
    r.L = L;
    r.seq.assign(s, s + L);
    r.seq.push_back(0);
    r.qual.assign(q, q + L);
    return r;
}
/* PrimarySeq::revcom = reverse seq and qual, then complement (src/PrimarySeq.cpp:28-38) */
Read revcomRead(const Read& r) {
    Read o;
    o.name = r.name;
    o.L = r.L;
    o.seq.resize(r.L + 1, 0);
    o.qual.resize(r.L);
    for (int64_t i = 0; i < r.L; ++i) {
        o.seq[i] = complement(r.seq[r.L - 1 - i]);
        o.qual[i] = r.qual[r.L - 1 - i];
    }
    return o;
}

/* ------------------------------------------------------------------ A10 */
/* SMEM object and its incremental evaluation (src/SMEM.h:48-54,140-177; src/SMEM.cpp:19-40) */
struct SmemW { int64_t from, to, fwdStart, revStart, size; double logP; };

static SmemW newSmem(const Db& db, const Read& r, int64_t from, int64_t to, int64_t fs, int64_t rs, int64_t size) {
    SmemW m{from, to, fs, rs, size, 0};
    m.logP = 0;
    for (int64_t i = from; i < to; ++i) m.logP += loglik(db, r.seq[i]);
    return m;
}
static void smemFwdExt(const Db& db, const Read& r, SmemW& m) {
    fwdExt(db, m.fwdStart, m.revStart, m.size, r.seq[m.to]);   /* seq[L] is the 0 terminator */
    if (m.to < r.L) m.logP += loglik(db, r.seq[m.to]);
    m.to++;
}
static void smemBackExt(const Db& db, const Read& r, SmemW& m) {
    backExt(db, m.fwdStart, m.revStart, m.size, m.from > 0 ? r.seq[m.from - 1] : 0);
    if (m.from > 0) m.logP += loglik(db, r.seq[m.from - 1]);
    m.from--;
}
static double smemLoglik(const Db& db, const Read& r, const SmemW& m, int64_t from, int64_t to) {
    double ll = 0;
    if (from <= m.from && m.to <= to) {
        ll += m.logP;
        for (int64_t i = from; i < m.from; ++i) ll += loglik(db, r.seq[i]);
        for (int64_t i = m.to; i < to; ++i) ll += loglik(db, r.seq[i]);
    } else {
        for (int64_t i = from; i < to; ++i) ll += loglik(db, r.seq[i]);
    }
    return ll;
}
static double evalueOf(const Db& db, double ll) { return (double)db.n * std::exp(ll); }

/* ------------------------------------------------------------------ A11 */
/* SMEM_LIST::findFwdBackSMEMS (src/SMEM.cpp:64-117) */
static std::vector<SmemW> findFwdBack(const Db& db, const Read& r, int64_t& from, int64_t& to,
                                      int64_t minLen, double maxE, int64_t minSize = 1) {
    std::vector<SmemW> curr, prev;
    nt16_t b = r.seq[from];
    to = from + 1;
    if (isAmbiguous(b)) return curr;
    SmemW smem = newSmem(db, r, from, to, db.Cc[b], db.Cc[complement(b)], db.Cc[b + 1] - db.Cc[b]);
    SmemW smem0 = smem;
    while (smem.size >= minSize) {
        smemFwdExt(db, r, smem);
        if (smem.size != smem0.size && smem0.to >= minLen &&
            (!std::isfinite(maxE) || evalueOf(db, smemLoglik(db, r, smem0, 0, smem0.to)) <= maxE))
            prev.push_back(smem0);
        smem0 = smem;
    }
    to = smem.to;
    if (prev.empty()) { from--; return curr; }
    std::reverse(prev.begin(), prev.end());
    while (from >= 0 && !prev.empty()) {
        int64_t s0 = -1, nValid = 0;
        for (SmemW& m : prev) {
            smem0 = m;
            smemBackExt(db, r, m);
            if (m.size != smem0.size && smem0.size >= minSize && nValid == 0 &&
                smem0.to - smem0.from >= minLen && evalueOf(db, smem0.logP) <= maxE)
                curr.push_back(smem0);
            if (m.size >= minSize && m.size != s0) { s0 = m.size; nValid++; }
            else m.size = 0;
        }
        prev.erase(std::remove_if(prev.begin(), prev.end(), [](const SmemW& m) { return !(m.size > 0); }),
                   prev.end());
        from--;
    }
    return curr;
}

/* ------------------------------------------------------------------ A12 */
/* SMEM_LIST::findBackFwdSMEMS (src/SMEM.cpp:119-172) */
static std::vector<SmemW> findBackFwd(const Db& db, const Read& r, int64_t& from, int64_t& to,
                                      int64_t minLen, double maxE, int64_t minSize = 1) {
    const uint64_t L = (uint64_t)r.L;
    std::vector<SmemW> curr, prev;
    nt16_t b = r.seq[from];
    to = from + 1;
    if (isAmbiguous(b)) return curr;
    SmemW smem = newSmem(db, r, from, to, db.Cc[b], db.Cc[complement(b)], db.Cc[b + 1] - db.Cc[b]);
    SmemW smem0 = smem;
    while (smem.size >= minSize) {
        smemBackExt(db, r, smem);
        if (smem.size != smem0.size && L - (uint64_t)smem0.from >= (uint64_t)minLen &&
            (!std::isfinite(maxE) || evalueOf(db, smemLoglik(db, r, smem0, smem0.from, (int64_t)L)) <= maxE))
            prev.push_back(smem0);
        smem0 = smem;
    }
    from = smem.from;
    if (prev.empty()) { to++; return curr; }
    std::reverse(prev.begin(), prev.end());
    while ((uint64_t)to <= L && !prev.empty()) {
        int64_t s0 = -1, nValid = 0;
        for (SmemW& m : prev) {
            smem0 = m;
            smemFwdExt(db, r, m);
            if (m.size != smem0.size && smem0.size >= minSize && nValid == 0 &&
                smem0.to - smem0.from >= minLen && evalueOf(db, smem0.logP) <= maxE)
                curr.push_back(smem0);
            if (m.size >= minSize && m.size != s0) { s0 = m.size; nValid++; }
            else m.size = 0;
        }
        prev.erase(std::remove_if(prev.begin(), prev.end(), [](const SmemW& m) { return !(m.size > 0); }),
                   prev.end());
        to++;
    }
    return curr;
}

/* ------------------------------------------------------------------ A13 */
/* SMEM_LIST::findAllSMEMS (src/SMEM.cpp:186-215).  An empty read yields no
 * SMEMs (the reference asserts from < L). */
std::vector<Smem> findAllSMEMS(const Db& db, const Read& r, int64_t minLen, double maxE) {
    std::vector<Smem> out;
    const int64_t L = r.L;
    if (L == 0) return out;
    std::vector<SmemW> curr;
    auto add = [&](const std::vector<SmemW>& v) { curr.insert(curr.end(), v.begin(), v.end()); };
    int64_t from = L / 2;
    int64_t to = from + 1;
    add(findFwdBack(db, r, from, to, minLen, maxE));
    while (from > 0 || to < L) {
        int64_t from0 = to;
        if (from0 < L) add(findFwdBack(db, r, from0, to, minLen, maxE));
        int64_t to0 = --from + 1;
        if (from >= 0) add(findBackFwd(db, r, from, to0, minLen, maxE));
    }
    for (const SmemW& m : curr) out.push_back(Smem{m.from, m.to, m.fwdStart, m.revStart, m.size, m.logP});
    return out;
}

/* SMEM::findMEM / SMEM_LIST::findMEMS (src/SMEM.cpp:42-62,174-184) */
std::vector<Smem> findMEMS(const Db& db, const Read& r, int64_t minLen, double maxE) {
    std::vector<Smem> out;
    const int64_t L = r.L;
    for (int64_t from = 0, to = 1; from < L; from = to + 1) {
        nt16_t b = r.seq[from];
        SmemW mem = newSmem(db, r, from, from + 1, db.Cc[b], db.Cc[complement(b)], db.Cc[b + 1] - db.Cc[b]);
        SmemW mem0 = mem;
        while (smemFwdExt(db, r, mem), mem.size > 0) mem0 = mem;
        to = mem0.to;
        mem0 = mem;
        while (smemBackExt(db, r, mem), mem.size > 0) mem0 = mem;
        from = mem0.from;
        if (mem0.to - mem0.from >= minLen && evalueOf(db, mem0.logP) <= maxE)
            out.push_back(Smem{mem0.from, mem0.to, mem0.fwdStart, mem0.revStart, mem0.size, mem0.logP});
    }
    return out;
}

/* ------------------------------------------------------------------ A14/A15 */
/* SMEM::getSeeds (src/SMEM.cpp:217-269) */
void getSeeds(const Db& db, const Read& r, const Smem& m, std::vector<Seed>& out) {
    const int64_t len = m.to - m.from;
    for (int64_t i = 0; i < m.size; ++i) {
        {
            int64_t bd = accessSA(db, m.fwdStart + i);
            int64_t tid = getTid(db, bd);
            int64_t start = getLoc(db, tid, bd);
            if (getStrand(db, tid, bd) == FWD)
                out.push_back(Seed{m.from, m.from + len, start, start + len, tid, FWD, m.logP});
        }
        {
            int64_t bd = accessSA(db, m.revStart + i);
            int64_t tid = getTid(db, bd);
            int64_t start = getLoc(db, tid, bd);
            if (getStrand(db, tid, bd) == FWD) {
                int64_t f = r.L - m.to;
                out.push_back(Seed{f, f + len, start, start + len, tid, REV, m.logP});
            }
        }
    }
}

/* SMEM_LIST::findSeeds (src/SMEM.cpp:271-305) */
std::vector<Seed> findSeeds(const Db& db, const Read& r, int64_t minLen, double maxE, int64_t maxNSeed) {
    std::vector<Smem> smems = findAllSMEMS(db, r, minLen, maxE);
    std::sort(smems.begin(), smems.end(), [](const Smem& a, const Smem& b) { return a.logP < b.logP; });
    std::vector<Seed> all;
    for (const Smem& m : smems) {     /* SAcached only saves work; values are identical */
        getSeeds(db, r, m, all);
        if ((uint64_t)all.size() >= (uint64_t)maxNSeed) break;
    }
    removeRedundant(all);
    return all;
}

/* ------------------------------------------------------------------ A16/A17 */
/* SeedPair operator< (src/SeedPair.h:225-232) */
static bool seedLess(const Seed& l, const Seed& r) {
    return l.from != r.from ? l.from < r.from :
           l.to != r.to ? l.to < r.to :
           l.tid != r.tid ? l.tid < r.tid :
           l.strand != r.strand ? l.strand < r.strand :
           l.start != r.start ? l.start < r.start : l.end < r.end;
}
static bool containing(const Seed& l, const Seed& r) {   /* src/SeedPair.h:171-175 */
    return l.tid == r.tid && l.strand == r.strand && l.from <= r.from && l.to >= r.to &&
           l.start <= r.start && l.end >= r.end;
}
/* SeedPair::removeRedundant (src/SeedPair.cpp:71-88) */
void removeRedundant(std::vector<Seed>& seeds) {
    if (seeds.size() <= 1) return;
    std::sort(seeds.begin(), seeds.end(), seedLess);
    for (size_t i = 0; i < seeds.size() - 1; ++i)
        for (size_t j = i + 1; j < seeds.size(); ++j)
            if (containing(seeds[i], seeds[j]) || containing(seeds[j], seeds[i])) {
                seeds[i] = Seed{0, 0, 0, 0, -1, UNK, 0};
                break;
            }
    seeds.erase(std::remove_if(seeds.begin(), seeds.end(), [](const Seed& s) { return s.to - s.from == 0; }),
                seeds.end());
}

/* ------------------------------------------------------------------ A18 */
static int64_t locDist(int64_t ls, int64_t le, int64_t rs, int64_t re) {   /* src/Loc.h:85-91 */
    bool ov = ls < re && le > rs;
    return ov ? 0 : ls < rs ? rs - le + 1 : ls - re + 1;
}
static bool compatible(const Seed& l, const Seed& r) {   /* src/SeedPair.h:156-158 */
    return l.to <= r.from && l.end <= r.start && l.tid == r.tid && l.strand == r.strand;
}
static int64_t nMismatch(const Seed& l, const Seed& r) {
    if (!compatible(l, r)) return INT64_MAX;
    return std::min(locDist(l.from, l.to, r.from, r.to), locDist(l.start, l.end, r.start, r.end));
}
static int64_t nIndel(const Seed& l, const Seed& r) {
    if (!compatible(l, r)) return INT64_MAX;
    return (r.from - l.to) - (r.start - l.end);
}
/* SeedChain::getChains / dfsSeeds (src/SeedChain.cpp:40-73) */
static void dfs(const std::vector<Seed>& in, size_t i, std::vector<std::vector<Seed>>& out,
                std::vector<size_t>& stack, std::vector<bool>& seen, int64_t maxMM, int64_t maxIndel) {
    if (seen[i]) return;
    stack.push_back(i);
    seen[i] = true;
    bool leaf = true;
    for (size_t j = i + 1; j < in.size(); ++j) {
        if (!seen[j] && nMismatch(in[i], in[j]) <= maxMM && std::abs(nIndel(in[i], in[j])) <= maxIndel) {
            leaf = false;
            dfs(in, j, out, stack, seen, maxMM, maxIndel);
        }
    }
    if (leaf) {
        std::vector<Seed> chain;
        for (size_t k : stack) chain.push_back(in[k]);
        out.push_back(chain);
    }
    stack.pop_back();
}
std::vector<std::vector<Seed>> getChains(const std::vector<Seed>& seeds, int64_t maxMM, int64_t maxIndel) {
    std::vector<bool> seen(seeds.size());
    std::vector<std::vector<Seed>> out;
    for (size_t i = 0; i < seeds.size(); ++i) {
        std::vector<size_t> stack;
        dfs(seeds, i, out, stack, seen, maxMM, maxIndel);
    }
    return out;
}

/* ------------------------------------------------------------------ A19-A24 */
/* ScoreScheme (src/ScoreScheme.cpp:20-24, src/ScoreScheme.h:81-96) */
static double score(const Params& P, nt16_t a, nt16_t b) {
    return ((a & b) && a != N && b != N) ? P.match : -P.mismatch;
}
static double gapPenalty(const Params& P, uint32_t l) { return l == 0 ? 0 : P.gapO + P.gapE * l; }

enum { OP_M = 0, OP_I = 1, OP_D = 2, OP_S = 4, OP_EQ = 7, OP_X = 8 };

/* dense DP state of one Alignment (src/Alignment.h:366-368: three (qLen+1)x(tLen+1) matrices) */
struct Dense {
    int64_t rows = 0, cols = 0;
    std::vector<double> M, I, D;
    int hazard = 0;
    void init(int64_t r, int64_t c) {
        rows = r; cols = c;
        M.assign(r * c, -kInf); I.assign(r * c, -kInf); D.assign(r * c, -kInf);
    }
    /* reads outside the matrix are undefined in the reference; flagged and read as -inf */
    double get(const std::vector<double>& X, int64_t i, int64_t j) {
        if (i < 0 || j < 0 || i >= rows || j >= cols) { hazard = 1; return -kInf; }
        return X[i * cols + j];
    }
    double& at(std::vector<double>& X, int64_t i, int64_t j) { return X[i * cols + j]; }
};

struct AlnCtx {
    const Params* P;
    Aln* a;
    Dense d;
    const std::vector<nt16_t>* query;
    int64_t cells = 0;
};

static inline nt16_t tgt(const Aln& a, int64_t t) {
    const std::vector<nt16_t>& s = *a.target;
    return (t >= 0 && t < (int64_t)s.size()) ? s[t] : 0;
}

/* calculateScoresNW / calculateScoresSW (src/Alignment.cpp:72-133) */
static void rectDP(AlnCtx& c, int64_t from, int64_t to, int64_t start, int64_t end) {
    Aln& a = *c.a;
    const double o = -(c.P->gapO + c.P->gapE), e = -c.P->gapE;
    const bool sw = c.P->mode == 1;
    for (int64_t t = start; t < end; ++t) {
        int64_t j = t - a.tStart + 1;
        for (int64_t q = from; q < to; ++q) {
            int64_t i = q - a.qFrom + 1;
            double s = score(*c.P, (*c.query)[q], tgt(a, t));
            double m = std::max({c.d.at(c.d.M, i - 1, j - 1) + s, c.d.at(c.d.D, i - 1, j - 1) + s,
                                 c.d.at(c.d.I, i - 1, j - 1) + s});
            if (sw) m = std::max(m, 0.0);
            c.d.at(c.d.M, i, j) = m;
            c.d.at(c.d.I, i, j) = std::max(c.d.at(c.d.M, i - 1, j) + o, c.d.at(c.d.I, i - 1, j) + e);
            c.d.at(c.d.D, i, j) = std::max(c.d.at(c.d.M, i, j - 1) + o, c.d.at(c.d.D, i, j - 1) + e);
            c.cells++;
        }
    }
}

/* calculateScores(SeedPair) (src/Alignment.cpp:135-171) */
static void seedDP(AlnCtx& c, const Seed& sd) {
    Aln& a = *c.a;
    const double o = -(c.P->gapO + c.P->gapE);
    const double s = c.P->match;
    const int64_t len = sd.to - sd.from;
    int64_t bi = sd.from - a.qFrom + 1, bj = sd.start - a.tStart + 1;
    for (int64_t x = 0; x < len; ++x)
        for (int64_t y = 0; y < len; ++y) c.d.at(c.d.M, bi + x, bj + y) = -kInf;
    for (int64_t k = 0; k < len; ++k) {
        int64_t i = sd.from + k - a.qFrom + 1, j = sd.start + k - a.tStart + 1;
        if (k == 0)
            c.d.at(c.d.M, i, j) = std::max({c.d.at(c.d.M, i - 1, j - 1) + s, c.d.at(c.d.D, i - 1, j - 1) + s,
                                            c.d.at(c.d.I, i - 1, j - 1) + s, 0.0});
        else
            c.d.at(c.d.M, i, j) = c.d.at(c.d.M, i - 1, j - 1) + s;
        c.cells++;
    }
    int64_t i = sd.to - a.qFrom + 1, j = sd.end - a.tStart + 1;
    const int64_t qLen = a.qTo - a.qFrom, tLen = a.tEnd - a.tStart;
    if (i <= qLen && j <= tLen) { c.d.at(c.d.M, i, j) = c.d.at(c.d.M, i - 1, j - 1) + s; c.cells++; }
    if (i <= qLen)
        for (int64_t y = 0; y < c.d.cols; ++y) c.d.at(c.d.I, i, y) = c.d.at(c.d.M, i - 1, y) + o;
    if (j <= tLen)
        for (int64_t x = 0; x < c.d.rows; ++x) c.d.at(c.d.D, x, j) = c.d.at(c.d.M, x, j - 1) + o;
}

/* matchMax / insMax / delMax (src/Alignment.h:577-597), including the matchMax quirk */
static uint32_t matchMax(double match, double ins, double del) {
    double mx = match;
    uint32_t s = OP_M;
    if (ins > mx) { mx = del; s = OP_I; }
    if (del > mx) s = OP_D;
    return s;
}
static uint32_t insMax(double match, double ins) { return match > ins ? OP_M : OP_I; }
static uint32_t delMax(double match, double del) { return match > del ? OP_M : OP_D; }

/* backTraceNW (src/Alignment.cpp:173-233) */
static void backTraceNW(AlnCtx& c) {
    Aln& a = *c.a;
    Dense& d = c.d;
    int64_t alnTo = d.rows - 1;
    double best = -kInf;
    int64_t alnEnd = 0;
    uint32_t s = OP_M;
    for (int64_t j = 1; j < d.cols; ++j) {
        if (d.at(d.M, alnTo, j) > best) { best = d.at(d.M, alnTo, j); alnEnd = j; s = OP_M; }
        if (d.at(d.I, alnTo, j) > best) { best = d.at(d.I, alnTo, j); alnEnd = j; s = OP_I; }
    }
    a.alnScore = best;
    a.alnTo = alnTo + a.qFrom;
    a.alnEnd = alnEnd + a.tStart;
    const double o = c.P->gapO, e = c.P->gapE;
    a.path.clear();
    int64_t i = a.alnTo - a.qFrom, j = a.alnEnd - a.tStart;
    while (i > 0) {
        a.path.push_back(s);
        switch (s) {
        case OP_M: s = matchMax(d.get(d.M, i - 1, j - 1), d.get(d.I, i - 1, j - 1), d.get(d.D, i - 1, j - 1)); i--; j--; break;
        case OP_I: s = insMax(d.get(d.M, i - 1, j) - o - e, d.get(d.I, i - 1, j) - e); i--; break;
        case OP_D: s = delMax(d.get(d.M, i, j - 1) - o - e, d.get(d.D, i, j - 1) - e); j--; break;
        default: break;
        }
        if (d.hazard && j < -1) break;   /* undefined territory: stop walking */
    }
    a.alnFrom = a.qFrom + i;
    a.alnStart = a.tStart + j;
    std::reverse(a.path.begin(), a.path.end());
}

/* backTraceSW (src/Alignment.cpp:235-276); maxCoeff = first maximum in column-major order */
static void backTraceSW(AlnCtx& c) {
    Aln& a = *c.a;
    Dense& d = c.d;
    double best = -kInf;
    int64_t bi = 0, bj = 0;
    bool first = true;
    for (int64_t j = 0; j < d.cols; ++j)
        for (int64_t i = 0; i < d.rows; ++i) {
            double v = d.at(d.M, i, j);
            if (first || v > best) { best = v; bi = i; bj = j; first = false; }
        }
    a.alnScore = best;
    a.alnTo = bi + a.qFrom;
    a.alnEnd = bj + a.tStart;
    const double o = c.P->gapO, e = c.P->gapE;
    a.path.clear();
    int64_t i = a.alnTo - a.qFrom, j = a.alnEnd - a.tStart;
    uint32_t s = OP_M;
    while (s != OP_M || d.get(d.M, i, j) > 0) {
        a.path.push_back(s);
        switch (s) {
        case OP_M: s = matchMax(d.get(d.M, i - 1, j - 1), d.get(d.I, i - 1, j - 1), d.get(d.D, i - 1, j - 1)); i--; j--; break;
        case OP_I: s = insMax(d.get(d.M, i - 1, j) - o - e, d.get(d.I, i - 1, j) - e); i--; break;
        case OP_D: s = delMax(d.get(d.M, i, j - 1) - o - e, d.get(d.D, i, j - 1) - e); j--; break;
        default: break;
        }
        if (d.hazard && (i < -1 || j < -1)) break;   /* undefined territory: stop walking */
    }
    a.alnFrom = a.qFrom + i;
    a.alnStart = a.tStart + j;
    std::reverse(a.path.begin(), a.path.end());
}

static void initMatrices(AlnCtx& c) {   /* initNW / initSW (src/Alignment.h:599-613) */
    Aln& a = *c.a;
    const int64_t qLen = a.qTo - a.qFrom, tLen = a.tEnd - a.tStart;
    c.d.init(qLen + 1, tLen + 1);
    for (int64_t j = 0; j < c.d.cols; ++j) c.d.at(c.d.M, 0, j) = 0;
    if (c.P->mode == 0) {
        for (int64_t i = 0; i < c.d.rows; ++i) c.d.at(c.d.I, i, 0) = -gapPenalty(*c.P, (uint32_t)i);
    } else {
        for (int64_t i = 0; i < c.d.rows; ++i) c.d.at(c.d.M, i, 0) = 0;
    }
}

/* Alignment(read, rcRead, mtg, chain) + calculateScores(chain) + backTrace
 * (src/Alignment.h:84-94, src/Alignment.cpp:46-70,420-427) */
Aln alignChain(const Db& db, const Params& P, const Read& r, const Read& rc, const std::vector<Seed>& chain) {
    Aln a;
    a.read = &r; a.rcRead = &rc;
    a.tid = (int32_t)chain.front().tid;
    a.target = &db.chroms[a.tid];
    a.qStrand = chain.front().strand;
    a.qFrom = 0; a.qTo = r.L;
    const int64_t cFrom = chain.front().from, cTo = chain.back().to;
    const int64_t cStart = chain.front().start, cEnd = chain.back().end;
    const double kIndel = 1 + 0.1;   /* 1 + MAX_INDEL_RATE */
    a.tStart = std::max<int64_t>(0, (int64_t)(cStart - kIndel * (cFrom - a.qFrom)));
    a.tEnd = std::min<int64_t>(db.chromLen[a.tid] + 1, (int64_t)(cEnd + kIndel * (a.qTo - cTo)));
    AlnCtx c;
    c.P = &P; c.a = &a;
    c.query = a.qStrand == FWD ? &r.seq : &rc.seq;
    initMatrices(c);
    const int64_t qTo = a.qTo;
    if (a.qFrom < cFrom && a.tStart < cStart) rectDP(c, a.qFrom, cFrom, a.tStart, cStart);
    for (size_t k = 0; k < chain.size(); ++k) {
        seedDP(c, chain[k]);
        if (k + 1 < chain.size()) rectDP(c, chain[k].to, chain[k + 1].from, chain[k].end, chain[k + 1].start);
    }
    if (cTo < qTo && cEnd < a.tEnd) rectDP(c, cTo, qTo, cEnd, a.tEnd);
    if (P.mode == 0) backTraceNW(c); else backTraceSW(c);
    a.hazard = c.d.hazard;
    a.dpCells = c.cells;
    return a;
}

Aln alignFull(const Params& P, const Read& r, const Read& rc, const std::vector<nt16_t>& target, int mode) {
    Params Q = P;
    Q.mode = mode;
    Aln a;
    a.read = &r; a.rcRead = &rc; a.tid = 0; a.target = &target; a.qStrand = FWD;
    a.qFrom = 0; a.qTo = r.L; a.tStart = 0; a.tEnd = (int64_t)target.size() - 1;  /* target carries a 0 terminator */
    AlnCtx c;
    c.P = &Q; c.a = &a; c.query = &r.seq;
    initMatrices(c);
    rectDP(c, a.qFrom, a.qTo, 0, a.tEnd - a.tStart);
    if (mode == 0) backTraceNW(c); else backTraceSW(c);
    a.hazard = c.d.hazard;
    a.dpCells = c.cells;
    return a;
}

/* ------------------------------------------------------------------ A25 */
/* getAlnCigar (src/Alignment.cpp:278-303) */
std::string cigarString(const Aln& a) {
    static const char kOps[] = "MIDNSHP=XB";
    std::string s;
    auto put = [&](uint32_t len, uint32_t op) { s += std::to_string(len); s += kOps[op]; };
    int64_t clip5 = a.alnFrom - a.qFrom, clip3 = a.qTo - a.alnTo;
    if ((int32_t)clip5 > 0) put((uint32_t)clip5, OP_S);
    if (!a.path.empty()) {
        uint32_t op = a.path[0], len = 1;
        for (size_t i = 1; i < a.path.size(); ++i) {
            if (a.path[i] != op) { put(len, op); len = 1; }
            else len++;
            op = a.path[i];
        }
        put(len, op);
    }
    if ((int32_t)clip3 > 0) put((uint32_t)clip3, OP_S);
    return s;
}

/* getAlnMDTag (src/Alignment.cpp:305-353) */
std::string mdTag(const Aln& a) {
    const std::vector<nt16_t>& query = a.qStrand == FWD ? a.read->seq : a.rcRead->seq;
    std::string md;
    uint32_t prev = (uint32_t)-1, len = 0;
    int64_t i = a.alnFrom, j = a.alnStart;
    for (size_t k = 0; k < a.path.size(); ++k) {
        nt16_t qb = (i >= 0 && i < (int64_t)query.size()) ? query[i] : 0;
        nt16_t tb = tgt(a, j);
        uint32_t op = a.path[k];
        if (op == OP_M) op = (qb & tb) ? OP_EQ : OP_X;
        switch (op) {
        case OP_EQ: i++; j++; len++; break;
        case OP_X: md += std::to_string(len); len = 0; md.push_back(decodeBase(tb)); i++; j++; break;
        case OP_I:
            if (prev != OP_I) md += std::to_string(len);
            len = 0;
            if (op != prev) md.push_back('^');
            md.push_back(decodeBase(qb));
            i++;
            break;
        case OP_D: j++; break;
        default: break;
        }
        prev = op;
    }
    md += std::to_string(len);
    return md;
}

/* ------------------------------------------------------------------ A28 */
/* Alignment::evaluate (src/Alignment.cpp:446-489), phredQ2P (src/QualStr.h:109-112) */
void evaluate(const Params& P, Aln& a) {
    if (a.path.empty()) return;
    a.log10P = 0;
    const Read& q = a.qStrand == FWD ? *a.read : *a.rcRead;
    for (int64_t i = a.qFrom; i < a.alnFrom; ++i) a.log10P += -P.clip;
    int64_t i = a.alnFrom, j = a.alnStart;
    for (size_t k = 0; k < a.path.size(); ++k) {
        uint32_t s = a.path[k];
        double qv = (i >= 0 && i < q.L) ? (double)q.qual[i] : 0.0;
        switch (s) {
        case OP_M:
            if (q.seq[i] & tgt(a, j)) a.log10P += ::log10(1 - ::exp(std::log(10.0) * qv / -10.0));
            else a.log10P += qv / -10.0;
            i++; j++;
            break;
        case OP_I:
            if (k == 0 || a.path[k - 1] != OP_I) a.log10P += qv / -10.0 - P.gapO;
            a.log10P += -P.gapE;
            i++;
            break;
        case OP_D:
            if (k == 0 || a.path[k - 1] != OP_D) a.log10P += qv / -10.0 - P.gapO;
            a.log10P += -P.gapE;
            j++;
            break;
        default: break;
        }
    }
    for (int64_t x = a.alnTo; x < a.qTo; ++x) a.log10P += -P.clip;
}

/* ------------------------------------------------------------------ A29 */
/* Eigen sum of a VectorXd with 2-wide packets (SSE2): two packet accumulators
 * over multiples of 4, one more packet, horizontal add, then the scalar tail. */
static double eigenSum(const std::vector<double>& v) {
    const size_t n = v.size();
    if (n == 0) return 0;
    const size_t aSize2 = (n / 4) * 4, aSize = (n / 2) * 2;
    if (aSize == 0) return v[0];
    double p0a = v[0], p0b = v[1];
    if (aSize > 2) {
        double p1a = v[2], p1b = v[3];
        for (size_t k = 4; k < aSize2; k += 4) {
            p0a += v[k]; p0b += v[k + 1];
            p1a += v[k + 2]; p1b += v[k + 3];
        }
        p0a += p1a; p0b += p1b;
        if (aSize > aSize2) { p0a += v[aSize2]; p0b += v[aSize2 + 1]; }
    }
    double res = p0a + p0b;
    for (size_t k = aSize; k < n; ++k) res += v[k];
    return res;
}

static double alnLoglik(const Aln& a) { return a.log10P / ::log10(M_E); }

/* Alignment::calcMapQ (src/Alignment.cpp:491-512) */
void calcMapQ(std::vector<Aln>& alns) {
    if (alns.empty()) return;
    const size_t n = alns.size();
    std::vector<double> pr(n);
    for (size_t i = 0; i < n; ++i) pr[i] = alnLoglik(alns[i]);
    double maxV = *std::max_element(pr.begin(), pr.end());
    double scale = (maxV != -kInf && maxV < -256) ? -256 - maxV : 0;
    for (size_t i = 0; i < n; ++i) pr[i] = std::exp(pr[i] + scale);
    double sum = eigenSum(pr);
    for (size_t i = 0; i < n; ++i) {
        double postP = pr[i] / sum;
        alns[i].postP = postP;
        double mq = -10 * ::log10(1 - postP);
        alns[i].mapQ = (uint8_t)std::min(::round(mq), 250.0);
    }
}

/* ------------------------------------------------------------------ A30 */
struct Pair { const Aln* f; const Aln* r; int32_t isize; double postP; uint8_t mapQ; };

/* normal pdf as boost::math::pdf(normal(m, s), x) computes it */
static double normPdf(double m, double s, double x) {
    double ex = x - m;
    ex *= -ex;
    ex /= 2 * s * s;
    double r = std::exp(ex);
    r /= s * 2.506628274631000502415765284811045253e+00;
    return r;
}
/* PairingScheme::pr (src/PairingScheme.cpp:18-25) */
static double pairPr(const Params& P, double x) {
    if (P.insMean == 0) return 1;
    if (P.insMin <= x && x <= P.insMax) return normPdf(P.insMean, P.insSD, x);
    return 0;
}
static double pairLog10lik(const Params& P, const Pair& p) {
    return p.f->log10P + p.r->log10P + std::log10(pairPr(P, p.isize));
}
static double pairLoglik(const Params& P, const Pair& p) {
    return alnLoglik(*p.f) + alnLoglik(*p.r) + std::log(pairPr(P, p.isize));
}

/* AlignmentPE::getPairs (src/Alignment.cpp:573-581), AlignmentPE::sort (src/Alignment.h:635-639)
 * and AlignmentPE::filter (src/Alignment.cpp:556-571) over two filtered, evaluated lists */
static std::vector<Pair> candidatePairs(const Params& P, const std::vector<Aln>& a1, const std::vector<Aln>& a2) {
    std::vector<Pair> pairs;
    for (const Aln& f : a1)
        for (const Aln& r : a2)
            if ((f.qStrand & r.qStrand) == 0) {
                int32_t isize = f.tid != r.tid ? -1 :
                    (int32_t)(std::max(f.alnEnd, r.alnEnd) - std::min(f.alnStart, r.alnStart));
                pairs.push_back(Pair{&f, &r, isize, 0, 0xff});
            }
    std::sort(pairs.begin(), pairs.end(),
              [&](const Pair& l, const Pair& r) { return pairLog10lik(P, l) > pairLog10lik(P, r); });
    auto rm = [&](std::function<bool(const Pair&)> f) {
        pairs.erase(std::remove_if(pairs.begin(), pairs.end(), f), pairs.end());
    };
    rm([&](const Pair& p) { return !(P.insMin <= p.isize && p.isize <= P.insMax); });
    if (P.noDiscordant) rm([](const Pair& p) { return !(p.f->tid == p.r->tid && (p.f->qStrand & p.r->qStrand) == 0); });
    if (P.noTailOver) rm([](const Pair& p) {
        return (p.f->alnStart <= p.r->alnStart && p.f->alnEnd > p.r->alnStart) ||
               (p.r->alnStart <= p.f->alnStart && p.r->alnEnd > p.f->alnStart); });
    if (P.noContain) rm([](const Pair& p) {
        return (p.f->alnStart <= p.r->alnStart && p.f->alnEnd >= p.r->alnEnd) ||
               (p.r->alnStart <= p.f->alnStart && p.r->alnEnd >= p.f->alnEnd); });
    if (P.noOverlap) rm([](const Pair& p) { return p.f->alnStart < p.r->alnEnd && p.f->alnEnd > p.r->alnStart; });
    if (P.maxNPair > 0 && (int64_t)pairs.size() > P.maxNPair) pairs.erase(pairs.begin() + P.maxNPair, pairs.end());
    return pairs;
}

void pairAlignments(const Params& P, const std::vector<Aln>& a1, const std::vector<Aln>& a2,
                    std::vector<PairOut>& out) {
    out.clear();
    for (const Pair& p : candidatePairs(P, a1, a2))
        out.push_back(PairOut{(int64_t)(p.f - a1.data()), (int64_t)(p.r - a2.data()), p.isize, pairLog10lik(P, p),
                              pairLoglik(P, p)});
}

/* ------------------------------------------------------------------ A31 */
struct Rec {
    std::string qname; int flag; int32_t tid; int64_t pos; int mapq; std::string cigar;
    int32_t isize; const Read* q;
    bool hasTags = false;
    long as = 0; std::string md; float xh = 0; long nh = 0, xn = 0; float xp = 0;
};

static void fmtFloat(std::string& s, float f) {
    char buf[64];
    snprintf(buf, sizeof buf, "%g", (double)f);
    s += buf;
}

/* SAM text of one record, the field layout htslib's sam_format1 writes */
static std::string samLine(const Db& db, const Rec& r) {
    std::string s = r.qname;
    s += '\t'; s += std::to_string(r.flag);
    s += '\t'; s += r.tid >= 0 ? db.names[r.tid] : std::string("*");
    s += '\t'; s += std::to_string(r.pos + 1);
    s += '\t'; s += std::to_string(r.mapq);
    s += '\t'; s += r.cigar.empty() ? std::string("*") : r.cigar;
    /* RNEXT / PNEXT: every record is built with the BAM constructor's mate defaults, mtid -1
     * and mpos 0 (src/SAMtools/BAM.h:61-63,73-75; AlignmentPE::exportFwdBAM / exportRevBAM set
     * only MAPQ, ISIZE and flags, src/Alignment.h:512-529), and sam_format1 prints mpos + 1 */
    s += "\t*\t1\t"; s += std::to_string(r.isize);
    s += '\t';
    static const char kNt16[17] = "=ACMGRSVTWYHKDBN";
    if (r.q->L == 0) s += '*';
    for (int64_t i = 0; i < r.q->L; ++i) s += kNt16[r.q->seq[i] & 15];
    s += '\t';
    if (r.q->L == 0 || r.q->qual[0] == 0xff) s += '*';
    else for (int64_t i = 0; i < r.q->L; ++i) s += (char)(r.q->qual[i] + 33);
    if (r.hasTags) {
        s += "\tAS:i:"; s += std::to_string(r.as);
        s += "\tMD:Z:"; s += r.md;
        s += "\tXH:f:"; fmtFloat(s, r.xh);
        s += "\tNH:i:"; s += std::to_string(r.nh);
        s += "\tXN:i:"; s += std::to_string(r.xn);
        s += "\tXP:f:"; fmtFloat(s, r.xp);
    }
    s += '\n';
    return s;
}

/* exportBAM (src/Alignment.cpp:429-444) */
static Rec exportRec(const Aln& a) {
    Rec r;
    const Read* q = a.qStrand == FWD ? a.read : a.rcRead;
    r.qname = q->name;
    r.flag = a.qStrand != REV ? 0 : 16;
    r.tid = a.tid; r.pos = (int32_t)a.alnStart; r.mapq = a.mapQ;
    r.cigar = cigarString(a);
    r.isize = 0; r.q = q;
    r.hasTags = true;
    r.as = (long)(int)std::round(a.alnScore);
    r.md = mdTag(a);
    r.xh = (float)a.log10P;
    return r;
}

/* output(ALIGN_LIST...) (src/msgseqtk-align.cpp:538-557) */
static std::string outputAlns(const Db& db, const std::vector<Aln>& alns, uint32_t maxReport, int extraFlag) {
    size_t numReport = maxReport == 0 ? alns.size() : std::min((size_t)maxReport, alns.size());
    std::string s;
    for (size_t i = 0; i < numReport; ++i) {
        Rec r = exportRec(alns[i]);
        if (i > 0) r.flag |= 256;
        r.flag |= extraFlag;
        r.nh = (long)numReport; r.xn = (long)alns.size(); r.xp = (float)alns[i].postP;
        s += samLine(db, r);
    }
    return s;
}
static std::string outputUnmapped(const Db& db, const Read& q, int flag) {
    Rec r;
    r.qname = q.name; r.flag = flag; r.tid = -1; r.pos = 0; r.mapq = 0; r.isize = 0; r.q = &q;
    return samLine(db, r);
}

/* Alignment::sort (src/Alignment.h:629-633) */
static void sortAlns(std::vector<Aln>& v) {
    std::sort(v.begin(), v.end(), [](const Aln& l, const Aln& r) { return l.log10P > r.log10P; });
}
/* Alignment::filter (src/Alignment.cpp:548-555) */
static void filterAlns(std::vector<Aln>& v, double minScoreRate) {
    if (v.empty()) return;
    const double thr = (double)v.front().read->L * minScoreRate;
    v.erase(std::remove_if(v.begin(), v.end(), [&](const Aln& a) { return a.alnScore < thr; }), v.end());
}

static std::vector<Aln> buildAlignments(const Db& db, const Params& P, const Read& r, const Read& rc,
                                        const std::vector<std::vector<Seed>>& chains) {
    std::vector<Aln> v;
    v.reserve(chains.size());
    for (const auto& ch : chains) v.push_back(alignChain(db, P, r, rc, ch));
    return v;
}

/* ------------------------------------------------------------------ A32 */
/* main_SE per-read task (src/msgseqtk-align.cpp:607-649) */
std::string mapSE(const Db& db, const Params& P, const Read& read) {
    return mapSE(db, P, read,
                 [&](const Read& r) { return findSeeds(db, r, P.minLen, P.maxEvalue, P.maxNSeed); }, getChains);
}

std::string mapSE(const Db& db, const Params& P, const Read& read, const SeedFn& seedFn, const ChainFn& chainFn) {
    Read rc = revcomRead(read);
    std::vector<Seed> seeds = seedFn(read);
    if (seeds.empty()) return outputUnmapped(db, read, 4);
    const int64_t L = read.L;
    const int64_t maxMM = (int64_t)std::ceil(L * 0.2), maxIndel = (int64_t)std::ceil(L * 0.1);
    auto chains = chainFn(seeds, maxMM, maxIndel);
    std::vector<Aln> alns = buildAlignments(db, P, read, rc, chains);
    /* the alignments point into read/rc, which live in this frame */
    filterAlns(alns, P.minScoreRate);
    if (alns.empty()) return outputUnmapped(db, read, 4);
    for (Aln& a : alns) evaluate(P, a);
    calcMapQ(alns);
    sortAlns(alns);
    return outputAlns(db, alns, (uint32_t)P.maxReport, 0);
}

/* AlignmentPE::calcMapQ (src/Alignment.cpp:514-539) */
static void calcMapQPE(const Params& P, std::vector<Pair>& v) {
    if (v.empty()) return;
    const size_t n = v.size();
    std::vector<double> pr(n);
    for (size_t i = 0; i < n; ++i) pr[i] = pairLoglik(P, v[i]);
    double maxV = *std::max_element(pr.begin(), pr.end());
    double scale = (maxV != -kInf && maxV < -256) ? -256 - maxV : 0;
    for (size_t i = 0; i < n; ++i) pr[i] = std::exp(pr[i] + scale);
    double sum = eigenSum(pr);
    for (size_t i = 0; i < n; ++i) {
        double postP = sum > 0 ? pr[i] / sum : 1.0 / (double)n;
        v[i].postP = postP;
        double mq = -10 * ::log10(1 - postP);
        v[i].mapQ = (uint8_t)std::min(::round(mq), 250.0);
    }
}

/* main_PE per-pair task (src/msgseqtk-align.cpp:679-754) */
std::string mapPE(const Db& db, const Params& P, const Read& r1, const Read& r2) {
    return mapPE(db, P, r1, r2,
                 [&](const Read& r) { return findSeeds(db, r, P.minLen, P.maxEvalue, P.maxNSeed); }, getChains);
}

std::string mapPE(const Db& db, const Params& P, const Read& r1, const Read& r2, const SeedFn& seedFn,
                  const ChainFn& chainFn) {
    Read rc1 = revcomRead(r1), rc2 = revcomRead(r2);
    std::vector<Seed> s1 = seedFn(r1);
    std::vector<Seed> s2 = seedFn(r2);
    if (s1.empty() && s2.empty())
        return outputUnmapped(db, r1, 4 | 8 | 1 | 64) + outputUnmapped(db, r2, 4 | 8 | 1 | 128);
    auto ch1 = chainFn(s1, (int64_t)std::ceil(r1.L * 0.2), (int64_t)std::ceil(r1.L * 0.1));
    auto ch2 = chainFn(s2, (int64_t)std::ceil(r2.L * 0.2), (int64_t)std::ceil(r2.L * 0.1));
    std::vector<Aln> a1 = buildAlignments(db, P, r1, rc1, ch1);
    std::vector<Aln> a2 = buildAlignments(db, P, r2, rc2, ch2);
    filterAlns(a1, P.minScoreRate);
    filterAlns(a2, P.minScoreRate);
    if (a1.empty() && a2.empty())
        return outputUnmapped(db, r1, 4 | 8 | 1 | 64) + outputUnmapped(db, r2, 4 | 8 | 1 | 128);
    for (Aln& a : a1) evaluate(P, a);
    for (Aln& a : a2) evaluate(P, a);
    /* getPairs, AlignmentPE::sort, AlignmentPE::filter */
    std::vector<Pair> pairs = candidatePairs(P, a1, a2);
    std::string out;
    if (!pairs.empty()) {
        calcMapQPE(P, pairs);
        size_t numReport = P.maxReport == 0 ? pairs.size() : std::min((size_t)P.maxReport, pairs.size());
        for (size_t i = 0; i < numReport; ++i) {
            const Pair& p = pairs[i];
            /* exportFwdBAM / exportRevBAM (src/Alignment.h:512-539) */
            Rec f = exportRec(*p.f), r = exportRec(*p.r);
            f.mapq = p.mapQ; r.mapq = p.mapQ;
            f.isize = p.f->alnStart < p.r->alnStart ? p.isize : -p.isize;
            r.isize = p.f->alnStart < p.r->alnStart ? -p.isize : p.isize;
            f.flag = (p.f->qStrand != REV ? 0 : 16) | 1 | 2 | 64;
            r.flag = (p.r->qStrand != REV ? 0 : 16) | 1 | 2 | 128;
            if (i > 0) { f.flag |= 256; r.flag |= 256; }
            f.nh = r.nh = (long)numReport;
            f.xn = r.xn = (long)pairs.size();
            f.xp = r.xp = (float)p.postP;
            out += samLine(db, f);
            out += samLine(db, r);
        }
    } else if (!P.noMixed) {
        if (!a1.empty()) { calcMapQ(a1); sortAlns(a1); out += outputAlns(db, a1, (uint32_t)P.maxReport, 1 | 64); }
        if (!a2.empty()) { calcMapQ(a2); sortAlns(a2); out += outputAlns(db, a2, (uint32_t)P.maxReport, 1 | 128); }
    }
    return out;
}

} // namespace orc
