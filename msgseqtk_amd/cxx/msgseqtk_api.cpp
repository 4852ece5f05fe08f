/*
 * msgseqtk_api.cpp -- the C++ drop-in face (include/msgseqtk/msgseqtk_api.h) over the C-ABI of
 * libmsgs_hip.so.  Per-read calls are device batches of one; nothing here computes a mapping
 * result on the host except the bookkeeping the reference also does there (filter, sort order,
 * MD-tag parsing).  Built into libmsgs_cxx.so.
 */
#include "msgseqtk/msgseqtk_api.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <regex>
#include <stdexcept>

#include "msgs_hip.h"

namespace EGriceLab {
namespace MSGseqTK {

namespace {

const nt16_t kComp[16] = {0, 8, 4, 12, 2, 10, 6, 14, 1, 9, 5, 13, 3, 11, 7, 15};
const char kSym[17] = "$ACMGRSVTWYHKDBN";

int deviceId() {
    const char* e = std::getenv("MSGS_DEVICE");
    return e ? std::atoi(e) : 0;
}

msgs_params params(int64_t minLen, double maxEvalue, int64_t maxNSeed) {
    msgs_params p;
    msgs_params_default(&p);
    p.minLen = minLen;
    p.maxEvalue = maxEvalue;
    p.maxNSeed = maxNSeed;
    return p;
}

/* one read as a batch of one */
struct OneRead {
    DNAseq seq;
    QualStr qual;
    int64_t offs[2] = {0, 0};
    string name;
    msgs_reads r{};
    explicit OneRead(const PrimarySeq& s) : seq(s.getSeq()), qual(s.getQual()), name(s.getName()) {
        if (seq.empty()) seq.push_back(0), offs[1] = 0;
        else offs[1] = (int64_t)s.length();
        if (qual.size() < seq.size()) qual.resize(seq.size(), 30);
        r.seq = seq.data();
        r.qual = qual.data();
        r.offs = offs;
        r.names = name.c_str();
        r.n = 1;
    }
};

}  // namespace

/* ------------------------------------------------------------------ device database */

class Device {
public:
    msgs_ctx* ctx = nullptr;
    explicit Device(int dev) {
        if (msgs_ctx_create(dev, &ctx) != MSGS_OK) throw std::runtime_error("msgs_ctx_create: no usable GPU");
    }
    ~Device() { msgs_ctx_destroy(ctx); }
    void check(int rc, const char* what) const {
        if (rc != MSGS_OK) throw std::runtime_error(string(what) + ": " + msgs_last_error(ctx));
    }
    static std::shared_ptr<Device> build(const vector<string>& names, const vector<DNAseq>& seqs) {
        auto d = std::make_shared<Device>(deviceId());
        DNAseq all;
        vector<int64_t> lens;
        string joined;
        for (size_t t = 0; t < seqs.size(); ++t) {
            all += seqs[t];
            lens.push_back((int64_t)seqs[t].size());
            joined += names[t];
            joined += '\n';
        }
        if (all.empty()) all.push_back(0);
        d->check(msgs_db_build(d->ctx, all.data(), lens.data(), (int32_t)lens.size(), joined.c_str()), "msgs_db_build");
        return d;
    }
};

/* a device context without a database (getChains needs none) */
static msgs_ctx* chainContext() {
    static std::mutex mu;
    static std::unique_ptr<Device> d;
    std::lock_guard<std::mutex> g(mu);
    if (!d) d.reset(new Device(deviceId()));
    return d->ctx;
}

/* ------------------------------------------------------------------ alphabet, reads */

nt16_t DNAalphabet::encode(char s) {
    switch (s) {
    case '0': return A; case '1': return C; case '2': return G; case '3': return T; case '4': return N;
    default: break;
    }
    const char u = (s >= 'a' && s <= 'z') ? (char)(s - 32) : s;
    switch (u) {
    case 'A': return A; case 'B': return B; case 'C': return C; case 'D': return D; case 'G': return G;
    case 'H': return H; case 'K': return K; case 'M': return M; case 'N': return N; case 'R': return R;
    case 'S': return S; case 'T': return T; case 'U': return T; case 'V': return V; case 'W': return W;
    case 'Y': return Y;
    default: return 0;
    }
}
char DNAalphabet::decode(nt16_t b) { return b < 16 ? kSym[b] : '='; }
nt16_t DNAalphabet::complement(nt16_t b) { return b < 16 ? kComp[b] : 0; }
nt16_t DNAalphabet::toBasic(nt16_t b) {
    if (b == 0 || b >= 15) return b == 15 ? 15 : 0;
    nt16_t h = 8;
    while (!(b & h)) h >>= 1;
    return h;
}
int DNAalphabet::toInt(nt16_t b) { return b == A ? 1 : b == C ? 2 : b == G ? 3 : b == T ? 4 : 0; }

DNAseq dna::encode(const string& s) {
    DNAseq d(s.size(), 0);
    for (size_t i = 0; i < s.size(); ++i) d[i] = DNAalphabet::encode(s[i]);
    return d;
}
string dna::decode(const DNAseq& d) {
    string s(d.size(), ' ');
    for (size_t i = 0; i < d.size(); ++i) s[i] = DNAalphabet::decode(d[i]);
    return s;
}
DNAseq dna::revcom(const DNAseq& d) {
    DNAseq r(d.rbegin(), d.rend());
    for (nt16_t& b : r) b = DNAalphabet::complement(b);
    return r;
}

QualStr quality::encode(const string& q, uint8_t shift) {
    QualStr r(q.size(), 0);
    for (size_t i = 0; i < q.size(); ++i) r[i] = quality::encode(q[i], shift);
    return r;
}

PrimarySeq::PrimarySeq(const DNAseq& seq, const string& name, const string& desc, const string& qStr, uint8_t qShift)
    : seq(seq), name(name), desc(desc), qual(quality::encode(qStr, qShift)) {
    if (qual.size() != seq.size()) qual.resize(seq.size(), qShift);   /* fixQual */
}
PrimarySeq::PrimarySeq(const string& seqStr, const string& name, const string& desc, const string& qStr,
                       uint8_t qShift)
    : PrimarySeq(dna::encode(seqStr), name, desc, qStr, qShift) {}
PrimarySeq PrimarySeq::revcom() const {
    PrimarySeq r(*this);
    r.seq = dna::revcom(seq);
    std::reverse(r.qual.begin(), r.qual.end());
    return r;
}

/* ------------------------------------------------------------------ MetaGenome */

MetaGenome::MetaGenome() = default;
MetaGenome::~MetaGenome() = default;

MetaGenome& MetaGenome::addChrom(const string& name, const DNAseq& seq) {
    bdStart.push_back(BDSize());
    names.push_back(name);
    seqs.push_back(seq);
    dev.reset();
    return *this;
}

int64_t MetaGenome::BDSize() const {
    int64_t n = 0;
    for (const DNAseq& s : seqs) n += 2 * ((int64_t)s.size() + 1);
    return n;
}

DNAseq MetaGenome::getBDSeq() const {
    DNAseq bd;
    bd.reserve(BDSize());
    for (const DNAseq& s : seqs) {
        DNAseq b(s);
        for (nt16_t& x : b) x = DNAalphabet::toBasic(x);
        bd += b;
        bd.push_back(0);
        bd += dna::revcom(b);
        bd.push_back(0);
    }
    return bd;
}

int64_t MetaGenome::getTid(int64_t pos) const {
    /* the first chromosome whose BD range holds pos (src/MetaGenome.h:474-479) */
    for (size_t t = 0; t < seqs.size(); ++t)
        if (bdStart[t] <= pos && pos < bdStart[t] + 2 * ((int64_t)seqs[t].size() + 1)) return (int64_t)t;
    return -1;
}
GLoc::STRAND MetaGenome::getStrand(int64_t tid, int64_t pos) const {
    const int64_t len = 2 * ((int64_t)seqs[tid].size() + 1);
    return (bdStart[tid] <= pos && pos < bdStart[tid] + len / 2) ? GLoc::FWD : GLoc::REV;
}
int64_t MetaGenome::getLoc(int64_t tid, int64_t pos) const {
    const int64_t len = 2 * ((int64_t)seqs[tid].size() + 1);
    return getStrand(tid, pos) == GLoc::FWD ? pos - bdStart[tid] : bdStart[tid] + len - pos - 1;
}

std::shared_ptr<Device> MetaGenome::device() const {
    if (!dev) dev = Device::build(names, seqs);
    return dev;
}

namespace {
template <class V> bool rd(std::istream& in, V& v) { return (bool)in.read((char*)&v, sizeof(V)); }
bool rdStr(std::istream& in, string& s) {
    uint64_t n = 0;
    if (!rd(in, n) || n > (1u << 20)) return false;
    s.resize(n);
    return (bool)in.read(&s[0], (std::streamsize)n);
}
}  // namespace

/* DB.mtg after its program header: genomes with their chromosomes' names and lengths */
std::istream& MetaGenome::load(std::istream& in) {
    names.clear(); seqs.clear(); bdStart.clear(); dev.reset();
    uint64_t ng = 0;
    if (!rd(in, ng)) { in.setstate(std::ios::badbit); return in; }
    for (uint64_t g = 0; g < ng; ++g) {
        string id, gname;
        uint64_t nc = 0;
        if (!rdStr(in, id) || !rdStr(in, gname) || !rd(in, nc)) { in.setstate(std::ios::badbit); return in; }
        for (uint64_t c = 0; c < nc; ++c) {
            string cname;
            int64_t len = 0;
            if (!rdStr(in, cname) || !rd(in, len) || len < 0) { in.setstate(std::ios::badbit); return in; }
            bdStart.push_back(BDSize());
            names.push_back(cname);
            seqs.push_back(DNAseq((size_t)len, 0));
        }
    }
    return in;
}

/* DB.mgs: each chromosome's len + 1 nt16 bytes, the last one 0 */
std::istream& MetaGenome::loadSeq(std::istream& in) {
    dev.reset();
    for (DNAseq& s : seqs) {
        in.read((char*)&s[0], (std::streamsize)s.size());
        char term = 1;
        in.read(&term, 1);
        if (!in || term != 0) { in.setstate(std::ios::badbit); return in; }
    }
    return in;
}

/* ------------------------------------------------------------------ FMDIndex */

FMDIndex::FMDIndex() = default;
FMDIndex::~FMDIndex() = default;

FMDIndex::FMDIndex(const DNAseq& bd, bool, int) {
    /* chromosomes from the BD text: each is seq + 0 + revcom(seq) + 0 */
    vector<string> names;
    vector<DNAseq> seqs;
    size_t pos = 0;
    while (pos < bd.size()) {
        const size_t z = bd.find((nt16_t)0, pos);
        if (z == DNAseq::npos) throw std::invalid_argument("input seq is not bi-directional");
        const size_t len = z - pos;
        if (z + len + 1 >= bd.size() || bd[z + len + 1] != 0 ||
            bd.compare(z + 1, len, dna::revcom(bd.substr(pos, len))) != 0)
            throw std::invalid_argument("input seq is not bi-directional");
        names.push_back("chr" + std::to_string(seqs.size()));
        seqs.push_back(bd.substr(pos, len));
        pos = z + len + 2;
    }
    for (const DNAseq& s : seqs)
        for (nt16_t b : s)
            if (b != 15 && DNAalphabet::isAmbiguous(b)) throw std::invalid_argument("input seq does not allow IUPAC-extended bases");
    dev = Device::build(names, seqs);
}

std::shared_ptr<Device> FMDIndex::device(const MetaGenome* mtg) const {
    if (dev) return dev;
    if (!mtg) throw std::runtime_error("FMDIndex: no BD text bound (pass the MetaGenome of the database)");
    auto d = mtg->device();
    if (loadedBwt) {
        /* the file's index must be the one of these sequences */
        const int64_t n = msgs_db_length(d->ctx);
        DNAseq bwt((size_t)n, 0);
        int64_t B[256], C[256];
        d->check(msgs_db_copy(d->ctx, nullptr, bwt.data(), nullptr), "msgs_db_copy");
        d->check(msgs_db_counts(d->ctx, B, C), "msgs_db_counts");
        if (bwt != *loadedBwt || std::memcmp(B, loadedB, sizeof B) != 0 || std::memcmp(C, loadedC, sizeof C) != 0)
            throw std::runtime_error("FMDIndex: the loaded index is not the index of this MetaGenome");
    }
    dev = d;
    return dev;
}

int64_t FMDIndex::length() const {
    if (!dev && loadedBwt) return (int64_t)loadedBwt->size();
    return msgs_db_length(device()->ctx);
}
int64_t FMDIndex::getBaseCount(nt16_t b) const {
    if (!dev && loadedBwt) return loadedB[b];
    int64_t B[256], C[256];
    device()->check(msgs_db_counts(dev->ctx, B, C), "msgs_db_counts");
    return B[b];
}
int64_t FMDIndex::getCumCount(nt16_t b) const {
    if (!dev && loadedBwt) return loadedC[b];
    int64_t B[256], C[256];
    device()->check(msgs_db_counts(dev->ctx, B, C), "msgs_db_counts");
    return C[b];
}
double FMDIndex::loglik(nt16_t b) const {
    return std::log((double)getBaseCount(DNAalphabet::toBasic(b))) - std::log((double)length());
}
int64_t FMDIndex::backExt(int64_t& p, int64_t& q, int64_t& s, nt16_t b) const {
    int64_t pqs[3] = {p, q, s}, ret = 0;
    const std::shared_ptr<Device> d = device();
    d->check(msgs_ext_batch(d->ctx, pqs, &b, 0, &ret, 1), "msgs_ext_batch");
    p = pqs[0]; q = pqs[1]; s = pqs[2];
    return ret;
}
int64_t FMDIndex::fwdExt(int64_t& p, int64_t& q, int64_t& s, nt16_t b) const {
    int64_t pqs[3] = {p, q, s}, ret = 0;
    const std::shared_ptr<Device> d = device();
    d->check(msgs_ext_batch(d->ctx, pqs, &b, 1, &ret, 1), "msgs_ext_batch");
    p = pqs[0]; q = pqs[1]; s = pqs[2];
    return ret;
}
int64_t FMDIndex::LF(nt16_t b, int64_t i) const {
    int64_t r = 0;
    const std::shared_ptr<Device> d = device();
    d->check(msgs_rank_batch(d->ctx, &b, &i, &r, 1), "msgs_rank_batch");
    return getCumCount(b) + r;
}
int64_t FMDIndex::accessSA(int64_t i) const {
    int64_t v = 0;
    const std::shared_ptr<Device> d = device();
    d->check(msgs_sa_batch(d->ctx, &i, 1, &v), "msgs_sa_batch");
    return v;
}
nt16_t FMDIndex::accessBWT(int64_t i) const {
    if (!dev && loadedBwt) return (*loadedBwt)[i];
    const DNAseq s = getSeq();   /* BWT[i] = BD[SA[i] - 1] */
    const int64_t sa = accessSA(i);
    return sa == 0 ? 0 : s[sa - 1];
}
DNAseq FMDIndex::getSeq() const {
    const std::shared_ptr<Device> d = device();
    DNAseq bd((size_t)msgs_db_length(d->ctx), 0);
    d->check(msgs_db_copy(d->ctx, bd.data(), nullptr, nullptr), "msgs_db_copy");
    return bd;
}
int64_t FMDIndex::count(const DNAseq& pattern) const {
    int64_t offs[2] = {0, (int64_t)pattern.size()}, pqs[3];
    const nt16_t z = 0;
    const std::shared_ptr<Device> d = device();
    d->check(msgs_search_batch(d->ctx, pattern.empty() ? &z : pattern.data(), offs, 1, 0, pqs), "msgs_search_batch");
    return std::max<int64_t>(pqs[2], 0);
}
vector<GLoc> FMDIndex::locateAll(const DNAseq& pattern, GLoc::STRAND strand) const {
    vector<GLoc> locs;
    if (pattern.empty()) return locs;
    int64_t offs[2] = {0, (int64_t)pattern.size()}, pqs[3];
    const std::shared_ptr<Device> d = device();
    const bool fwd = strand == GLoc::FWD;
    d->check(msgs_search_batch(d->ctx, pattern.data(), offs, 1, fwd ? 1 : 2, pqs), "msgs_search_batch");
    const int64_t s = pqs[2], first = fwd ? pqs[0] : pqs[1];
    if (s <= 0) return locs;
    vector<int64_t> rows(s), sa(s);
    for (int64_t j = 0; j < s; ++j) rows[j] = first + j;
    d->check(msgs_sa_batch(d->ctx, rows.data(), s, sa.data()), "msgs_sa_batch");
    for (int64_t j = 0; j < s; ++j) locs.push_back(GLoc(sa[j], sa[j] + (int64_t)pattern.size(), -1, strand));
    return locs;
}

/* the body of DB.fmd (after its program header), decoded by the library */
std::istream& FMDIndex::load(std::istream& in) {
    std::string body((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    char tmpl[] = "/tmp/msgs_fmdXXXXXX";
    const int fd = mkstemp(tmpl);
    if (fd < 0) { in.setstate(std::ios::badbit); return in; }
    FILE* f = fdopen(fd, "wb");
    const bool ok = f && fwrite(body.data(), 1, body.size(), f) == body.size();
    if (f) fclose(f);
    int64_t n = 0;
    int rc = ok ? msgs_fmd_read(tmpl, 0, &n, nullptr, loadedB, loadedC, nullptr, nullptr, nullptr, nullptr) : MSGS_E_ARG;
    if (rc == MSGS_OK) {
        loadedBwt = std::make_shared<DNAseq>((size_t)n, 0);
        rc = msgs_fmd_read(tmpl, 0, &n, loadedBwt->data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    }
    std::remove(tmpl);
    if (rc != MSGS_OK) { loadedBwt.reset(); in.setstate(std::ios::badbit); return in; }
    dev.reset();
    in.clear(std::ios::eofbit);
    return in;
}

/* ------------------------------------------------------------------ seeding */

const double SMEM::MAX_EVALUE = 0.01;

SMEM::SMEM(const PrimarySeq* seq, const MetaGenome* mtg, const FMDIndex* fmdidx, int64_t from, int64_t to,
           int64_t fwdStart, int64_t revStart, int64_t size)
    : seq(seq), mtg(mtg), fmdidx(fmdidx), from(from), to(to), fwdStart(fwdStart), revStart(revStart), size(size) {
    evaluate();
}

SMEM& SMEM::evaluate() {
    logP = 0;
    for (int64_t i = from; i < to; ++i) logP += fmdidx->loglik(seq->getBase(i));
    return *this;
}

SeedList SMEM::getSeeds() const {
    SeedList seeds;
    if (size <= 0) return seeds;
    const std::shared_ptr<Device> d = fmdidx->device(mtg);
    vector<int64_t> rows(2 * size), bd(2 * size);
    for (int64_t i = 0; i < size; ++i) { rows[2 * i] = fwdStart + i; rows[2 * i + 1] = revStart + i; }
    d->check(msgs_sa_batch(d->ctx, rows.data(), 2 * size, bd.data()), "msgs_sa_batch");
    for (int64_t i = 0; i < 2 * size; ++i) {
        const int64_t tid = mtg->getTid(bd[i]);
        if (mtg->getStrand(tid, bd[i]) != GLoc::FWD) continue;   /* forward target strands only */
        const int64_t start = mtg->getLoc(tid, bd[i]);
        if (i % 2 == 0) seeds.push_back(SeedPair(from, start, length(), tid, GLoc::FWD, logP));
        else seeds.push_back(SeedPair((int64_t)seq->length() - to, start, length(), tid, GLoc::REV, logP));
    }
    return seeds;
}

double SMEM_LIST::loglik() const {
    double ll = 0;
    for (const SMEM& m : *this) ll += m.loglik();
    return ll;
}

namespace {
SMEM_LIST smemCall(bool mems, const PrimarySeq* seq, const MetaGenome* mtg, const FMDIndex* fmdidx, int64_t minLen,
                   double maxEvalue) {
    const std::shared_ptr<Device> d = fmdidx->device(mtg);
    OneRead r(*seq);
    msgs_params p = params(minLen, maxEvalue, SMEM::MAX_NSEED);
    int64_t cnt = 0, need = 0, cap = 64;
    vector<int64_t> rows;
    vector<double> lp;
    for (int attempt = 0; attempt < 2; ++attempt) {
        rows.assign(5 * cap, 0);
        lp.assign(cap, 0);
        const int rc = mems ? msgs_mems_batch(d->ctx, &r.r, &p, &cnt, rows.data(), lp.data(), nullptr, cap, &need)
                            : msgs_smem_batch(d->ctx, &r.r, &p, &cnt, rows.data(), lp.data(), cap, &need);
        if (rc == MSGS_E_CAP) { cap = need; continue; }
        d->check(rc, mems ? "msgs_mems_batch" : "msgs_smem_batch");
        break;
    }
    SMEM_LIST out;
    for (int64_t k = 0; k < cnt; ++k)
        out.push_back(SMEM(seq, mtg, fmdidx, rows[5 * k], rows[5 * k + 1], rows[5 * k + 2], rows[5 * k + 3],
                           rows[5 * k + 4], lp[k]));
    return out;
}
}  // namespace

SMEM_LIST SMEM_LIST::findAllSMEMS(const PrimarySeq* seq, const MetaGenome* mtg, const FMDIndex* fmdidx, int64_t minLen,
                                  double maxEvalue) {
    return smemCall(false, seq, mtg, fmdidx, minLen, maxEvalue);
}
SMEM_LIST SMEM_LIST::findMEMS(const PrimarySeq* seq, const MetaGenome* mtg, const FMDIndex* fmdidx, int64_t minLen,
                              double maxEvalue) {
    return smemCall(true, seq, mtg, fmdidx, minLen, maxEvalue);
}
SeedList SMEM_LIST::findSeeds(const PrimarySeq* seq, const MetaGenome* mtg, const FMDIndex* fmdidx, int64_t minLen,
                              double maxEvalue, int64_t maxNSeed) {
    const std::shared_ptr<Device> d = fmdidx->device(mtg);
    OneRead r(*seq);
    msgs_params p = params(minLen, maxEvalue, maxNSeed);
    int64_t cnt = 0, need = 0, cap = 256;
    vector<int64_t> rows;
    vector<double> lp;
    for (int attempt = 0; attempt < 2; ++attempt) {
        rows.assign(6 * cap, 0);
        lp.assign(cap, 0);
        const int rc = msgs_seed_batch(d->ctx, &r.r, &p, &cnt, rows.data(), lp.data(), cap, &need);
        if (rc == MSGS_E_CAP) { cap = need; continue; }
        d->check(rc, "msgs_seed_batch");
        break;
    }
    SeedList out;
    for (int64_t k = 0; k < cnt; ++k) {
        const int64_t* x = &rows[6 * k];
        out.push_back(SeedPair(x[0], x[2], x[1] - x[0], x[4], (GLoc::STRAND)x[5], lp[k]));
    }
    return out;
}

vector<SeedChain> SeedChain::getChains(const SeedList& seeds, int64_t maxMismatch, int64_t maxIndel) {
    msgs_ctx* ctx = chainContext();
    vector<int64_t> rows(6 * std::max<size_t>(1, seeds.size()));
    for (size_t i = 0; i < seeds.size(); ++i) {
        const SeedPair& s = seeds[i];
        int64_t* x = &rows[6 * i];
        x[0] = s.getFrom(); x[1] = s.getTo(); x[2] = s.getStart(); x[3] = s.getEnd(); x[4] = s.getTid(); x[5] = s.getStrand();
    }
    int64_t offs[2] = {0, (int64_t)seeds.size()}, chains = 0, need = 0;
    vector<int64_t> out(4 * seeds.size() + 16);
    int rc = msgs_chain_batch(ctx, rows.data(), offs, 1, &maxMismatch, &maxIndel, &chains, out.data(),
                              (int64_t)out.size(), &need);
    if (rc == MSGS_E_CAP) {
        out.assign(need + 1, 0);
        rc = msgs_chain_batch(ctx, rows.data(), offs, 1, &maxMismatch, &maxIndel, &chains, out.data(),
                              (int64_t)out.size(), &need);
    }
    if (rc != MSGS_OK) throw std::runtime_error(string("msgs_chain_batch: ") + msgs_last_error(ctx));
    vector<SeedChain> res;
    for (int64_t c = 0, k = 0; c < chains; ++c) {
        SeedChain ch;
        const int64_t n = out[k++];
        for (int64_t t = 0; t < n; ++t) ch.push_back(seeds[out[k++]]);
        res.push_back(ch);
    }
    return res;
}

/* ------------------------------------------------------------------ alignment */

ScoreScheme Alignment::ss;

double Alignment::loglik() const { return log10P / ::log10(M_E); }

int32_t Alignment::mdTag2alnQLen(const string& md) {
    /* src/Alignment.cpp:30-31,402-418: the leading number, then per mismatch base one query base
     * or per '^'-run its bases, each followed by a number */
    static const std::regex lead("^\\d+"), main_("([A-Z]|\\^[A-Z]+)(\\d+)");
    std::smatch m;
    auto it = md.cbegin();
    if (!std::regex_search(it, md.cend(), m, lead)) return 0;
    int32_t len = std::stoi(m.str(0));
    for (it = m[0].second; std::regex_search(it, md.cend(), m, main_); it = m[0].second) {
        len += m.length(1) == 1 ? 1 : (int32_t)m.length(1) - 1;
        len += std::stoi(m.str(2));
    }
    return len;
}

ALIGN_LIST Alignment::buildAlignments(const PrimarySeq* read, const PrimarySeq* rcRead, const MetaGenome& mtg,
                                      const ChainList& chains, MODE alnMode) {
    (void)rcRead;   /* the device makes the reverse complement */
    ALIGN_LIST out;
    if (chains.empty()) return out;
    const std::shared_ptr<Device> d = mtg.device();
    OneRead r(*read);
    msgs_params p;
    msgs_params_default(&p);
    p.mode = alnMode == LOCAL ? 1 : 0;
    p.match = ss.match; p.mismatch = ss.mismatch; p.gapO = ss.gapO; p.gapE = ss.gapE; p.clip = ss.clip;
    vector<int64_t> chainOffs = {0, (int64_t)chains.size()}, seedOffs(1, 0), rows;
    for (const SeedChain& ch : chains) {
        for (const SeedPair& s : ch) {
            rows.insert(rows.end(), {s.getFrom(), s.getTo(), s.getStart(), s.getEnd(), s.getTid(), (int64_t)s.getStrand()});
        }
        seedOffs.push_back((int64_t)rows.size() / 6);
    }
    const size_t nj = chains.size();
    vector<int64_t> r10(10 * nj);
    vector<double> sc(nj), lp(nj);
    const int64_t cap = 16 * (int64_t)(read->length() + 64) * (int64_t)nj + 1024;
    string cig((size_t)cap, '\0'), md((size_t)cap, '\0');
    int64_t need = 0;
    d->check(msgs_dp_batch(d->ctx, &r.r, chainOffs.data(), seedOffs.data(), rows.data(), &p, r10.data(), sc.data(),
                           lp.data(), &cig[0], cap, &md[0], cap, &need),
             "msgs_dp_batch");
    size_t cp = 0, mp = 0;
    for (size_t k = 0; k < nj; ++k) {
        Alignment a;
        a.read = read;
        a.tid = (int32_t)r10[10 * k];
        a.qStrand = (GLoc::STRAND)r10[10 * k + 1];
        a.tStart = r10[10 * k + 2]; a.tEnd = r10[10 * k + 3];
        a.alnFrom = r10[10 * k + 4]; a.alnTo = r10[10 * k + 5];
        a.alnStart = r10[10 * k + 6]; a.alnEnd = r10[10 * k + 7];
        a.alnScore = sc[k];
        a.log10P = lp[k];
        const size_t ce = cig.find('\n', cp), me = md.find('\n', mp);
        a.cigar = cig.substr(cp, ce - cp);
        a.md = md.substr(mp, me - mp);
        cp = ce + 1; mp = me + 1;
        out.push_back(a);
    }
    return out;
}

ALIGN_LIST& Alignment::filter(ALIGN_LIST& alnList, double minScoreRate) {
    if (alnList.empty()) return alnList;
    const double thr = (double)alnList.front().read->length() * minScoreRate;
    alnList.erase(std::remove_if(alnList.begin(), alnList.end(), [&](const Alignment& a) { return a.alnScore < thr; }),
                  alnList.end());
    return alnList;
}

ALIGN_LIST& Alignment::calcMapQ(ALIGN_LIST& alnList) {
    if (alnList.empty()) return alnList;
    vector<double> ll(alnList.size()), pp(alnList.size());
    vector<uint8_t> mq(alnList.size());
    for (size_t i = 0; i < alnList.size(); ++i) ll[i] = alnList[i].loglik();
    msgs_util_mapq(ll.data(), (int64_t)ll.size(), 0, pp.data(), mq.data());
    for (size_t i = 0; i < alnList.size(); ++i) { alnList[i].postP = pp[i]; alnList[i].mapQ = mq[i]; }
    return alnList;
}

ALIGN_LIST& Alignment::sort(ALIGN_LIST& alnList) {
    /* std::sort by log10lik decreasingly (src/Alignment.h:629-633), the library's restated
     * libstdc++ sort on the negated keys (the same comparisons) */
    vector<double> key(alnList.size());
    vector<int32_t> perm(alnList.size());
    for (size_t i = 0; i < alnList.size(); ++i) { key[i] = -alnList[i].log10P; perm[i] = (int32_t)i; }
    msgs_util_sort_by_key(key.data(), perm.data(), (int64_t)perm.size());
    ALIGN_LIST sorted;
    sorted.reserve(alnList.size());
    for (int32_t i : perm) sorted.push_back(alnList[i]);
    alnList.swap(sorted);
    return alnList;
}

/* ---- PairingScheme / AlignmentPE (src/PairingScheme.cpp:14-25, src/Alignment.cpp:514-581,
 * src/Alignment.h:635-639) ---- */
const double PairingScheme::DEFAULT_MEAN_INSERT = 200;
const double PairingScheme::DEFAULT_CV_INSERT = 0.1;
const double PairingScheme::DEFAULT_OUTLIER_DIST = 10;
PairingScheme AlignmentPE::ps;

double PairingScheme::pr(double x) const {
    msgs_params p;
    msgs_params_default(&p);
    p.insMean = m;
    p.insSD = s;
    p.insMin = min;
    p.insMax = max;
    return msgs_util_insert_pr(&p, x);
}

int32_t AlignmentPE::getInsertSize(const Alignment& lhs, const Alignment& rhs) {
    if (lhs.tid != rhs.tid) return -1;
    return (int32_t)(std::max(lhs.alnEnd, rhs.alnEnd) - std::min(lhs.alnStart, rhs.alnStart));
}

PAIR_LIST AlignmentPE::getPairs(const ALIGN_LIST& fwdAlnList, const ALIGN_LIST& revAlnList) {
    PAIR_LIST out;
    out.reserve(fwdAlnList.size() * revAlnList.size());
    for (const Alignment& f : fwdAlnList)
        for (const Alignment& r : revAlnList)
            if ((f.qStrand & r.qStrand) == 0) out.push_back(AlignmentPE(&f, &r));
    return out;
}

PAIR_LIST& AlignmentPE::filter(PAIR_LIST& v, bool noDiscordant, bool noTailOver, bool noContain, bool noOverlap,
                               int64_t maxNPair) {
    auto rm = [&v](auto pred) { v.erase(std::remove_if(v.begin(), v.end(), pred), v.end()); };
    rm([](const AlignmentPE& q) { return !(ps.getMin() <= q.getInsertSize() && q.getInsertSize() <= ps.getMax()); });
    if (noDiscordant) rm([](const AlignmentPE& q) { return !q.isConcordant(); });
    if (noTailOver) rm([](const AlignmentPE& q) { return q.isTailOver(); });
    if (noContain) rm([](const AlignmentPE& q) { return q.isContained(); });
    if (noOverlap) rm([](const AlignmentPE& q) { return q.isOverlap(); });
    if (maxNPair > 0 && (int64_t)v.size() > maxNPair) v.erase(v.begin() + maxNPair, v.end());
    return v;
}

PAIR_LIST& AlignmentPE::calcMapQ(PAIR_LIST& v) {
    if (v.empty()) return v;
    vector<double> ll(v.size()), pp(v.size());
    vector<uint8_t> mq(v.size());
    for (size_t i = 0; i < v.size(); ++i) ll[i] = v[i].loglik();
    msgs_util_mapq(ll.data(), (int64_t)ll.size(), 1, pp.data(), mq.data());
    for (size_t i = 0; i < v.size(); ++i) { v[i].postP = pp[i]; v[i].mapQ = mq[i]; }
    return v;
}

PAIR_LIST& AlignmentPE::sort(PAIR_LIST& v) {
    /* std::sort by log10lik decreasingly, as Alignment::sort */
    vector<double> key(v.size());
    vector<int32_t> perm(v.size());
    for (size_t i = 0; i < v.size(); ++i) { key[i] = -v[i].log10lik(); perm[i] = (int32_t)i; }
    msgs_util_sort_by_key(key.data(), perm.data(), (int64_t)perm.size());
    PAIR_LIST sorted;
    sorted.reserve(v.size());
    for (int32_t i : perm) sorted.push_back(v[i]);
    v.swap(sorted);
    return v;
}

}  // namespace MSGseqTK
}  // namespace EGriceLab
