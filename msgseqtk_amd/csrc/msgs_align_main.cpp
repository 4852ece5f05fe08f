/*
 * msgs_align_main.cpp -- msgseqtk-align-amd, a drop-in for msgseqtk-align
 * (src/msgseqtk-align.cpp) that maps through libmsgs_hip.so.
 *
 *   msgseqtk-align-amd <DB> <READ-FILE> [MATE-FILE] -o <OUT.sam|OUT.bam> [options]
 *
 * DB is the reference's database prefix (DB.mtg + DB.mgs) or a FASTA file of
 * chromosomes.  Reads are FASTA or FASTQ (by extension, SeqIO::guessFormat,
 * src/SeqIO.cpp:119-131), parsed like SeqIO::nextFastaSeq / nextFastqSeq
 * (src/SeqIO.cpp:53-89); FASTA reads get quality 33 (PrimarySeq::fixQual with the
 * default shift).  Options and defaults follow src/msgseqtk-align.cpp:59-400; reads
 * are mapped in batches of --batch reads (pairs) on one GPU.
 */
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "../../include/msgs_hip.h"
#include "msgs_bgzf.h"

namespace {

uint8_t sym2base(char s) {
    switch (s) {
    case '0': return 1; case '1': return 2; case '2': return 4; case '3': return 8; case '4': return 15;
    default: break;
    }
    char u = (s >= 'a' && s <= 'z') ? char(s - 32) : s;
    switch (u) {
    case 'A': return 1; case 'B': return 14; case 'C': return 2; case 'D': return 13; case 'G': return 4;
    case 'H': return 11; case 'K': return 12; case 'M': return 3; case 'N': return 15; case 'R': return 5;
    case 'S': return 6; case 'T': return 8; case 'U': return 8; case 'V': return 7; case 'W': return 9;
    case 'Y': return 10;
    default: return 0;
    }
}

bool endsWith(const std::string& s, const std::string& e) {
    return s.size() >= e.size() && s.compare(s.size() - e.size(), e.size(), e) == 0;
}
enum Fmt { UNK, FASTA, FASTQ };
Fmt guessFormat(const std::string& fn) {
    if (endsWith(fn, ".fasta") || endsWith(fn, ".fa") || endsWith(fn, ".fas") || endsWith(fn, ".fna")) return FASTA;
    if (endsWith(fn, ".fastq") || endsWith(fn, ".fq")) return FASTQ;
    return UNK;
}

struct Rec {
    std::string name;
    std::vector<uint8_t> seq, qual;
};

/* SeqIO::nextFastaSeq / nextFastqSeq */
struct Reader {
    std::istream& in;
    Fmt fmt;
    bool next(Rec& r) {
        int c = in.peek();
        if (c == EOF) return false;
        char head = fmt == FASTA ? '>' : '@';
        if (c != head) {
            if (c == EOF) return false;
            throw std::runtime_error(fmt == FASTA ? "input is not a valid FASTA format" : "input is not a valid FASTQ format");
        }
        in.get();
        in >> r.name;
        while (::isspace(in.peek()) && in.peek() != '\n') in.get();
        std::string desc, line, seq, qual;
        std::getline(in, desc);
        if (fmt == FASTA) {
            while (in.peek() != EOF && in.peek() != '>') { std::getline(in, line); seq += line; }
        } else {
            std::getline(in, seq);
            std::getline(in, line);
            std::getline(in, qual);
        }
        r.seq.resize(seq.size());
        for (size_t i = 0; i < seq.size(); ++i) r.seq[i] = sym2base(seq[i]);
        r.qual.assign(seq.size(), 33);   /* fixQual: missing qualities take the shift value */
        for (size_t i = 0; i < qual.size() && i < seq.size(); ++i)
            r.qual[i] = std::max<uint8_t>((uint8_t)(qual[i] - 33), 2);   /* quality::encode */
        return true;
    }
};

void usage(const char* prog) {
    std::cerr << "Usage:    " << prog << "  <DB> <READ-FILE> [MATE-FILE] <-o ALIGN-OUT> [other-options]\n"
              << "DB  STR                          : database prefix (DB.mtg + DB.mgs) or FASTA of chromosomes\n"
              << "Options:    -o  FILE             : SAM output file\n"
              << "            --global / --local   : Needleman-Wunsch (default) / Smith-Waterman DP\n"
              << "            --match --mis-match --gap-open --gap-ext  DBL : scores [1 3 5 2]\n"
              << "            -k/--max-report INT  : [1]    -s/--min-score DBL : [0.65]\n"
              << "            -m/--mean-ins -d/--sd-ins -I/--min-ins -X/--max-ins DBL\n"
              << "            --no-mixed --no-discordant --no-tail-over --no-contain --no-overlap --max-npair INT\n"
              << "            --min-seed INT [17]  --max-evalue DBL [0.01, 0 = no limit]  --max-nseed INT [500]\n"
              << "            --batch INT          : reads (pairs) per GPU batch [262144]\n";
}

std::vector<std::string> g_names;

/* The header msgseqtk-align builds (src/msgseqtk-align.cpp:509-520): the BAMheader of the
 * chromosomes, the @HD tag set by setHDTag("SQ", "unsorted") (the reference's key, "SQ"), and the
 * @PG tag "ID:MSGseqTK VN:v1.2.3 PN:<argv[0]> CL:<argv[0]> <options>" with its fields separated by
 * spaces (ProgEnv's progName / progVer, CommandOptions::getProg / getCmdStr,
 * src/util/CommandOptions.cpp:31-55).  The text layout is htslib 1.9's (the reference leaves
 * the htslib version open): sam_hdr_change_HD puts "@HD\tVN:1.6\t<key>:<val>" in front of the
 * text and BAMheader::addTag appends "@PG\t<value>"; a SAM file gets that text followed by one
 * @SQ line per target (sam_hdr_write adds them when the text has none), a BAM file the text and
 * the binary target list (bam_hdr_write). */
struct Header {
    std::string text;
    std::vector<std::string> names;
    std::vector<int64_t> lens;
};

bool makeHeader(msgs_ctx* ctx, const std::string& prog, const std::string& cmd, Header& h) {
    h.text = "@HD\tVN:1.6\tSQ:unsorted\n@PG\tID:MSGseqTK VN:v1.2.3 PN:" + prog + " CL:" + cmd + "\n";
    const int32_t n = msgs_db_nchrom(ctx);
    char buf[4096];
    for (int32_t t = 0; t < n; ++t) {
        int64_t len = 0;
        if (msgs_db_chrom(ctx, t, buf, sizeof buf, &len) != 0) return false;
        h.names.push_back(buf);
        h.lens.push_back(len);
    }
    return true;
}

/* records out: SAM text, or BAM through BGZF */
struct Sink {
    bool bam = false;
    std::ofstream sam;
    FILE* f = nullptr;
    msgs_bgzf::Writer* z = nullptr;

    bool open(const std::string& fn, bool asBam) {
        bam = asBam;
        if (!bam) { sam.open(fn); return (bool)sam; }
        f = std::fopen(fn.c_str(), "wb");
        if (!f) return false;
        z = new msgs_bgzf::Writer(f);
        return true;
    }
    bool header(const Header& h) {
        if (!bam) {
            sam << h.text;
            for (size_t t = 0; t < h.names.size(); ++t) sam << "@SQ\tSN:" << h.names[t] << "\tLN:" << h.lens[t] << "\n";
            return (bool)sam;
        }
        auto i32 = [this](int64_t v) { const int32_t x = (int32_t)v; z->write(&x, 4); };
        z->write("BAM\1", 4);
        i32((int64_t)h.text.size());
        z->write(h.text.data(), h.text.size());
        i32((int64_t)h.names.size());
        for (size_t t = 0; t < h.names.size(); ++t) {
            i32((int64_t)h.names[t].size() + 1);
            z->write(h.names[t].c_str(), h.names[t].size() + 1);
            i32(h.lens[t]);
        }
        z->flush();
        return z->ok();
    }
    /* SAM text, or BAM records (each int32 block_size + block_size bytes) */
    bool records(const char* p, int64_t n) {
        if (!bam) { sam.write(p, n); return (bool)sam; }
        int64_t k = 0;
        while (k < n) {
            if (n - k < 4) return false;
            uint32_t bs;
            std::memcpy(&bs, p + k, 4);
            if ((int64_t)bs > n - k - 4) return false;
            z->reserveRecord(4 + (size_t)bs);
            z->write(p + k, 4 + (size_t)bs);
            k += 4 + bs;
        }
        return z->ok();
    }
    bool close() {
        if (!bam) { sam.close(); return !sam.fail(); }
        z->close();
        const bool ok = z->ok();
        delete z;
        z = nullptr;
        return std::fclose(f) == 0 && ok;
    }
};

bool loadFastaDb(msgs_ctx* ctx, const std::string& fn) {
    std::ifstream in(fn);
    if (!in) { std::cerr << "Unable to open '" << fn << "'\n"; return false; }
    Reader rd{in, FASTA};
    Rec r;
    std::vector<uint8_t> seqs;
    std::vector<int64_t> lens;
    std::string names;
    while (rd.next(r)) {
        seqs.insert(seqs.end(), r.seq.begin(), r.seq.end());
        lens.push_back((int64_t)r.seq.size());
        names += r.name + "\n";
    }
    if (lens.empty()) { std::cerr << "no sequences in '" << fn << "'\n"; return false; }
    if (msgs_db_build(ctx, seqs.data(), lens.data(), (int32_t)lens.size(), names.c_str()) != 0) {
        std::cerr << "database build failed: " << msgs_last_error(ctx) << "\n";
        return false;
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> pos;
    std::string outFn, cmd;
    msgs_params p;
    msgs_params_default(&p);
    bool hasM = false, hasD = false, hasI = false, hasX = false;
    double meanIns = 0, sdIns = 0, minIns = 0, maxIns = 0;
    int64_t batch = 262144;
    for (int i = 0; i < argc; ++i) { if (i) cmd += ' '; cmd += argv[i]; }
    auto val = [&](int& i) -> const char* {
        if (i + 1 >= argc) { std::cerr << argv[i] << " needs a value\n"; std::exit(EXIT_FAILURE); }
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "-h" || a == "--help") { usage(argv[0]); return EXIT_SUCCESS; }
        else if (a == "--version") { std::cout << "msgseqtk-align-amd v1.2.3\n"; return EXIT_SUCCESS; }
        else if (a == "-o") outFn = val(i);
        else if (a == "--global") p.mode = 0;
        else if (a == "--local") p.mode = 1;
        else if (a == "--match") p.match = std::atof(val(i));
        else if (a == "--mis-match") p.mismatch = std::atof(val(i));
        else if (a == "--gap-open") p.gapO = std::atof(val(i));
        else if (a == "--gap-ext") p.gapE = std::atof(val(i));
        else if (a == "-k" || a == "--max-report") p.maxReport = std::atoi(val(i));
        else if (a == "-s" || a == "--min-score") p.minScoreRate = std::atof(val(i));
        else if (a == "-m" || a == "--mean-ins") { meanIns = std::atof(val(i)); hasM = true; }
        else if (a == "-d" || a == "--sd-ins") { sdIns = std::atof(val(i)); hasD = true; }
        else if (a == "-I" || a == "--min-ins") { minIns = std::atof(val(i)); hasI = true; }
        else if (a == "-X" || a == "--max-ins") { maxIns = std::atof(val(i)); hasX = true; }
        else if (a == "--no-mixed") p.noMixed = 1;
        else if (a == "--no-discordant") p.noDiscordant = 1;
        else if (a == "--no-tail-over") p.noTailOver = 1;
        else if (a == "--no-contain") p.noContain = 1;
        else if (a == "--no-overlap") p.noOverlap = 1;
        else if (a == "--max-npair") p.maxNPair = std::atol(val(i));
        else if (a == "--min-seed") p.minLen = std::atol(val(i));
        else if (a == "--max-evalue") p.maxEvalue = std::atof(val(i));
        else if (a == "--max-nseed") p.maxNSeed = std::atol(val(i));
        else if (a == "--batch") batch = std::atol(val(i));
        else if (a == "-p" || a == "--process") (void)val(i);
        else if (a == "-v") {}
        else if (!a.empty() && a[0] == '-') { std::cerr << "unknown option " << a << "\n"; usage(argv[0]); return EXIT_FAILURE; }
        else pos.push_back(a);
    }
    if (pos.size() < 2 || pos.size() > 3 || outFn.empty()) { usage(argv[0]); return EXIT_FAILURE; }
    /* option checks and the insert model as msgseqtk-align sets it (src/msgseqtk-align.cpp:262-400) */
    if (!(p.minScoreRate <= 1)) { std::cerr << "-s|-min-score must no greater than 1\n"; return EXIT_FAILURE; }
    if (!(0 < p.minLen)) { std::cerr << "--min-seed must be non-negative\n"; return EXIT_FAILURE; }
    if (!(p.maxEvalue >= 0)) { std::cerr << "--max-evalue must be non-negative\n"; return EXIT_FAILURE; }
    if (p.maxEvalue == 0) p.maxEvalue = std::numeric_limits<double>::infinity();
    if (!(p.maxNSeed > 0)) { std::cerr << "--max-nseed must be positive\n"; return EXIT_FAILURE; }
    if (batch <= 0) { std::cerr << "--batch must be positive\n"; return EXIT_FAILURE; }
    double m = 200, s = 20;   /* PairingScheme() (src/PairingScheme.h) */
    if (hasM) { if (!(meanIns >= 0)) { std::cerr << "-m/--mean-ins must non-netative\n"; return EXIT_FAILURE; } m = meanIns; }
    if (hasD) { if (!(sdIns > 0)) { std::cerr << "-d/--sd-ins must be positive\n"; return EXIT_FAILURE; } s = sdIns; }
    double mn = 0, mx = 400;
    if (m > 0 && s > 0) { mn = m - s * 10; mx = m + s * 10; }   /* updateRange */
    if (hasI) { if (!(minIns >= 0)) { std::cerr << "-I/--min-ins must be non-negative\n"; return EXIT_FAILURE; } mn = minIns; }
    if (hasX) { if (!(maxIns > 0)) { std::cerr << "-X/--max-ins must be positive\n"; return EXIT_FAILURE; } mx = maxIns; }
    if (m == 0) s = 1;          /* updateParam */
    if (s == 0) s = m * 0.1;
    p.insMean = m; p.insSD = s; p.insMin = mn; p.insMax = mx;

    const bool paired = pos.size() == 3;
    Fmt fmt = guessFormat(pos[1]);
    if (fmt == UNK) { std::cerr << "Unrecognized sequence format for file '" << pos[1] << "'\n"; return EXIT_FAILURE; }
    /* output mode by suffix (src/msgseqtk-align.cpp:472-481, src/MSGseqTKConst.h:31-32) */
    const bool asBam = endsWith(outFn, ".bam");
    if (!asBam && !endsWith(outFn, ".sam")) { std::cerr << "Unable to determine SAM/BAM output mode\n"; return EXIT_FAILURE; }
    std::ifstream in1(pos[1]), in2;
    if (!in1) { std::cerr << "Unable to open '" << pos[1] << "' " << std::strerror(errno) << "\n"; return EXIT_FAILURE; }
    if (paired) {
        in2.open(pos[2]);
        if (!in2) { std::cerr << "Unable to open '" << pos[2] << "' " << std::strerror(errno) << "\n"; return EXIT_FAILURE; }
    }
    msgs_ctx* ctx = nullptr;
    if (msgs_ctx_create(0, &ctx) != 0) { std::cerr << "no usable GPU\n"; return EXIT_FAILURE; }
    bool ok = guessFormat(pos[0]) == FASTA ? loadFastaDb(ctx, pos[0]) : (msgs_db_load(ctx, pos[0].c_str()) == 0);
    if (!ok) { std::cerr << "Unable to load database: " << msgs_last_error(ctx) << "\n"; msgs_ctx_destroy(ctx); return EXIT_FAILURE; }
    Header hdr;
    Sink out;
    if (asBam && msgs_ctx_set(ctx, "record_format", 1) != 0) { std::cerr << msgs_last_error(ctx) << "\n"; return EXIT_FAILURE; }
    if (!makeHeader(ctx, argv[0], cmd, hdr) || !out.open(outFn, asBam) || !out.header(hdr)) {
        std::cerr << "Unable to write '" << outFn << "'\n";
        msgs_ctx_destroy(ctx);
        return EXIT_FAILURE;
    }

    Reader r1{in1, fmt}, r2{in2, fmt};
    uint64_t total = 0;
    int rc = 0;
    try {
        while (true) {
            std::vector<Rec> a, b;
            Rec x, y;
            while ((int64_t)a.size() < batch && r1.next(x)) {
                if (paired) {
                    if (!r2.next(y)) break;
                    if (x.name != y.name) {   /* trimNameExt (src/msgseqtk-align.cpp:668-674) */
                        std::string nx = x.name.size() > 2 ? x.name.substr(0, x.name.size() - 2) : x.name;
                        std::string ny = y.name.size() > 2 ? y.name.substr(0, y.name.size() - 2) : y.name;
                        if (nx != ny) {
                            std::cerr << "Error: unmatched read ID between " << x.name << " and " << y.name << "\n";
                            msgs_ctx_destroy(ctx);
                            return EXIT_FAILURE;
                        }
                        x.name = nx; y.name = ny;
                    }
                    b.push_back(y);
                }
                a.push_back(x);
            }
            if (a.empty()) break;
            auto pack = [](const std::vector<Rec>& v, std::vector<uint8_t>& seq, std::vector<uint8_t>& qual,
                           std::vector<int64_t>& offs, std::string& names, msgs_reads& rs) {
                offs.assign(1, 0);
                for (const Rec& r : v) {
                    seq.insert(seq.end(), r.seq.begin(), r.seq.end());
                    qual.insert(qual.end(), r.qual.begin(), r.qual.end());
                    offs.push_back((int64_t)seq.size());
                    names += r.name + "\n";
                }
                seq.push_back(0); qual.push_back(0);
                rs.seq = seq.data(); rs.qual = qual.data(); rs.offs = offs.data(); rs.names = names.c_str();
                rs.n = (int64_t)v.size();
            };
            std::vector<uint8_t> s1, q1, s2, q2;
            std::vector<int64_t> o1, o2;
            std::string n1, n2;
            msgs_reads m1, m2;
            pack(a, s1, q1, o1, n1, m1);
            if (paired) pack(b, s2, q2, o2, n2, m2);
            int64_t need = 0;
            std::vector<int64_t> lo(a.size() + 1);
            std::string text(1024 * a.size() + 4096, '\0');
            rc = msgs_map_batch(ctx, &m1, paired ? &m2 : nullptr, &p, &text[0], (int64_t)text.size(), lo.data(), &need);
            if (rc == MSGS_E_CAP) {
                text.assign(need + 1, '\0');
                rc = msgs_map_batch(ctx, &m1, paired ? &m2 : nullptr, &p, &text[0], (int64_t)text.size(), lo.data(), &need);
            }
            if (rc) { std::cerr << "mapping failed: " << msgs_last_error(ctx) << "\n"; break; }
            if (!out.records(text.data(), need)) { std::cerr << "Unable to write '" << outFn << "'\n"; rc = 1; break; }
            total += a.size();
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        rc = 1;
    }
    if (!out.close() && !rc) { std::cerr << "Unable to write '" << outFn << "'\n"; rc = 1; }
    std::cerr << "Read alignment finished. Total processed reads: " << total << "\n";
    msgs_ctx_destroy(ctx);
    return rc ? EXIT_FAILURE : EXIT_SUCCESS;
}
