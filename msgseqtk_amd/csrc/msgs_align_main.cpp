/*
 * msgs_align_main.cpp -- msgseqtk-align-amd, a drop-in for msgseqtk-align
 * (src/msgseqtk-align.cpp) that maps through libmsgs_hip.so.
 *
 *   msgseqtk-align-amd <DB> <READ-FILE> [MATE-FILE] -o <OUT.sam> [options]
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

int writeHeader(msgs_ctx* ctx, std::ostream& out, const std::string& cmd) {
    out << "@HD\tVN:1.6\tSO:unsorted\n";
    int32_t n = msgs_db_nchrom(ctx);
    char buf[4096];
    for (int32_t t = 0; t < n; ++t) {
        int64_t len = 0;
        if (msgs_db_chrom(ctx, t, buf, sizeof buf, &len) != 0) return 1;
        out << "@SQ\tSN:" << buf << "\tLN:" << len << "\n";
    }
    out << "@PG\tID:msgseqtk-align\tVN:v1.2.3\tPN:msgseqtk-align-amd\tCL:" << cmd << "\n";
    return 0;
}

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
    if (!endsWith(outFn, ".sam")) { std::cerr << "Unable to determine SAM output mode (only .sam is written)\n"; return EXIT_FAILURE; }
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
    std::ofstream out(outFn);
    if (!out || writeHeader(ctx, out, cmd)) { std::cerr << "Unable to write '" << outFn << "'\n"; return EXIT_FAILURE; }

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
            out.write(text.data(), need);
            total += a.size();
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        rc = 1;
    }
    std::cerr << "Read alignment finished. Total processed reads: " << total << "\n";
    msgs_ctx_destroy(ctx);
    return rc ? EXIT_FAILURE : EXIT_SUCCESS;
}
