/*
 * msgs_fmd.cpp -- the reference's FMD-index file DB.fmd, read and written on the host:
 * msgs_fmd_read / msgs_fmd_write (include/msgs_hip.h).
 *
 * Layout (FMDIndex::save / load, src/FMDIndex.cpp:143-161), optionally preceded by the
 * program-info header of every database file ("MSGseqTK" and three int32 version numbers,
 * src/util/ProgEnv.cpp:24-64, src/util/VersionSequence.cpp:35-46):
 *   int64 B[256], int64 C[256]                      base counts, cumulative counts
 *   WaveletTreeRRR of the BWT                       src/libsds/src/WaveletTreeRRR.cpp:14-48
 *     size_t n, sigma, height, wid, min, max, sample_rate, nOCC, OCC[nOCC]
 *     height x BitSeqRRR, level L holding bit (height-1-L) of the symbols in the order of a
 *     stable partition by their higher bits (src/libsds/include/WaveletTreeRRR.h:232-330)
 *   size_t len + int64 gapSA[len]                   SA rows of the '\0' suffixes
 *   BitSeqRRR SAidx                                 rows whose SA value is sampled
 *   size_t len + int64 SAsampled[len]
 * BitSeqRRR (src/libsds/src/BitSeqRRR.cpp:308-360): size_t n, ones; BitStr32 C (4-bit class =
 * popcount of each 15-bit block), BitStr32 O (each block's index among the 15-bit values of its
 * class, in ceil(log2 C(15, class)) bits); size_t nC, wC, nO, wO; BitStr64 Csampled (ones before
 * every 32nd block), BitStr64 Osampled (O position of every 32nd block); size_t nCsampled,
 * wCsampled, nOsampled, wOsampled, sample_rate.  BitStr<T> (src/libsds/include/BitStr.h:386-405):
 * 8 bytes of width (the 4-byte int and 4 bytes after it), size_t bits, size_t words, words.  Bits
 * are numbered from the least significant bit of word 0; a block's value has bit j = bit 15b + j.
 * The values of a class are numbered in the lexicographic order of their set-bit positions
 * (lowest position first), as BitSeqRRR::TableOffset enumerates them (src/libsds/src/BitSeqRRR.cpp:
 * 38-62).  Everything here is written from that format description; nothing is linked from the
 * reference.
 */
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/msgs_hip.h"

namespace {

constexpr int kBlock = 15;
constexpr size_t kSample = 32;

/* number of bits of n (0 for 0) */
uint32_t nbits(uint64_t n) {
    uint32_t b = 0;
    for (; n; n >>= 1) ++b;
    return b;
}

/* class tables: width of a class's offsets, the values of each class in order, and each
 * value's index inside its class */
struct RrrTables {
    uint32_t width[kBlock + 1];
    std::vector<uint16_t> values[kBlock + 1];
    std::vector<uint16_t> index;   /* by 15-bit value */
    RrrTables() : index(1u << kBlock) {
        for (int k = 0; k <= kBlock; ++k) {
            /* k-subsets of {0..14} in lexicographic order of their sorted positions */
            std::vector<int> pos(k);
            for (int i = 0; i < k; ++i) pos[i] = i;
            while (true) {
                uint32_t v = 0;
                for (int i = 0; i < k; ++i) v |= 1u << pos[i];
                index[v] = (uint16_t)values[k].size();
                values[k].push_back((uint16_t)v);
                int i = k - 1;
                while (i >= 0 && pos[i] == kBlock - k + i) --i;
                if (i < 0) break;
                ++pos[i];
                for (int t = i + 1; t < k; ++t) pos[t] = pos[t - 1] + 1;
            }
            width[k] = nbits(values[k].size() - 1);
        }
    }
};
const RrrTables& tables() {
    static const RrrTables t;
    return t;
}

/* a bit string of W-bit words */
template <class T>
struct Bits {
    std::vector<T> w;
    uint64_t nB = 0;
    static constexpr int W = 8 * sizeof(T);
    void resize(uint64_t bits) { nB = bits; w.assign((bits + W - 1) / W, 0); }
    uint64_t get(uint64_t start, uint32_t len) const {
        if (len == 0) return 0;
        uint64_t v = 0;
        for (uint32_t k = 0; k < len && start + k < nB; ++k)
            v |= (uint64_t)((w[(start + k) / W] >> ((start + k) % W)) & 1) << k;
        return v;
    }
    void set(uint64_t start, uint32_t len, uint64_t v) {
        for (uint32_t k = 0; k < len && start + k < nB; ++k) {
            const uint64_t b = start + k;
            const T m = (T)1 << (b % W);
            if ((v >> k) & 1) w[b / W] |= m; else w[b / W] &= (T)~m;
        }
    }
};

struct Reader {
    std::istream& in;
    bool ok = true;
    template <class V> V val() {
        V v{};
        in.read((char*)&v, sizeof(V));
        if (!in) ok = false;
        return v;
    }
    template <class T> void bits(Bits<T>& b) {
        (void)val<uint64_t>();            /* width field, written as 8 bytes */
        b.nB = val<uint64_t>();
        const uint64_t n = val<uint64_t>();
        if (!ok || n > (uint64_t)1 << 40 || n != (b.nB + Bits<T>::W - 1) / Bits<T>::W) { ok = false; return; }
        b.w.resize(n);
        in.read((char*)b.w.data(), n * sizeof(T));
        if (!in) ok = false;
    }
    template <class V> void array(std::vector<V>& a) {
        const uint64_t n = val<uint64_t>();
        if (!ok || n > (uint64_t)1 << 40) { ok = false; return; }
        a.resize(n);
        in.read((char*)a.data(), n * sizeof(V));
        if (!in) ok = false;
    }
};

struct Writer {
    std::ostream& out;
    template <class V> void val(V v) { out.write((const char*)&v, sizeof(V)); }
    template <class T> void bits(const Bits<T>& b) {
        const int32_t wid = Bits<T>::W;
        const int32_t pad = 0;
        val(wid); val(pad);
        val<uint64_t>(b.nB);
        val<uint64_t>(b.w.size());
        out.write((const char*)b.w.data(), b.w.size() * sizeof(T));
    }
    template <class V> void array(const std::vector<V>& a) {
        val<uint64_t>(a.size());
        out.write((const char*)a.data(), a.size() * sizeof(V));
    }
};

/* one RRR bit sequence */
struct Rrr {
    uint64_t n = 0, ones = 0;
    Bits<uint32_t> C, O;
    uint64_t nC = 0, wC = 0, nO = 0, wO = 0;
    Bits<uint64_t> Cs, Os;
    uint64_t nCs = 0, wCs = 0, nOs = 0, wOs = 0, sample = 0;

    bool read(Reader& r) {
        n = r.val<uint64_t>(); ones = r.val<uint64_t>();
        r.bits(C); r.bits(O);
        nC = r.val<uint64_t>(); wC = r.val<uint64_t>(); nO = r.val<uint64_t>(); wO = r.val<uint64_t>();
        r.bits(Cs); r.bits(Os);
        nCs = r.val<uint64_t>(); wCs = r.val<uint64_t>(); nOs = r.val<uint64_t>(); wOs = r.val<uint64_t>();
        sample = r.val<uint64_t>();
        return r.ok && wC == nbits(kBlock) && nC == (n + kBlock - 1) / kBlock && C.nB == nC * wC && O.nB == nO * wO;
    }
    void write(Writer& w) const {
        w.val(n); w.val(ones);
        w.bits(C); w.bits(O);
        w.val(nC); w.val(wC); w.val(nO); w.val(wO);
        w.bits(Cs); w.bits(Os);
        w.val(nCs); w.val(wCs); w.val(nOs); w.val(wOs);
        w.val(sample);
    }
    /* the bits, one byte per bit */
    bool decode(std::vector<uint8_t>& out) const {
        const RrrTables& T = tables();
        out.assign(n, 0);
        uint64_t pos = 0, cnt = 0;
        for (uint64_t b = 0; b < nC; ++b) {
            const uint32_t c = (uint32_t)C.get(b * wC, (uint32_t)wC);
            if (c > kBlock) return false;
            const uint64_t off = O.get(pos, T.width[c]);
            pos += T.width[c];
            if (off >= T.values[c].size()) return false;
            const uint32_t v = T.values[c][off];
            for (int j = 0; j < kBlock && b * kBlock + j < n; ++j) out[b * kBlock + j] = (v >> j) & 1;
            cnt += c;
        }
        return pos == nO && cnt == ones;
    }
    /* built from bits (BitSeqRRR(bstr, 32): build_basic + build_sampled) */
    void encode(const std::vector<uint8_t>& bits) {
        const RrrTables& T = tables();
        n = bits.size();
        wC = nbits(kBlock);
        nC = (n + kBlock - 1) / kBlock;
        C.resize(nC * wC);
        std::vector<uint16_t> val(nC);
        ones = 0; nO = 0; wO = 1;
        for (uint64_t b = 0; b < nC; ++b) {
            uint32_t v = 0;
            for (int j = 0; j < kBlock && b * kBlock + j < n; ++j) v |= (uint32_t)(bits[b * kBlock + j] & 1) << j;
            val[b] = (uint16_t)v;
            const uint32_t c = (uint32_t)__builtin_popcount(v);
            C.set(b * wC, (uint32_t)wC, c);
            ones += c;
            nO += T.width[c];
        }
        O.resize(nO * wO);
        for (uint64_t b = 0, p = 0; b < nC; ++b) {
            const uint32_t c = (uint32_t)__builtin_popcount(val[b]);
            O.set(p, T.width[c], T.index[val[b]]);
            p += T.width[c];
        }
        sample = kSample;
        nCs = nC / sample + 2;
        wCs = nbits(ones);
        Cs.resize(nCs * wCs);
        uint64_t sum = 0;
        for (uint64_t b = 0; b < nC; ++b) {
            if (b % sample == 0) Cs.set(b / sample * wCs, (uint32_t)wCs, sum);
            sum += __builtin_popcount(val[b]);
        }
        for (uint64_t i = (nC + sample - 1) / sample; i < nCs; ++i) Cs.set(i * wCs, (uint32_t)wCs, sum);
        nOs = (nC + sample - 1) / sample;
        wOs = nbits(nO);
        Os.resize(nOs * wOs);
        for (uint64_t b = 0, p = 0; b < nC; ++b) {
            if (b % sample == 0) Os.set(b / sample * wOs, (uint32_t)wOs, p);
            p += T.width[__builtin_popcount(val[b])];
        }
    }
};

constexpr uint64_t kMaxSym = 15;   /* nt16 */

struct Fmd {
    int64_t B[256] = {}, C[256] = {};
    uint64_t n = 0, sigma = 0, height = 0, wid = 0, mn = 0, mx = 0, sample = 0;
    std::vector<uint64_t> occ;
    std::vector<Rrr> levels;
    std::vector<int64_t> gapSA, saSampled;
    Rrr saIdx;
};

bool readHeader(Reader& r) {
    char name[8];
    r.in.read(name, 8);
    int32_t ver[3];
    r.in.read((char*)ver, sizeof ver);
    if (!r.in || std::memcmp(name, "MSGseqTK", 8) != 0) return false;
    /* the reference refuses files of a newer version than its own (v1.2.3) */
    const int64_t v = (int64_t)ver[0] * 1000000 + (int64_t)ver[1] * 1000 + ver[2];
    return v <= 1002003;
}

bool readFmd(std::istream& in, bool header, Fmd& f) {
    Reader r{in};
    if (header && !readHeader(r)) return false;
    in.read((char*)f.B, sizeof f.B);
    in.read((char*)f.C, sizeof f.C);
    f.n = r.val<uint64_t>(); f.sigma = r.val<uint64_t>();
    f.height = r.val<uint64_t>(); f.wid = r.val<uint64_t>(); f.mn = r.val<uint64_t>(); f.mx = r.val<uint64_t>();
    f.sample = r.val<uint64_t>();
    r.array(f.occ);
    if (!r.ok || f.height == 0 || f.height > 8 || f.mx > kMaxSym || f.occ.size() != f.mx + 2) return false;
    f.levels.resize(f.height);
    for (Rrr& l : f.levels)
        if (!l.read(r) || l.n != f.n) return false;
    r.array(f.gapSA);
    if (!f.saIdx.read(r)) return false;
    r.array(f.saSampled);
    return r.ok;
}

/* the BWT from the wavelet-tree levels: symbol i's bit at level L sits at the position of i in
 * the stable partition of the symbols by their higher bits */
bool decodeBwt(const Fmd& f, std::vector<uint8_t>& bwt) {
    const uint64_t n = f.n, H = f.height;
    std::vector<std::vector<uint8_t>> lv(H);
    for (uint64_t L = 0; L < H; ++L)
        if (!f.levels[L].decode(lv[L])) return false;
    /* start of each prefix group at each level, from the symbol counts */
    std::vector<uint64_t> cnt(1u << H, 0);
    for (uint64_t s = 0; s + 1 < f.occ.size() && s < cnt.size(); ++s) cnt[s] = f.occ[s + 1] - f.occ[s];
    std::vector<std::vector<uint64_t>> start(H), used(H);
    for (uint64_t L = 0; L < H; ++L) {
        const uint64_t groups = 1ull << L;
        start[L].assign(groups, 0);
        used[L].assign(groups, 0);
        uint64_t acc = 0;
        for (uint64_t g = 0; g < groups; ++g) {
            start[L][g] = acc;
            for (uint64_t s = 0; s < cnt.size(); ++s)
                if ((s >> (H - L)) == g) acc += cnt[s];
        }
    }
    bwt.assign(n, 0);
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t g = 0;
        for (uint64_t L = 0; L < H; ++L) {
            const uint64_t p = start[L][g] + used[L][g]++;
            if (p >= n) return false;
            g = (g << 1) | lv[L][p];
        }
        bwt[i] = (uint8_t)g;
    }
    /* the counts must agree with the file's */
    std::vector<uint64_t> seen(cnt.size(), 0);
    for (uint8_t b : bwt) seen[b]++;
    return seen == cnt;
}

}  // namespace

extern "C" int msgs_fmd_read(const char* path, int32_t header, int64_t* n, uint8_t* bwt, int64_t* B256, int64_t* C256,
                             int64_t* gapSA, int64_t* nGap, int64_t* saSampled, int64_t* nSampled) {
    if (!path || !n) return MSGS_E_ARG;
    try {
        std::ifstream in(path, std::ios::binary);
        if (!in) return MSGS_E_ARG;
        Fmd f;
        if (!readFmd(in, header != 0, f)) return MSGS_E_DB;
        *n = (int64_t)f.n;
        if (nGap) *nGap = (int64_t)f.gapSA.size();
        if (nSampled) *nSampled = (int64_t)f.saSampled.size();
        if (B256) std::memcpy(B256, f.B, sizeof f.B);
        if (C256) std::memcpy(C256, f.C, sizeof f.C);
        if (gapSA) std::memcpy(gapSA, f.gapSA.data(), f.gapSA.size() * 8);
        if (saSampled) std::memcpy(saSampled, f.saSampled.data(), f.saSampled.size() * 8);
        if (bwt) {
            std::vector<uint8_t> b;
            if (!decodeBwt(f, b)) return MSGS_E_DB;
            std::memcpy(bwt, b.data(), b.size());
        }
        return MSGS_OK;
    } catch (const std::bad_alloc&) {
        return MSGS_E_NOMEM;
    }
}

/* The file FMDIndex(bdSeq, true, 32) + save would write for a BD text with this BWT and SA
 * (src/FMDIndex.cpp:29-116,143-151,406-480): counts, the wavelet tree of the BWT, gapSA =
 * SA[0 .. numGaps), SAidx marking rows with i % 32 == 0 or BWT[i] == 0, their SA values. */
int fmdWrite(const std::string& path, bool header, const uint8_t* bwt, const int64_t* sa, int64_t n) {
    try {
        std::ofstream out(path, std::ios::binary | std::ios::trunc);
        if (!out) return MSGS_E_ARG;
        Writer w{out};
        if (header) {
            out.write("MSGseqTK", 8);
            const int32_t ver[3] = {1, 2, 3};
            out.write((const char*)ver, sizeof ver);
        }
        int64_t B[256] = {}, C[256] = {};
        for (int64_t i = 0; i < n; ++i) B[bwt[i]]++;
        int64_t S = 0;
        for (int b = 0; b <= 16; ++b) { C[b] = S; S += B[b]; }
        out.write((const char*)B, sizeof B);
        out.write((const char*)C, sizeof C);
        /* wavelet tree over symbols 0..15: height 4, OCC = cumulative counts shifted by one */
        const uint64_t H = nbits(kMaxSym);
        w.val<uint64_t>(n); w.val<uint64_t>(kMaxSym + 1);
        w.val<uint64_t>(H); w.val<uint64_t>(8); w.val<uint64_t>(0); w.val<uint64_t>(kMaxSym);
        w.val<uint64_t>(kSample);
        std::vector<uint64_t> occ(kMaxSym + 2, 0);
        for (int64_t i = 0; i < n; ++i) occ[bwt[i] + 1]++;
        for (uint64_t i = 1; i < occ.size(); ++i) occ[i] += occ[i - 1];
        w.array(occ);
        /* level L: bit (H-1-L) of each symbol, symbols in the stable order of their top L bits */
        std::vector<uint64_t> cnt(1u << H, 0);
        for (int64_t i = 0; i < n; ++i) cnt[bwt[i]]++;
        std::vector<uint8_t> lv(n);
        for (uint64_t L = 0; L < H; ++L) {
            const uint64_t groups = 1ull << L;
            std::vector<uint64_t> pos(groups, 0);
            uint64_t acc = 0;
            for (uint64_t g = 0; g < groups; ++g) {
                pos[g] = acc;
                for (uint64_t s = 0; s < cnt.size(); ++s)
                    if ((s >> (H - L)) == g) acc += cnt[s];
            }
            for (int64_t i = 0; i < n; ++i) {
                const uint64_t s = bwt[i];
                lv[pos[s >> (H - L)]++] = (uint8_t)((s >> (H - 1 - L)) & 1);
            }
            Rrr r;
            r.encode(lv);
            r.write(w);
        }
        std::vector<int64_t> gap(sa, sa + B[0]);
        w.array(gap);
        /* the sampled values array has (n + 1) / 32 + numGaps entries, as sampleSA(SA, bwt)
         * sizes it (src/FMDIndex.cpp:461-480): rows past its end are not kept, unused entries
         * stay zero */
        std::vector<uint8_t> mark(n, 0);
        std::vector<int64_t> sampled((size_t)((n + 1) / 32 + B[0]), 0);
        size_t j = 0;
        for (int64_t i = 0; i < n; ++i)
            if (i % 32 == 0 || bwt[i] == 0) {
                mark[i] = 1;
                if (j < sampled.size()) sampled[j] = sa[i];
                ++j;
            }
        Rrr idx;
        idx.encode(mark);
        idx.write(w);
        w.array(sampled);
        out.flush();
        return out ? MSGS_OK : MSGS_E_ARG;
    } catch (const std::bad_alloc&) {
        return MSGS_E_NOMEM;
    }
}

extern "C" int msgs_fmd_write_arrays(const char* path, int32_t header, const uint8_t* bwt, const int64_t* sa, int64_t n) {
    if (!path || n <= 0 || !bwt || !sa) return MSGS_E_ARG;
    for (int64_t i = 0; i < n; ++i)
        if (bwt[i] > 15 || sa[i] < 0 || sa[i] >= n) return MSGS_E_ARG;
    return fmdWrite(path, header != 0, bwt, sa, n);
}
