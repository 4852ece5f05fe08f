/*
 * msgs_samfmt.cpp -- SAM text fields (host code, SSSE3 for SEQ / QUAL).
 *
 * putG6 prints a float the way printf("%g") prints it after widening to double (the C
 * standard's rule for %g with precision 6): round the exact value to six significant
 * digits (half to even), exponent X of that rounded value; style f with 5 - X decimals
 * when -4 <= X < 6, else style e with five decimals and a two-digit minimum exponent;
 * trailing fractional zeros and a bare decimal point removed.  Values in [1e-20, 1e6)
 * are rounded with exact 128-bit integer arithmetic on the float's m * 2^e; the rest
 * (zero, subnormals, very small or large values, inf, nan) go through std::to_chars with
 * chars_format::general and precision 6, which the standard defines as printf's %.6g.
 */
#include "msgs_samfmt.h"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstring>
#include <tmmintrin.h>

namespace msgs_sam {

char* putInt(char* p, int64_t v) { return std::to_chars(p, p + 24, v).ptr; }

static char* g6Library(char* p, float f) {
    return std::to_chars(p, p + 48, (double)f, std::chars_format::general, 6).ptr;
}

struct Pow10Table {
    unsigned __int128 p[31];   /* 10^0 .. 10^30 */
    double d[28];              /* 10^-21 .. 10^6 (nearest doubles; only used for an estimate) */
    Pow10Table() {
        unsigned __int128 r = 1;
        for (int s = 0; s <= 30; ++s) { p[s] = r; r *= 10; }
        for (int k = -21; k <= 6; ++k) d[k + 21] = std::pow(10.0, k);
    }
};
static const Pow10Table kP10;

/* floor(log10(a)) up to one off near powers of ten, for 1e-20 <= a < 1e6 */
static int log10Estimate(double a) {
    int lo = 0, hi = 27;       /* d[lo] <= a < d[hi] */
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (kP10.d[mid] <= a) lo = mid; else hi = mid;
    }
    return lo - 21;
}

char* putG6(char* p, float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t ex = (u >> 23) & 0xffu, mant = u & 0x7fffffu;
    if (ex == 0 || ex == 0xffu) return g6Library(p, f);   /* zero, subnormal, inf, nan */
    const double a = std::fabs((double)f);
    if (!(a >= 1e-20 && a < 1e6)) return g6Library(p, f);
    const uint64_t m = mant | 0x800000u;
    const int sh = 150 - (int)ex;                          /* a = m / 2^sh, sh > 0 here */
    int k = log10Estimate(a);
    uint64_t q = 0;
    for (int tries = 0; tries < 4; ++tries) {
        const int s = 5 - k;                               /* q ~ a * 10^s in [1e5, 1e6) */
        if (s < 0 || s > 30) return g6Library(p, f);
        const unsigned __int128 num = (unsigned __int128)m * kP10.p[s];
        const unsigned __int128 qt = num >> sh, rem = num - (qt << sh);
        if (qt < 100000u) { --k; continue; }
        if (qt >= 1000000u) { ++k; continue; }
        const unsigned __int128 half = (unsigned __int128)1 << (sh - 1);
        q = (uint64_t)qt;
        if (rem > half || (rem == half && (q & 1))) ++q;
        if (q == 1000000u) { q = 100000u; ++k; }
        break;
    }
    if (q == 0) return g6Library(p, f);
    if (u >> 31) *p++ = '-';
    char d[6];
    for (int t = 5; t >= 0; --t) { d[t] = (char)('0' + q % 10); q /= 10; }
    int nd = 6;                                            /* significant digits kept */
    while (nd > 1 && d[nd - 1] == '0') --nd;
    const int X = k;
    if (X >= -4 && X < 6) {
        if (X >= 0) {
            for (int t = 0; t <= X; ++t) *p++ = d[t];
            if (nd > X + 1) { *p++ = '.'; for (int t = X + 1; t < nd; ++t) *p++ = d[t]; }
        } else {
            *p++ = '0'; *p++ = '.';
            for (int t = 0; t < -X - 1; ++t) *p++ = '0';
            for (int t = 0; t < nd; ++t) *p++ = d[t];
        }
    } else {
        *p++ = d[0];
        if (nd > 1) { *p++ = '.'; for (int t = 1; t < nd; ++t) *p++ = d[t]; }
        *p++ = 'e';
        int x = X;
        *p++ = x < 0 ? '-' : '+';
        if (x < 0) x = -x;
        if (x >= 100) { *p++ = (char)('0' + x / 100); x %= 100; }
        *p++ = (char)('0' + x / 10);
        *p++ = (char)('0' + x % 10);
    }
    return p;
}

static const char kNt16[17] = "=ACMGRSVTWYHKDBN";
static const char kNt16c[17] = "=TGKCYSBAWRDMHVN";   /* complement of each code */

__attribute__((target("ssse3"))) char* putSeq(char* p, const uint8_t* s, int64_t L, bool rc) {
    const __m128i tab = _mm_loadu_si128((const __m128i*)(rc ? kNt16c : kNt16));
    const __m128i lo = _mm_set1_epi8(0x0f);
    int64_t i = 0;
    if (!rc) {
        for (; i + 16 <= L; i += 16) {
            const __m128i v = _mm_and_si128(_mm_loadu_si128((const __m128i*)(s + i)), lo);
            _mm_storeu_si128((__m128i*)(p + i), _mm_shuffle_epi8(tab, v));
        }
        for (; i < L; ++i) p[i] = kNt16[s[i] & 15];
    } else {
        const __m128i rev = _mm_setr_epi8(15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0);
        for (; i + 16 <= L; i += 16) {
            __m128i v = _mm_loadu_si128((const __m128i*)(s + L - 16 - i));
            v = _mm_and_si128(_mm_shuffle_epi8(v, rev), lo);
            _mm_storeu_si128((__m128i*)(p + i), _mm_shuffle_epi8(tab, v));
        }
        for (; i < L; ++i) p[i] = kNt16c[s[L - 1 - i] & 15];
    }
    return p + L;
}

__attribute__((target("ssse3"))) char* putQual(char* p, const uint8_t* q, int64_t L, bool rc) {
    const __m128i k33 = _mm_set1_epi8(33);
    int64_t i = 0;
    if (!rc) {
        for (; i + 16 <= L; i += 16)
            _mm_storeu_si128((__m128i*)(p + i), _mm_add_epi8(_mm_loadu_si128((const __m128i*)(q + i)), k33));
        for (; i < L; ++i) p[i] = (char)(q[i] + 33);
    } else {
        const __m128i rev = _mm_setr_epi8(15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0);
        for (; i + 16 <= L; i += 16) {
            const __m128i v = _mm_shuffle_epi8(_mm_loadu_si128((const __m128i*)(q + L - 16 - i)), rev);
            _mm_storeu_si128((__m128i*)(p + i), _mm_add_epi8(v, k33));
        }
        for (; i < L; ++i) p[i] = (char)(q[L - 1 - i] + 33);
    }
    return p + L;
}

/* ------------------------------------------------------------------ BAM records */

int32_t reg2bin(int64_t beg, int64_t end) {
    /* the smallest bin of the UCSC binning scheme (levels of 2^14, 2^17, ... 2^29 bp) holding
     * [beg, end): level l's bins start at (8^l - 1) / 7 */
    --end;
    int s = 14, t = ((1 << 15) - 1) / 7;
    for (int l = 5; l > 0; --l, s += 3, t -= 1 << (3 * l))
        if ((beg >> s) == (end >> s)) return (int32_t)(t + (beg >> s));
    return 0;
}

size_t bamBound(const BamFields& f) {
    return 36 + (size_t)f.nameLen + 1 + 4 * (size_t)f.nCigar + (size_t)(f.L + 1) / 2 + (size_t)f.L +
           (f.tags ? 7 * 4 + (size_t)f.mdLen + 4 + 14 : 0);
}

static inline char* put32(char* p, uint32_t v) { std::memcpy(p, &v, 4); return p + 4; }

/* bam_aux_update_int's choice for a new tag: 'i' below INT16_MIN, 's' below INT8_MIN, 'c'
 * below 0, 'C' below UINT8_MAX, 'S' below UINT16_MAX, else 'I' (strict bounds as htslib has
 * them: 255 is stored as 'S') */
static char* putAuxInt(char* p, const char* tag, int64_t v) {
    *p++ = tag[0]; *p++ = tag[1];
    if (v < -32768) { *p++ = 'i'; return put32(p, (uint32_t)(int32_t)v); }
    if (v < -128) { *p++ = 's'; const int16_t x = (int16_t)v; std::memcpy(p, &x, 2); return p + 2; }
    if (v < 0) { *p++ = 'c'; *p++ = (char)(int8_t)v; return p; }
    if (v < 255) { *p++ = 'C'; *p++ = (char)(uint8_t)v; return p; }
    if (v < 65535) { *p++ = 'S'; const uint16_t x = (uint16_t)v; std::memcpy(p, &x, 2); return p + 2; }
    *p++ = 'I';
    return put32(p, (uint32_t)v);
}

static char* putAuxFloat(char* p, const char* tag, float v) {
    *p++ = tag[0]; *p++ = tag[1]; *p++ = 'f';
    std::memcpy(p, &v, 4);
    return p + 4;
}

char* putBam(char* p, const BamFields& f) {
    static const uint8_t kComp[16] = {0, 8, 4, 12, 2, 10, 6, 14, 1, 9, 5, 13, 3, 11, 7, 15};
    char* start = p;
    p += 4;   /* block_size, below */
    int64_t rlen = 0;
    for (int k = 0; k < f.nCigar; ++k) {
        const uint32_t op = f.cigar[k] & 15u;
        if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) rlen += f.cigar[k] >> 4;   /* M D N = X */
    }
    p = put32(p, (uint32_t)f.tid);
    p = put32(p, (uint32_t)(int32_t)f.pos);
    p = put32(p, (uint32_t)reg2bin(f.pos, f.pos + rlen) << 16 | (uint32_t)(f.mapq & 255) << 8 |
                     (uint32_t)(f.nameLen + 1));
    p = put32(p, (uint32_t)f.flag << 16 | (uint32_t)f.nCigar);
    p = put32(p, (uint32_t)f.L);
    p = put32(p, 0xffffffffu);   /* next refID -1 */
    p = put32(p, 0u);            /* next pos 0 */
    p = put32(p, (uint32_t)f.isize);
    std::memcpy(p, f.name, f.nameLen); p += f.nameLen;
    *p++ = 0;
    if (f.nCigar) { std::memcpy(p, f.cigar, 4 * (size_t)f.nCigar); p += 4 * (size_t)f.nCigar; }
    /* SEQ: two codes per byte, the first in the high nibble, 0 after an odd last base */
    for (int64_t i = 0; i < f.L; i += 2) {
        const uint8_t a = f.rc ? kComp[f.seq[f.L - 1 - i] & 15] : (uint8_t)(f.seq[i] & 15);
        const uint8_t b = i + 1 < f.L ? (f.rc ? kComp[f.seq[f.L - 2 - i] & 15] : (uint8_t)(f.seq[i + 1] & 15)) : 0;
        *p++ = (char)(a << 4 | b);
    }
    for (int64_t i = 0; i < f.L; ++i) *p++ = (char)(f.rc ? f.qual[f.L - 1 - i] : f.qual[i]);
    if (f.tags) {
        p = putAuxInt(p, "AS", f.as);
        *p++ = 'M'; *p++ = 'D'; *p++ = 'Z';
        std::memcpy(p, f.md, f.mdLen); p += f.mdLen;
        *p++ = 0;
        p = putAuxFloat(p, "XH", f.xh);
        p = putAuxInt(p, "NH", f.nh);
        p = putAuxInt(p, "XN", f.xn);
        p = putAuxFloat(p, "XP", f.xp);
    }
    put32(start, (uint32_t)(p - start - 4));
    return p;
}

}  // namespace msgs_sam

/* test hook (no reference counterpart): putG6 of n floats, each result NUL-terminated in
 * out[16 * i .. 16 * i + 16) */
extern "C" void msgs_util_format_g6(const float* v, int64_t n, char* out) {
    for (int64_t i = 0; i < n; ++i) {
        char* e = msgs_sam::putG6(out + 16 * i, v[i]);
        *e = 0;
    }
}
