/*
 * msgs_samfmt.h -- SAM text fields written through a raw cursor (host code).
 *
 * The record layout is htslib's sam_format1 as called by the reference's output path
 * (src/msgseqtk-align.cpp:538-593, src/SAMtools/BAM.cpp:20-56); float tags are set with
 * BAM::setAux(double) -> float (src/SAMtools/BAM.h:618-620) and printed by htslib as
 * "%g" of the float widened to double.
 */
#pragma once
#include <cstdint>
#include <string>

namespace msgs_sam {

/* decimal integer */
char* putInt(char* p, int64_t v);
/* printf("%g", (double)f): six significant digits, rounded half-to-even on the exact
 * binary value, trailing zeros removed */
char* putG6(char* p, float f);
/* SEQ of an nt16 read (reverse complement when rc) and QUAL (+33, reversed when rc) */
char* putSeq(char* p, const uint8_t* s, int64_t L, bool rc);
char* putQual(char* p, const uint8_t* q, int64_t L, bool rc);

/* One alignment record's fields, as the reference's BAM constructor and output() set them
 * (src/SAMtools/BAM.cpp:20-56, src/Alignment.cpp:429-444, src/msgseqtk-align.cpp:538-593):
 * no mate (mtid -1, mpos 0), SEQ / QUAL of the read or of its reverse complement. */
struct BamFields {
    const char* name; int nameLen;
    int flag; int32_t tid; int64_t pos; int mapq;
    const uint32_t* cigar; int nCigar;      /* BAM CIGAR words (length << 4 | op) */
    int32_t isize;
    const uint8_t* seq; const uint8_t* qual; int64_t L; bool rc;
    bool tags;                              /* AS, MD, XH, NH, XN, XP (mapped records) */
    int64_t as; const char* md; int mdLen; float xh; int64_t nh, xn; float xp;
};
/* bytes putBam writes at most */
size_t bamBound(const BamFields& f);
/* the record as bam_write1 puts it in the BGZF stream: block_size, the 32-byte core, the read
 * name with one NUL, CIGAR, 4-bit SEQ, QUAL and the tags (integer tags in the smallest type
 * bam_aux_update_int picks, float tags as 'f', MD as 'Z') */
char* putBam(char* p, const BamFields& f);
/* hts_reg2bin(beg, end, 14, 5) */
int32_t reg2bin(int64_t beg, int64_t end);

/* an output text with a write cursor: grows its string in large steps; finish() trims it */
struct Out {
    std::string* buf = nullptr;
    size_t n = 0;
    size_t size() const { return n; }
    char* reserve(size_t k) {
        if (n + k > buf->size()) buf->resize(std::max(n + k, 2 * buf->size() + ((size_t)1 << 16)));
        return &(*buf)[n];
    }
    void commit(const char* end) { n = (size_t)(end - buf->data()); }
    void finish() { buf->resize(n); }
};

}  // namespace msgs_sam
