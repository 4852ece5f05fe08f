/*
 * msgs_bgzf.h -- BGZF output for the driver's BAM files (host code, zlib).
 *
 * The block rules of htslib 1.9's BGZF writer, which the reference's SAMfile writes BAM
 * through (sam_open(path, "wb"), src/SAMtools/SAMfile.h:33-35; the reference leaves the
 * htslib version open): data fill 0xff00-byte uncompressed blocks; a record that does not fit
 * in the current block's remainder first closes the block (bgzf_flush_try before each
 * bam_write1); the BAM header is followed by a flush (bam_hdr_write); each block is raw
 * deflate at zlib's default level (window 2^15, memory level 8, default strategy) inside a
 * gzip member whose extra field is the 'BC' subfield holding the member size - 1, then the
 * CRC-32 and length of the uncompressed data; closing writes an empty block (the 28-byte BGZF
 * end-of-file marker of the SAM specification).
 */
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <zlib.h>

namespace msgs_bgzf {

constexpr size_t kBlock = 0xff00;            /* uncompressed bytes per block */
constexpr size_t kMaxMember = 65536;         /* a BGZF member's size limit */

class Writer {
public:
    explicit Writer(FILE* f) : f_(f) { buf_.reserve(kBlock); }
    bool ok() const { return ok_; }
    /* bgzf_write: fill blocks of kBlock bytes, compressing each as it fills */
    void write(const void* data, size_t n) {
        const uint8_t* p = (const uint8_t*)data;
        while (n) {
            const size_t k = std::min(n, kBlock - buf_.size());
            buf_.insert(buf_.end(), p, p + k);
            p += k; n -= k;
            if (buf_.size() == kBlock) flush();
        }
    }
    /* bgzf_flush_try: close the block first when `need` more bytes would not fit */
    void reserveRecord(size_t need) {
        if (buf_.size() + need > kBlock) flush();
    }
    void flush() {
        if (!buf_.empty()) emit(buf_.data(), buf_.size());
        buf_.clear();
    }
    /* flush, then the empty end-of-file block */
    void close() {
        flush();
        emit(nullptr, 0);
        if (std::fflush(f_) != 0) ok_ = false;
    }

private:
    void emit(const uint8_t* src, size_t n) {
        uint8_t out[kMaxMember];
        z_stream zs;
        std::memset(&zs, 0, sizeof zs);
        if (deflateInit2(&zs, Z_DEFAULT_COMPRESSION, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) { ok_ = false; return; }
        zs.next_in = const_cast<Bytef*>(src);
        zs.avail_in = (uInt)n;
        zs.next_out = out + 18;
        zs.avail_out = (uInt)(kMaxMember - 18 - 8);
        const int rc = deflate(&zs, Z_FINISH);
        const size_t body = zs.total_out;
        deflateEnd(&zs);
        if (rc != Z_STREAM_END) { ok_ = false; return; }   /* 0xff00 bytes always fit */
        const size_t total = 18 + body + 8;
        static const uint8_t kHead[16] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 'B', 'C', 2, 0};
        std::memcpy(out, kHead, 16);
        out[16] = (uint8_t)((total - 1) & 0xff);
        out[17] = (uint8_t)((total - 1) >> 8);
        const uint32_t crc = (uint32_t)crc32(crc32(0L, Z_NULL, 0), src, (uInt)n);
        const uint32_t len = (uint32_t)n;
        for (int k = 0; k < 4; ++k) out[18 + body + k] = (uint8_t)(crc >> (8 * k));
        for (int k = 0; k < 4; ++k) out[22 + body + k] = (uint8_t)(len >> (8 * k));
        if (std::fwrite(out, 1, total, f_) != total) ok_ = false;
    }

    FILE* f_;
    std::vector<uint8_t> buf_;
    bool ok_ = true;
};

}  // namespace msgs_bgzf
