"""A BAM reader for the tests: BGZF members (header, 'BC' size field, CRC-32 and length
checked), the BAM header, and records printed back as SAM text the way htslib's sam_format1
prints them (integer tags as ':i:', float tags with '%g')."""
import gzip
import struct
import zlib

EOF_BLOCK = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")
OPS = "MIDNSHP=XB"
NT16 = "=ACMGRSVTWYHKDBN"
INT_TYPES = {"c": "<b", "C": "<B", "s": "<h", "S": "<H", "i": "<i", "I": "<I"}


def blocks(raw):
    """the BGZF members of a file: (member bytes, uncompressed data)"""
    out, k = [], 0
    while k < len(raw):
        assert raw[k:k + 4] == b"\x1f\x8b\x08\x04" and raw[k + 12:k + 16] == b"BC\x02\x00", k
        size = struct.unpack_from("<H", raw, k + 16)[0] + 1
        member = raw[k:k + size]
        data = zlib.decompress(member[18:-8], -15)
        crc, isize = struct.unpack_from("<II", member, size - 8)
        assert crc == zlib.crc32(data) and isize == len(data)
        assert gzip.decompress(member) == data
        out.append((member, data))
        k += size
    return out


def read(path):
    """(header text, [(name, length)], [record bytes], blocks)"""
    raw = open(path, "rb").read()
    bl = blocks(raw)
    data = b"".join(d for _, d in bl)
    assert data[:4] == b"BAM\x01"
    lt = struct.unpack_from("<i", data, 4)[0]
    text = data[8:8 + lt].decode()
    k = 8 + lt
    nref = struct.unpack_from("<i", data, k)[0]
    k += 4
    refs = []
    for _ in range(nref):
        ln = struct.unpack_from("<i", data, k)[0]
        name = data[k + 4:k + 4 + ln - 1].decode()
        lref = struct.unpack_from("<i", data, k + 4 + ln)[0]
        refs.append((name, lref))
        k += 8 + ln
    recs = []
    while k < len(data):
        bs = struct.unpack_from("<i", data, k)[0]
        recs.append(data[k:k + 4 + bs])
        k += 4 + bs
    assert k == len(data)
    return text, refs, recs, bl


def tags(rec):
    """[(tag, type, value)] of a record"""
    bs, tid, pos, bmn, fnc, lseq, mtid, mpos, tlen = struct.unpack_from("<9i", rec, 0)
    lname, nc = bmn & 255, fnc & 0xffff
    k = 36 + lname + 4 * nc + (lseq + 1) // 2 + lseq
    out = []
    while k < len(rec):
        tg, ty = rec[k:k + 2].decode(), chr(rec[k + 2])
        k += 3
        if ty in INT_TYPES:
            v = struct.unpack_from(INT_TYPES[ty], rec, k)[0]
            k += struct.calcsize(INT_TYPES[ty])
        elif ty == "f":
            v = struct.unpack_from("<f", rec, k)[0]
            k += 4
        elif ty == "Z":
            e = rec.index(b"\x00", k)
            v = rec[k:e].decode()
            k = e + 1
        else:
            raise ValueError(ty)
        out.append((tg, ty, v))
    return out


def sam_line(rec, refs):
    bs, tid, pos, bmn, fnc, lseq, mtid, mpos, tlen = struct.unpack_from("<9i", rec, 0)
    lname, mapq, flag, nc = bmn & 255, (bmn >> 8) & 255, (fnc >> 16) & 0xffff, fnc & 0xffff
    k = 36
    name = rec[k:k + lname - 1].decode()
    assert rec[k + lname - 1] == 0
    k += lname
    cig = struct.unpack_from("<%dI" % nc, rec, k)
    k += 4 * nc
    seq = rec[k:k + (lseq + 1) // 2]
    k += (lseq + 1) // 2
    qual = rec[k:k + lseq]
    f = [name, str(flag), refs[tid][0] if tid >= 0 else "*", str(pos + 1), str(mapq),
         "".join("%d%s" % (c >> 4, OPS[c & 15]) for c in cig) or "*",
         "*" if mtid < 0 else ("=" if mtid == tid else refs[mtid][0]), str(mpos + 1), str(tlen)]
    if lseq:
        f.append("".join(NT16[(seq[i >> 1] >> (4 * (1 - (i & 1)))) & 15] for i in range(lseq)))
        f.append("*" if qual[0] == 0xff else "".join(chr(q + 33) for q in qual))
    else:
        f += ["*", "*"]
    for tg, ty, v in tags(rec):
        if ty in INT_TYPES:
            f.append("%s:i:%d" % (tg, v))
        elif ty == "f":
            f.append("%s:f:%g" % (tg, v))
        else:
            f.append("%s:Z:%s" % (tg, v))
    return "\t".join(f)
