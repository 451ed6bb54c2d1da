"""A BAM writer written from the SAM/BAM format specification (SAMv1, sections 4.1-4.2), independent
of the product's own encoder, for testing the product's BAM reader.

BGZF: each block is a gzip member with FEXTRA set and one 'BC' subfield holding the block size minus
one; the payload is raw DEFLATE of at most 64 KiB; CRC32 and ISIZE close it; an empty block ends the
file.  Blocks here are cut at a fixed uncompressed size, so records straddle block boundaries.

BAM: "BAM\\1", the header text, the reference list, then per record block_size, refID, pos,
l_read_name, mapq, bin, n_cigar_op, flag, l_seq, next_refID, next_pos, tlen, the NUL-terminated
name, the CIGAR words (len << 4 | op), 4-bit packed bases, qualities, then the aux fields.
"""
import re
import struct
import zlib

OPS = "MIDNSHP=X"
CODES = "=ACMGRSVTWYHKDBN"
EOF_BLOCK = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")


def bgzf_blocks(data, block_size):
    out = bytearray()
    for i in range(0, len(data), block_size):
        chunk = data[i:i + block_size]
        c = zlib.compressobj(6, zlib.DEFLATED, -15)
        comp = c.compress(chunk) + c.flush()
        bsize = 12 + 6 + len(comp) + 8   # gzip header with XLEN, extra subfield, payload, trailer
        out += struct.pack("<BBBBIBBH", 0x1F, 0x8B, 8, 4, 0, 0, 0xFF, 6)
        out += b"BC" + struct.pack("<HH", 2, bsize - 1)
        out += comp
        out += struct.pack("<II", zlib.crc32(chunk) & 0xFFFFFFFF, len(chunk))
    out += EOF_BLOCK
    return bytes(out)


def reg2bin(beg, end):
    end -= 1
    if beg >> 14 == end >> 14:
        return ((1 << 15) - 1) // 7 + (beg >> 14)
    if beg >> 17 == end >> 17:
        return ((1 << 12) - 1) // 7 + (beg >> 17)
    if beg >> 20 == end >> 20:
        return ((1 << 9) - 1) // 7 + (beg >> 20)
    if beg >> 23 == end >> 23:
        return ((1 << 6) - 1) // 7 + (beg >> 23)
    if beg >> 26 == end >> 26:
        return ((1 << 3) - 1) // 7 + (beg >> 26)
    return 0


def aux_field(tag, typ, value):
    b = tag.encode() + typ.encode()
    if typ == "A":
        return b + value.encode()
    if typ in "cCsSiI":
        return b + struct.pack("<" + {"c": "b", "C": "B", "s": "h", "S": "H", "i": "i", "I": "I"}[typ], value)
    if typ == "f":
        return b + struct.pack("<f", value)
    if typ in "ZH":
        return b + value.encode() + b"\0"
    if typ == "B":
        sub, vals = value
        fmt = {"c": "b", "C": "B", "s": "h", "S": "H", "i": "i", "I": "I", "f": "f"}[sub]
        return b + sub.encode() + struct.pack("<I", len(vals)) + struct.pack("<%d%s" % (len(vals), fmt), *vals)
    raise ValueError(typ)


def record(r):
    cig = [(int(n), OPS.index(o)) for n, o in re.findall(r"(\d+)([MIDNSHP=X])", r["cigar"])]
    seq = r["seq"]
    ref_len = sum(n for n, o in cig if OPS[o] in "MDN=X")
    name = r["name"].encode() + b"\0"
    body = struct.pack("<iiBBHHHiiii", r["tid"], r["pos"], len(name), r["mapq"],
                       reg2bin(r["pos"], r["pos"] + max(ref_len, 1)), len(cig), r["flag"], len(seq),
                       r.get("mate_tid", -1), r.get("mate_pos", -1), r.get("tlen", 0))
    body += name
    body += b"".join(struct.pack("<I", (n << 4) | o) for n, o in cig)
    packed = bytearray()
    for i in range(0, len(seq), 2):
        hi = CODES.index(seq[i])
        lo = CODES.index(seq[i + 1]) if i + 1 < len(seq) else 0
        packed.append((hi << 4) | lo)
    body += bytes(packed)
    body += bytes(r["qual"])
    for tag, typ, value in r.get("aux", []):
        body += aux_field(tag, typ, value)
    return struct.pack("<i", len(body)) + body


def write_bam(path, refs, reads, read_groups=(), block_size=977):
    """refs: [(name, sequence)]; reads: dicts with name, tid, pos, flag, mapq, cigar, seq, qual,
    tlen, aux [(tag, type, value)]."""
    text = "@HD\tVN:1.6\tSO:coordinate\n"
    text += "".join("@SQ\tSN:%s\tLN:%d\n" % (n, len(s)) for n, s in refs)
    text += "".join("@RG\tID:%s\tSM:x\n" % g for g in read_groups)
    data = b"BAM\1" + struct.pack("<i", len(text)) + text.encode() + struct.pack("<i", len(refs))
    for n, s in refs:
        nm = n.encode() + b"\0"
        data += struct.pack("<i", len(nm)) + nm + struct.pack("<i", len(s))
    data += b"".join(record(r) for r in reads)
    with open(path, "wb") as f:
        f.write(bgzf_blocks(data, block_size))


def write_fasta(path, refs):
    with open(path, "w") as f:
        for n, s in refs:
            f.write(">%s\n" % n)
            for i in range(0, len(s), 60):
                f.write(s[i:i + 60] + "\n")
