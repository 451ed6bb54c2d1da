"""The product's BAM reader (bsq_sample_read) against BAM files written by tests/bam_writer.py, an
encoder written from the SAMv1 specification rather than from the product's own writer: BGZF blocks
with their BC fields and the EOF block, records straddling blocks, every CIGAR operator, and aux
fields of every type (the reader must take X1/XL/XR from any integer type, XP and RG from strings,
and skip the rest)."""
import re

import numpy as np
import pytest

import breseq_amd as B
from bam_writer import CODES, OPS, write_bam, write_fasta

INT_TYPES = "cCsSiI"


def _reads(rng, refs, n):
    reads = []
    for i in range(n):
        tid = int(rng.integers(len(refs)))
        L = len(refs[tid][1])
        ops = []
        if rng.random() < 0.2:
            ops.append((int(rng.integers(1, 9)), "S"))
        if rng.random() < 0.05:
            ops.append((int(rng.integers(1, 4)), "H"))
        for k in range(int(rng.integers(1, 4))):
            ops.append((int(rng.integers(5, 60)), str(rng.choice(list("M=X")))))
            if k < 2 and rng.random() < 0.5:
                ops.append((int(rng.integers(1, 5)), str(rng.choice(list("IDNP")))))
        if rng.random() < 0.2:
            ops.append((int(rng.integers(1, 9)), "S"))
        qlen = sum(n for n, o in ops if o in "MIS=X")
        span = sum(n for n, o in ops if o in "MDN=X")
        pos = int(rng.integers(0, L - span))
        seq = "".join(rng.choice(list("ACGTN"), qlen, p=[0.24, 0.24, 0.24, 0.24, 0.04]))
        aux = [("AS", "i", int(rng.integers(0, 300)))]
        if rng.random() < 0.7:
            aux.append(("X1", str(rng.choice(list(INT_TYPES))), int(rng.integers(1, 4))))
        if rng.random() < 0.6:
            aux.append(("XL", str(rng.choice(list(INT_TYPES))), int(rng.integers(0, 20))))
            aux.append(("XR", str(rng.choice(list(INT_TYPES))), int(rng.integers(0, 20))))
        if rng.random() < 0.3:
            aux.append(("XP", "Z", str(rng.choice(["FR", "RF", "FF", "RR", "F"]))))
        if rng.random() < 0.8:
            aux.append(("RG", "Z", str(rng.choice(["lib1", "lib2", "unknown"]))))
        if rng.random() < 0.3:
            aux.append(("ZA", "A", "q"))
            aux.append(("ZF", "f", 1.5))
            aux.append(("ZB", "B", ("s", [1, -2, 3])))
            aux.append(("ZH", "H", "1AE301"))
        rng.shuffle(aux)
        reads.append(dict(name="r%d" % i, tid=tid, pos=pos, flag=int(rng.choice([0, 16, 1 | 2 | 64, 1 | 2 | 128 | 16])),
                          mapq=int(rng.integers(0, 61)), cigar="".join("%d%s" % (n, o) for n, o in ops), seq=seq,
                          qual=[int(q) for q in rng.integers(2, 42, qlen)], tlen=int(rng.integers(-500, 500)),
                          aux=aux))
    reads.sort(key=lambda r: (r["tid"], r["pos"]))
    return reads


@pytest.mark.parametrize("block_size", [977, 65280])
def test_reader_reproduces_the_specification_encoder(tmp_path, block_size):
    rng = np.random.default_rng(block_size)
    refs = [("chrA", "".join(rng.choice(list("ACGT"), 5000))), ("plasmid", "".join(rng.choice(list("ACGT"), 1200)))]
    reads = _reads(rng, refs, 1500)
    bam, fa = str(tmp_path / "x.bam"), str(tmp_path / "x.fasta")
    write_bam(bam, refs, reads, read_groups=("lib1", "lib2"), block_size=block_size)
    write_fasta(fa, refs)
    s = B.Sample.read(bam, fa)
    assert s.reference_sequences() == [(n, q) for n, q in refs]
    a = s.arrays()
    n = len(reads)
    assert a["pos"].size == n
    groups = ["lib1", "lib2"]
    xp = {"FR": 0, "RF": 1, "FF": 2}
    cig_all, seq_all, qual_all = [], [], []
    for i, r in enumerate(reads):
        tags = {t: v for t, _, v in r["aux"]}
        assert a["tid"][i] == r["tid"] and a["pos"][i] == r["pos"] and a["flag"][i] == r["flag"], i
        assert a["mapq"][i] == r["mapq"] and a["tlen"][i] == r["tlen"], i
        assert a["x1"][i] == tags.get("X1", -1) and a["xl"][i] == tags.get("XL", -1) and a["xr"][i] == tags.get("XR", -1)
        rg = tags.get("RG")
        assert a["rg"][i] == (groups.index(rg) if rg in groups else -1), (i, rg)
        assert a["xp"][i] == xp.get(tags.get("XP"), -1), (i, tags.get("XP"))
        cig_all += [(int(k) << 4) | OPS.index(o) for k, o in re.findall(r"(\d+)([MIDNSHP=X])", r["cigar"])]
        seq_all += [CODES.index(c) for c in r["seq"]]
        qual_all += r["qual"]
    assert np.array_equal(a["cigar"], np.array(cig_all, dtype=np.uint32))
    assert np.array_equal(a["seq"], np.array(seq_all, dtype=np.uint8))
    assert np.array_equal(a["qual"], np.array(qual_all, dtype=np.uint8))
    assert s.read_names() == [r["name"] for r in reads]


def test_reader_rejects_a_corrupted_block(tmp_path):
    rng = np.random.default_rng(1)
    refs = [("chrA", "".join(rng.choice(list("ACGT"), 3000)))]
    bam, fa = str(tmp_path / "x.bam"), str(tmp_path / "x.fasta")
    write_bam(bam, refs, _reads(rng, refs, 300), block_size=977)
    write_fasta(fa, refs)
    raw = bytearray(open(bam, "rb").read())
    bsize = int.from_bytes(raw[16:18], "little") + 1
    raw[bsize - 8] ^= 0xFF   # the first block's CRC32
    open(bam, "wb").write(bytes(raw))
    with pytest.raises(B.BsqError):
        B.Sample.read(bam, fa)
