"""Hand-built alignment sets for known-answer tests (bsq_reference + bsq_alignments structs)."""
import ctypes as C
import re

import numpy as np

from breseq_amd import _lib as L

XP = {"FR": 0, "RF": 1, "FF": 2}
OPS = {"M": 0, "I": 1, "D": 2, "N": 3, "S": 4, "H": 5, "P": 6, "=": 7, "X": 8}
BAM = {"A": 1, "C": 2, "G": 4, "T": 8, "N": 15}


class Built:
    """Keeps the numpy buffers alive while the C structs point at them."""

    def __init__(self, refs, reads):
        self.names = [n.encode() for n, _ in refs]
        self.seqs = [s.encode() for _, s in refs]
        self.lengths = np.array([len(s) for _, s in refs], dtype=np.int64)
        self.c_names = (C.c_char_p * len(refs))(*self.names)
        self.c_seqs = (C.c_char_p * len(refs))(*self.seqs)
        self.ref = L.Reference(len(refs), self.c_names, self.lengths.ctypes.data_as(L.c_i64p), self.c_seqs)
        n = len(reads)
        self.tid = np.array([r.get("tid", 0) for r in reads], dtype=np.int32)
        self.pos = np.array([r["pos"] for r in reads], dtype=np.int32)
        self.flag = np.array([r.get("flag", 0) for r in reads], dtype=np.uint16)
        self.mapq = np.array([r.get("mapq", 42) for r in reads], dtype=np.uint8)
        cig, coff = [], [0]
        seq, qual, soff = [], [], [0]
        for r in reads:
            for ln, op in re.findall(r"(\d+)([MIDNSHP=X])", r["cigar"]):
                cig.append((int(ln) << 4) | OPS[op])
            coff.append(len(cig))
            seq.extend(BAM[b] for b in r["seq"])
            q = r.get("qual", [30] * len(r["seq"]))
            assert len(q) == len(r["seq"])
            qual.extend(q)
            soff.append(len(seq))
        self.cigar = np.array(cig, dtype=np.uint32)
        self.cigar_off = np.array(coff, dtype=np.int64)
        self.seq = np.array(seq, dtype=np.uint8)
        self.qual = np.array(qual, dtype=np.uint8)
        self.seq_off = np.array(soff, dtype=np.int64)
        self.x1 = np.array([r.get("x1", -1) for r in reads], dtype=np.int32)
        self.xl = np.array([r.get("xl", -1) for r in reads], dtype=np.int32)
        self.xr = np.array([r.get("xr", -1) for r in reads], dtype=np.int32)
        self.rg = np.array([r.get("rg", -1) for r in reads], dtype=np.int32)
        self.tlen = np.array([r.get("tlen", 0) for r in reads], dtype=np.int32)
        self.xp = np.array([XP.get(r.get("xp"), -1) for r in reads], dtype=np.int8)
        names = [r.get("name", "r%d" % i).encode() for i, r in enumerate(reads)]
        self.name_off = np.concatenate([[0], np.cumsum([len(x) for x in names])]).astype(np.int64)
        self.name_chars = b"".join(names)
        p = lambda a, t: a.ctypes.data_as(t)
        self.aln = L.Alignments(n, p(self.tid, L.c_i32p), p(self.pos, L.c_i32p), p(self.flag, L.c_u16p),
                                p(self.mapq, L.c_u8p), p(self.cigar_off, L.c_i64p), p(self.cigar, L.c_u32p),
                                p(self.seq_off, L.c_i64p), p(self.seq, L.c_u8p), p(self.qual, L.c_u8p),
                                p(self.x1, L.c_i32p), p(self.xl, L.c_i32p), p(self.xr, L.c_i32p), p(self.rg, L.c_i32p),
                                p(self.tlen, L.c_i32p), p(self.xp, L.c_i8p), p(self.name_off, L.c_i64p),
                                self.name_chars)
        self.base = (C.c_uint32 * 1)(0)
        self.count = (C.c_uint32 * 1)(1)
        self.rf_names = (C.c_char_p * 1)(b"reads")
        self.read_files = L.ReadFiles(1, self.base, self.count, 1, self.rf_names)

    def error_count_args(self, quality_max=42):
        a = L.ErrorCountArgs()
        a.ref = C.pointer(self.ref)
        a.aln = C.pointer(self.aln)
        a.read_files = self.read_files
        self._groups = (C.c_int32 * len(self.names))(*([0] * len(self.names)))
        a.coverage_group = self._groups
        a.n_coverage_groups = 1
        a.call_target = None
        a.read_set_max = 1
        a.quality_max = quality_max
        a.do_coverage = 1
        a.do_errors = 1
        a.output_dir = None
        a.error_rates_file = None
        return a


def bin_index(S, rs, ref, obs, q):
    """Count-table index of covariates (read_set, ref_base, obs_base, quality), error_count.cpp:477-497."""
    return rs + S * ref + 5 * S * obs + 25 * S * q


IDX = {"A": 0, "C": 1, "G": 2, "T": 3, ".": 4}


class BuiltSample:
    """Adapter giving a Built set the Sample interface the argument builders use."""

    def __init__(self, built, n_read_groups=0):
        self.b = built
        self.ref = C.pointer(built.ref)
        self.aln = C.pointer(built.aln)
        self.n_read_groups = n_read_groups

    @property
    def n_targets(self):
        return len(self.b.names)

    def target_lengths(self):
        return [int(x) for x in self.b.lengths]

    def max_quality(self):
        return int(self.b.qual.max()) if self.b.qual.size else 0

    def read_files(self):
        return self.b.read_files


COMP = {"A": "T", "C": "G", "G": "C", "T": "A", "N": "N"}


def random_alignments(seed, n_targets=2, target_len=400, n_reads=600, max_len=60, run_len=(3, 20), clip_len=(1, 6)):
    """Alignments with every CIGAR shape the pileup distinguishes: soft clips, insertions and
    deletions (also adjacent ones), reference skips, pads before insertions, =/X runs, N read bases,
    both strands, redundant reads and reads ending at a target end.  Sorted by (target, position).
    run_len and clip_len bound the lengths of aligned runs and soft clips."""
    rng = np.random.default_rng(seed)
    refs = []
    for t in range(n_targets):
        s = rng.choice(list("ACGT"), size=target_len)
        s[rng.random(target_len) < 0.01] = "N"
        refs.append(("t%d" % t, "".join(s)))
    reads = []
    for _ in range(n_reads):
        t = int(rng.integers(n_targets))
        ref = refs[t][1]
        ops = []
        if rng.random() < 0.3:
            ops.append(("S", int(rng.integers(*clip_len))))
        if rng.random() < 0.05:
            ops.append(("I", int(rng.integers(1, 3))))
        n_runs = 1 if rng.random() < 0.3 else int(rng.integers(2, 5))
        for k in range(n_runs):
            m = "M" if rng.random() < 0.8 else ("=" if rng.random() < 0.5 else "X")
            ops.append((m, int(rng.integers(*run_len))))
            if k == n_runs - 1:
                break
            u = rng.random()
            if u < 0.35:
                ops.append(("D", int(rng.integers(1, 4))))
                if rng.random() < 0.2:
                    ops.append(("D", int(rng.integers(1, 3))))
            elif u < 0.7:
                ops.append(("I", int(rng.integers(1, 4))))
                if rng.random() < 0.2:
                    ops.append(("I", int(rng.integers(1, 3))))
            elif u < 0.8:
                ops.append(("P", 1))
                ops.append(("I", int(rng.integers(1, 3))))
            elif u < 0.9:
                ops.append(("N", int(rng.integers(2, 8))))
            else:
                ops.append(("I", 1))
                ops.append(("D", 1))
        if rng.random() < 0.3:
            ops.append(("S", int(rng.integers(*clip_len))))
        span = sum(l for o, l in ops if o in "MDN=X")
        if span >= target_len:
            continue
        pos = int(rng.integers(0, target_len - span + 1))
        if rng.random() < 0.03:
            pos = target_len - span   # ends at the last base of the target
        seq = []
        rp = pos
        for o, l in ops:
            if o in "M=X":
                for k in range(l):
                    b = ref[rp + k]
                    if o == "X" or (o == "M" and rng.random() < 0.05):
                        b = str(rng.choice([c for c in "ACGT" if c != b]))
                    seq.append(b)
                rp += l
            elif o in "IS":
                seq.extend(str(c) for c in rng.choice(list("ACGT"), size=l))
            elif o in "DN":
                rp += l
        seq = ["N" if rng.random() < 0.01 else b for b in seq]
        q = [int(x) for x in rng.integers(2, 42, size=len(seq))]
        d = dict(tid=t, pos=pos, cigar="".join("%d%s" % (l, o) for o, l in ops), seq="".join(seq), qual=q,
                 flag=16 if rng.random() < 0.5 else 0, mapq=42 if rng.random() < 0.9 else int(rng.integers(0, 42)))
        if rng.random() < 0.08:
            d["x1"] = int(rng.integers(2, 4))
        if rng.random() < 0.5:
            d["xl"] = int(rng.integers(0, 6))
            d["xr"] = int(rng.integers(0, 6))
        reads.append(d)
    reads.sort(key=lambda r: (r["tid"], r["pos"]))
    return refs, reads
