"""Hand-built alignment sets for known-answer tests (bsq_reference + bsq_alignments structs)."""
import ctypes as C
import re

import numpy as np

from breseq_amd import _lib as L

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
        p = lambda a, t: a.ctypes.data_as(t)
        self.aln = L.Alignments(n, p(self.tid, L.c_i32p), p(self.pos, L.c_i32p), p(self.flag, L.c_u16p),
                                p(self.mapq, L.c_u8p), p(self.cigar_off, L.c_i64p), p(self.cigar, L.c_u32p),
                                p(self.seq_off, L.c_i64p), p(self.seq, L.c_u8p), p(self.qual, L.c_u8p),
                                p(self.x1, L.c_i32p), p(self.xl, L.c_i32p), p(self.xr, L.c_i32p), p(self.rg, L.c_i32p))
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
