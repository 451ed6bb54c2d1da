"""SIMULATE-READS (bsq_simulate_reads) against the reference's own reproducibility test,
tests/simulate_reads_paired: lambda (48,502 bp, linear), paired-end, 25 pairs of 100 bp, fragments
N(350, 40), seed 1, the command's default quality and indel-error model.  The FASTQ must match the
expected files byte for byte (tests/golden/simreads_lambda.json.gz, written from the reference's
test data by tests/golden/make_simreads_fixture.py).  Further checks: the single-end mode's
placement and name fields, and the quality model's bounds."""
import gzip
import json
import os

import pytest

import breseq_amd as B
from breseq_amd import _lib as L

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "simreads_lambda.json.gz")


def fixture():
    with gzip.open(GOLDEN, "rt") as f:
        return json.load(f)


def simulate(seq, **kw):
    """(FASTQ text of read 1, of read 2 or None) through bsq_simulate_reads."""
    out = B.api.simulate_reads_fastq(seq, **kw)
    return out if isinstance(out, tuple) else (out, None)


def test_paired_end_matches_the_reference_fixture():
    fx = fixture()
    o = fx["options"]
    t1, t2 = simulate(fx["reference"], mode=1, circular=1 if fx["circular"] else 0, n_reads=o["n_reads"],
                      read_length=o["read_length"], fragment_mean=o["fragment_mean"],
                      fragment_stdev=o["fragment_stdev"], seed=o["seed"])
    assert t1 == fx["expected_1"]
    assert t2 == fx["expected_2"]


def records(text):
    ln = text.splitlines()
    return [ln[i:i + 4] for i in range(0, len(ln), 4)]


def test_single_end_reads_lie_inside_a_linear_sequence():
    seq = fixture()["reference"]
    t1, t2 = simulate(seq, mode=0, n_reads=400, read_length=100, seed=7, quality_start=32, quality_end=32,
                      quality_stdev=6)
    assert t2 is None
    rec = records(t1)
    assert len(rec) == 400
    for i, (name, bases, plus, qual) in enumerate(rec):
        assert name == "@READ-%d" % (i + 1)
        strand, left = plus[1:].split("\t")
        assert strand in ("[strand]:1", "[strand]:-1")
        left = int(left.split(":")[1])
        assert 1 <= left <= len(seq) - 100 + 1
        assert len(bases) == 100 and len(qual) == 100
        assert all(2 <= ord(c) - 33 <= 41 for c in qual)


def test_quality_model_without_scatter_is_exact():
    seq = fixture()["reference"]
    # no scatter: every quality is the quadratic decline's center, start 40 to end 30 over 51 cycles
    t1, _ = simulate(seq, mode=0, n_reads=5, read_length=51, seed=3, quality_start=40, quality_end=30,
                     quality_stdev=0, indel_error_rate_per_million=0)
    want = "".join(chr(33 + 40 - (10 * c * c) // (50 * 50)) for c in range(51))
    for r in records(t1):
        assert r[3] == want


def test_bad_arguments_are_refused():
    with pytest.raises(L.BsqError):
        simulate("ACGT" * 10, mode=0, n_reads=1, read_length=100)   # read longer than a linear sequence
    with pytest.raises(L.BsqError):
        simulate("ACGT" * 100, mode=0, n_reads=1, read_length=50, quality_min=30, quality_max=20)


# ---------------------------------------------------------------- aligned samples from the simulator

import numpy as np  # noqa: E402

OPS = "MIDNSHP=X"
CODE = {1: "A", 2: "C", 4: "G", 8: "T", 15: "N"}
COMP = {"A": "T", "C": "G", "G": "C", "T": "A", "N": "N"}


def lambda_mutations(ref):
    return [(1000, 1, "G" if ref[1000] != "G" else "A"), (5000, 0, "TTA"), (9000, 3, ""), (20000, 1, "C" if ref[20000] != "C" else "T")]


def test_aligned_sample_walks_back_to_the_reference():
    ref = fixture()["reference"]
    muts = lambda_mutations(ref)
    opts = dict(mode="single", n_reads=12000, read_length=100, seed=5, quality_start=32, quality_end=32,
                quality_stdev=6)
    s = B.Sample.simulate_breseq(ref, name="lambda", mutations=muts, **opts)
    a = s.arrays()
    n = a["pos"].size
    assert n == 12000 and np.all(np.diff(a["pos"]) >= 0)
    # the same reads as FASTQ, from the mutated sequence: each record's bases, in read direction
    genome = ref[:1000] + muts[0][2] + ref[1001:5000] + "TTA" + ref[5000:9000] + ref[9003:20000] + muts[3][2] + ref[20001:]
    fq = records(B.api.simulate_reads_fastq(genome, **opts))
    by_name = {r[0][1:]: r for r in fq}
    mism = total = 0
    alt_at = {1000: 0, 20000: 0}
    cover = {1000: 0, 20000: 0}
    ins_seen = del_seen = 0
    names = s.read_names()
    ref_codes = np.frombuffer(ref.encode(), dtype=np.uint8)
    lut = np.zeros(16, dtype=np.uint8)
    for k, c in CODE.items():
        lut[k] = ord(c)
    for i in range(n):
        cig = [(int(c) >> 4, OPS[int(c) & 15]) for c in a["cigar"][a["cigar_off"][i]:a["cigar_off"][i + 1]]]
        codes = a["seq"][a["seq_off"][i]:a["seq_off"][i + 1]]
        seq = lut[codes].tobytes().decode()
        f = by_name[names[i]]
        rev = bool(a["flag"][i] & 16)
        assert f[1] == ("".join(COMP[c] for c in reversed(seq)) if rev else seq)
        assert sum(l for l, o in cig if o in "MI") == len(seq)
        p, q = int(a["pos"][i]), 0
        if len(cig) == 1 and not any(p <= x < p + len(seq) for x in alt_at):
            total += len(seq)
            mism += int(np.count_nonzero(lut[codes] != ref_codes[p:p + len(seq)]))
            continue
        for l, o in cig:
            if o == "M":
                for k in range(l):
                    total += 1
                    mism += seq[q + k] != ref[p + k]
                    if p + k in alt_at:
                        cover[p + k] += 1
                        alt_at[p + k] += seq[q + k] != ref[p + k]
                p += l
                q += l
            elif o == "I":
                ins_seen += p == 5000 and l == 3
                q += l
            elif o == "D":
                del_seen += p == 9000 and l == 3
                p += l
    # substitution errors at Q ~32 plus the planted SNPs
    assert mism / total < 0.004
    for k in alt_at:
        assert cover[k] > 10 and alt_at[k] >= 0.9 * cover[k], (k, alt_at[k], cover[k])
    assert ins_seen > 10 and del_seen > 10
