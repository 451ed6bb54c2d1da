"""SIMULATE-READS (bsq_simulate_reads) against the reference's own reproducibility test,
tests/simulate_reads_paired: lambda (48,502 bp, linear), paired-end, 25 pairs of 100 bp, fragments
N(350, 40), seed 1, the command's default quality and indel-error model.  The FASTQ must match the
expected files byte for byte (tests/golden/simreads_lambda.json.gz, written from the reference's
test data by tests/golden/make_simreads_fixture.py).  Further checks: the single-end mode's
placement and name fields, and the quality model's bounds."""
import ctypes as C
import gzip
import json
import os

import pytest

from breseq_amd import _lib as L

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "simreads_lambda.json.gz")


def fixture():
    with gzip.open(GOLDEN, "rt") as f:
        return json.load(f)


class SimArgs(C.Structure):
    _fields_ = [("mode", C.c_int32), ("circular", C.c_int32), ("n_reads", C.c_uint32), ("coverage", C.c_uint32),
                ("read_length", C.c_uint32), ("fragment_mean", C.c_uint32), ("fragment_stdev", C.c_uint32),
                ("seed", C.c_uint32), ("quality_start", C.c_uint32), ("quality_end", C.c_uint32),
                ("quality_stdev", C.c_uint32), ("quality_min", C.c_uint32), ("quality_max", C.c_uint32),
                ("indel_error_rate_per_million", C.c_uint32)]


def simulate(seq, **kw):
    lib = L.lib()
    lib.bsq_simreads_defaults.argtypes = [C.POINTER(SimArgs)]
    lib.bsq_simulate_reads.argtypes = [C.c_char_p, C.c_int64, C.POINTER(SimArgs), C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_int64), C.POINTER(C.c_void_p), C.POINTER(C.c_int64),
                                       C.POINTER(L.Status)]
    a = SimArgs()
    assert lib.bsq_simreads_defaults(C.byref(a)) == 0
    for k, v in kw.items():
        setattr(a, k, v)
    p1, p2 = C.c_void_p(), C.c_void_p()
    n1, n2 = C.c_int64(), C.c_int64()
    st = L.Status()
    b = seq.encode()
    rc = lib.bsq_simulate_reads(b, len(b), C.byref(a), C.byref(p1), C.byref(n1), C.byref(p2), C.byref(n2), C.byref(st))
    if rc != 0:
        raise L.BsqError(rc, st.message.decode())
    t1 = C.string_at(p1, n1.value).decode()
    lib.bsq_free_text(p1)
    t2 = None
    if p2.value:
        t2 = C.string_at(p2, n2.value).decode()
        lib.bsq_free_text(p2)
    return t1, t2


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
