"""CPU tests of libbsq.so's host side: the C ABI surface, the synthetic-sample generator, BAM/FASTA
round trips and the coverage-distribution fit.  No GPU call is made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import breseq_amd as B
from breseq_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bsq.h")

TINY = dict(seed=3, ref_length=6000, gc=0.45, read_length=80, coverage=12.0, paired=True, insert_mean=250,
            insert_sd=30, qual_mean=30, qual_sd=6, qual_min=2, qual_max=41, n_snps=5, n_indels=3, indel_max=4,
            redundant_fraction=0.05, mapq=42)


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(bsq_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert len(names) >= 30, names
    lib = L.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_simulator_is_deterministic_and_seeded():
    a = B.Sample.simulate(**TINY)
    b = B.Sample.simulate(**TINY)
    c = B.Sample.simulate(**dict(TINY, seed=4))
    xa, xb, xc = a.arrays(), b.arrays(), c.arrays()
    for k in xa:
        assert np.array_equal(xa[k], xb[k]), k
    assert a.reference_sequences() == b.reference_sequences()
    assert a.reference_sequences() != c.reference_sequences()
    assert len(xa["pos"]) > 100
    # sorted by position, as a coordinate-sorted BAM is
    assert np.all(np.diff(xa["pos"]) >= 0)


def test_simulated_reads_follow_their_cigars():
    s = B.Sample.simulate(**TINY)
    x = s.arrays()
    ref = s.reference_sequences()[0][1]
    code = {1: "A", 2: "C", 4: "G", 8: "T", 15: "N"}
    matches = total = 0
    for i in range(0, len(x["pos"]), 7):
        cig = x["cigar"][x["cigar_off"][i]: x["cigar_off"][i + 1]]
        seq = x["seq"][x["seq_off"][i]: x["seq_off"][i + 1]]
        qlen = sum(int(c >> 4) for c in cig if (c & 15) in (0, 1, 4, 7, 8))
        assert qlen == len(seq)
        rp, qp = int(x["pos"][i]), 0
        for c in cig:
            op, ln = int(c & 15), int(c >> 4)
            if op in (0, 7, 8):
                for k in range(ln):
                    total += 1
                    matches += code[int(seq[qp + k])] == ref[rp + k]
                rp += ln
                qp += ln
            elif op in (1, 4):
                qp += ln
            elif op in (2, 3):
                rp += ln
    # sequencing errors and planted variants only: nearly every aligned base matches
    assert matches / total > 0.97


def test_bam_and_fasta_round_trip(tmp_path):
    s = B.Sample.simulate(**TINY)
    bam = str(tmp_path / "reads.bam")
    fa = str(tmp_path / "ref.fasta")
    s.write_bam(bam)
    s.write_fasta(fa)
    with open(bam, "rb") as f:
        head = f.read(4)
    assert head[:2] == b"\x1f\x8b"   # BGZF is gzip framed
    r = B.Sample.read(bam, fa)
    assert r.reference_sequences() == s.reference_sequences()
    xs, xr = s.arrays(), r.arrays()
    for k in xs:
        assert np.array_equal(xs[k], xr[k]), k


def test_fit_coverage_distribution_matches_scipy_quantile():
    st = pytest.importorskip("scipy.stats")
    mu, size, L_seq = 100.0, 40.0, 4.6e6
    p = size / (size + mu)
    k = np.arange(0, 400)
    hist = np.round(st.nbinom.pmf(k, size, p) * L_seq).astype(np.uint32)
    fit = B.fit_coverage_distribution(hist, L_seq)
    assert abs(fit["nb_fit_mu"] - mu) / mu < 0.02, fit
    assert abs(fit["nb_fit_size"] - size) / size < 0.10, fit
    # deletion cutoff = qnbinom(0.05 / sqrt(L)) of the fitted distribution (coverage_distribution.cpp:377-394)
    fs, fm = fit["nb_fit_size"], fit["nb_fit_mu"]
    want = st.nbinom.ppf(0.05 / np.sqrt(L_seq), fs, fs / (fs + fm))
    assert fit["deletion_coverage_propagation_cutoff"] == pytest.approx(want, abs=0.0)


def test_error_paths_return_codes_not_exits():
    lib = L.lib()
    st = L.Status()
    out = np.zeros(8, dtype=np.uint8)
    rc = lib.bsq_sequence_trims(None, 4, out.ctypes.data_as(L.c_u8p), C.byref(st))
    assert rc != 0 and st.message
    buf = C.create_string_buffer(4)
    assert lib.bsq_format_double(1.5, 3, 1, buf, 4) < 0   # buffer too small is reported


def test_number_formatting_rules():
    lib = L.lib()
    buf = C.create_string_buffer(64)

    def f(v, places, sci):
        assert lib.bsq_format_double(v, places, sci, buf, 64) > 0
        return buf.value.decode()

    assert f(float("nan"), 3, 1) == "NA"
    assert f(1.2345e-120, 3, 1) == "1.234e-120" or f(1.2345e-120, 3, 1) == "1.235e-120"
    assert f(0.5, 3, 1) == "5.000e-01"
    assert f(55.25, 1, 0) == "55.2"
    assert f(1e-05, 5, 1) == "1.00000e-05"


def test_number_formatting_matches_printf_rounding():
    """bsq_format_double prints %.*e / %.*f digits; checked against Python's correctly rounded
    formatting (the same digits glibc printf gives) on random values, halfway cases and extremes."""
    lib = L.lib()
    buf = C.create_string_buffer(512)
    rng = np.random.default_rng(7)
    vals = list(rng.uniform(-1, 1, 3000) * 10.0 ** rng.integers(-300, 300, 3000))
    vals += [0.5, 1.5, 2.5, 0.125, 1e-5, 9.9995e-1, 0.0, -0.0, 1e300, 5e-324, 123456.5, 0.05, 0.15, 0.25]
    vals += list(rng.uniform(0, 1, 2000))
    n = 0
    for v in vals:
        for places in (0, 1, 3, 5, 8):
            for sci in (0, 1):
                if not sci and abs(v) > 1e30:
                    continue
                got = lib.bsq_format_double(float(v), places, sci, buf, 512)
                assert got > 0
                want = ("%.*e" if sci else "%.*f") % (places, v)
                if sci and len(want) >= 3 and want[-3] == "0" and want[-4] in "+-":
                    want = want[:-3] + want[-2:]
                assert buf.value.decode() == want, (v, places, sci, buf.value, want)
                n += 1
    assert n > 40000


def test_general_formatting_matches_printf_g():
    """error_rates.tab values are printed as %g (6 significant digits); the %.*g mode of
    bsq_format_double is checked against Python's %g on random log10 probabilities and extremes."""
    lib = L.lib()
    buf = C.create_string_buffer(128)
    rng = np.random.default_rng(11)
    vals = list(-rng.uniform(0, 8, 5000)) + list(np.log10(rng.uniform(1e-9, 1, 3000)))
    vals += [0.0, -0.0, 1e-5, 1e-4, 123456.0, 1234567.0, 1e16, 5e-324, -2.5, 0.30103, 1e-300]
    for v in vals:
        for prec in (6, 3, 1):
            assert lib.bsq_format_double(float(v), prec, 2, buf, 128) > 0
            assert buf.value.decode() == "%.*g" % (prec, v), (v, prec, buf.value)
