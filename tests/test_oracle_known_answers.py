"""Known-answer tests that pin the oracle independently of its own code.

The expected values are worked out by hand from the reference's definitions (htslib pileup
columns as breseq reads them, error_count.cpp:854-986, calculate_trims.cpp:67-138) or come from
closed forms and independent libraries (scipy), never by re-running the oracle's formula.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from aln_builder import IDX, Built, bin_index

REF = "ACGTACGTACGTACGTACGT"


def column(b, pos0):
    return [tuple(x) for x in O.pileup_column(C_ptr(b), 0, pos0)]


def C_ptr(b):
    import ctypes
    return ctypes.pointer(b.aln)


# ---------------------------------------------------------------- pileup columns (htslib semantics)

def test_pileup_simple_match_and_soft_clip():
    b = Built([("r", REF)], [dict(pos=2, cigar="2S5M", seq="TTGTACG")])
    # the read covers columns 2..6; qpos skips the two soft-clipped bases
    assert column(b, 1) == []
    assert column(b, 2) == [(0, 2, 0, 0)]
    assert column(b, 6) == [(0, 6, 0, 0)]
    assert column(b, 7) == []


def test_pileup_deletion_qpos_is_next_base_and_indel_on_last_match():
    b = Built([("r", REF)], [dict(pos=0, cigar="5M2D5M", seq="ACGTAACGTA")])
    assert column(b, 3) == [(0, 3, 0, 0)]
    assert column(b, 4) == [(0, 4, -2, 0)]           # last base before the deletion
    assert column(b, 5) == [(0, 5, 0, 1)]            # deleted column: qpos of the next aligned base
    assert column(b, 6) == [(0, 5, 0, 1)]
    assert column(b, 7) == [(0, 5, 0, 0)]
    assert column(b, 11) == [(0, 9, 0, 0)]


def test_pileup_insertion_length_on_preceding_base():
    b = Built([("r", REF)], [dict(pos=0, cigar="3M2I5M", seq="ACGTTTACG" + "T")])
    assert column(b, 2) == [(0, 2, 2, 0)]
    assert column(b, 3) == [(0, 5, 0, 0)]
    assert column(b, 7) == [(0, 9, 0, 0)]


def test_pileup_merged_deletions_insertions_pads_and_skips():
    b = Built([("r", REF)], [
        dict(pos=0, cigar="3M1D2D3M", seq="ACGCGT"),     # adjacent D ops: merged at the run start
        dict(pos=0, cigar="2M1I2I2M", seq="ACTTTGT"),    # adjacent I ops: summed
        dict(pos=0, cigar="2M1P1I2M", seq="ACTGT"),      # pad then insertion
        dict(pos=0, cigar="3M4N3M", seq="ACGCGT"),       # reference skip behaves as a deletion
    ])
    col2 = column(b, 2)
    assert col2[0] == (0, 2, -3, 0)
    assert col2[3] == (3, 2, 0, 0)
    col1 = column(b, 1)
    assert col1[1] == (1, 1, 3, 0)
    assert col1[2] == (2, 1, 1, 0)
    col3 = column(b, 3)
    assert col3[0] == (0, 3, 0, 1)                   # first deleted column, inside the merged run
    assert col3[3] == (3, 3, 0, 1)                   # N column: is_del, qpos of the next base
    assert column(b, 4)[0] == (0, 3, 0, 1)
    assert column(b, 6)[0] == (0, 3, 0, 0)


def test_pileup_skips_unmapped_and_keeps_file_order():
    b = Built([("r", REF)], [dict(pos=0, cigar="4M", seq="ACGT"), dict(pos=0, cigar="4M", seq="ACGT", flag=4),
                             dict(pos=1, cigar="3M", seq="CGT")])
    assert [x[0] for x in column(b, 2)] == [0, 2]


# ---------------------------------------------------------------- error_count (error_count.cpp:854-986)

def counts_of(b, Q=42):
    args = b.error_count_args(quality_max=Q)
    hists, counts, log10p, texts = O.error_count(args)
    return hists, counts, log10p, texts


def nonzero(counts):
    return {i: int(c) for i, c in enumerate(counts) if c}


def test_error_count_forward_read_with_mismatch():
    # ref ACGTACGT..., read CGTT at column 1: last base is a T read over a reference A
    b = Built([("r", REF)], [dict(pos=1, cigar="4M", seq="CGTT", qual=[30, 31, 32, 33])])
    hists, counts, _, _ = counts_of(b)
    exp = {}
    def add(ref, obs, q):
        k = bin_index(1, 0, IDX[ref], IDX[obs], q)
        exp[k] = exp.get(k, 0) + 1
    add("C", "C", 30), add("G", "G", 31), add("T", "T", 32), add("A", "T", 33)
    # '..' pairs: quality of the next read base, not past the last aligned base
    add(".", ".", 31), add(".", ".", 32), add(".", ".", 33)
    assert nonzero(counts) == exp
    assert list(hists[0]) == [len(REF) - 4, 4]


def test_error_count_reverse_read_complements_and_uses_own_quality_for_pairs():
    b = Built([("r", REF)], [dict(pos=1, cigar="4M", seq="CGTT", qual=[30, 31, 32, 33], flag=16)])
    _, counts, _, _ = counts_of(b)
    exp = {}
    def add(ref, obs, q):
        k = bin_index(1, 0, IDX[ref], IDX[obs], q)
        exp[k] = exp.get(k, 0) + 1
    # read-strand space: both bases complemented
    add("G", "G", 30), add("C", "C", 31), add("A", "A", 32), add("T", "A", 33)
    # '..' pairs on the reverse strand read the quality at qpos itself (qpos + 1 - 1)
    add(".", ".", 30), add(".", ".", 31), add(".", ".", 32)
    assert nonzero(counts) == exp


def test_error_count_one_base_deletion_and_insertion_observations():
    # deletion after column 2: 'ref .' observation against the NEXT reference base, with the
    # quality of the next read base; insertion after column 1: '. obs' with the inserted base's quality
    b = Built([("r", REF)], [dict(pos=0, cigar="3M1D3M", seq="ACGACG", qual=[20, 21, 22, 23, 24, 25]),
                             dict(pos=0, cigar="2M1I3M", seq="ACTGTA", qual=[10, 11, 12, 13, 14, 15])])
    _, counts, _, _ = counts_of(b)
    got = nonzero(counts)
    k_del = bin_index(1, 0, IDX["T"], IDX["."], 23)   # ref[3] = T, quality of read base 3
    k_ins = bin_index(1, 0, IDX["."], IDX["T"], 12)   # inserted T, quality 12
    assert got.get(k_del) == 1
    assert got.get(k_ins) == 1
    # the column before the deletion contributes no '..' pair (that would use quality 23)
    assert bin_index(1, 0, 4, 4, 23) not in got
    assert got.get(bin_index(1, 0, 4, 4, 22)) == 1      # pair at column 1 of the deletion read


def test_error_count_redundant_reads_mark_columns_and_are_not_counted():
    b = Built([("r", REF)], [dict(pos=0, cigar="4M", seq="ACGT"), dict(pos=2, cigar="4M", seq="GTAC", x1=2)])
    hists, counts, _, _ = counts_of(b)
    # columns 2,3 carry a redundant read: excluded from the histogram
    # columns 0,1: 1 unique read; 4,5: 0 unique (redundant only) -> excluded too
    h = list(hists[0])
    assert h[1] == 2
    assert h[0] == len(REF) - 6
    assert int(counts.sum()) == 4 + 3     # only the unique read's 4 bases + 3 pairs


def test_error_count_log10_table_is_laplace_ratio():
    b = Built([("r", REF)], [dict(pos=0, cigar="8M", seq="ACGTACGA", qual=[30] * 8)])
    _, counts, log10p, texts = counts_of(b)
    S, Q = 1, 42
    for i in np.flatnonzero(counts):
        rs, ref, q = i % S, (i // S) % 5, i // (25 * S)
        tot = sum(int(counts[bin_index(S, rs, ref, o, q)]) for o in range(5))
        assert math.isclose(log10p[i], math.log10(counts[i] + 1) - math.log10(tot + 5), rel_tol=0, abs_tol=1e-15)
    # text: 6 significant digits, as an ostream prints a double
    line = [l for l in texts["error_rates"].splitlines()[2:] if l.split("\t")[:4] == ["0", "A", "A", "30"]][0]
    assert line.split("\t")[4] == "%g" % log10p[bin_index(S, 0, 0, 0, 30)]


# ---------------------------------------------------------------- trims (calculate_trims.cpp:67-138)

def test_trims_homopolymer_known_answer():
    left, right = O.sequence_trims("AAAA")
    assert list(left) == [4, 3, 2, 1]
    assert list(right) == [1, 2, 3, 4]


def test_trims_dinucleotide_and_cap():
    left, right = O.sequence_trims("ACACG")
    # position 0: unit 2 repeats twice -> 4; position 1: 'CAC' unit 2 does not repeat, unit 1 no -> 1
    assert list(left[:2]) == [4, 3]
    assert list(right[:4]) == [1, 2, 3, 4]
    left, right = O.sequence_trims("A" * 400)
    assert left.max() == 255 and right.max() == 255


def test_trims_brute_force_tiny():
    rng = np.random.default_rng(3)
    for _ in range(50):
        s = "".join(rng.choice(list("AC"), rng.integers(1, 12)))
        n = len(s)
        # direct definition: longest prefix of s[p:] that is a perfect tandem repeat (>= 2 copies)
        ext = []
        for p in range(n):
            best = 1
            for size in range(1, 19):
                k = 1
                while p + (k + 1) * size <= n and s[p:p + size] * (k + 1) == s[p:p + (k + 1) * size]:
                    k += 1
                if k > 1:
                    best = max(best, k * size)
            ext.append(min(best, 255))
        exp_l = [max([ext[p] - (x - p) for p in range(x + 1) if x - p < ext[p]] + [0]) for x in range(n)]
        exp_r = [max([x - p + 1 for p in range(x + 1) if x - p < ext[p]] + [0]) for x in range(n)]
        left, right = O.sequence_trims(s)
        assert list(left) == exp_l, s
        assert list(right) == exp_r, s


# ---------------------------------------------------------------- statistics

def test_fisher_against_scipy():
    stats = pytest.importorskip("scipy.stats")
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b, c, d = (int(x) for x in rng.integers(0, 40, 4))
        exp = stats.fisher_exact([[a, b], [c, d]], alternative="two-sided").pvalue
        got = O.fisher(a, b, c, d)
        assert math.isclose(got, min(exp, 1.0), rel_tol=1e-9, abs_tol=1e-12), (a, b, c, d, got, exp)


def test_ks_less_against_scipy_exact_tie_free():
    stats = pytest.importorskip("scipy.stats")
    rng = np.random.default_rng(2)
    for _ in range(40):
        m, n = (int(x) for x in rng.integers(3, 40, 2))
        v = rng.permutation(m + n).astype(float)
        x, y = v[:m], v[m:] + 0.5 * rng.integers(0, 2)
        exp = stats.ks_2samp(x, y, alternative="less", method="exact").pvalue
        got = O.ks_less(x, y)
        assert math.isclose(got, exp, rel_tol=1e-7, abs_tol=1e-12), (m, n, got, exp)


def test_pure_genotype_closed_form():
    rng = np.random.default_rng(4)
    for _ in range(200):
        lps = rng.normal(-50, 30, 5)
        best, score = O.pure_genotype(lps, 10)
        assert best == int(np.argmax(lps))
        others = np.delete(lps, best)
        m = others.max()
        exp = lps[best] - (m + math.log10(sum(10 ** (o - m) for o in others)))
        assert math.isclose(score, exp, rel_tol=1e-12, abs_tol=1e-9)
    best, score = O.pure_genotype(np.zeros(5), 0)
    assert best == 5 and math.isnan(score)


def test_em_fixed_point_satisfies_kkt():
    """The EM result maximizes the mixture likelihood over the simplex: at the optimum the
    gradient component Sum_i r_ib / s_i equals n on the support and is <= n off it."""
    rng = np.random.default_rng(5)
    for trial in range(20):
        n = int(rng.integers(5, 60))
        lp = rng.normal(-3, 2, (n, 5))
        truth = rng.integers(0, 3, n)
        lp[np.arange(n), truth] = -0.01
        lpmax = lp.max(axis=1)
        r = 10 ** (lp - lpmax[:, None])
        obs = lp.argmax(axis=1)
        f, ll, it = O.fit_alleles(r, lpmax, obs, tol=1e-12)
        assert abs(f.sum() - 1) < 1e-12
        s = r @ f
        grad = (r / s[:, None]).sum(axis=0)
        if it < 50:
            # converged: the reported likelihood (at the previous iterate) matches the final one
            assert math.isclose(ll, float(np.sum(np.log10(s) + lpmax)), rel_tol=1e-9, abs_tol=1e-9)
            for b in range(5):
                if f[b] > 1e-3:
                    assert abs(grad[b] - n) < 1e-3 * n, (trial, b, grad[b], n)
                else:
                    assert grad[b] <= n * (1 + 1e-3), (trial, b, grad[b], n)


def test_em_respects_held_out_allele():
    rng = np.random.default_rng(6)
    n = 30
    lp = rng.normal(-2, 1, (n, 5))
    lpmax = lp.max(axis=1)
    r = 10 ** (lp - lpmax[:, None])
    f, ll, it = O.fit_alleles(r, lpmax, lp.argmax(axis=1), allowed=(1, 0, 1, 1, 1))
    assert f[1] == 0.0
    assert abs(f.sum() - 1) < 1e-12
