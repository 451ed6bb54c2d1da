"""Pins of the oracle's identify_mutations, statistics and table-text functions.

Every expected value here is worked out by hand from the reference's definitions (cited per test),
from a closed form solved independently (scipy brentq), or by brute-force enumeration; none re-runs
the oracle's own formula.  Each test names the planted mistake it catches.
"""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as O
from aln_builder import IDX, Built, BuiltSample, bin_index
from breseq_amd import api
from parity_util import parse_gd

COMP = {"A": "T", "C": "G", "G": "C", "T": "A", ".": "."}
S1, Q42 = 1, 42


def sym_table(eps, S=S1, Q=Q42):
    """log10 error table with P(obs | ref) = 1 - 4 eps on the diagonal and eps elsewhere."""
    t = np.empty(S * 25 * Q)
    for q in range(Q):
        for ref in range(5):
            for obs in range(5):
                for rs in range(S):
                    t[bin_index(S, rs, ref, obs, q)] = math.log10(1 - 4 * eps) if ref == obs else math.log10(eps)
    return t


def run_identify(b, table, prop, seed=None, mode=api.CONSENSUS, **kw):
    s = BuiltSample(b)
    args = api.make_identify_args(s, table, S1, Q42, prop, seed, mode=mode, **kw)
    gd, _ = O.identify_mutations(args)
    return parse_gd(gd)


def entries(parsed, kind):
    return [(p, kv) for t, _, p, kv in parsed if t == kind]


def read_over(ref, pos, length, subs=None, strand=0, qual=30, **kw):
    """An ungapped read of `length` bases from 0-based `pos`, bases from the reference except the
    substitutions {column: base}; qual is one value or a {column: quality} map over a base of 30."""
    seq = list(ref[pos:pos + length])
    for c, b in (subs or {}).items():
        seq[c - pos] = b
    if isinstance(qual, dict):
        q = [qual.get(pos + i, 30) for i in range(length)]
    else:
        q = [qual] * length
    return dict(pos=pos, cigar="%dM" % length, seq="".join(seq), qual=q, flag=16 if strand else 0, **kw)


REF30 = "ACGTTGCAAGATAGCTTACGGATCCATGCA"   # column 10 (position 11) is 'A'


# ---------------------------------------------------------------- identify callback, one column

def test_identify_column_tallies_by_hand():
    """pos_info and total_cov of one column (identify_mutations.cpp:1561-1657, 1693-1700, 1885-1900):
    an N read is skipped entirely, a trimmed read, a redundant read and a read whose base quality is
    below base_quality_cutoff (3) are left out of pos_info; the rest count per strand.  Planted
    mistakes caught: counting the q=2 read, the XL-trimmed read or the X1=2 read, or swapping the
    strand slots."""
    c, V = 10, "C"
    assert REF30[c] == "A"
    reads = [
        read_over(REF30, 6, 8, {c: V}),                       # r0 forward V
        read_over(REF30, 6, 8, {c: V}),                       # r1 forward V
        read_over(REF30, 6, 8, {c: V}, strand=1),             # r2 reverse V
        read_over(REF30, 6, 8, {c: V}, strand=1, qual={c: 2}),  # r3 reverse V, quality 2: left out
        read_over(REF30, 6, 8, {c: "N"}),                     # r4 N at the column: skipped
        read_over(REF30, 6, 8, {c: V}, xl=5),                 # r5 qpos 4 -> q1 = 5 <= XL: trimmed
        read_over(REF30, 6, 8, {c: V}, x1=2),                 # r6 redundant
        read_over(REF30, 6, 8),                               # r7 forward A (reference)
        read_over(REF30, 7, 8, {c: V}, strand=1),             # r8 reverse V
        read_over(REF30, 7, 8, strand=1),                     # r9 reverse A
    ]
    b = Built([("r", REF30)], reads)
    ra = entries(run_identify(b, sym_table(1e-3), [0.0]), "RA")
    at = [(p, kv) for p, kv in ra if p[1] == "11" and p[2] == "0"]
    assert len(at) == 1
    p, kv = at[0]
    assert p[3:5] == ("A", "C")
    assert kv["new_cov"] == "2/2"      # r0 r1 / r2 r8
    assert kv["ref_cov"] == "1/1"      # r7 / r9
    assert kv["total_cov"] == "3/3"
    assert kv["major_base"] == "C" and kv["minor_base"] == "A"
    assert kv["major_cov"] == "2/2" and kv["minor_cov"] == "1/1"


def test_identify_insert_slots_by_hand():
    """The insert-slot loop (identify_mutations.cpp:1359-1366, 1598-1599) and the slot quality of
    alignment_position_to_covariates (error_count.cpp:1049-1105): slots run while a UNIQUE read has
    a longer insertion (a redundant 3-base insertion opens no slot 3); a read without an insertion
    observes '.' with the quality of its next base on the forward strand (qpos + 1) and of its own
    base on the reverse strand (qpos + 1 - 1); a forward read whose last aligned base is the column
    has no next base (past q_end) and is left out.  Planted mistakes caught: own quality on the
    forward strand (n5 would count), qpos + 1 on the reverse strand (e7 would drop), no q_end test
    (e6 would count), slots from redundant reads (a slot-3 entry)."""
    c = 10
    ins = lambda strand=0, n=2, **kw: dict(pos=6, cigar="5M%dI5M" % n,
                                           seq=REF30[6:11] + "G" * n + REF30[11:16], flag=16 if strand else 0,
                                           **kw)
    reads = [
        ins(), ins(), ins(),                                  # u0-u2 forward, GG inserted after column 10
        ins(strand=1),                                        # u3 reverse
        ins(n=3, x1=2),                                       # r4 redundant, GGG: opens no slot 3
        read_over(REF30, 6, 10, qual={c + 1: 2}),             # n5 forward, no insertion, next base q 2
        read_over(REF30, 6, 5),                               # e6 forward, last aligned base at the column
        read_over(REF30, 6, 5, strand=1),                     # e7 reverse, last aligned base at the column
    ]
    b = Built([("r", REF30)], reads)
    ra = entries(run_identify(b, sym_table(1e-3), [0.0]), "RA")
    slots = {int(p[2]): (p, kv) for p, kv in ra if p[1] == "11" and p[2] != "0"}
    assert sorted(slots) == [1, 2]
    for k in (1, 2):
        p, kv = slots[k]
        assert p[3:5] == (".", "G")
        assert kv["new_cov"] == "3/1"    # u0-u2 / u3
        assert kv["ref_cov"] == "0/1"    # n5 out (quality 2), e6 out (past q_end) / e7 in
        assert kv["total_cov"] == "3/2"


# ---------------------------------------------------------------- MC and UN runs

def cov_reads(ref_len_reads):
    out = []
    for pos, length, n, kw in ref_len_reads:
        for _ in range(n):
            out.append(dict(pos=pos, length=length, **kw))
    return out


def build_targets(spec):
    """spec: [(name, length, [(pos0, read length, copies, extra fields)])] -> Built (reads sorted)."""
    rng = np.random.default_rng(7)
    refs, reads = [], []
    for t, (name, L, rl) in enumerate(spec):
        seq = "".join(rng.choice(list("ACGT"), L))
        refs.append((name, seq))
        for r in cov_reads(rl):
            d = read_over(seq, r["pos"], r["length"], **{k: v for k, v in r.items() if k not in ("pos", "length")})
            d["tid"] = t
            reads.append(d)
    reads.sort(key=lambda r: (r["tid"], r["pos"]))
    return Built(refs, reads)


MC_SPEC = [
    # a: unique 10 on 1-20 and 41-60, nothing unique on 21-40; redundant reads on 21-23 and 37-40
    ("a", 60, [(0, 20, 10, {}), (20, 3, 1, {"x1": 2}), (36, 4, 1, {"x1": 2}), (40, 20, 10, {})]),
    # b: one unique read on 16-30 between two blocks of ten: unique 1 <= cutoff but total 1 > seed 0
    ("b", 50, [(0, 15, 10, {}), (15, 15, 1, {}), (30, 20, 10, {})]),
    # c: uncovered to the end of the target
    ("c", 30, [(0, 20, 10, {})]),
    # d: the coverage fit failed (propagation cutoff < 0)
    ("d", 20, [(0, 20, 10, {})]),
]


def mc_un(seed_b=0.0):
    b = build_targets(MC_SPEC)
    parsed = run_identify(b, sym_table(1e-3), [1.0, 1.0, 1.0, -1.0], [0.0, seed_b, 0.0, 0.0])
    mc = {p[0]: (p, kv) for p, kv in entries(parsed, "MC")}
    un = sorted((p[0], int(p[1]), int(p[2])) for p, kv in entries(parsed, "UN"))
    ra_d = [p for p, kv in entries(parsed, "RA") if p[0] == "d"]
    return mc, un, ra_d


def test_mc_runs_by_hand():
    """check_deletion_completion (identify_mutations.cpp:2262-2344) and at_target_end (:2117-2164),
    worked through by hand.
    a: the run opens at 21 (unique 0 <= 1) and closes at 41 (unique 10): MC 21-40.  Redundant
    coverage: 21, 22, 23 > 0 before any zero, so the redundant start advances to 23; 24 is zero; 37
    is the first non-zero after it, so the redundant end is 37: start_range 23-21 = 2, end_range
    40-37 = 3.  Outside/inside coverages are the unique coverage at 20, 21, 40 and 41.
    b: unique 1 on 16-30 never has total <= seed 0: no MC.
    c: the run 21-30 closes at position 31 = L+1 with NaN coverage: right_outside_cov NA.
    d: failed fit: one whole-target MC with NA outside coverages, and no RA from its columns.
    Planted mistakes caught: first instead of last redundant start (start_range 0), keeping the
    redundant end updated (end_range 0), '<' for the seed test, '<' for the propagation test."""
    mc, un, ra_d = mc_un()
    assert sorted(mc) == ["a", "c", "d"]
    p, kv = mc["a"]
    assert p[1:5] == ("21", "40", "2", "3")
    assert (kv["left_outside_cov"], kv["left_inside_cov"], kv["right_inside_cov"], kv["right_outside_cov"]) == \
        ("10", "0", "0", "10")
    p, kv = mc["c"]
    assert p[1:5] == ("21", "30", "0", "0")
    assert (kv["left_outside_cov"], kv["left_inside_cov"], kv["right_inside_cov"], kv["right_outside_cov"]) == \
        ("10", "0", "0", "NA")
    p, kv = mc["d"]
    assert p[1:5] == ("1", "20", "0", "0")
    assert (kv["left_outside_cov"], kv["left_inside_cov"], kv["right_inside_cov"], kv["right_outside_cov"]) == \
        ("NA", "0", "0", "NA")
    assert ra_d == []


def test_mc_seed_cutoff_by_hand():
    """With seed cutoff 1 on target b, total 1 on 16-30 reaches the seed: MC 16-30, inside coverage 1
    at both ends, outside 10 (identify_mutations.cpp:2279-2281)."""
    mc, _, _ = mc_un(seed_b=1.0)
    p, kv = mc["b"]
    assert p[1:5] == ("16", "30", "0", "0")
    assert (kv["left_outside_cov"], kv["left_inside_cov"], kv["right_inside_cov"], kv["right_outside_cov"]) == \
        ("10", "1", "1", "10")


def test_un_runs_by_hand():
    """update_unknown_intervals (identify_mutations.cpp:2972-3007): maximal runs of columns whose
    consensus score is below the mutation cutoff (10).  Columns with no counted read have no score;
    one read of quality 30 at eps 1e-3 scores log10((1-4e)/e) - log10 4 - log10(160) < 10; ten reads
    score far above 10.  So: a 21-40, b 16-30, c 21-30 (closed at L+1), nothing on the failed-fit
    target d (its columns are not scored)."""
    _, un, _ = mc_un()
    assert un == [("a", 21, 40), ("b", 16, 30), ("c", 21, 30)]


# ---------------------------------------------------------------- read-base likelihoods

def test_read_base_likelihoods_limits_and_strand():
    """fill_read_base_likelihoods (identify_mutations.cpp:3359-3384).  MAPQ 0: the read is
    uninformative, every hypothesis has probability 1/5.  MAPQ 250 (incorrect 1e-25): log10 of the
    table entry itself, P(obs | ref = b).  A reverse-strand read looks up the complemented
    observation under complemented hypotheses.  Planted mistakes caught: no complement on the
    reverse strand, correct/incorrect swapped, a missing uniform term at MAPQ 0."""
    rng = np.random.default_rng(8)
    S, Q = 2, 40
    t = np.log10(rng.uniform(1e-4, 1.0, S * 25 * Q))
    for rs, q, obs in [(0, 5, 0), (1, 33, 2), (1, 39, 4), (0, 0, 3)]:
        for strand in (1, -1):
            lp, r, mx = O.read_base_likelihoods(t, S, Q, rs, q, obs, strand, 0)
            assert np.allclose(lp, math.log10(0.2), rtol=0, atol=1e-15)
            assert np.all(r == 1.0) and mx == lp.max()
            lp, r, mx = O.read_base_likelihoods(t, S, Q, rs, q, obs, strand, 250)
            for b in range(5):
                if strand == 1:
                    want = t[bin_index(S, rs, b, obs, q)]
                else:
                    cb, co = IDX[COMP["ACGT."[b]]], IDX[COMP["ACGT."[obs]]]
                    want = t[bin_index(S, rs, cb, co, q)]
                assert math.isclose(lp[b], want, rel_tol=1e-12, abs_tol=1e-12)
            assert mx == lp.max() and r.max() == 1.0
            assert np.allclose(r, 10.0 ** (lp - mx), rtol=1e-14)


# ---------------------------------------------------------------- presence score and profile bounds

def two_allele_column(n, k, eps, mapq):
    """k reads observing C and n-k observing A, all forward, under sym_table(eps): per read the
    max-normalized likelihood rows and their maxima, plus the two per-read probabilities a (the
    observed base's own hypothesis) and b (any other hypothesis)."""
    inc = 10.0 ** (-mapq / 10.0)
    cor = 1.0 - inc
    a = cor * (1 - 4 * eps) + inc / 5
    bb = cor * eps + inc / 5
    rows, lpmax, obs = [], [], []
    for i in range(n):
        o = 1 if i < k else 0
        pr = np.full(5, bb)
        pr[o] = a
        lp = np.log10(pr)
        rows.append(10.0 ** (lp - lp.max()))
        lpmax.append(lp.max())
        obs.append(o)
    return np.array(rows), np.array(lpmax), np.array(obs), a, bb


@pytest.mark.parametrize("n,k", [(20, 6), (40, 3), (60, 45)])
def test_presence_score_and_bounds_closed_form(n, k):
    """variant_presence_score (identify_mutations.cpp:3329-3344) and write_RA_frequency_bounds
    (:3094-3217) on a two-allele column with symmetric errors, against the binomial-mixture closed
    form solved with scipy: L(f) = k log10(f a + (1-f) b) + (n-k) log10(f b + (1-f) a); the null fit
    without C puts all weight on A: L0 = k log10 b + (n-k) log10 a; the bounds solve
    L(f) = max L - 0.587566 on each side of the maximum (the drop is the reference's constant 0.587566).
    Planted mistakes caught: a drop of 0.5, bounds from the wrong side, a null fit that keeps C."""
    opt = pytest.importorskip("scipy.optimize")
    stats = pytest.importorskip("scipy.stats")
    # the reference derives its constant from chi2_1(0.90) = 2.705543 (identify_mutations.cpp:3169-3173);
    # 2.705543 / (2 ln 10) is 0.5875013, and the constant as written, 0.587566, is what it uses
    assert abs(stats.chi2.ppf(0.9, 1) / 2 / math.log(10) - 0.587566) < 1e-4
    eps, mapq, log10_len = 1e-3, 60, math.log10(5e6)
    r, lpmax, obs, a, b = two_allele_column(n, k, eps, mapq)
    L = lambda f: k * math.log10(f * a + (1 - f) * b) + (n - k) * math.log10(f * b + (1 - f) * a)
    dL = lambda f: k * (a - b) / (f * a + (1 - f) * b) + (n - k) * (b - a) / (f * b + (1 - f) * a)
    fh = opt.brentq(dL, 1e-9, 1 - 1e-9, xtol=1e-15)
    Lmax = L(fh)
    L0 = k * math.log10(b) + (n - k) * math.log10(a)
    g = lambda f: L(f) - (Lmax - 0.587566)
    lo = opt.brentq(g, 1e-12, fh, xtol=1e-14)
    hi = opt.brentq(g, fh, 1 - 1e-12, xtol=1e-14)
    f, ll, score, lower, upper = O.presence_and_bounds(r, lpmax, obs, 1, log10_ref_length=log10_len)
    assert abs(f[1] - fh) < 1e-5
    assert abs(ll - Lmax) < 1e-5
    # the null EM stops once no frequency moves by 1e-6, a little short of the limit f_A = 1
    assert abs(score - (Lmax - L0 - log10_len)) < 1e-3
    # bisection to width <= 1e-6: lower is the interval's upper end, upper its lower end
    assert lo <= lower + 1e-9 and lower - lo < 2e-6, (lower, lo)
    assert upper <= hi + 1e-9 and hi - upper < 2e-6, (upper, hi)


# ---------------------------------------------------------------- KS with ties

def ks_brute(x, y):
    """P-value of the one-sided statistic max(Fy - Fx) by enumerating every assignment of the pooled
    sorted values to the two samples (sizes kept), the statistic evaluated only where the pooled
    value changes (ties folded): the permutation distribution the lattice count of
    ks_exact_allowed_paths (stats.cpp:2190-2211) sums, computed here in exact rationals."""
    pooled = sorted(list(x) + list(y))
    m, n = len(y), len(x)
    ends = [i + 1 for i in range(len(pooled)) if i + 1 == len(pooled) or pooled[i + 1] != pooled[i]]

    def stat(ys):
        best = None
        for e in ends:
            cy = sum(1 for i in ys if i < e)
            d = Fraction(cy, m) - Fraction(e - cy, n)
            best = d if best is None or d > best else best
        return best

    obs_ys = []
    used = [False] * len(pooled)
    # the observed labelling: within a tie group the split is what matters, not the order
    for v in y:
        for i, w in enumerate(pooled):
            if w == v and not used[i]:
                used[i] = True
                obs_ys.append(i)
                break
    d_obs = stat(obs_ys)
    hits = total = 0
    for ys in itertools.combinations(range(len(pooled)), m):
        total += 1
        if stat(ys) >= d_obs:
            hits += 1
    return hits / total


def test_ks_less_ties_brute_force():
    """ks_test_two_sample_less (stats.cpp:2221-2284) with tied values, as base qualities always are,
    against the brute-force permutation p-value.  Planted mistakes caught: testing the boundary
    inside a tie group (boundary[] all true), '>' for '>=' in the lattice test, x and y swapped."""
    rng = np.random.default_rng(9)
    cases = 0
    for _ in range(60):
        nx = int(rng.integers(1, 8))
        ny = int(rng.integers(1, 14 - nx + 1))
        if nx + ny > 14:
            continue
        vals = rng.integers(0, int(rng.integers(2, 6)), nx + ny).astype(float)
        x, y = vals[:nx], vals[nx:]
        got = O.ks_less(x, y)
        want = ks_brute(x, y)
        assert math.isclose(got, want, rel_tol=1e-12, abs_tol=1e-12), (list(x), list(y), got, want)
        cases += 1
    assert cases >= 40


def test_ks_less_asymptotic_branch():
    """At nx*ny >= 10000 the reference returns exp(-2 D^2 nx ny / (nx + ny)) (stats.cpp:2279-2283),
    D = max over pooled values of Fy - Fx with ties folded (computed here from searchsorted ECDFs).
    Planted mistake caught: the exact lattice used above the threshold, or D from unfolded ties."""
    rng = np.random.default_rng(10)
    for nx, ny in [(100, 100), (150, 90), (400, 30)]:
        x = rng.integers(20, 40, nx).astype(float)
        y = rng.integers(22, 41, ny).astype(float)
        u = np.unique(np.concatenate([x, y]))
        fx = np.searchsorted(np.sort(x), u, side="right") / nx
        fy = np.searchsorted(np.sort(y), u, side="right") / ny
        d = max(0.0, float(np.max(fy - fx)))
        want = math.exp(-2.0 * d * d * nx * ny / (nx + ny))
        assert math.isclose(O.ks_less(x, y), want, rel_tol=1e-12), (nx, ny)


# ---------------------------------------------------------------- base_qual_error_prob text

def test_base_qual_error_prob_text_by_hand():
    """write_base_qual_only_prob_table (error_count.cpp:697-785) for one forward read CGTT over CGTA
    (qualities 30-33): counts are summed into a table whose innermost index is the reference base,
    normalized over the reference base for each (quality, observed base), and printed in cells
    labelled b1 b2 (b1 over the reference-base loop) that read the table at reference b2, observed b1.
    So the one A->T observation at quality 33 prints as 1 in column "TA", and "AT" is NA.  Zero
    cells print NA; values use 6 significant digits.  Planted mistakes caught: the natural
    (ref, obs) layout, normalizing over the observed base."""
    ref = "ACGTACGTACGTACGTACGT"
    b = Built([("r", ref)], [dict(pos=1, cigar="4M", seq="CGTT", qual=[30, 31, 32, 33]),
                             dict(pos=1, cigar="1M", seq="G", qual=[30])])
    args = b.error_count_args(quality_max=Q42)
    _, _, _, texts = O.error_count(args)
    lines = texts["base_qual_0"].splitlines()
    header = lines[0].split("\t")
    cols = [x + y for x in "ACGT." for y in "ACGT."]
    assert header == ["quality"] + cols
    rows = {int(l.split("\t")[0]): dict(zip(cols, l.split("\t")[1:])) for l in lines[1:]}
    assert sorted(rows) == list(range(Q42))
    want = {
        # (ref C, obs C, 30) and (ref C, obs G, 30): observed C -> ref C only; observed G -> ref C only
        30: {"CC": "1", "GC": "1"},
        # (G,G,31) and the '..' pair at quality 31
        31: {"GG": "1", "..": "1"},
        32: {"TT": "1", "..": "1"},
        # (A,T,33): observed T, reference A -> cell "TA"; the '..' pair at 33
        33: {"TA": "1", "..": "1"},
    }
    for q in range(Q42):
        for col in cols:
            assert rows[q][col] == want.get(q, {}).get(col, "NA"), (q, col, rows[q][col])


def test_base_qual_error_prob_normalization_digits():
    """Three observations at one (quality, observed base) from three reference bases normalize to
    thirds, printed with 6 significant digits (0.333333) in the cells of observed base b1."""
    ref = "ACGTACGTAC"
    # observed G at quality 20 over reference A (pos 0), C (pos 1) and G (pos 2)
    b = Built([("r", ref)], [dict(pos=0, cigar="1M", seq="G", qual=[20]), dict(pos=1, cigar="1M", seq="G", qual=[20]),
                             dict(pos=2, cigar="1M", seq="G", qual=[20])])
    _, _, _, texts = O.error_count(b.error_count_args(quality_max=Q42))
    lines = texts["base_qual_0"].splitlines()
    cols = [x + y for x in "ACGT." for y in "ACGT."]
    row = dict(zip(cols, lines[1 + 20].split("\t")[1:]))
    assert (row["GA"], row["GC"], row["GG"], row["GT"], row["G."]) == ("0.333333", "0.333333", "0.333333", "NA", "NA")


def test_preprocess_read_starts_by_hand():
    """preprocess_stage (error_count.cpp:159-166, 187-190, 209-228): per scored target, the share of
    (position, strand) pairs at positions without redundant reads where no unique read has its first
    query base aligned with a stranded aligned end >= required_junction_read_end_min_coordinate."""
    import breseq_amd as B
    from aln_builder import Built, BuiltSample

    rng = np.random.default_rng(3)
    ref = "".join(rng.choice(list("ACGT"), 400))
    F, R = 0, 16

    def read(pos, cigar, flag=F, x1=-1):
        import re as _re
        qlen = sum(int(n) for n, o in _re.findall(r"(\d+)([MIS=X])", cigar))
        return dict(pos=pos, cigar=cigar, seq="A" * qlen, flag=flag, x1=x1)

    reads = [
        read(10, "97M3S"),            # forward, first base aligned at 10, end 97 >= 95: start (10, +)
        read(20, "3S97M"),            # starts with a clip: no start
        read(30, "90M10S"),           # forward end 90 < 95: no start
        read(40, "90M10S", R),        # reverse: stranded end = 100 - 1 + 1 = 100: start (40, -)
        read(40, "10S90M", R),        # reverse: stranded end = 100 - 11 + 1 = 90 < 95: no start
        read(50, "2D98M2S"),          # a leading deletion: first base at 52, end 98: start (52, +)
        read(60, "100M", F, x1=2),    # redundant: no start, and columns 60..159 are not counted
        read(70, "100M"),             # starts inside the redundant stretch: not counted
        read(200, "100M", R),         # start (200, -)
        read(200, "100M"),            # start (200, +)
        read(210, "1I99M"),           # an insertion first: no start
    ]
    reads.sort(key=lambda r: r["pos"])
    s = BuiltSample(Built([("t0", ref)], reads))
    args = B.api.make_error_count_args(s, preprocess_stage=True, unmatched_end_minimum_read_length=50,
                                       require_match_fraction=0.9)
    _, _, _, texts = O.error_count(args)
    # required end for 100-base reads: 100 - floor((100 - 50) * (1 - 0.9)) = 95
    # redundant columns 60..159 drop out: 400 - 100 = 300 positions, 600 (position, strand) pairs;
    # starts at (10,+), (40,-), (52,+), (200,-), (200,+) -- none inside 60..159
    assert texts["no_pos_hash_pr"][0] == (600 - 5) / 600
    # a factor that makes every end qualify adds (30,+) and (40,-) of "10S90M" (same pair) only
    args2 = B.api.make_error_count_args(s, preprocess_stage=True, unmatched_end_minimum_read_length=50,
                                        require_match_fraction=0.0)
    _, _, _, texts2 = O.error_count(args2)
    # required end 100 - floor(50 * 1.0) = 50: (30,+) qualifies, the reverse 10S90M shares (40,-)
    assert texts2["no_pos_hash_pr"][0] == (600 - 6) / 600
    # without preprocess_stage nothing is reported
    _, _, _, texts3 = O.error_count(B.api.make_error_count_args(s))
    assert np.isnan(texts3["no_pos_hash_pr"][0])


# ---------------------------------------------------------------- alignment trims (XL / XR)

# A reference whose only tandem repeat is the AAAA run at 3..6.  Its SequenceTrims, worked out from
# calculate_trims.cpp:67-138: the longest repeat starting at each position is 1 everywhere except
# 4, 3, 2 at 3, 4, 5; right trims scatter o + 1 over o < m from each position and left trims
# m - o over o <= m.
TRIM_REF = "ACGAAAATCG"
TRIM_LEFT = [1, 1, 1, 4, 3, 2, 1, 1, 1, 1]
TRIM_RIGHT = [1, 1, 1, 1, 2, 3, 4, 1, 1, 1]

# (pos, cigar, seq, flag) -> XL, XR by hand (calculate_trims.cpp:183-285): max(genomic trim at the
# first / last reference base, edge repeat of the aligned bases; the right edge on the reverse
# complement) + the soft-clipped bases on that side.
TRIM_READS = [
    # genomic 1 / 1; aligned "AAC": left repeat AA = 2, right (GTT) 1; plus the 2 clipped bases
    (0, "2S3M", "AAAAC", 0, 4, 1),
    # ends at 6 (right trim 4); reverse complement of GATT is AATC: AA = 2
    (2, "2M1D2M", "GATT", 0, 1, 4),
    # starts at 3 (left 4), ends at 6 (right 4); CAAT has no repeat at either end
    (3, "4M", "CAAT", 16, 4, 4),
    # hard clips are not part of the read: AAA gives 3 / 3; genomic 4 (at 3) / 3 (at 5)
    (3, "2H3M1H", "AAA", 0, 4, 3),
    (4, "3M", "CGT", 0, 3, 4),
    # GCACAC: none on the left; its reverse complement GTGTGC starts with GT GT = 4
    (4, "6M", "GCACAC", 16, 3, 4),
    # covers no reference base: not trimmed (-1)
    (5, "5S", "AAAAA", 0, -1, -1),
    # unmapped: -1
    (5, "3M", "AAA", 4, -1, -1),
    (7, "2M", "TC", 0, 1, 1),
    # aligned "TCG" ends at 9; two clipped bases on the right
    (7, "3M2S", "TCGTT", 0, 1, 3),
]


def trim_sample():
    reads = [dict(pos=p, cigar=c, seq=s, flag=f) for p, c, s, f, _, _ in TRIM_READS]
    return BuiltSample(Built([("t0", TRIM_REF)], reads))


def test_alignment_trims_by_hand():
    left, right = O.sequence_trims(TRIM_REF)
    assert list(left) == TRIM_LEFT and list(right) == TRIM_RIGHT
    s = trim_sample()
    xl, xr = O.alignment_trims(s.ref, s.aln)
    assert list(xl) == [r[4] for r in TRIM_READS]
    assert list(xr) == [r[5] for r in TRIM_READS]


def test_simulator_trims_equal_oracle():
    # the sample generator writes XL / XR from its own restatement of get_alignment_trims
    s = api.Sample.simulate(seed=5, ref_length=20000, read_length=100, coverage=30.0, paired=True, n_snps=10,
                            n_indels=2)
    xl, xr = O.alignment_trims(s.ref, s.aln)
    a = s.arrays()
    assert np.array_equal(a["xl"], xl) and np.array_equal(a["xr"], xr)
