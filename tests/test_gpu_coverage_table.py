"""BAM2COV coverage tables through the GPU entry point (bsq_session_coverage_table) against the
oracle: the hand-worked tables, random alignments with every CIGAR shape, read groups, regions that
start or end inside reads, no-read regions and downsampling, and a simulated genome."""
import sys

import numpy as np
import pytest

import breseq_amd as B
import bam2cov_golden
import oracle_lib as O
from aln_builder import Built, BuiltSample, random_alignments
from test_coverage_table import READS, REF

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def both(sess, s, **kw):
    args = B.api.coverage_table_args(s, **kw)
    want = O.coverage_table(s.ref, s.aln, args)
    got = sess.coverage_table(**kw)
    return got, want


CASES = [
    dict(resolution=0),
    dict(resolution=5, start=1, end=10),
    dict(resolution=0, start=8, end=10),
    dict(start=3, end=4),
    dict(start=4, end=5, total_only=True, csv=True, reference_average=12.5),
    dict(start=4, end=5, per_read_group=True),
]


def test_hand_worked_tables():
    say("[gpu] coverage table, hand-worked cases")
    s = BuiltSample(Built([("chr", REF)], READS), n_read_groups=2)
    sess = B.Session().load(s)
    for kw in CASES:
        got, want = both(sess, s, **kw)
        assert got == want, kw


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_alignments(seed):
    say("[gpu] coverage table, random alignments seed", seed)
    refs, reads = random_alignments(seed, n_targets=2, target_len=500, n_reads=700)
    rng = np.random.default_rng(100 + seed)
    for r in reads:
        r["rg"] = int(rng.integers(-1, 3))
    s = BuiltSample(Built(refs, reads), n_read_groups=3)
    sess = B.Session().load(s)
    n = 0
    for t in range(2):
        for kw in [dict(), dict(resolution=0), dict(resolution=7), dict(start=1, end=37, resolution=0),
                   dict(start=120, end=131, resolution=0, total_only=True),
                   dict(start=450, end=500, resolution=3, per_read_group=True),
                   dict(start=2, end=499, resolution=11, per_read_group=True, total_only=True, csv=True),
                   dict(start=200, end=200, per_read_group=True, n_read_groups=1)]:
            got, want = both(sess, s, target=t, **kw)
            assert got == want, (t, kw)
            n += want.count("\n")
    say("[gpu]   rows compared", n)


def test_regions_beyond_the_reads():
    say("[gpu] coverage table, regions with no reads")
    refs, reads = random_alignments(5, n_targets=1, target_len=600, n_reads=200)
    reads = [r for r in reads if r["pos"] < 300 and r["pos"] + 80 < 400]
    s = BuiltSample(Built(refs, reads))
    sess = B.Session().load(s)
    for kw in [dict(start=450, end=600, resolution=0), dict(start=450, end=600, resolution=40),
               dict(start=380, end=600, resolution=9), dict(resolution=13)]:
        got, want = both(sess, s, **kw)
        assert got == want, kw


def test_simulated_genome():
    say("[gpu] coverage table, simulated 60 kb genome at 40x")
    s = B.Sample.simulate(seed=31, ref_length=60000, read_length=100, coverage=40.0, redundant_fraction=0.05,
                          n_snps=10, n_indels=4)
    sess = B.Session().load(s)
    for kw in [dict(), dict(resolution=0, start=1000, end=3000), dict(start=59000, end=60000, resolution=0)]:
        got, want = both(sess, s, **kw)
        assert got == want, kw


@pytest.mark.parametrize("case", sorted(bam2cov_golden.CASES))
def test_reference_golden_tables(case):
    say("[gpu] coverage table, breseq golden", case)
    refs, reads, n_rg, kw, want = bam2cov_golden.CASES[case]()
    s = BuiltSample(Built(refs, reads), n_read_groups=n_rg)
    sess = B.Session().load(s)
    assert sess.coverage_table(**kw) == want
