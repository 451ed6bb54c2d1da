"""Session loading (bsq_session_load): the alignment invariants it checks are reported for the
first offending alignment in input order, also when the packing runs over many threads; unmapped
records, records without a target and records that cover no reference base are skipped."""
import sys

import numpy as np
import pytest

import breseq_amd as B
from aln_builder import Built, BuiltSample, random_alignments

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(refs, reads):
    s = BuiltSample(Built(refs, reads))
    sess = B.Session()
    try:
        sess.load(s)
        return s, sess
    except Exception:
        sess.close()
        raise


def many_reads(seed=3):
    # enough records for the load to split them into many chunks
    return random_alignments(seed, n_targets=2, target_len=20000, n_reads=60000, run_len=(10, 40))


def test_invariants_are_reported_for_the_first_offender():
    say("[gpu] session load: invariants")
    refs, reads = many_reads()
    n = len(reads)
    # an alignment out of coordinate order near the end
    bad = [dict(r) for r in reads]
    bad[n - 50] = dict(bad[n - 50], pos=0, tid=0)
    with pytest.raises(B.BsqError, match="not sorted"):
        load(refs, bad)
    # a CIGAR whose query length differs from SEQ, after that disorder: the disorder comes first
    bad2 = [dict(r) for r in bad]
    bad2[n - 10] = dict(bad2[n - 10], seq=bad2[n - 10]["seq"] + "A", qual=bad2[n - 10]["qual"] + [30])
    with pytest.raises(B.BsqError, match="not sorted"):
        load(refs, bad2)
    # the same CIGAR error before the disorder is reported instead
    bad3 = [dict(r) for r in bad]
    bad3[100] = dict(bad3[100], seq=bad3[100]["seq"] + "A", qual=bad3[100]["qual"] + [30])
    with pytest.raises(B.BsqError, match="CIGAR query length"):
        load(refs, bad3)
    # an alignment past its reference's end
    bad4 = [dict(r) for r in reads]
    last = max(i for i, r in enumerate(bad4) if r["tid"] == 1)
    bad4[last] = dict(bad4[last], pos=20000 - 5)
    with pytest.raises(B.BsqError, match="extends past"):
        load(refs, bad4)


def test_skipped_records_keep_the_others_intact():
    say("[gpu] session load: skipped records")
    refs, reads = many_reads(4)
    rng = np.random.default_rng(1)
    mixed = []
    for r in reads:
        mixed.append(r)
        u = rng.random()
        if u < 0.01:
            mixed.append(dict(r, flag=4))                     # unmapped
        elif u < 0.02:
            mixed.append(dict(r, cigar="%dS" % len(r["seq"])))   # no reference base covered
    s_all, sess = load(refs, mixed)
    try:
        xl, xr = sess.alignment_trims()
    finally:
        sess.close()
    s_ref, sess2 = load(refs, reads)
    try:
        xl2, xr2 = sess2.alignment_trims()
    finally:
        sess2.close()
    got = np.array([xl[i] for i in range(len(mixed)) if xl[i] >= 0])
    assert got.size == len(reads) and np.array_equal(got, xl2)
    assert np.array_equal(np.array([xr[i] for i in range(len(mixed)) if xr[i] >= 0]), xr2)
