"""GPU parity on the shapes the default samples do not reach, against the oracle:

  * long reads (runs of 60-420 aligned bases, soft clips up to 40): alignments longer than the
    grouped error_count wavefronts take (three 168-base alignments per wavefront), and aligned runs
    above 256 bases, which the per-lane loops cover in more than one pass;
  * simulated 2x300 bp pairs;
  * columns of about 9,000 reads: deeper than the scorer's read lists fit in LDS (they move to
    device scratch), and deep enough that the column pass's FP32 certification bound must use the
    full recursive-summation term."""
import sys

import numpy as np
import pytest

import breseq_amd as B
import oracle_lib as O
from aln_builder import Built, BuiltSample, random_alignments
from parity_util import compare_gd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _both(s, mode=B.CONSENSUS, prop=3.0, session=None):
    args = B.api.make_error_count_args(s)
    h_o, c_o, l_o, texts = O.error_count(args)
    res = session.error_count() if session is not None else B.api.error_count(s)
    for g in range(len(h_o)):
        assert np.array_equal(res.histograms[g], h_o[g]), "histogram %d" % g
    assert np.array_equal(res.counts, c_o.astype(np.uint64)), "covariate counts"
    assert res.rates_text == texts["error_rates"]
    table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
    props = [prop] * s.n_targets
    gd_o, _ = O.identify_mutations(B.api.make_identify_args(s, table, args.read_set_max, args.quality_max, props,
                                                            mode=mode))
    if session is not None:
        gd_g = session.identify_mutations(log10_prob=table, read_set_max=args.read_set_max,
                                          quality_max=args.quality_max, propagation_cutoffs=props, mode=mode)
    else:
        gd_g = B.api.identify_mutations(s, table, args.read_set_max, args.quality_max, props, mode=mode)
    return compare_gd(gd_g, gd_o)


@pytest.mark.parametrize("seed,mode", [(4, B.CONSENSUS), (5, B.POLYMORPHISM)])
def test_long_read_cigar_shapes(seed, mode):
    say("[gpu] long-read cigar shapes seed", seed, mode)
    refs, reads = random_alignments(seed, n_targets=2, target_len=4000, n_reads=700, run_len=(60, 420),
                                    clip_len=(1, 40))
    runs = [int(x[:-1]) for r in reads for x in __import__("re").findall(r"\d+[M=X]", r["cigar"])]
    assert max(runs) > 256 and sum(1 for r in reads if len(r["seq"]) > 168) > 100
    s = BuiltSample(Built(refs, reads))
    n, same, errs = _both(s, mode=mode)
    say("[gpu]   entries", n, "identical", same, errs)
    assert n > 10


def test_simulated_300bp_pairs():
    say("[gpu] simulated 2x300 bp pairs")
    s = B.Sample.simulate(seed=21, ref_length=30000, read_length=300, coverage=60.0, paired=True, insert_mean=700,
                          insert_sd=80, qual_mean=32, qual_sd=6, n_snps=15, n_indels=5)
    n, same, errs = _both(s, prop=5.0)
    say("[gpu]   entries", n, "identical", same, errs)
    assert n > 10


def test_deep_columns():
    say("[gpu] ~9,000-read columns")
    s = B.Sample.simulate(seed=22, ref_length=4000, read_length=100, coverage=9000.0, paired=False, n_snps=4,
                          n_indels=1, redundant_fraction=0.002)
    sess = B.Session().load(s)
    try:
        n, same, errs = _both(s, prop=2000.0, session=sess)
    finally:
        sess.close()
    say("[gpu]   entries", n, "identical", same, errs)
    assert n >= 4
