"""DP_candidate_regions.csv through the GPU identify_mutations entry point (the accumulator runs
beside the column kernels) against the oracle, with ra_mc_evidence.gd checked on the same run."""
import sys

import pytest

import breseq_amd as B
import oracle_lib as O
from aln_builder import Built, BuiltSample
from parity_util import compare_gd
from test_dp_regions import SETS, random_pair_reads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _both(s, dp, **kw):
    args = B.api.make_error_count_args(s)
    _, _, l_o, texts = O.error_count(args)
    table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
    prop = [3.0] * s.n_targets
    iargs = B.api.make_identify_args(s, table, args.read_set_max, args.quality_max, prop, dp=dp, **kw)
    gd_o, dp_o = O.identify_with_dp(iargs)
    gd_g, dp_g = B.api.identify_mutations(s, table, args.read_set_max, args.quality_max, prop, dp=dp, with_dp=True,
                                          **kw)
    return gd_o, dp_o, gd_g, dp_g


@pytest.mark.parametrize("seed", [1, 2])
def test_identify_dp_regions_match_oracle(seed):
    say("[gpu] identify with DP groups, random pairs seed", seed)
    refs, reads = random_pair_reads(seed)
    s = BuiltSample(Built(refs, reads))
    dp = B.api.dp_settings(SETS, seed=2)
    gd_o, dp_o, gd_g, dp_g = _both(s, dp)
    assert dp_g == dp_o
    assert dp_o.count("\n") > 3
    n, same, errs = compare_gd(gd_g, gd_o)
    say("[gpu]   entries", n, "identical", same, errs)


def test_simulated_discordant_sample_session():
    say("[gpu] simulated discordant pairs through a session")
    s = B.Sample.simulate(seed=22, ref_length=30000, read_length=100, coverage=30.0, paired=True, insert_mean=300,
                          insert_sd=30, n_snps=8, n_indels=3, discordant_fraction=0.03, n_junctions=5)
    dp = B.api.dp_settings([dict(median=300.0, mad=20.0, r1_prefix=1, r2_prefix=2)], seed=3)
    sess = B.Session().load(s)
    res = sess.error_count()
    prop = [3.0] * s.n_targets
    iargs = B.api.make_identify_args(s, res.rates_table, res.read_set_max, res.quality_max, prop, dp=dp)
    gd_o, dp_o = O.identify_with_dp(iargs)
    gd_g, dp_g = sess.identify_mutations(log10_prob=res.rates_table, read_set_max=res.read_set_max,
                                         quality_max=res.quality_max, propagation_cutoffs=prop, dp=dp, with_dp=True)
    assert dp_g == dp_o
    assert dp_o.count("\n") > 2
    n, same, errs = compare_gd(gd_g, gd_o)
    say("[gpu]   entries", n, "identical", same, errs)
    assert n > 0
