"""GPU parity: the MI355X path against the oracle on seeded synthetic samples.

Integer outputs (coverage histograms, covariate count tables, error-rate text, trims, MC/UN
intervals, RA positions and coverage strings) must be bit-exact.  Printed FP64 fields may differ by
at most one unit in the last printed digit (device vs glibc log10/pow/lgamma ulps).
"""
import sys

import numpy as np
import pytest

import breseq_amd as B
import oracle_lib as O
from parity_util import compare_gd, make_sample, oracle_error_count, oracle_identify

pytestmark = pytest.mark.gpu

CACHE = {}


def session_for(name):
    if name not in CACHE:
        s = make_sample(name)
        sess = B.Session().load(s)
        CACHE[name] = (s, sess)
    return CACHE[name]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["se100", "pe150", "poly"])
def test_error_count_bit_exact(name):
    say("[gpu] error_count", name)
    s, sess = session_for(name)
    args, (h_o, c_o, l_o, texts) = oracle_error_count(s)
    res = sess.error_count()
    assert len(res.histograms) == len(h_o)
    for g in range(len(h_o)):
        assert np.array_equal(res.histograms[g], h_o[g]), "histogram of group %d differs" % g
    assert np.array_equal(res.counts, c_o.astype(np.uint64))
    assert np.array_equal(res.log10_prob, l_o)
    assert res.rates_text == texts["error_rates"]
    assert np.array_equal(res.rates_table, O.read_log10_prob_table(texts["error_rates"], len(l_o)))
    assert res.counts.sum() > 0


@pytest.mark.parametrize("name,mode", [("se100", B.CONSENSUS), ("pe150", B.CONSENSUS), ("poly", B.POLYMORPHISM)])
def test_identify_mutations_matches_oracle(name, mode):
    say("[gpu] identify", name, mode)
    s, sess = session_for(name)
    args, (h_o, c_o, l_o, texts) = oracle_error_count(s)
    table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
    prop = [5.0] * s.n_targets
    gd_o, _ = oracle_identify(s, table, args.read_set_max, args.quality_max, prop, mode=mode)
    gd_g = sess.identify_mutations(log10_prob=table, read_set_max=args.read_set_max, quality_max=args.quality_max,
                                   propagation_cutoffs=prop, mode=mode)
    n, same, errs = compare_gd(gd_g, gd_o)
    say("[gpu]   entries", n, "identical lines", same, "max field errors", errs)
    assert n > 10
    assert "RA\t" in gd_g


def test_screen_drops_nothing():
    say("[gpu] screen vs unscreened")
    s, sess = session_for("poly")
    args, (h_o, c_o, l_o, texts) = oracle_error_count(s)
    table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
    prop = [5.0] * s.n_targets
    kw = dict(log10_prob=table, read_set_max=args.read_set_max, quality_max=args.quality_max,
              propagation_cutoffs=prop, mode=B.POLYMORPHISM)
    a = sess.identify_mutations(**kw)
    b = sess.identify_mutations(screen=False, **kw)
    assert a == b


def test_sequence_trims_bit_exact():
    say("[gpu] trims")
    rng = np.random.default_rng(5)
    seq = "".join(rng.choice(list("ACGT"), 5000))
    seq = seq[:1000] + "A" * 300 + "CAG" * 40 + seq[1000:]   # long repeats, cap at 255
    l_g, r_g = B.calculate_trims(seq)
    l_o, r_o = O.sequence_trims(seq)
    assert np.array_equal(l_g, l_o)
    assert np.array_equal(r_g, r_o)


@pytest.mark.parametrize("name,mode", [("pe150", B.CONSENSUS), ("poly", B.POLYMORPHISM)])
def test_precision_audit(name, mode):
    """The column pass's FP32 sums against the FP64 sums in read order: every final call equals the
    exact one, read counts agree, and the relative error stays within the summation bound
    (n + 2) 2^-24 for columns of a few hundred reads."""
    say("[gpu] precision audit", name)
    s, sess = session_for(name)
    args, (h_o, c_o, l_o, texts) = oracle_error_count(s)
    table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
    rep = sess.precision_audit(table, args.read_set_max, args.quality_max, [5.0] * s.n_targets, mode=mode)
    say("[gpu]  ", rep)
    assert rep["columns"] > 0.9 * sum(s.target_lengths())
    assert rep["decision_mismatches"] == 0 and rep["count_mismatches"] == 0
    assert 0.0 < rep["max_rel_error"] < 1000 * 2.0 ** -24
