"""Launch shape of the column pass (bsq_session_set_column_launch) and its autotuner
(bsq_session_tune_columns).

* The genome diff does not depend on the launch: one block for the whole sample, grids that are
  and are not multiples of the 8 XCDs, more blocks than windows, blocks of 4 and of 8 wavefronts,
  and the likelihood table in LDS and in global memory all print the same text as the default.
* The tuner times at least the default and one other launch, keeps one no slower than the
  default, and a second session holding a sample of the same shape reuses the choice untimed.
"""
import sys

import pytest

import breseq_amd as B
import oracle_lib as O
from parity_util import make_sample, oracle_error_count

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def inputs(name):
    s = make_sample(name)
    args, (_, _, l_o, texts) = oracle_error_count(s)
    table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
    return s, table, args.read_set_max, args.quality_max, [8.0] * s.n_targets


@pytest.mark.parametrize("name,mode", [("se100", B.CONSENSUS), ("poly", B.POLYMORPHISM)])
def test_genome_diff_does_not_depend_on_launch(name, mode):
    s, table, S, Q, prop = inputs(name)
    sess = B.Session().load(s)
    try:
        ref = B.api.identify_mutations(s, table, S, Q, prop, session=sess, mode=mode)
        assert ref.count("\nRA\t") > 5
        for blocks, lds, waves in [(1, True, 8), (1, False, 4), (8, True, 4), (8, False, 8), (13, True, 8),
                                   (24, False, 4), (100000, True, 4), (100000, False, 8), (0, True, 4), (0, False, 0),
                                   (0, True, 0)]:
            say("[gpu] column launch", name, blocks, lds, waves)
            sess.set_column_launch(blocks, lds, waves)
            got = B.api.identify_mutations(s, table, S, Q, prop, session=sess, mode=mode)
            assert got == ref, (blocks, lds, waves)
    finally:
        sess.close()


def test_launch_arguments_are_checked():
    s, table, S, Q, prop = inputs("se100")
    sess = B.Session().load(s)
    try:
        with pytest.raises(B.BsqError):
            sess.set_column_launch(-1, True)
        with pytest.raises(B.BsqError):
            sess.set_column_launch(0, True, waves=2)
        with pytest.raises(B.BsqError):
            sess.tune_columns(table, S, Q, prop, repeats=0)
    finally:
        sess.close()


def test_tuner_keeps_a_timed_launch_and_reuses_it():
    s, table, S, Q, prop = inputs("pe150")
    say("[gpu] column tuner")
    plain = B.Session().load(s)
    try:
        want = B.api.identify_mutations(s, table, S, Q, prop, session=plain)
    finally:
        plain.close()
    first_sess = B.Session().load(s)
    try:
        first = first_sess.tune_columns(table, S, Q, prop, repeats=1)
        say("[gpu] tuned", first)
        assert first["from_cache"] == 0
        assert first["n_tried"] >= 2
        assert 0.0 < first["ms"] <= first["default_ms"]
        assert first["grid"] >= 1
        assert B.api.identify_mutations(s, table, S, Q, prop, session=first_sess) == want
    finally:
        first_sess.close()
    second_sess = B.Session().load(s)
    try:
        second = second_sess.tune_columns(table, S, Q, prop, repeats=1)
        assert second["from_cache"] == 1
        assert (second["blocks"], second["table_in_lds"], second["waves"]) == \
            (first["blocks"], first["table_in_lds"], first["waves"])
        assert B.api.identify_mutations(s, table, S, Q, prop, session=second_sess) == want
    finally:
        second_sess.close()
