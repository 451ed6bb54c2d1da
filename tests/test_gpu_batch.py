"""Batched and served runs (breseq_amd.batch): every sample's GD and DP text in the mix equals the
same sample run alone."""
import sys

import pytest

import breseq_amd as B
from breseq_amd import batch

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def jobs():
    dp = B.api.dp_settings([dict(median=350.0, mad=33.7, r1_prefix=1, r2_prefix=2)], seed=3)
    return [
        (B.Sample.simulate(seed=41, ref_length=40000, read_length=100, coverage=50.0, n_snps=10, n_indels=2),
         B.CONSENSUS, None),
        (B.Sample.simulate(seed=42, ref_length=60000, read_length=150, coverage=80.0, paired=True, n_snps=12,
                           n_indels=3, discordant_fraction=0.01, n_junctions=3), B.CONSENSUS, dp),
        (B.Sample.simulate(seed=43, ref_length=30000, read_length=150, coverage=150.0, paired=True, n_snps=15,
                           n_indels=3, variant_freq_min=0.05, variant_freq_max=0.5), B.POLYMORPHISM, dp),
        (B.Sample.simulate(seed=44, ref_length=25000, read_length=36, coverage=120.0, n_snps=8, n_indels=1),
         B.CONSENSUS, None),
    ]


def test_batch_equals_alone():
    say("[gpu] batch of four samples vs each alone")
    js = jobs()
    alone = [batch.run_pipeline(s, m, d) for s, m, d in js]
    mixed, wall = batch.run_batch(js, parallel=4)
    for a, b in zip(alone, mixed):
        assert a["gd"] == b["gd"]
        assert a["dp"] == b["dp"]
        assert a["gd"].count("\n") > 1
    say("[gpu]   batch wall s", round(wall, 3))


def test_server_equals_alone():
    say("[gpu] server with a small device-memory budget vs each alone")
    js = jobs()
    alone = [batch.run_pipeline(s, m, d) for s, m, d in js]
    # room for about two of the samples at a time
    budget = 2 * max(batch.device_bytes(s) for s, _, _ in js)
    srv = batch.Server(parallel=4, hbm_budget=budget)
    for i, (s, m, d) in enumerate(js):
        srv.submit(i, s, m, d)
    results, errors = srv.close()
    assert not errors
    for i, a in enumerate(alone):
        assert results[i]["gd"] == a["gd"]
        assert results[i]["dp"] == a["dp"]
        assert results[i]["latency_s"] > 0.0


def test_reused_session_equals_alone():
    say("[gpu] one session reused across samples of every shape vs each alone")
    js = jobs()
    alone = [batch.run_pipeline(s, m, d) for s, m, d in js]
    order = [0, 1, 2, 3, 2, 0, 3, 1]   # larger and smaller samples, both modes, one after another
    reused, _ = batch.run_batch([js[i] for i in order], parallel=1)
    for i, r in zip(order, reused):
        assert r["gd"] == alone[i]["gd"]
        assert r["dp"] == alone[i]["dp"]
