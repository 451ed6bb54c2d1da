"""The device genome-diff writer (gd_emit.hip) against the host writer and against printf.

* The device number formatter prints the text of printf's %.*e and %.*f (Python's % operator calls
  the C library's printf) for random doubles of every magnitude, exact ties, zeros and infinities.
* bsq_session_identify through the device writer and through the host writer give byte-identical
  genome diffs: consensus and polymorphism samples, several targets whose names are not in their
  index order, targets left out of the call set, a failed coverage fit (whole-target MC), seed
  cutoffs that keep or drop MC runs, and an MC with redundant-coverage ranges.  The host writer's
  text is compared with the oracle's by the other GPU tests."""
import ctypes as C
import sys

import numpy as np
import pytest

import breseq_amd as B
import oracle_lib as O
from aln_builder import Built, BuiltSample, random_alignments
from breseq_amd import _lib as L
from parity_util import make_sample, oracle_error_count

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_format(values, precision, scientific):
    v = np.ascontiguousarray(values, dtype=np.float64)
    n = v.size
    out = C.create_string_buffer(512 * n)
    lens = np.zeros(n, dtype=np.int32)
    st = L.Status()
    rc = L.lib().bsq_format_doubles_device(v.ctypes.data_as(L.c_f64p), n, precision, 1 if scientific else 0, out,
                                           lens.ctypes.data_as(L.c_i32p), C.byref(st))
    assert rc == 0, st.message
    raw = out.raw
    return [raw[512 * i: 512 * i + lens[i]].decode() for i in range(n)]


def test_device_formatter_matches_printf():
    say("[gpu] device formatter")
    rng = np.random.default_rng(3)
    bits = rng.integers(0, 2**63, size=60000, dtype=np.int64).astype(np.uint64)
    bits[rng.random(bits.size) < 0.5] |= np.uint64(1) << np.uint64(63)
    v = bits.view(np.float64)
    v = v[np.isfinite(v)]
    mid = rng.random(40000) * 10.0 ** rng.integers(-30, 30, size=40000)
    ties = np.ldexp(rng.integers(-100000, 100000, size=20000).astype(np.float64), -rng.integers(0, 12, size=20000))
    fixed = rng.integers(-2000000, 2000000, size=20000) / 1000.0
    special = np.array([0.0, -0.0, np.inf, -np.inf, 5e-324, 1.7976931348623157e308, 0.5, 2.5, 0.125, 9.5, 0.05])
    vals = np.concatenate([v, mid, ties, fixed, special])
    for prec in (0, 1, 3, 5, 8, 17):
        for sci in (True, False):
            sample = vals if sci else vals[np.abs(vals) < 1e30]
            got = device_format(sample, prec, sci)
            fmt = "%%.%d%s" % (prec, "e" if sci else "f")
            for x, g in zip(sample, got):
                want = fmt % x
                assert g == want, (x.hex(), prec, sci, g, want)
    assert device_format(np.array([np.nan]), 3, True) == ["NA"]


def both_writers(s, table, S, Q, prop, capacity=None, **kw):
    sess = B.Session().load(s)
    try:
        sess.set_gd_writer(False)
        host = B.api.identify_mutations(s, table, S, Q, prop, session=sess, **kw)
        sess.set_gd_writer(True)
        if capacity is not None:
            sess.set_gd_capacity(capacity)
        dev = B.api.identify_mutations(s, table, S, Q, prop, session=sess, **kw)
        if capacity is not None:
            # the session kept the grown lists: a second pass fits at once
            again = B.api.identify_mutations(s, table, S, Q, prop, session=sess, **kw)
            assert again == dev
    finally:
        sess.close()
    return host, dev


def counts(text):
    c = {"RA": 0, "MC": 0, "UN": 0}
    for ln in text.splitlines():
        if ln and not ln.startswith("#"):
            c[ln.split("\t")[0]] += 1
    return c


@pytest.mark.parametrize("name,mode", [("se100", B.CONSENSUS), ("poly", B.POLYMORPHISM)])
def test_writers_agree_on_simulated_samples(name, mode):
    say("[gpu] gd writers", name)
    s = make_sample(name)
    args, (_, _, l_o, texts) = oracle_error_count(s)
    table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
    host, dev = both_writers(s, table, args.read_set_max, args.quality_max, [8.0] * s.n_targets, mode=mode)
    assert dev == host
    assert counts(dev)["RA"] > 10


def test_writers_agree_on_targets_out_of_name_order():
    say("[gpu] gd writers: several targets, failed fit, seed cutoffs")
    refs, reads = random_alignments(5, n_targets=4, target_len=1500, n_reads=2400, run_len=(20, 60))
    names = ["zeta", "alpha", "Mid", "beta_2"]
    refs = [(names[i], sq) for i, (_, sq) in enumerate(refs)]
    s = BuiltSample(Built(refs, reads))
    args = B.api.make_error_count_args(s)
    _, _, l_o, texts = O.error_count(args)
    table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
    S, Q = args.read_set_max, args.quality_max
    cases = [
        dict(prop=[6.0, 4.0, -1.0, 9.0]),                                  # target "Mid": failed fit
        dict(prop=[12.0, 12.0, 12.0, 12.0], seed_cutoffs=[3.0, 0.0, 5.0, 1.0]),
        dict(prop=[6.0, -1.0, 6.0, -1.0], call_targets=[0, 1, 3]),         # "Mid" not called
        dict(prop=[6.0, 4.0, 5.0, 9.0], mode=B.POLYMORPHISM),
        dict(prop=[6.0, 4.0, -1.0, 9.0], skip_mc=True),
    ]
    seen = {"RA": 0, "MC": 0, "UN": 0}
    for c in cases:
        kw = dict(c)
        prop = kw.pop("prop")
        host, dev = both_writers(s, table, S, Q, prop, **kw)
        assert dev == host, kw
        for k, v in counts(dev).items():
            seen[k] += v
    assert seen["RA"] > 0 and seen["MC"] > 2 and seen["UN"] > 0, seen


def test_writer_mc_with_redundant_ranges():
    say("[gpu] gd writers: MC with redundant ranges")
    s = B.Sample.simulate(seed=31, ref_length=30000, read_length=100, coverage=40.0, paired=False, n_snps=5,
                          n_indels=1, redundant_fraction=0.05, n_missing=2, missing_length=500, missing_edge_reads=6)
    args, (_, _, l_o, texts) = oracle_error_count(s)
    table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
    host, dev = both_writers(s, table, args.read_set_max, args.quality_max, [15.0] * s.n_targets)
    assert dev == host
    assert counts(dev)["MC"] >= 1


def test_writer_lists_grow_when_a_pass_needs_more():
    say("[gpu] gd writers: entry lists smaller than the pass needs")
    s = make_sample("poly")
    args, (_, _, l_o, texts) = oracle_error_count(s)
    table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
    host, dev = both_writers(s, table, args.read_set_max, args.quality_max, [8.0] * s.n_targets, capacity=5,
                             mode=B.POLYMORPHISM)
    assert dev == host
    assert sum(counts(dev).values()) > 5
