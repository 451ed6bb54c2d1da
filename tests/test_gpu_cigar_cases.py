"""GPU parity on hand-built alignment sets covering every CIGAR shape the pileup distinguishes (soft
clips, adjacent insertions/deletions, reference skips, pads, =/X runs, N bases, both strands,
redundant reads, reads ending at a target end), against the oracle.  Integer outputs bit-exact;
GD lines identical up to one unit in the last printed digit of FP fields.  One read in ten has a
random MAPQ, so the likelihood table has ~40 MAPQ slots (~8,400 rows) and the column kernel reads
it from HBM; the simulated samples of the other GPU tests take the LDS-staged table."""
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


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_error_count_cigar_shapes_bit_exact(seed):
    say("[gpu] error_count cigar shapes seed", seed)
    refs, reads = random_alignments(seed, n_reads=900)
    b = Built(refs, reads)
    s = BuiltSample(b)
    args = B.api.make_error_count_args(s)
    h_o, c_o, l_o, texts = O.error_count(args)
    res = B.api.error_count(s)
    assert len(res.histograms) == len(h_o)
    for g in range(len(h_o)):
        assert np.array_equal(res.histograms[g], h_o[g]), g
    assert np.array_equal(res.counts, c_o.astype(np.uint64))
    assert res.rates_text == texts["error_rates"]
    assert res.counts.sum() > 0


@pytest.mark.parametrize("seed,mode", [(1, B.CONSENSUS), (2, B.POLYMORPHISM), (3, B.CONSENSUS)])
def test_identify_cigar_shapes_match_oracle(seed, mode):
    say("[gpu] identify cigar shapes seed", seed, mode)
    refs, reads = random_alignments(seed, n_reads=900)
    b = Built(refs, reads)
    s = BuiltSample(b)
    args = B.api.make_error_count_args(s)
    _, _, l_o, texts = O.error_count(args)
    table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
    prop = [3.0] * s.n_targets
    iargs = B.api.make_identify_args(s, table, args.read_set_max, args.quality_max, prop, mode=mode)
    gd_o, _ = O.identify_mutations(iargs)
    gd_g = B.api.identify_mutations(s, table, args.read_set_max, args.quality_max, prop, mode=mode)
    n, same, errs = compare_gd(gd_g, gd_o)
    say("[gpu]   entries", n, "identical", same, errs)
    assert n > 10


@pytest.mark.parametrize("seed", [1, 4])
def test_preprocess_read_starts_match_oracle(seed):
    """error_count's preprocess mode (coverage only, read starts per target) on every CIGAR shape,
    and on a simulated paired sample with soft-clipped and redundant reads."""
    say("[gpu] preprocess read starts seed", seed)
    if seed == 4:
        refs, reads = random_alignments(seed, n_targets=3, target_len=3000, n_reads=900, run_len=(30, 120),
                                        clip_len=(1, 30))
        s = BuiltSample(Built(refs, reads))
    else:
        s = B.Sample.simulate(seed=17, ref_length=40000, read_length=150, coverage=30.0, paired=True,
                              redundant_fraction=0.05)
    for frac in (0.9, 0.5):
        kw = dict(preprocess_stage=True, do_errors=False, require_match_fraction=frac)
        args = B.api.make_error_count_args(s, **kw)
        h_o, _, _, texts = O.error_count(args)
        res = B.api.error_count(s, **kw)
        for g in range(len(h_o)):
            assert np.array_equal(res.histograms[g], h_o[g])
        assert np.array_equal(res.no_pos_hash_pr, texts["no_pos_hash_pr"]), (res.no_pos_hash_pr, texts["no_pos_hash_pr"])
        assert np.all(res.no_pos_hash_pr < 1.0)


def test_deletion_next_to_the_last_base():
    """A deletion one base before the read's end: its quality comes from the base after it in read
    direction, which for a forward read lies past the end.  breseq reaches that base only for reads
    that are neither redundant nor trimmed at the deletion (identify_mutations.cpp:1591-1626); XL/XR
    as breseq writes them always trim the read's last base, so such reads are skipped, and both the
    product and the oracle must accept them.  Without XL/XR both must refuse them."""
    say("[gpu] deletion next to the last base")
    ref = "ACGTTGCAAGATAGCTTACGGATCCATGCAAGGCTTAACGTGACCTGA"
    reads = []
    for k in range(12):
        reads.append(dict(pos=2 + k, cigar="8M1D1M", seq=ref[2 + k:10 + k] + ref[11 + k], xl=1, xr=1))
        reads.append(dict(pos=3 + k, cigar="10M", seq=ref[3 + k:13 + k], xl=1, xr=1))
    reads.sort(key=lambda r: r["pos"])
    for trimmed in (True, False):
        rr = [dict(r) for r in reads]
        if not trimmed:
            for r in rr:
                r["xl"] = r["xr"] = -1
        s = BuiltSample(Built([("t0", ref)], rr))
        args = B.api.make_error_count_args(s)
        _, _, l_o, texts = O.error_count(args)
        table = O.read_log10_prob_table(texts["error_rates"], len(l_o))
        iargs = B.api.make_identify_args(s, table, args.read_set_max, args.quality_max, [1.0])
        if trimmed:
            gd_o, _ = O.identify_mutations(iargs)
            gd_g = B.api.identify_mutations(s, table, args.read_set_max, args.quality_max, [1.0])
            n, same, errs = compare_gd(gd_g, gd_o)
            assert n == same
        else:
            with pytest.raises(B.BsqError):
                O.identify_mutations(iargs)
            with pytest.raises(B.BsqError):
                B.api.identify_mutations(s, table, args.read_set_max, args.quality_max, [1.0])
