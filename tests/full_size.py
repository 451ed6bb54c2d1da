"""Helpers of the full-size parity tests: the oracle's fixtures of the benchmark workloads, the
product's pipeline on the same seeded sample, and the comparison."""
import gzip
import json
import os

import numpy as np

import breseq_amd as B
from breseq_amd import workloads
from parity_util import compare_gd, parse_gd

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    with gzip.open(os.path.join(GOLDEN, "oracle_%s.json.gz" % name), "rt") as f:
        return json.load(f)


def mode_of(name):
    return B.POLYMORPHISM if workloads.MODES[name] == "polymorphism" else B.CONSENSUS


def product_pipeline(sample, mode, out_dir):
    """bsq_session_pipeline on a loaded sample, writing breseq's output files into out_dir."""
    paired = len(sample.read_file_groups[0]) == 2
    dp = B.api.dp_settings([workloads.DP_SET], seed=workloads.DP_SEED) if paired else None
    sess = B.Session().load(sample)
    try:
        ec, summaries, gd, dp_text = sess.pipeline(mode=mode, dp=dp, ec_kw=dict(output_dir=out_dir))
        timing = sess.timing()
    finally:
        sess.close()
    files = {}
    for fn in sorted(os.listdir(out_dir)):
        with open(os.path.join(out_dir, fn)) as f:
            files[fn] = f.read()
    return ec, summaries, gd, dp_text, files, timing


def compare_with_fixture(fx, ec, summaries, gd, dp_text, files, read_file_names):
    """Integer outputs and texts bit-exact, GD fields to one unit in the last printed digit."""
    assert len(ec.histograms) == len(fx["histograms"])
    for g, h in enumerate(fx["histograms"]):
        assert np.array_equal(ec.histograms[g], np.array(h, dtype=np.uint32)), "histogram of group %d" % g
    assert np.array_equal(ec.counts, np.array(fx["counts"], dtype=np.uint64)), "covariate counts"
    assert ec.rates_text == fx["error_rates"], "error_rates.tab text"
    assert files["error_rates.tab"] == fx["error_rates"]
    assert len(fx["base_qual"]) == len(read_file_names)
    for k, text in fx["base_qual"].items():
        fn = "base_qual_error_prob.%s.tab" % read_file_names[int(k.split("_")[-1])]
        assert files[fn] == text, fn
    for k, text in fx["coverage_text"].items():
        g = int(k.split("_")[-1])
        assert files["%d.unique_only_coverage_distribution.tab" % g] == text
    fit = fx["fit"]
    s = summaries[0]
    assert s["deletion_coverage_propagation_cutoff"] == fit["deletion_coverage_propagation_cutoff"]
    assert s["nbinom_mean_parameter"] == fit["nb_fit_mu"] and s["nbinom_size_parameter"] == fit["nb_fit_size"]
    assert s["average"] == fit["average"] and s["variance"] == fit["variance"]
    n, same, errs = compare_gd(gd, fx["gd"])
    assert dp_text == fx["dp"], "DP_candidate_regions.csv"
    return n, same, errs


def required_cases(gd_text, polymorphism):
    """Counts of the evidence kinds a full-size workload must exercise."""
    e = parse_gd(gd_text)
    mc = [x for x in e if x[0] == "MC"]
    mc_ranges = [x for x in mc if int(x[2][3]) != 0 or int(x[2][4]) != 0]
    ra = [x for x in e if x[0] == "RA"]
    ra_insert = [x for x in ra if int(x[2][2]) != 0]
    interior = [x for x in ra if x[3].get("frequency_lower", "NA") not in ("NA",) and
                0.0 < float(x[3]["frequency_lower"]) and float(x[3].get("frequency_upper", "1")) < 1.0]
    out = {"RA": len(ra), "MC": len(mc), "UN": sum(1 for x in e if x[0] == "UN"), "MC_with_ranges": len(mc_ranges),
           "RA_insert_slot": len(ra_insert), "RA_interior_bounds": len(interior)}
    if not polymorphism:
        out.pop("RA_interior_bounds")
    return out


def c1_simreads_sample():
    """C1 made the way its configuration describes it: breseq's SIMULATE-READS (-m single -l 100
    -n 100000 --quality-start 32 --quality-end 32 --quality-stdev 6 --quality-min 2 --quality-max 41)
    on the lambda genome (48,502 bp, the reference's test data) with 20 SNPs and 2 small indels
    planted, aligned by construction."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "simreads_lambda.json.gz")
    with gzip.open(path, "rt") as f:
        ref = json.load(f)["reference"]
    rng = np.random.default_rng(2026)
    sites = []
    while len(sites) < 22:
        p = int(rng.integers(200, len(ref) - 200))
        if all(abs(p - q) >= 30 for q in sites):
            sites.append(p)
    muts = []
    for k, p in enumerate(sites):
        if k < 20:
            muts.append((p, 1, str(rng.choice([c for c in "ACGT" if c != ref[p]]))))
        elif k == 20:
            muts.append((p, 0, "".join(rng.choice(list("ACGT"), int(rng.integers(1, 6))))))
        else:
            muts.append((p, int(rng.integers(1, 6)), ""))
    muts.sort()
    return B.Sample.simulate_breseq(ref, name="NC_001416", mutations=muts, mode="single", n_reads=100000,
                                    read_length=100, seed=1, quality_start=32, quality_end=32, quality_stdev=6,
                                    quality_min=2, quality_max=41)

