"""Writes the oracle's results on the full-size workloads as fixtures, tests/golden/oracle_<NAME>.json.gz.

The oracle (oracle/liboracle.so, the CPU restatement of breseq's path) runs error_count, the
coverage-distribution fit and identify_mutations on each workload of breseq_amd.workloads, exactly
as breseq chains them: the error table identify reads is the error_rates.tab text read back, and
the deletion propagation cutoff is the fit's.  The GPU tests regenerate the same seeded sample on
the device box and compare the product's pipeline with these results.  The larger workloads take
minutes of single-threaded oracle time, which is why they are fixtures rather than live runs.

    python tests/golden/make_oracle_fixtures.py C2 C3
"""
import gzip
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import breseq_amd as B  # noqa: E402
import oracle_lib as O  # noqa: E402
from breseq_amd import api, workloads  # noqa: E402


def oracle_run(cfg, mode, log=print):
    """The oracle's stages 07-08 on the sample of generator settings cfg."""
    t0 = time.time()
    s = B.Sample.simulate(**cfg)
    log("  simulated: %d reads, %d aligned bases (%.1f s)" % (s.aln.contents.n_reads, s.aligned_bases(), time.time() - t0))
    t0 = time.time()
    args = api.make_error_count_args(s)
    hists, counts, log10p, texts = O.error_count(args)
    log("  error_count %.1f s" % (time.time() - t0))
    length = sum(s.target_lengths())
    fit = O.fit_coverage_distribution(hists[0], length)
    table = O.read_log10_prob_table(texts["error_rates"], len(log10p))
    dp = api.dp_settings([workloads.DP_SET], seed=workloads.DP_SEED) if cfg.get("paired") else None
    m = B.POLYMORPHISM if mode == "polymorphism" else B.CONSENSUS
    t0 = time.time()
    iargs = api.make_identify_args(s, table, args.read_set_max, args.quality_max,
                                   [fit["deletion_coverage_propagation_cutoff"]] * s.n_targets, mode=m, dp=dp)
    gd, dp_text = O.identify_with_dp(iargs)
    log("  identify_mutations %.1f s" % (time.time() - t0))
    return {
        "config": cfg,
        "mode": mode,
        "aligned_bases": s.aligned_bases(),
        "n_reads": s.aln.contents.n_reads,
        "histograms": [h.tolist() for h in hists],
        "counts": counts.tolist(),
        "error_rates": texts["error_rates"],
        "base_qual": {k: v for k, v in texts.items() if k.startswith("base_qual_")},
        "coverage_text": {k: v for k, v in texts.items() if k.startswith("coverage_")},
        "fit": fit,
        "gd": gd,
        "dp": dp_text,
    }


def path_of(name):
    return os.path.join(HERE, "oracle_%s.json.gz" % name)


def main(names):
    for name in names:
        print("[fixture] %s" % name, flush=True)
        res = oracle_run(getattr(workloads, name), workloads.MODES[name], log=lambda m: print(m, flush=True))
        with open(path_of(name), "wb") as raw, gzip.GzipFile(fileobj=raw, mode="wb", mtime=0, compresslevel=9) as f:
            f.write(json.dumps(res, sort_keys=True).encode())
        print("  wrote %s" % path_of(name), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["C2", "C3"])
