#!/usr/bin/env python
"""Benchmark of the MI355X pileup-scoring path (breseq error_count + coverage distribution +
identify_mutations).

One step = one pass of the whole path over one sample already resident in HBM, as breseq runs its
stages 07-08 (bsq_session_pipeline):
  error_count (coverage histogram + covariate counts) -> error_rates.tab round trip ->
  coverage-distribution fit (deletion propagation cutoff; on host threads beside the next step) ->
  identify_mutations (RA / MC / UN scoring) -> ra_mc_evidence.gd text, with the discordant-pair
  accumulator -> DP_candidate_regions.csv text (breseq runs it on every paired sample).
Metric: aligned bases/s (pileup entries: M/=/X bases + D positions), plus p50/p99 step latency.
Workload: 4.6 Mb reference (GC 50.8%), 100x of 2x150 bp paired reads, insert N(350,50),
qualities N(34,5), 500 SNPs + 50 indels of 1-5 bp, 0.2% discordant pairs at 40 junctions, three
400-bp deletions, consensus mode.  Synthetic, seeded (breseq_amd.workloads.C2); its results are
checked against the oracle's by the GPU tests.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0

from breseq_amd import workloads  # noqa: E402

C2 = workloads.C2
DP_SET = workloads.DP_SET

# Per-unit algorithmic bytes of the column kernels: 1 B base + 1 B quality per aligned base,
# 24 B of per-read metadata (position, CIGAR, flags, tags), and per reference column 1 B of
# reference plus the column kernel's output (identify: 1 B flags + 4 B unique + 4 B total).
BYTES_PER_BASE = 2
BYTES_PER_READ = 24
BYTES_PER_COLUMN_EC = 1
BYTES_PER_COLUMN_IM = 1 + 9


COUNTERS = os.path.join(ROOT, "profiles", "counters_4.6Mb_100x.json")


def traffic_of(kernel, cfg):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter runs (FETCH_SIZE and
    WRITE_SIZE in KiB per dispatch, each scaled by what the counter reads for a copy of known size),
    when they were taken on this workload; None otherwise."""
    try:
        with open(COUNTERS) as f:
            c = json.load(f)
    except (OSError, ValueError):
        return None
    if c.get("ref_length") != cfg["ref_length"] or c.get("coverage") != cfg["coverage"]:
        return None
    k = c.get("kernels", {}).get(kernel)
    if not k:
        return None
    return (k["fetch_kib"] * c["fetch_scale"] + k["write_kib"] * c["write_scale"]) * 1024.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return 0.0
    k = (len(xs) - 1) * q / 100.0
    f = int(k)
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


def run_step(sess, mode, dp):
    return sess.pipeline(mode=mode, dp=dp)


def batched_block(n_samples, parallel=8, rate=4.0, max_length=2_000_000, seed=7):
    """The batched and served workloads at a reduced size: n_samples mixed samples (reference
    lengths log-uniform in [50 kb, max_length], coverage U(20, 300), reads {36, 100, 150}, consensus
    and polymorphism alternating), run one at a time, `parallel` at a time (results checked
    identical), and served with Poisson arrivals at `rate` per second (latency from submission to
    finished GD).  Every run includes the session's load of the in-memory alignments."""
    import numpy as np
    from breseq_amd import batch
    t = time.time()
    jobs = batch.mixed_jobs(n_samples, seed, max_length=max_length)
    log("[bench]   %d samples simulated in %.1f s" % (n_samples, time.time() - t))
    nab = sum(s_.aligned_bases() for s_, _, _ in jobs)
    batch.run_pipeline(*jobs[0])
    seq, wall_seq = batch.run_batch(jobs, parallel=1)
    bat, wall_bat = batch.run_batch(jobs, parallel=parallel)
    same = all(x["gd"] == y["gd"] and x["dp"] == y["dp"] for x, y in zip(seq, bat))
    rng = np.random.default_rng(seed + 1)
    srv = batch.Server(parallel=parallel)
    t0 = time.perf_counter()
    for i, (s_, m, d) in enumerate(jobs):
        srv.submit(i, s_, m, d)
        time.sleep(float(rng.exponential(1.0 / rate)))
    results, errors = srv.close()
    wall_srv = time.perf_counter() - t0
    lat = [r["latency_s"] * 1e3 for r in results.values()]
    return {
        "workload": "%d samples (the batched workload's shape at a reduced size: reference lengths "
                    "log-uniform 50 kb-%d bp, coverage U(20,300), reads {36,100,150}, 150 bp paired, "
                    "consensus/polymorphism alternating)" % (n_samples, max_length),
        "aligned_bases": nab,
        "one_at_a_time_aligned_bases_per_s": nab / wall_seq,
        "batched": {"in_flight": parallel, "aligned_bases_per_s": nab / wall_bat, "speedup": wall_seq / wall_bat,
                    "results_identical": same},
        "served": {"arrivals_per_s": rate, "in_flight": parallel, "errors": len(errors),
                   "p50_ms": percentile(lat, 50) if lat else None, "p99_ms": percentile(lat, 99) if lat else None,
                   "offered_aligned_bases_per_s": nab / max(1e-9, wall_srv)},
    }


def end_to_end_block(B, cfg, dp, ref_length=1_000_000, runs=2):
    """The path from files: a sample of the bench's shape at ref_length bp written as BAM + FASTA
    (not timed), then read back (BGZF inflate and record decode on the host cores), loaded into a
    new session and run through the pipeline to the GD text.  Best of `runs` (the first also warms
    the file cache)."""
    import shutil
    import tempfile
    c = dict(cfg)
    c["ref_length"] = ref_length
    sample = B.Sample.simulate(**c)
    d = tempfile.mkdtemp(prefix="bsq_e2e_")
    try:
        bam, fa = os.path.join(d, "s.bam"), os.path.join(d, "s.fasta")
        t = time.time()
        sample.write_bam(bam)
        sample.write_fasta(fa)
        log("[bench]   BAM written in %.1f s (%.0f MB)" % (time.time() - t, os.path.getsize(bam) / 1e6))
        best = None
        for _ in range(runs):
            t0 = time.perf_counter()
            s = B.Sample.read(bam, fa)
            t1 = time.perf_counter()
            sess = B.Session().load(s)
            t2 = time.perf_counter()
            try:
                sess.pipeline(mode=B.CONSENSUS, dp=dp)
            finally:
                t3 = time.perf_counter()
                sess.close()
            r = dict(read_ms=(t1 - t0) * 1e3, load_ms=(t2 - t1) * 1e3, pipeline_ms=(t3 - t2) * 1e3, total_ms=(t3 - t0) * 1e3)
            if best is None or r["total_ms"] < best["total_ms"]:
                best = r
        nab = sample.aligned_bases()
        best.update(workload="%d bp reference at %.0fx, 2x150 bp paired, as BAM (%.0f MB) + FASTA files; a new "
                             "session per run" % (ref_length, c["coverage"], os.path.getsize(bam) / 1e6),
                    aligned_bases=nab, aligned_bases_per_s=nab / (best["total_ms"] / 1e3), runs=runs)
        return best
    finally:
        shutil.rmtree(d, ignore_errors=True)


def cpu_baseline(B, cfg, target_seconds):
    """The oracle (single-threaded C++ restatement of the reference path) on a bounded sample of
    the same workload: same generator settings on a shorter reference."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from breseq_amd import api
    # the oracle scores about 6.6e6 aligned bases (66 kb of reference at 100x) per second on one core
    length = int(min(max(60_000 * target_seconds, 100_000), cfg["ref_length"]))
    small = dict(cfg)
    small["ref_length"] = length
    small["n_snps"] = max(1, int(cfg["n_snps"] * length / cfg["ref_length"]))
    small["n_indels"] = max(1, int(cfg["n_indels"] * length / cfg["ref_length"]))
    small["n_missing"] = 1
    s = B.Sample.simulate(**small)
    nab = s.aligned_bases()
    t0 = time.perf_counter()
    args = api.make_error_count_args(s)
    hists, counts, log10p, texts = O.error_count(args)
    fit = O.fit_coverage_distribution(hists[0], sum(s.target_lengths()))
    table = O.read_log10_prob_table(texts["error_rates"], len(log10p))
    iargs = api.make_identify_args(s, table, args.read_set_max, args.quality_max,
                                   [fit["deletion_coverage_propagation_cutoff"]] * s.n_targets,
                                   dp=api.dp_settings([DP_SET], seed=workloads.DP_SEED))
    O.identify_with_dp(iargs)
    dt = time.perf_counter() - t0
    return {"value": nab / dt, "unit": "aligned bases/s", "cores": 1, "kind": "port",
            "sample": "%d bp reference, %.0fx 2x%d bp paired reads, same generator and mode (%d aligned bases, "
                      "%.1f s single-threaded)" % (length, cfg["coverage"], cfg["read_length"], nab, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ref-length", type=int, default=C2["ref_length"])
    ap.add_argument("--coverage", type=float, default=C2["coverage"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-audit", action="store_true", help="skip the FP32-vs-FP64 precision report")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--latency-runs", type=int, default=5,
                    help="submit-to-GD runs of the sample (new session, load, error_count, fit, identify)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched / served block")
    ap.add_argument("--no-end-to-end", action="store_true", help="skip the BAM-to-GD block")
    ap.add_argument("--batched-samples", type=int, default=16)
    ap.add_argument("--no-tune", action="store_true", help="keep the column pass's default launch")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the per-sample work "
              "shards by genome window; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)
    import breseq_amd as B

    cfg = dict(C2)
    cfg["ref_length"] = args.ref_length
    cfg["coverage"] = args.coverage
    log("[bench] simulating %d bp at %.0fx" % (cfg["ref_length"], cfg["coverage"]))
    t = time.time()
    sample = B.Sample.simulate(**cfg)
    log("[bench] simulated in %.1f s" % (time.time() - t))
    sess = B.Session().load(sample)
    dp = B.api.dp_settings([DP_SET], seed=workloads.DP_SEED)
    tm = sess.timing()
    nab, n_reads, n_cols = tm["aligned_bases"], tm["n_reads"], tm["n_positions"]
    log("[bench] loaded: %d aligned bases, %d reads, %d columns; pack %.0f ms, H2D %.0f ms"
        % (nab, n_reads, n_cols, tm["pack_ms"], tm["h2d_ms"]))
    warm = None
    for i in range(args.warmup):
        warm = run_step(sess, B.CONSENSUS, dp)
        log("[bench] warmup step %d done" % i)
    # the column pass's launch, chosen by the library's autotuner on this sample (untimed)
    column_launch = None
    if not args.no_tune:
        if warm is None:
            warm = run_step(sess, B.CONSENSUS, dp)
        t = time.perf_counter()
        column_launch = sess.tune_columns(warm[0].rates_table, warm[0].read_set_max, warm[0].quality_max,
                                          [x["deletion_coverage_propagation_cutoff"] for x in warm[1]],
                                          repeats=3, mode=B.CONSENSUS, dp=dp)
        column_launch["tune_ms"] = (time.perf_counter() - t) * 1000.0
        log("[bench] column launch: %d blocks of %d wavefronts, table in %s, %.3f ms (default %.3f ms), %d launches "
            "timed in %.0f ms"
            % (column_launch["grid"], column_launch["waves"], "LDS" if column_launch["table_in_lds"] else "global memory",
               column_launch["ms"],
               column_launch["default_ms"], column_launch["n_tried"], column_launch["tune_ms"]))
        run_step(sess, B.CONSENSUS, dp)
    step_ms = []
    ec_k, im_k = [], []
    host = {"ec_call": [], "im_call": [], "im_host_pre": [], "im_d2h": [], "im_host_post": [], "fit": [],
            "fit_wait": []}
    detail = None
    for i in range(args.steps):
        t0 = time.perf_counter()
        ec, summaries, gd, dp_text = run_step(sess, B.CONSENSUS, dp)
        t1 = time.perf_counter()
        imt = sess.timing()
        step_ms.append((t1 - t0) * 1000.0)
        ec_k.append(imt["pipeline_ec_kernel_ms"])
        im_k.append(imt["kernel_detail_ms"])
        host["ec_call"].append(imt["pipeline_ec_ms"])
        host["im_call"].append((t1 - t0) * 1000.0 - imt["pipeline_ec_ms"])
        host["im_host_pre"].append(imt["host_pre_ms"])
        host["im_d2h"].append(imt["d2h_ms"])
        host["im_host_post"].append(imt["host_post_ms"])
        host["fit"].append(imt["fit_ms"])
        host["fit_wait"].append(imt["fit_wait_ms"])
        detail = imt
    total_s = sum(step_ms) / 1000.0
    ms_per_step = sum(step_ms) / len(step_ms)
    value = nab * len(step_ms) / total_s
    n_ra = gd.count("\nRA\t")
    # dominant kernel and its roofline
    avg = lambda xs: sum(xs) / len(xs)
    k_ec = avg([d[0] for d in ec_k])
    k_cov = avg([d[1] for d in ec_k])
    k_col = avg([d[1] for d in im_k])
    k_ins = avg([d[2] for d in im_k])
    k_score = avg([d[3] for d in im_k])
    k_runs = avg([d[4] for d in im_k])
    kernels = {"k_ec_reads": k_ec, "k_cov_scan+hist": k_cov, "k_identify_columns": k_col,
               "k_insert_slots": k_ins, "k_score_candidates": k_score, "k_runs": k_runs}
    bytes_ec = BYTES_PER_BASE * nab + BYTES_PER_READ * n_reads + BYTES_PER_COLUMN_EC * n_cols
    bytes_im = BYTES_PER_BASE * nab + BYTES_PER_READ * n_reads + BYTES_PER_COLUMN_IM * n_cols
    col_kernels = {"k_ec_reads": (k_ec, bytes_ec), "k_identify_columns": (k_col, bytes_im)}
    dom = max(col_kernels, key=lambda k: col_kernels[k][0])
    dms, dbytes = col_kernels[dom]
    achieved = dbytes / (dms / 1000.0) / 1e9
    host_ms = {k: sum(v) / len(v) for k, v in host.items()}
    host_ms["im_runs_host_beside_scorer"] = avg([d[5] for d in im_k])
    host_ms["im_dp_regions_thread"] = avg([d[0] for d in im_k])
    host_ms["fit_beside_identify"] = host_ms.pop("fit")
    host_ms["identify_waited_for_fit"] = host_ms.pop("fit_wait")
    # submit-to-GD latency of the sample as a server runs it: the load of its in-memory alignments
    # (host packing and upload) into a session kept from the previous request, error_count, the
    # coverage fit, identify and the GD text; and once into a new session (cold)
    lat_ms = []
    t0 = time.perf_counter()
    s2 = B.Session()
    if column_launch is not None:
        # a server reuses the launch tuned for a sample of this shape
        s2.set_column_launch(column_launch["blocks"], column_launch["table_in_lds"], column_launch["waves"])
    s2.load(sample)
    run_step(s2, B.CONSENSUS, dp)
    cold_ms = (time.perf_counter() - t0) * 1000.0
    for i in range(args.latency_runs):
        t0 = time.perf_counter()
        s2.load(sample)
        run_step(s2, B.CONSENSUS, dp)
        lat_ms.append((time.perf_counter() - t0) * 1000.0)
    s2.close()
    log("[bench] sample latency (ms): cold %.1f, then %s" % (cold_ms, ", ".join("%.1f" % x for x in lat_ms)))
    out = {
        "metric": "aligned bases/sec through pileup scoring, 4.6Mb@100x; p50/p99 sample latency",
        "value": value,
        "unit": "aligned bases/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "p50_ms": percentile(lat_ms, 50) if lat_ms else None,
        "p99_ms": percentile(lat_ms, 99) if lat_ms else None,
        "latency": "p50_ms / p99_ms: submit-to-GD per sample as the server runs it (load of the in-memory "
                   "alignments into a session kept from the previous request, error_count, coverage fit, "
                   "identify, GD text; BAM decode not included), over %d runs; cold_sample_ms: the same "
                   "into a new session; step_p50_ms / step_p99_ms: one step on the resident sample" % len(lat_ms),
        "cold_sample_ms": cold_ms,
        "step_p50_ms": percentile(step_ms, 50),
        "step_p99_ms": percentile(step_ms, 99),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 bases/quals, int32 counts, f64 likelihoods",
        "data": "synthetic (seeded generator: random reference, planted variants, reads aligned by construction)",
        "config": {"workload": "consensus-mode pileup scoring of one sample (error_count, coverage fit, "
                               "identify_mutations): %d bp reference at %.0fx, 2x150 bp paired reads, insert "
                               "N(350,50), quals N(34,5), 500 SNPs + 50 indels, %.1f%% discordant pairs, %d "
                               "deletions" % (cfg["ref_length"], cfg["coverage"], 100.0 * cfg["discordant_fraction"],
                                              cfg["n_missing"]),
                   "aligned_bases": nab, "reads": n_reads, "ra_entries": n_ra,
                   "mc_entries": gd.count("\nMC\t"), "un_entries": gd.count("\nUN\t"),
                   "deletion_propagation_cutoff": summaries[0]["deletion_coverage_propagation_cutoff"],
                   "dp_regions": dp_text.count("\n") - 1,
                   "candidates": detail["candidates"]},
        "kernels_ms": kernels,
        "host_ms": host_ms,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic_of(dom, cfg),
                     "algorithmic_bytes": dbytes},
    }
    # per-position error of the FP32 column sums against the FP64 sums breseq forms, over every
    # column with reads (outside the timed steps)
    out["column_launch"] = column_launch
    if args.no_audit:
        out["fp32_vs_fp64"] = None
    else:
        log("[bench] precision audit")
        out["fp32_vs_fp64"] = sess.precision_audit(ec.rates_table, ec.read_set_max, ec.quality_max,
                                                   [x["deletion_coverage_propagation_cutoff"] for x in summaries],
                                                   mode=B.CONSENSUS, dp=dp)
    if args.no_batched:
        out["batched"] = None
    else:
        log("[bench] batched and served samples")
        out["batched"] = batched_block(args.batched_samples)
    if args.no_end_to_end:
        out["end_to_end"] = None
    else:
        log("[bench] BAM to GD")
        out["end_to_end"] = end_to_end_block(B, cfg, dp)
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle, 1 core)")
        out["cpu_baseline"] = cpu_baseline(B, cfg, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
