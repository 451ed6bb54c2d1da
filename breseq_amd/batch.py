"""Batched and served runs of the pileup path.

One sample's path (error_count, the coverage-distribution fit, identify_mutations; the order of
breseq_cmdline.cpp:2293-2471) leaves the GPU partly idle: the exact scorer is a latency chain over
few candidates, and the host formats results between the passes.  Several samples in flight at once
fill those gaps.  Each worker thread drives its own device-resident session, with its own HIP
streams, and loads the samples it runs into it one after another; the device runs the kernels of
different samples concurrently.

* ``run_batch`` runs a list of samples with at most ``parallel`` of them on the device at a time.
* ``Server`` admits requests from a queue while their estimated device memory fits a budget, runs
  up to ``parallel`` at once, and records each request's latency from submission to finished GD.

Every sample's results are those of the same sample run alone: a session reused for the next sample
keeps only its streams and device buffers, and the load replaces everything in them.
"""
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor

from . import api
from .api import CONSENSUS, Session


def device_bytes(sample):
    """Estimated device memory of a loaded sample and one identify pass: bases and qualities (padded
    per read), the per-read records, and the per-column arrays."""
    a = sample.arrays()
    n_reads = int(a["tid"].size)
    n_bases = int(a["seq"].size)
    n_pos = int(sum(sample.target_lengths())) + sample.n_targets
    return 2 * (n_bases + 8 * n_reads) + 96 * n_reads + 32 * n_pos + (64 << 20)


def run_pipeline(sample, mode=CONSENSUS, dp=None, session=None, **identify_kw):
    """One sample through the path.  The deletion propagation cutoff of each coverage group comes
    from the negative-binomial fit of its unique-only coverage histogram, over the group's total
    length (coverage_distribution.cpp:510-546).  Returns a dict with the GD text, the DP candidate
    regions (or None), the per-stage seconds, and the aligned bases."""
    t0 = time.perf_counter()
    own = session is None
    sess = session if session is not None else Session().load(sample)
    t1 = time.perf_counter()
    try:
        ec = sess.error_count()
        t2 = time.perf_counter()
        total_len = float(sum(sample.target_lengths()))
        fit = api.fit_coverage_distribution(ec.histograms[0], total_len)
        prop = [fit["deletion_coverage_propagation_cutoff"]] * sample.n_targets
        t3 = time.perf_counter()
        out = sess.identify_mutations(log10_prob=ec.rates_table, read_set_max=ec.read_set_max,
                                      quality_max=ec.quality_max, propagation_cutoffs=prop, mode=mode, dp=dp,
                                      with_dp=dp is not None, **identify_kw)
        t4 = time.perf_counter()
    finally:
        if own:
            sess.close()
    gd, dp_text = out if dp is not None else (out, None)
    return dict(gd=gd, dp=dp_text, load_s=t1 - t0, error_count_s=t2 - t1, fit_s=t3 - t2, identify_s=t4 - t3,
                total_s=time.perf_counter() - t0, aligned_bases=sample.aligned_bases())


class SessionPool:
    """One device session per worker thread, kept across the samples that thread runs: loading the
    next sample reuses the session's streams and every device buffer that is already large enough,
    so a served sample pays for its own data, not for device allocations."""

    def __init__(self):
        self._local = threading.local()
        self._all = []
        self._lock = threading.Lock()

    def get(self):
        sess = getattr(self._local, "session", None)
        if sess is None:
            sess = Session()
            self._local.session = sess
            with self._lock:
                self._all.append(sess)
        return sess

    def close(self):
        with self._lock:
            for sess in self._all:
                sess.close()
            self._all = []


def run_pooled(pool, sample, mode=CONSENSUS, dp=None, **identify_kw):
    """run_pipeline on the calling thread's session of `pool`, the sample loaded into it first."""
    t0 = time.perf_counter()
    sess = pool.get().load(sample)
    t1 = time.perf_counter()
    r = run_pipeline(sample, mode, dp, session=sess, **identify_kw)
    r["load_s"] = t1 - t0
    r["total_s"] = time.perf_counter() - t0
    return r


def run_batch(jobs, parallel=4, pool=None):
    """Runs jobs = [(sample, mode, dp), ...] with up to ``parallel`` in flight, each worker thread on
    its own reused session (a new one per sample when ``pool`` is False).  Returns (results in job
    order, wall seconds)."""
    own = pool is None
    pool = SessionPool() if own else pool
    t0 = time.perf_counter()
    try:
        with ThreadPoolExecutor(max_workers=max(1, parallel)) as ex:
            if pool is False:
                futs = [ex.submit(run_pipeline, s, m, d) for s, m, d in jobs]
            else:
                futs = [ex.submit(run_pooled, pool, s, m, d) for s, m, d in jobs]
            results = [f.result() for f in futs]
        return results, time.perf_counter() - t0
    finally:
        if own:
            pool.close()


class Server:
    """A request loop: ``submit`` queues a sample; a dispatcher admits queued samples in arrival
    order while the sum of their estimated device memory stays within ``hbm_budget`` bytes and fewer
    than ``parallel`` are running, and worker threads run them.  ``close`` waits for every admitted
    request.  Each finished request records its latency from submission to finished GD."""

    def __init__(self, parallel=4, hbm_budget=200 << 30):
        self.parallel = max(1, parallel)
        self.budget = int(hbm_budget)
        self.in_use = 0
        self.running = 0
        self.pending = queue.Queue()
        self.cv = threading.Condition()
        self.results = {}
        self.errors = {}
        self.pool = ThreadPoolExecutor(max_workers=self.parallel)
        self.sessions = SessionPool()
        self.closed = False
        self.dispatcher = threading.Thread(target=self._dispatch, daemon=True)
        self.dispatcher.start()

    def submit(self, key, sample, mode=CONSENSUS, dp=None):
        need = device_bytes(sample)
        if need > self.budget:
            raise ValueError("request %r needs %d bytes of device memory, above the budget" % (key, need))
        self.pending.put((key, sample, mode, dp, need, time.perf_counter()))

    def _dispatch(self):
        while True:
            item = self.pending.get()
            if item is None:
                return
            key, sample, mode, dp, need, t_submit = item
            with self.cv:
                self.cv.wait_for(lambda: self.running < self.parallel and self.in_use + need <= self.budget)
                self.running += 1
                self.in_use += need
            self.pool.submit(self._run, key, sample, mode, dp, need, t_submit)

    def _run(self, key, sample, mode, dp, need, t_submit):
        try:
            r = run_pooled(self.sessions, sample, mode, dp)
            r["latency_s"] = time.perf_counter() - t_submit
            with self.cv:
                self.results[key] = r
        except Exception as e:   # reported per request; the server keeps running
            with self.cv:
                self.errors[key] = e
        finally:
            with self.cv:
                self.running -= 1
                self.in_use -= need
                self.cv.notify_all()

    def close(self):
        """Stops admitting, waits for every queued and running request, returns (results, errors)."""
        if not self.closed:
            self.closed = True
            self.pending.put(None)
            self.dispatcher.join()
            self.pool.shutdown(wait=True)
            self.sessions.close()
        return self.results, self.errors


def mixed_jobs(n_samples, seed=7, min_length=50_000, max_length=5_000_000):
    """Samples shaped like the batched workload: reference lengths log-uniform in [min_length,
    max_length], coverage uniform in [20, 300], read lengths from {36, 100, 150} (150 paired, with
    discordant-pair regions), consensus and polymorphism modes alternating."""
    import numpy as np
    rng = np.random.default_rng(seed)
    dp = api.dp_settings([dict(median=350.0, mad=50.0 * 0.6745, r1_prefix=1, r2_prefix=2)], seed=3)
    out = []
    for i in range(n_samples):
        length = int(np.exp(rng.uniform(np.log(min_length), np.log(max_length))))
        cov = float(rng.uniform(20.0, 300.0))
        rl = int(rng.choice([36, 100, 150]))
        paired = rl == 150
        poly = i % 2 == 1
        n_var = max(2, length // 10_000)
        s = api.Sample.simulate(seed=1000 + i, ref_length=length, read_length=rl, coverage=cov, paired=paired,
                                n_snps=n_var, n_indels=max(1, n_var // 10),
                                variant_freq_min=0.05 if poly else 1.0, variant_freq_max=0.5 if poly else 1.0,
                                discordant_fraction=0.002 if paired else 0.0, n_junctions=4 if paired else 0)
        out.append((s, api.POLYMORPHISM if poly else CONSENSUS, dp if paired else None))
    return out


def main(argv=None):
    """Throughput of a mixed batch run one sample at a time and ``parallel`` at a time (results
    checked identical), then a served run with Poisson arrivals: one JSON line."""
    import argparse
    import json
    import numpy as np
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=16)
    ap.add_argument("--max-length", type=int, default=2_000_000)
    ap.add_argument("--parallel", type=int, default=4)
    ap.add_argument("--rate", type=float, default=4.0, help="served arrivals per second (Poisson)")
    ap.add_argument("--seed", type=int, default=7)
    a = ap.parse_args(argv)
    jobs = mixed_jobs(a.samples, a.seed, max_length=a.max_length)
    nab = sum(s.aligned_bases() for s, _, _ in jobs)
    run_pipeline(*jobs[0])   # warm-up: library load, first kernels
    seq, wall_seq = run_batch(jobs, parallel=1)
    bat, wall_bat = run_batch(jobs, parallel=a.parallel)
    same = all(x["gd"] == y["gd"] and x["dp"] == y["dp"] for x, y in zip(seq, bat))
    rng = np.random.default_rng(a.seed + 1)
    srv = Server(parallel=a.parallel)
    t0 = time.perf_counter()
    for i, (s, m, d) in enumerate(jobs):
        srv.submit(i, s, m, d)
        time.sleep(float(rng.exponential(1.0 / a.rate)))
    results, errors = srv.close()
    wall_srv = time.perf_counter() - t0
    lat = sorted(r["latency_s"] for r in results.values())
    pct = lambda q: float(np.percentile(lat, q)) * 1e3 if lat else None
    print(json.dumps({
        "workload": "%d samples: lengths log-uniform 50 kb-%d bp, coverage U(20,300), reads {36,100,150} "
                    "(150 paired), consensus/polymorphism alternating" % (a.samples, a.max_length),
        "aligned_bases": nab,
        "one_at_a_time": {"wall_s": wall_seq, "aligned_bases_per_s": nab / wall_seq},
        "batched": {"parallel": a.parallel, "wall_s": wall_bat, "aligned_bases_per_s": nab / wall_bat,
                    "speedup": wall_seq / wall_bat, "results_identical": same},
        "served": {"rate_per_s": a.rate, "parallel": a.parallel, "wall_s": wall_srv, "errors": len(errors),
                   "p50_ms": pct(50), "p99_ms": pct(99)},
        "per_sample_ms_one_at_a_time": [round(r["total_s"] * 1e3, 2) for r in seq],
    }), flush=True)


if __name__ == "__main__":
    main()
