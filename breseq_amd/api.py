"""Python mirror of breseq's pileup-scoring interface on the MI355X engine.

Function names and argument meanings follow the reference entry points:
  error_count(...)            error_count.h:41-52
  identify_mutations(...)     identify_mutations.h:46-60
  test_RA_evidence(...)       identify_mutations.h:63-67
  dp_candidate_regions(...)   identify_mutations.cpp:883-908, 1395-1457, 2358-2422, 2550-2573
  calculate_trims(...)        calculate_trims.h:34 (SequenceTrims, calculate_trims.cpp:67-138)
  fit_coverage_distribution   coverage_distribution.cpp:345-399
Errors raise BsqError with the C ABI's code instead of ending the process.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import BsqError  # noqa: F401  (re-exported)

CONSENSUS = "consensus"
POLYMORPHISM = "polymorphism"


def _status():
    return L.Status()


def _np(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


class Sample:
    """A reference plus sorted alignments (bsq_sample)."""

    def __init__(self, handle, read_file_groups):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self.read_file_groups = read_file_groups
        self._keep = []

    @classmethod
    def simulate(cls, seed=1, ref_length=48502, gc=0.5, read_length=100, coverage=100.0, paired=False,
                 insert_mean=350.0, insert_sd=50.0, qual_mean=32.0, qual_sd=6.0, qual_min=2, qual_max=41,
                 n_snps=20, n_indels=2, indel_max=5, variant_freq_min=1.0, variant_freq_max=1.0,
                 redundant_fraction=0.01, mapq=42, discordant_fraction=0.0, n_junctions=0, n_missing=0,
                 missing_length=0, missing_edge_reads=0):
        cfg = L.SimConfig(seed, ref_length, gc, read_length, coverage, 1 if paired else 0, insert_mean, insert_sd,
                          qual_mean, qual_sd, qual_min, qual_max, n_snps, n_indels, indel_max, variant_freq_min,
                          variant_freq_max, redundant_fraction, mapq, discordant_fraction, n_junctions, n_missing,
                          missing_length, missing_edge_reads)
        h = C.c_void_p()
        st = _status()
        L.check(L.lib().bsq_simulate(C.byref(cfg), C.byref(h), C.byref(st)), st)
        groups = [["reads_1", "reads_2"]] if paired else [["reads_1"]]
        return cls(h, groups)

    @classmethod
    def simulate_breseq(cls, sequence, name="ref", mutations=(), mapq=42, **options):
        """Reads as breseq's SIMULATE-READS makes them (options as bsq_simreads_args fields, the
        command's defaults otherwise) from `sequence` with the mutations (pos, ref_len, alt)
        applied, aligned to `sequence` by construction."""
        a = simreads_args(**options)
        muts = (L.SimMutation * max(1, len(mutations)))()
        keep = []
        for k, (pos, ref_len, alt) in enumerate(mutations):
            b = alt.encode()
            keep.append(b)
            muts[k] = L.SimMutation(int(pos), int(ref_len), b)
        seq = sequence.encode()
        h = C.c_void_p()
        st = _status()
        L.check(L.lib().bsq_simulate_reads_sample(name.encode(), seq, len(seq), muts, len(mutations), C.byref(a),
                                                  int(mapq), C.byref(h), C.byref(st)), st)
        groups = [["reads_1", "reads_2"]] if a.mode == 1 else [["reads_1"]]
        return cls(h, groups)

    @classmethod
    def read(cls, bam_path, fasta_path, read_file_groups=None):
        h = C.c_void_p()
        st = _status()
        L.check(L.lib().bsq_sample_read(bam_path.encode(), fasta_path.encode(), C.byref(h), C.byref(st)), st)
        s = cls(h, read_file_groups)
        if s.read_file_groups is None:
            n = L.lib().bsq_sample_n_read_groups(s._h)
            s.read_file_groups = [["rg%d" % i] for i in range(max(n, 1))]
        return s

    def write_bam(self, path):
        st = _status()
        L.check(L.lib().bsq_sample_write_bam(self._h, path.encode(), C.byref(st)), st)

    def write_fasta(self, path):
        st = _status()
        L.check(L.lib().bsq_sample_write_fasta(self._h, path.encode(), C.byref(st)), st)

    @property
    def ref(self):
        return L.lib().bsq_sample_reference(self._h)

    @property
    def n_read_groups(self):
        """@RG lines of the BAM header (0 for a simulated sample)."""
        return int(L.lib().bsq_sample_n_read_groups(self._h))

    @property
    def aln(self):
        return L.lib().bsq_sample_alignments(self._h)

    def close(self):
        if self._h:
            L.lib().bsq_sample_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # numpy views of the alignment arrays (valid while the sample lives)
    def arrays(self):
        a = self.aln.contents
        n = a.n_reads
        nb = int(_np(a.seq_off, n + 1, np.int64)[-1]) if n else 0
        nc = int(_np(a.cigar_off, n + 1, np.int64)[-1]) if n else 0
        return dict(
            tid=_np(a.tid, n, np.int32), pos=_np(a.pos, n, np.int32), flag=_np(a.flag, n, np.uint16),
            mapq=_np(a.mapq, n, np.uint8), cigar_off=_np(a.cigar_off, n + 1, np.int64),
            cigar=_np(a.cigar, nc, np.uint32), seq_off=_np(a.seq_off, n + 1, np.int64),
            seq=_np(a.seq, nb, np.uint8), qual=_np(a.qual, nb, np.uint8), x1=_np(a.x1, n, np.int32),
            xl=_np(a.xl, n, np.int32), xr=_np(a.xr, n, np.int32), rg=_np(a.rg, n, np.int32),
            tlen=_np(a.tlen, n, np.int32), xp=_np(a.xp, n, np.int8))

    def read_names(self):
        a = self.aln.contents
        n = a.n_reads
        off = _np(a.name_off, n + 1, np.int64)
        raw = C.string_at(a.names, int(off[-1])) if n else b""
        return [raw[off[i]:off[i + 1]].decode() for i in range(n)]

    def reference_sequences(self):
        r = self.ref.contents
        out = []
        for i in range(r.n_targets):
            out.append((r.names[i].decode(), C.string_at(r.seqs[i], r.lengths[i]).decode()))
        return out

    @property
    def n_targets(self):
        return self.ref.contents.n_targets

    def target_lengths(self):
        r = self.ref.contents
        return [r.lengths[i] for i in range(r.n_targets)]

    def max_quality(self):
        """Largest base quality in the sample (Summary max_qual); computed once."""
        if getattr(self, "_max_qual", None) is None:
            q = self.arrays()["qual"]
            self._max_qual = int(q.max()) if q.size else 0
        return self._max_qual

    def aligned_bases(self):
        a = self.arrays()
        ops = a["cigar"] & 0xF
        lens = a["cigar"] >> 4
        refc = np.isin(ops, [0, 2, 3, 7, 8])
        mapped = (a["flag"] & 4) == 0
        per = np.add.reduceat(np.where(refc, lens, 0).astype(np.int64), a["cigar_off"][:-1]) if len(ops) else np.zeros(0)
        n = len(a["pos"])
        has = a["cigar_off"][1:] > a["cigar_off"][:-1]
        per = np.where(has, per[: n], 0)
        return int(per[mapped].sum())

    def read_files(self):
        """bsq_read_files for this sample's read-group -> read-file partition."""
        groups = self.read_file_groups or [["reads_1"]]
        base, count, names = [], [], []
        for g in groups:
            base.append(len(names))
            count.append(len(g))
            names.extend(g)
        b = (C.c_uint32 * len(base))(*base)
        c = (C.c_uint32 * len(count))(*count)
        nm = (C.c_char_p * len(names))(*[n.encode() for n in names])
        self._keep.extend([b, c, nm])
        return L.ReadFiles(len(groups), b, c, len(names), nm)


@dataclass
class ErrorCountResult:
    histograms: list
    counts: np.ndarray
    log10_prob: np.ndarray
    read_set_max: int
    quality_max: int
    rates_table: np.ndarray = None    # the table identify_mutations reads back from error_rates.tab
    rates_text: str = ""
    texts: dict = field(default_factory=dict)
    no_pos_hash_pr: np.ndarray = None   # preprocess mode, per target


class Session:
    """A device-resident sample on one GPU (bsq_session)."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        st = _status()
        L.check(L.lib().bsq_session_create(device, C.byref(self._h), C.byref(st)), st)
        self.sample = None

    def load(self, sample):
        st = _status()
        rf = sample.read_files()
        L.check(L.lib().bsq_session_load(self._h, sample.ref, sample.aln, C.byref(rf), C.byref(st)), st)
        self.sample = sample
        return self

    def timing(self):
        t = L.Timing()
        L.lib().bsq_session_timing(self._h, C.byref(t))
        d = {k: getattr(t, k) for k, _ in L.Timing._fields_}
        d["kernel_detail_ms"] = list(t.kernel_detail_ms)
        return d

    def close(self):
        if self._h:
            L.lib().bsq_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error_count(self, **kw):
        return error_count(self.sample, session=self, **kw)

    def identify_mutations(self, **kw):
        return identify_mutations(self.sample, session=self, **kw)

    def pipeline(self, mode=None, dp=None, coverage_groups=None, targeted_sequencing=False,
                 propagation_cutoff_option=0.0, seed_cutoff_option=0.0, ec_kw=None, **im_kw):
        """error_count -> coverage-distribution stage -> identify_mutations on the loaded sample, as
        breseq runs stages 07-08 (bsq_session_pipeline).  Returns (ErrorCountResult, per-target
        coverage summaries, GD text, DP_candidate_regions.csv text)."""
        s = self.sample
        n = s.n_targets
        ec_args = make_error_count_args(s, coverage_groups=coverage_groups, **(ec_kw or {}))
        cov = coverage_stage_args(s.target_lengths(), coverage_groups, None, targeted_sequencing)
        im_args = make_identify_args(s, np.zeros(1), ec_args.read_set_max, ec_args.quality_max, [0.0] * n,
                                     mode=mode if mode is not None else CONSENSUS, dp=dp, **im_kw)
        summ = (L.CoverageSummary * n)()
        ec_res = C.c_void_p()
        im_res = C.c_void_p()
        st = _status()
        L.check(L.lib().bsq_session_pipeline(self._h, C.byref(ec_args), C.byref(cov), float(propagation_cutoff_option),
                                             float(seed_cutoff_option), C.byref(im_args), C.byref(ec_res), summ,
                                             C.byref(im_res), C.byref(st)), st)
        ec = _ec_result(ec_res, ec_args)
        gd, dp_text = _im_texts(im_res)
        return ec, [_summary_dict(x) for x in summ], gd, dp_text

    def precision_audit(self, log10_prob, read_set_max, quality_max, propagation_cutoffs, **kw):
        """FP32 column sums against breseq's FP64 ones over one identify pass
        (bsq_session_precision_audit); arguments as for identify_mutations."""
        args = make_identify_args(self.sample, log10_prob, read_set_max, quality_max, propagation_cutoffs, **kw)
        rep = L.PrecisionReport()
        st = _status()
        L.check(L.lib().bsq_session_precision_audit(self._h, C.byref(args), C.byref(rep), C.byref(st)), st)
        return {k: getattr(rep, k) for k, _ in L.PrecisionReport._fields_}

    def set_gd_writer(self, on_device=True):
        """Print later genome diffs on the GPU (the default) or assemble them on the host."""
        st = _status()
        L.check(L.lib().bsq_session_set_gd_writer(self._h, 1 if on_device else 0, C.byref(st)), st)
        return self

    def set_gd_capacity(self, entries):
        """Entries the device writer's lists hold before a pass needs more (default 4096)."""
        st = _status()
        L.check(L.lib().bsq_session_set_gd_capacity(self._h, int(entries), C.byref(st)), st)
        return self

    def set_column_launch(self, blocks=0, table_in_lds=True, waves=0):
        """Launch shape of later column passes: blocks (0: the default rule) of `waves` wavefronts
        (4 or 8; 0: 8), and the likelihood table in LDS when it fits, or in global memory."""
        st = _status()
        L.check(L.lib().bsq_session_set_column_launch(self._h, int(blocks), 1 if table_in_lds else 0, int(waves),
                                                      C.byref(st)), st)
        return self

    def tune_columns(self, log10_prob, read_set_max, quality_max, propagation_cutoffs, repeats=3, **kw):
        """Time the column pass under candidate launches on the loaded sample, set the fastest on
        the session and return the choice (bsq_session_tune_columns); identify arguments as for
        identify_mutations.  A sample of a shape tuned before in this process reuses that choice."""
        args = make_identify_args(self.sample, log10_prob, read_set_max, quality_max, propagation_cutoffs, **kw)
        out = L.ColumnLaunch()
        st = _status()
        L.check(L.lib().bsq_session_tune_columns(self._h, C.byref(args), int(repeats), C.byref(out), C.byref(st)), st)
        return {k: getattr(out, k) for k, _ in L.ColumnLaunch._fields_}

    def alignment_trims(self):
        """XL / XR of every alignment of the loaded sample, in its order (get_alignment_trims,
        calculate_trims.cpp:239-285): two int32 arrays, -1 where the record was not packed
        (unmapped, no target, no reference base covered)."""
        n = int(self.sample.aln.contents.n_reads)
        xl = np.zeros(n, dtype=np.int32)
        xr = np.zeros(n, dtype=np.int32)
        st = _status()
        L.check(L.lib().bsq_session_alignment_trims(self._h, xl.ctypes.data_as(L.c_i32p), xr.ctypes.data_as(L.c_i32p),
                                                    n, C.byref(st)), st)
        return xl, xr

    def coverage_table(self, target=0, start=0, end=0, resolution=600, total_only=False, per_read_group=False,
                       csv=False, reference_average=None, n_read_groups=None):
        """BAM2COV table of one region (coverage_output::table, coverage_output.cpp:190-283) from the
        loaded sample: start/end 1-based inclusive, both 0 for the whole target; reference_average
        adds the reference_unique_average_cov line (--show-average)."""
        args = coverage_table_args(self.sample, target, start, end, resolution, total_only, per_read_group, csv,
                                   reference_average, n_read_groups)
        out = C.c_void_p()
        n = C.c_int64()
        st = _status()
        L.check(L.lib().bsq_session_coverage_table(self._h, C.byref(args), C.byref(out), C.byref(n), C.byref(st)), st)
        try:
            return C.string_at(out, n.value).decode()
        finally:
            L.lib().bsq_free_text(out)


def simreads_args(**options):
    """bsq_simreads_args with the defaults of SIMULATE-READS; mode may be given as 'single' or
    'paired-end'."""
    a = L.SimReadsArgs()
    L.lib().bsq_simreads_defaults(C.byref(a))
    for k, v in options.items():
        if k == "mode" and isinstance(v, str):
            v = {"single": 0, "paired-end": 1, "paired": 1}[v]
        setattr(a, k, int(v))
    return a


def simulate_reads_fastq(sequence, **options):
    """FASTQ text of SIMULATE-READS (a pair of texts in paired-end mode)."""
    a = simreads_args(**options)
    p1, p2 = C.c_void_p(), C.c_void_p()
    n1, n2 = C.c_int64(), C.c_int64()
    st = _status()
    seq = sequence.encode()
    L.check(L.lib().bsq_simulate_reads(seq, len(seq), C.byref(a), C.byref(p1), C.byref(n1), C.byref(p2), C.byref(n2),
                                       C.byref(st)), st)
    try:
        t1 = C.string_at(p1, n1.value).decode()
        t2 = C.string_at(p2, n2.value).decode() if a.mode == 1 else None
    finally:
        L.lib().bsq_free_text(p1)
        if a.mode == 1:
            L.lib().bsq_free_text(p2)
    return (t1, t2) if a.mode == 1 else t1


def coverage_table_args(sample, target=0, start=0, end=0, resolution=600, total_only=False, per_read_group=False,
                        csv=False, reference_average=None, n_read_groups=None):
    """bsq_coverage_table_args for a region of a sample; the read-group count defaults to the
    sample's @RG lines."""
    n_rg = sample.n_read_groups if n_read_groups is None else n_read_groups
    return L.CoverageTableArgs(int(target), int(start), int(end), int(resolution), 1 if total_only else 0,
                               1 if per_read_group else 0, int(n_rg), 1 if csv else 0,
                               0 if reference_average is None else 1,
                               0.0 if reference_average is None else float(reference_average))


def dp_settings(sets, seed=3):
    """bsq_dp_settings from one dict per read file set: paired, median, mad, r1_prefix, r2_prefix
    (identify_mutations.cpp:893-908; seed = Settings::discordant_pair_seed)."""
    n = len(sets)
    paired = (C.c_uint8 * n)(*[1 if x.get("paired", True) else 0 for x in sets])
    med = (C.c_double * n)(*[float(x["median"]) for x in sets])
    mad = (C.c_double * n)(*[float(x["mad"]) for x in sets])
    r1 = (C.c_uint32 * n)(*[int(x.get("r1_prefix", 2 * i + 1)) for i, x in enumerate(sets)])
    r2 = (C.c_uint32 * n)(*[int(x.get("r2_prefix", 2 * i + 2)) for i, x in enumerate(sets)])
    d = L.DpSettings(n, paired, med, mad, r1, r2, seed)
    d._keep = (paired, med, mad, r1, r2)
    return d


def dp_candidate_regions(sample, dp, call_targets=None, propagation_cutoffs=None):
    """DP_candidate_regions.csv text for a sample (host accumulator, no GPU)."""
    call = _target_arrays(sample, call_targets)
    prop = None
    if propagation_cutoffs is not None:
        prop = (C.c_double * sample.n_targets)(*[float(x) for x in propagation_cutoffs])
    out = C.c_void_p()
    n = C.c_int64()
    st = _status()
    L.check(L.lib().bsq_dp_candidate_regions(sample.ref, sample.aln, call, prop, C.byref(dp) if dp is not None else None,
                                             C.byref(out), C.byref(n), C.byref(st)), st)
    try:
        return C.string_at(out, n.value).decode()
    finally:
        L.lib().bsq_free_text(out)


def _target_arrays(sample, call_targets):
    n = sample.n_targets
    call = (C.c_uint8 * n)(*[1 if (call_targets is None or i in call_targets) else 0 for i in range(n)])
    return call


def make_error_count_args(sample, read_set_max=None, quality_max=None, coverage_groups=None, call_targets=None,
                          output_dir=None, error_rates_file=None, do_coverage=True, do_errors=True,
                          preprocess_stage=False, unmatched_end_minimum_read_length=50, require_match_fraction=0.9):
    """bsq_error_count_args; the preprocess settings follow Settings (settings.cpp:1272, 1308-1309):
    unmatched_end_length_factor = 1 - require_match_fraction."""
    n = sample.n_targets
    if quality_max is None:
        quality_max = sample.max_quality() + 1
    rf = sample.read_files()
    if read_set_max is None:
        read_set_max = rf.n_files
    groups = coverage_groups if coverage_groups is not None else [0] * n
    g = (C.c_int32 * n)(*groups)
    call = _target_arrays(sample, call_targets)
    args = L.ErrorCountArgs()
    args.ref = sample.ref
    args.aln = sample.aln
    args.read_files = rf
    args.coverage_group = g
    args.n_coverage_groups = max(groups) + 1 if groups else 1
    args.call_target = call
    args.read_set_max = read_set_max
    args.quality_max = quality_max
    args.do_coverage = 1 if do_coverage else 0
    args.do_errors = 1 if do_errors else 0
    args.output_dir = output_dir.encode() if output_dir else None
    args.error_rates_file = error_rates_file.encode() if error_rates_file else None
    args.preprocess_stage = 1 if preprocess_stage else 0
    args.unmatched_end_minimum_read_length = unmatched_end_minimum_read_length
    args.unmatched_end_length_factor = 1 - require_match_fraction
    args._keep = (g, call, rf)
    return args


def error_count(sample, session=None, **kw):
    """error_count (error_count.cpp:50-68) on the GPU.  Returns ErrorCountResult."""
    args = make_error_count_args(sample, **kw)
    lib = L.lib()
    res = C.c_void_p()
    st = _status()
    if session is not None:
        L.check(lib.bsq_session_error_count(session._h, C.byref(args), C.byref(res), C.byref(st)), st)
    else:
        L.check(lib.bsq_error_count(C.byref(args), C.byref(res), C.byref(st)), st)
    return _ec_result(res, args)


def _ec_result(res, args):
    """ErrorCountResult of a bsq_ec_result handle, which it releases."""
    lib = L.lib()
    try:
        hists = []
        for gi in range(args.n_coverage_groups):
            p = L.c_u32p()
            n = C.c_int64()
            lib.bsq_ec_result_histogram(res, gi, C.byref(p), C.byref(n))
            hists.append(_np(p, n.value, np.uint32).copy())
        p64 = C.POINTER(C.c_uint64)()
        n = C.c_int64()
        lib.bsq_ec_result_counts(res, C.byref(p64), C.byref(n))
        counts = _np(p64, n.value, np.uint64).copy()
        pd = L.c_f64p()
        lib.bsq_ec_result_log10_prob(res, C.byref(pd), C.byref(n))
        log10p = _np(pd, n.value, np.float64).copy()
        lib.bsq_ec_result_rates_table(res, C.byref(pd), C.byref(n))
        rates = _np(pd, n.value, np.float64).copy()
        t = C.c_char_p()
        lib.bsq_ec_result_rates_text(res, C.byref(t), C.byref(n))
        text = C.string_at(t, n.value).decode()
        lib.bsq_ec_result_no_pos_hash_pr(res, C.byref(pd), C.byref(n))
        npr = _np(pd, n.value, np.float64).copy()
    finally:
        lib.bsq_ec_result_free(res)
    return ErrorCountResult(hists, counts, log10p, args.read_set_max, args.quality_max, rates, text,
                            no_pos_hash_pr=npr)


def mode_settings(mode):
    """(mutation_cutoff, polymorphism_cutoff, precision_decimal, precision_places) of a mode
    (settings.cpp:857-950)."""
    if mode == POLYMORPHISM:
        return 10.0, 2.0, 1e-6, 8
    return 10.0, 10.0, 1e-6, 3


def make_identify_args(sample, log10_prob, read_set_max, quality_max, propagation_cutoffs, seed_cutoffs=None,
                       mode=CONSENSUS, call_targets=None, base_quality_cutoff=3, skip_mc=False, gd_file=None,
                       dp=None, dp_file=None):
    n = sample.n_targets
    mut, poly, dec, places = mode_settings(mode)
    table = np.ascontiguousarray(log10_prob, dtype=np.float64)
    prop = (C.c_double * n)(*[float(x) for x in propagation_cutoffs])
    seed = (C.c_double * n)(*[float(x) for x in (seed_cutoffs if seed_cutoffs is not None else [0.0] * n)])
    call = _target_arrays(sample, call_targets)
    rf = sample.read_files()
    args = L.IdentifyArgs()
    args.ref = sample.ref
    args.aln = sample.aln
    args.read_files = rf
    args.call_target = call
    args.log10_prob_table = table.ctypes.data_as(L.c_f64p)
    args.n_bins = table.size
    args.read_set_max = read_set_max
    args.quality_max = quality_max
    args.deletion_propagation_cutoff = prop
    args.deletion_seed_cutoff = seed
    args.mutation_cutoff = mut
    args.polymorphism_cutoff = poly
    args.polymorphism_precision_decimal = dec
    args.polymorphism_precision_places = places
    args.base_quality_cutoff = base_quality_cutoff
    args.skip_missing_coverage_prediction = 1 if skip_mc else 0
    args.gd_file = gd_file.encode() if gd_file else None
    args.dp = C.pointer(dp) if dp is not None else None
    args.dp_file = dp_file.encode() if dp_file else None
    args._keep = (table, prop, seed, call, rf, dp)
    return args


def identify_mutations(sample, log10_prob, read_set_max, quality_max, propagation_cutoffs, session=None,
                       screen=True, with_dp=False, **kw):
    """identify_mutations (identify_mutations.cpp:48-88) on the GPU.  Returns ra_mc_evidence.gd text,
    or (gd text, DP_candidate_regions.csv text) with with_dp."""
    args = make_identify_args(sample, log10_prob, read_set_max, quality_max, propagation_cutoffs, **kw)
    lib = L.lib()
    res = C.c_void_p()
    st = _status()
    if session is not None:
        fn = lib.bsq_session_identify if screen else lib.bsq_session_identify_unscreened
        L.check(fn(session._h, C.byref(args), C.byref(res), C.byref(st)), st)
    else:
        if not screen:
            raise ValueError("screen=False needs a session")
        L.check(lib.bsq_identify_mutations(C.byref(args), C.byref(res), C.byref(st)), st)
    text, dp_text = _im_texts(res)
    return (text, dp_text) if with_dp else text


def _im_texts(res):
    """(GD text, DP CSV text) of a bsq_im_result handle, which it releases."""
    lib = L.lib()
    try:
        p = C.c_char_p()
        n = C.c_int64()
        lib.bsq_im_result_gd_text(res, C.byref(p), C.byref(n))
        text = C.string_at(p, n.value).decode()
        lib.bsq_im_result_dp_text(res, C.byref(p), C.byref(n))
        dp_text = C.string_at(p, n.value).decode()
    finally:
        lib.bsq_im_result_free(res)
    return text, dp_text


def ra_test_settings(polymorphism_prediction=False, **overrides):
    s = L.RaTestSettings()
    L.lib().bsq_ra_test_settings_default(C.byref(s), 1 if polymorphism_prediction else 0)
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


def test_RA_evidence(gd_text, sample=None, settings=None, polymorphism_prediction=False):
    """test_RA_evidence (identify_mutations.cpp:687-749) on GD text; returns the new GD text."""
    s = settings or ra_test_settings(polymorphism_prediction)
    out = C.c_void_p()
    n = C.c_int64()
    st = _status()
    ref = sample.ref if sample is not None else None
    L.check(L.lib().bsq_test_ra_evidence(gd_text.encode(), ref, C.byref(s), C.byref(out), C.byref(n), C.byref(st)), st)
    try:
        return C.string_at(out, n.value).decode()
    finally:
        L.lib().bsq_free_text(out)


def calculate_trims(seq):
    """SequenceTrims of one sequence on the GPU: (left, right) uint8 arrays."""
    b = seq.encode() if isinstance(seq, str) else bytes(seq)
    out = np.zeros(2 * len(b), dtype=np.uint8)
    st = _status()
    L.check(L.lib().bsq_sequence_trims(b, len(b), out.ctypes.data_as(L.c_u8p), C.byref(st)), st)
    return out[: len(b)], out[len(b):]


def fit_coverage_distribution(hist, sequence_length):
    """Negative-binomial fit of a unique-only coverage histogram and the deletion propagation
    cutoff (coverage_distribution.cpp:345-399)."""
    h = np.ascontiguousarray(hist, dtype=np.uint32)
    out = L.CoverageFit()
    st = _status()
    L.check(L.lib().bsq_fit_coverage_distribution(h.ctypes.data_as(L.c_u32p), h.size, float(sequence_length),
                                                  C.byref(out), C.byref(st)), st)
    return {k: getattr(out, k) for k, _ in L.CoverageFit._fields_}


def _summary_dict(x):
    return {k: getattr(x, k) for k, _ in L.CoverageSummary._fields_}


def coverage_stage_args(target_lengths, coverage_groups=None, histograms=None, targeted_sequencing=False):
    n = len(target_lengths)
    groups = list(coverage_groups) if coverage_groups is not None else [0] * n
    G = max(groups) + 1 if groups else 1
    lens = (C.c_int64 * n)(*[int(x) for x in target_lengths])
    g = (C.c_int32 * n)(*groups)
    hs = [np.ascontiguousarray(h, dtype=np.float64) for h in (histograms or [])]
    hp = (L.c_f64p * max(G, 1))(*[h.ctypes.data_as(L.c_f64p) for h in hs]) if hs else None
    hl = (C.c_int64 * max(G, 1))(*[h.size for h in hs]) if hs else None
    a = L.CoverageStageArgs()
    a.n_targets = n
    a.target_lengths = lens
    a.coverage_group = g
    a.n_coverage_groups = G
    a.histograms = hp
    a.histogram_lengths = hl
    a.targeted_sequencing = 1 if targeted_sequencing else 0
    a._keep = (lens, g, hs, hp, hl)
    return a


def analyze_unique_coverage_distributions(target_lengths, histograms, coverage_groups=None, targeted_sequencing=False):
    """CoverageDistribution::analyze_unique_coverage_distributions (coverage_distribution.cpp:510-633):
    per-target summaries from per-group histograms (n[0..max], rows j >= 1)."""
    a = coverage_stage_args(target_lengths, coverage_groups, histograms, targeted_sequencing)
    out = (L.CoverageSummary * len(target_lengths))()
    st = _status()
    L.check(L.lib().bsq_analyze_unique_coverage_distributions(C.byref(a), out, C.byref(st)), st)
    return [_summary_dict(x) for x in out]


def read_coverage_histogram(text):
    """read_coverage_histogram (coverage_distribution.cpp:34-65) of a coverage distribution file's text."""
    b = text.encode()
    n = C.c_int64()
    st = _status()
    L.check(L.lib().bsq_read_coverage_histogram(b, len(b), None, 0, C.byref(n), C.byref(st)), st)
    out = np.zeros(n.value, dtype=np.float64)
    L.check(L.lib().bsq_read_coverage_histogram(b, len(b), out.ctypes.data_as(L.c_f64p), out.size, C.byref(n),
                                                C.byref(st)), st)
    return out


def error_rates_text_to_table(text):
    """read_log10_prob_table (error_count.cpp:629-654): the values identify_mutations uses."""
    lines = text.split("\n")
    vals = []
    for ln in lines[2:]:
        if not ln:
            continue
        vals.append(float(ln.split("\t")[-1]))
    return np.array(vals, dtype=np.float64)
