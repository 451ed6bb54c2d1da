"""ctypes declarations of the C ABI in include/bsq.h (libbsq.so, built in-tree).

Loading fails loudly: there is no CPU fallback for any entry point of this package.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbsq.so")

c_i32p = C.POINTER(C.c_int32)
c_u32p = C.POINTER(C.c_uint32)
c_i64p = C.POINTER(C.c_int64)
c_u8p = C.POINTER(C.c_uint8)
c_i8p = C.POINTER(C.c_int8)
c_u16p = C.POINTER(C.c_uint16)
c_f64p = C.POINTER(C.c_double)
c_charpp = C.POINTER(C.c_char_p)

ERR_NAMES = {1: "ARG", 2: "IO", 3: "FORMAT", 4: "DEVICE", 5: "LIMIT", 6: "DATA"}


class Status(C.Structure):
    _fields_ = [("code", C.c_int32), ("message", C.c_char * 480)]


class Reference(C.Structure):
    _fields_ = [("n_targets", C.c_int32), ("names", c_charpp), ("lengths", c_i64p), ("seqs", c_charpp)]


class Alignments(C.Structure):
    _fields_ = [
        ("n_reads", C.c_int64),
        ("tid", c_i32p),
        ("pos", c_i32p),
        ("flag", c_u16p),
        ("mapq", c_u8p),
        ("cigar_off", c_i64p),
        ("cigar", c_u32p),
        ("seq_off", c_i64p),
        ("seq", c_u8p),
        ("qual", c_u8p),
        ("x1", c_i32p),
        ("xl", c_i32p),
        ("xr", c_i32p),
        ("rg", c_i32p),
        ("tlen", c_i32p),
        ("xp", c_i8p),
        ("name_off", c_i64p),
        ("names", C.c_char_p),
    ]


class ReadFiles(C.Structure):
    _fields_ = [
        ("n_groups", C.c_int32),
        ("base", c_u32p),
        ("count", c_u32p),
        ("n_files", C.c_int32),
        ("names", c_charpp),
    ]


class ErrorCountArgs(C.Structure):
    _fields_ = [
        ("ref", C.POINTER(Reference)),
        ("aln", C.POINTER(Alignments)),
        ("read_files", ReadFiles),
        ("coverage_group", c_i32p),
        ("n_coverage_groups", C.c_int32),
        ("call_target", c_u8p),
        ("read_set_max", C.c_int32),
        ("quality_max", C.c_int32),
        ("do_coverage", C.c_int32),
        ("do_errors", C.c_int32),
        ("output_dir", C.c_char_p),
        ("error_rates_file", C.c_char_p),
        ("preprocess_stage", C.c_int32),
        ("unmatched_end_minimum_read_length", C.c_uint32),
        ("unmatched_end_length_factor", C.c_double),
    ]


class DpSettings(C.Structure):
    _fields_ = [
        ("n_sets", C.c_int32),
        ("paired", c_u8p),
        ("median", c_f64p),
        ("mad", c_f64p),
        ("r1_prefix", c_u32p),
        ("r2_prefix", c_u32p),
        ("seed", C.c_int32),
    ]


class IdentifyArgs(C.Structure):
    _fields_ = [
        ("ref", C.POINTER(Reference)),
        ("aln", C.POINTER(Alignments)),
        ("read_files", ReadFiles),
        ("call_target", c_u8p),
        ("log10_prob_table", c_f64p),
        ("n_bins", C.c_int64),
        ("read_set_max", C.c_int32),
        ("quality_max", C.c_int32),
        ("deletion_propagation_cutoff", c_f64p),
        ("deletion_seed_cutoff", c_f64p),
        ("mutation_cutoff", C.c_double),
        ("polymorphism_cutoff", C.c_double),
        ("polymorphism_precision_decimal", C.c_double),
        ("polymorphism_precision_places", C.c_uint32),
        ("base_quality_cutoff", C.c_int32),
        ("skip_missing_coverage_prediction", C.c_int32),
        ("gd_file", C.c_char_p),
        ("dp", C.POINTER(DpSettings)),
        ("dp_file", C.c_char_p),
    ]


class RaTestSettings(C.Structure):
    _fields_ = [
        ("polymorphism_prediction", C.c_int32),
        ("mutation_log10_e_value_cutoff", C.c_double),
        ("polymorphism_log10_e_value_cutoff", C.c_double),
        ("consensus_frequency_cutoff", C.c_double),
        ("polymorphism_frequency_cutoff", C.c_double),
        ("consensus_minimum_variant_coverage", C.c_double),
        ("consensus_minimum_total_coverage", C.c_double),
        ("consensus_minimum_variant_coverage_each_strand", C.c_double),
        ("consensus_minimum_total_coverage_each_strand", C.c_double),
        ("polymorphism_minimum_variant_coverage", C.c_double),
        ("polymorphism_minimum_total_coverage", C.c_double),
        ("polymorphism_minimum_variant_coverage_each_strand", C.c_double),
        ("polymorphism_minimum_total_coverage_each_strand", C.c_double),
        ("polymorphism_fisher_strand_p_value_cutoff", C.c_double),
        ("polymorphism_ks_quality_p_value_cutoff", C.c_double),
        ("consensus_reject_indel_homopolymer_length", C.c_int32),
        ("consensus_reject_surrounding_homopolymer_length", C.c_int32),
        ("polymorphism_reject_indel_homopolymer_length", C.c_int32),
        ("polymorphism_reject_surrounding_homopolymer_length", C.c_int32),
        ("polymorphism_no_indels", C.c_int32),
    ]


class CoverageFit(C.Structure):
    _fields_ = [
        ("nb_fit_size", C.c_double),
        ("nb_fit_mu", C.c_double),
        ("average", C.c_double),
        ("variance", C.c_double),
        ("relative_variance", C.c_double),
        ("deletion_coverage_propagation_cutoff", C.c_double),
        ("censor_start", C.c_uint32),
        ("censor_end", C.c_uint32),
    ]


class ColumnLaunch(C.Structure):
    _fields_ = [
        ("blocks", C.c_int32),
        ("table_in_lds", C.c_int32),
        ("grid", C.c_int32),
        ("n_tried", C.c_int32),
        ("from_cache", C.c_int32),
        ("waves", C.c_int32),
        ("ms", C.c_double),
        ("default_ms", C.c_double),
    ]


class Timing(C.Structure):
    _fields_ = [
        ("pack_ms", C.c_double),
        ("h2d_ms", C.c_double),
        ("kernel_ms", C.c_double),
        ("d2h_ms", C.c_double),
        ("host_post_ms", C.c_double),
        ("host_pre_ms", C.c_double),
        ("aligned_bases", C.c_int64),
        ("candidates", C.c_int64),
        ("n_reads", C.c_int64),
        ("n_positions", C.c_int64),
        ("kernel_detail_ms", C.c_double * 8),
        ("fit_ms", C.c_double),
        ("fit_wait_ms", C.c_double),
        ("pipeline_ec_ms", C.c_double),
        ("pipeline_ec_kernel_ms", C.c_double * 2),
    ]


class PrecisionReport(C.Structure):
    _fields_ = [
        ("columns", C.c_int64),
        ("max_abs_error", C.c_double),
        ("p99_abs_error", C.c_double),
        ("mean_abs_error", C.c_double),
        ("max_rel_error", C.c_double),
        ("columns_redecided", C.c_int64),
        ("decision_mismatches", C.c_int64),
        ("count_mismatches", C.c_int64),
    ]


class CoverageSummary(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "deletion_coverage_propagation_cutoff", "nbinom_size_parameter", "nbinom_mean_parameter",
        "nbinom_prob_parameter", "nbinom_variance", "nbinom_relative_variance", "average", "variance",
        "relative_variance")]


class CoverageStageArgs(C.Structure):
    _fields_ = [
        ("n_targets", C.c_int32),
        ("target_lengths", C.POINTER(C.c_int64)),
        ("coverage_group", C.POINTER(C.c_int32)),
        ("n_coverage_groups", C.c_int32),
        ("histograms", C.POINTER(C.POINTER(C.c_double))),
        ("histogram_lengths", C.POINTER(C.c_int64)),
        ("targeted_sequencing", C.c_int32),
    ]


class CoverageTableArgs(C.Structure):
    _fields_ = [
        ("target", C.c_int32),
        ("start_1", C.c_int64),
        ("end_1", C.c_int64),
        ("resolution", C.c_uint32),
        ("total_only", C.c_int32),
        ("per_read_group", C.c_int32),
        ("n_read_groups", C.c_int32),
        ("csv", C.c_int32),
        ("show_average", C.c_int32),
        ("reference_average", C.c_double),
    ]


class SimReadsArgs(C.Structure):
    """bsq_simreads_args: the options of SIMULATE-READS."""
    _fields_ = [("mode", C.c_int32), ("circular", C.c_int32), ("n_reads", C.c_uint32), ("coverage", C.c_uint32),
                ("read_length", C.c_uint32), ("fragment_mean", C.c_uint32), ("fragment_stdev", C.c_uint32),
                ("seed", C.c_uint32), ("quality_start", C.c_uint32), ("quality_end", C.c_uint32),
                ("quality_stdev", C.c_uint32), ("quality_min", C.c_uint32), ("quality_max", C.c_uint32),
                ("indel_error_rate_per_million", C.c_uint32)]


class SimMutation(C.Structure):
    _fields_ = [("pos", C.c_int64), ("ref_len", C.c_int32), ("alt", C.c_char_p)]


class SimConfig(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("ref_length", C.c_int64),
        ("gc", C.c_double),
        ("read_length", C.c_int32),
        ("coverage", C.c_double),
        ("paired", C.c_int32),
        ("insert_mean", C.c_double),
        ("insert_sd", C.c_double),
        ("qual_mean", C.c_double),
        ("qual_sd", C.c_double),
        ("qual_min", C.c_int32),
        ("qual_max", C.c_int32),
        ("n_snps", C.c_int32),
        ("n_indels", C.c_int32),
        ("indel_max", C.c_int32),
        ("variant_freq_min", C.c_double),
        ("variant_freq_max", C.c_double),
        ("redundant_fraction", C.c_double),
        ("mapq", C.c_int32),
        ("discordant_fraction", C.c_double),
        ("n_junctions", C.c_int32),
        ("n_missing", C.c_int32),
        ("missing_length", C.c_int32),
        ("missing_edge_reads", C.c_int32),
    ]


class BsqError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("bsq error %s (%d): %s" % (ERR_NAMES.get(code, "?"), code, message))
        self.code = code


def _decl(lib, name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


_LIB = None


def lib():
    """The loaded libbsq.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libbsq.so is not built (%s); run `python -m breseq_amd.build` on a machine with hipcc" % LIB_PATH
        )
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    st = P(Status)
    _decl(L, "bsq_simulate", C.c_int, [P(SimConfig), P(vp), st])
    _decl(L, "bsq_simreads_defaults", C.c_int, [P(SimReadsArgs)])
    _decl(L, "bsq_simulate_reads", C.c_int, [C.c_char_p, C.c_int64, P(SimReadsArgs), P(vp), c_i64p, P(vp), c_i64p, st])
    _decl(L, "bsq_simulate_reads_sample", C.c_int, [C.c_char_p, C.c_char_p, C.c_int64, P(SimMutation), C.c_int32,
                                                     P(SimReadsArgs), C.c_int32, P(vp), st])
    _decl(L, "bsq_sample_read", C.c_int, [C.c_char_p, C.c_char_p, P(vp), st])
    _decl(L, "bsq_sample_write_bam", C.c_int, [vp, C.c_char_p, st])
    _decl(L, "bsq_sample_write_fasta", C.c_int, [vp, C.c_char_p, st])
    _decl(L, "bsq_sample_reference", P(Reference), [vp])
    _decl(L, "bsq_sample_alignments", P(Alignments), [vp])
    _decl(L, "bsq_sample_n_read_groups", C.c_int32, [vp])
    _decl(L, "bsq_sample_read_group", C.c_char_p, [vp, C.c_int32])
    _decl(L, "bsq_sample_free", None, [vp])
    _decl(L, "bsq_session_create", C.c_int, [C.c_int32, P(vp), st])
    _decl(L, "bsq_session_load", C.c_int, [vp, P(Reference), P(Alignments), P(ReadFiles), st])
    _decl(L, "bsq_session_error_count", C.c_int, [vp, P(ErrorCountArgs), P(vp), st])
    _decl(L, "bsq_session_identify", C.c_int, [vp, P(IdentifyArgs), P(vp), st])
    _decl(L, "bsq_session_identify_unscreened", C.c_int, [vp, P(IdentifyArgs), P(vp), st])
    _decl(L, "bsq_session_coverage_table", C.c_int, [vp, P(CoverageTableArgs), P(vp), c_i64p, st])
    _decl(L, "bsq_session_timing", C.c_int, [vp, P(Timing)])
    _decl(L, "bsq_session_precision_audit", C.c_int, [vp, P(IdentifyArgs), P(PrecisionReport), st])
    _decl(L, "bsq_session_alignment_trims", C.c_int, [vp, c_i32p, c_i32p, C.c_int64, st])
    _decl(L, "bsq_session_set_gd_writer", C.c_int, [vp, C.c_int32, st])
    _decl(L, "bsq_session_set_gd_capacity", C.c_int, [vp, C.c_int32, st])
    _decl(L, "bsq_session_set_column_launch", C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, st])
    _decl(L, "bsq_session_tune_columns", C.c_int, [vp, P(IdentifyArgs), C.c_int32, P(ColumnLaunch), st])
    _decl(L, "bsq_format_doubles_device", C.c_int, [c_f64p, C.c_int64, C.c_int32, C.c_int32, C.c_char_p, c_i32p, st])
    _decl(L, "bsq_session_destroy", None, [vp])
    _decl(L, "bsq_error_count", C.c_int, [P(ErrorCountArgs), P(vp), st])
    _decl(L, "bsq_identify_mutations", C.c_int, [P(IdentifyArgs), P(vp), st])
    _decl(L, "bsq_ec_result_histogram", C.c_int, [vp, C.c_int32, P(c_u32p), c_i64p])
    _decl(L, "bsq_ec_result_counts", C.c_int, [vp, P(C.POINTER(C.c_uint64)), c_i64p])
    _decl(L, "bsq_ec_result_log10_prob", C.c_int, [vp, P(c_f64p), c_i64p])
    _decl(L, "bsq_ec_result_rates_table", C.c_int, [vp, P(c_f64p), c_i64p])
    _decl(L, "bsq_ec_result_rates_text", C.c_int, [vp, P(C.c_char_p), c_i64p])
    _decl(L, "bsq_ec_result_no_pos_hash_pr", C.c_int, [vp, P(c_f64p), c_i64p])
    _decl(L, "bsq_ec_result_free", None, [vp])
    _decl(L, "bsq_im_result_gd_text", C.c_int, [vp, P(C.c_char_p), c_i64p])
    _decl(L, "bsq_im_result_counts", C.c_int, [vp, c_i64p, c_i64p, c_i64p])
    _decl(L, "bsq_im_result_dp_text", C.c_int, [vp, P(C.c_char_p), c_i64p])
    _decl(L, "bsq_dp_candidate_regions", C.c_int,
          [P(Reference), P(Alignments), c_u8p, c_f64p, P(DpSettings), P(C.c_void_p), c_i64p, st])
    _decl(L, "bsq_im_result_free", None, [vp])
    _decl(L, "bsq_ra_test_settings_default", None, [P(RaTestSettings), C.c_int32])
    _decl(L, "bsq_test_ra_evidence", C.c_int, [C.c_char_p, P(Reference), P(RaTestSettings), P(C.c_void_p), c_i64p, st])
    _decl(L, "bsq_free_text", None, [C.c_void_p])
    _decl(L, "bsq_sequence_trims", C.c_int, [C.c_char_p, C.c_int64, c_u8p, st])
    _decl(L, "bsq_sequence_trims_static", C.c_int, [C.c_char_p, C.c_int64, c_u8p, st])
    _decl(L, "bsq_selftest_division", C.c_int,
          [C.c_int32, C.c_uint64, C.c_uint64, P(C.c_uint64), P(C.c_uint64), st])
    _decl(L, "bsq_format_double", C.c_int32, [C.c_double, C.c_uint32, C.c_int32, C.c_char_p, C.c_int32])
    _decl(L, "bsq_fit_coverage_distribution", C.c_int, [c_u32p, C.c_int64, C.c_double, P(CoverageFit), st])
    _decl(L, "bsq_analyze_unique_coverage_distributions", C.c_int, [P(CoverageStageArgs), P(CoverageSummary), st])
    _decl(L, "bsq_read_coverage_histogram", C.c_int, [C.c_char_p, C.c_int64, c_f64p, C.c_int64, c_i64p, st])
    _decl(L, "bsq_session_pipeline", C.c_int,
          [vp, P(ErrorCountArgs), P(CoverageStageArgs), C.c_double, C.c_double, P(IdentifyArgs), P(vp),
           P(CoverageSummary), P(vp), st])
    _LIB = L
    return L


def check(rc, status):
    if rc != 0:
        raise BsqError(rc, status.message.decode(errors="replace"))
