/*
 * bsq.h -- C ABI of the MI355X pileup-scoring engine (breseq_amd).
 *
 * Drop-in boundary for breseq's pileup-scoring path.  Every entry point takes plain C structs,
 * pointers and sizes; nothing here names a torch or HIP type.  Each function cites the breseq
 * interface (src/breseq/..., file:line) whose behaviour it replaces.
 *
 * Conventions
 *   - Positions in structs are 0-based unless a name ends in _1.
 *   - Every call returns BSQ_OK (0) or an error code and fills *status (may be NULL).  Errors are
 *     returned, never raised with exit(): breseq's ASSERT/ERROR (common.h:111-193) ends the
 *     process, which a serving loop cannot afford.
 *   - Result objects are owned by the library and released with the matching *_free call.
 */
#ifndef BSQ_H
#define BSQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSQ_ABI_VERSION 1

enum bsq_code {
  BSQ_OK = 0,
  BSQ_ERR_ARG = 1,      /* malformed argument or unsupported option                    */
  BSQ_ERR_IO = 2,       /* file could not be opened, read or written                   */
  BSQ_ERR_FORMAT = 3,   /* BAM / FASTA / table text not in the expected format         */
  BSQ_ERR_DEVICE = 4,   /* HIP runtime failure or no device                            */
  BSQ_ERR_LIMIT = 5,    /* input exceeds a documented size limit                       */
  BSQ_ERR_DATA = 6      /* data violates an invariant breseq asserts (e.g. covariate
                           value above its declared maximum, error_count.cpp:484-491)  */
};

typedef struct bsq_status {
  int32_t code;
  char message[480];
} bsq_status;

/* ------------------------------------------------------------------------------------------ */
/* Input data model                                                                           */
/* ------------------------------------------------------------------------------------------ */

/* Reference sequences, in BAM header (target id) order.  Replaces the FASTA loaded by
 * pileup_base::pileup_base (pileup_base.cpp:61-88) and reference_sequence (:36-44). */
typedef struct bsq_reference {
  int32_t n_targets;
  const char* const* names;   /* seq_id of each target                                       */
  const int64_t* lengths;     /* length of each target                                       */
  const char* const* seqs;    /* upper-case bases of each target (lengths[i] chars)          */
} bsq_reference;

/* Sorted alignments, one record per BAM record in file order, as the fields breseq reads
 * through alignment_wrapper / pileup_wrapper (alignment.h:43-416). */
typedef struct bsq_alignments {
  int64_t n_reads;
  const int32_t* tid;         /* reference target id                                          */
  const int32_t* pos;         /* 0-based leftmost aligned reference position (BAM POS)       */
  const uint16_t* flag;       /* BAM FLAG                                                     */
  const uint8_t* mapq;        /* MAPQ                                                         */
  const int64_t* cigar_off;   /* n_reads+1 offsets into cigar                                 */
  const uint32_t* cigar;      /* BAM CIGAR words (len<<4 | op)                                */
  const int64_t* seq_off;     /* n_reads+1 offsets into seq / qual                            */
  const uint8_t* seq;         /* one BAM 4-bit base code per byte (A=1 C=2 G=4 T=8 N=15)     */
  const uint8_t* qual;        /* phred base qualities                                         */
  const int32_t* x1;          /* X1:i redundancy tag, -1 when absent (alignment.cpp:52-57)   */
  const int32_t* xl;          /* XL:i left trim tag, -1 when absent (alignment.h:389-410)    */
  const int32_t* xr;          /* XR:i right trim tag, -1 when absent                         */
  const int32_t* rg;          /* index of the RG:Z tag in the header @RG list, -1 when absent */
  /* Pair fields, read only by the discordant-pair accumulator (NULL when no read group is
   * paired; identify_mutations.cpp:1399-1457). */
  const int32_t* tlen;        /* BAM TLEN (alignment.h:284 insert_size)                       */
  const int8_t* xp;           /* XP:Z pair orientation (resolve_alignments.cpp:1272):
                                 0 "FR", 1 "RF", 2 "FF", -1 absent or any other value          */
  const int64_t* name_off;    /* n_reads+1 offsets into names                                 */
  const char* names;          /* read names, concatenated without terminators                 */
} bsq_alignments;

/* Read group -> read file resolution (alignment.h:533-591, read_group_index_map +
 * read_file_partition).  n_groups == 0 resolves every record to read file 0. */
typedef struct bsq_read_files {
  int32_t n_groups;
  const uint32_t* base;       /* flat index of each group's first read file                   */
  const uint32_t* count;      /* number of read files in each group (1 or 2)                  */
  int32_t n_files;
  const char* const* names;   /* read file base names, for base_qual_error_prob.#.tab         */
} bsq_read_files;

/* ------------------------------------------------------------------------------------------ */
/* error_count  (error_count.h:41-52, error_count.cpp:50-68)                                  */
/* ------------------------------------------------------------------------------------------ */

typedef struct bsq_error_count_args {
  const bsq_reference* ref;
  const bsq_alignments* aln;
  bsq_read_files read_files;
  const int32_t* coverage_group;   /* per target; Settings::seq_id_to_coverage_group           */
  int32_t n_coverage_groups;
  const uint8_t* call_target;      /* per target 1/0: call_mutations_seq_id_set (NULL = all)   */
  int32_t read_set_max;            /* covariate "read_set=N"                                   */
  int32_t quality_max;             /* covariate "quality=max_qual+1"                           */
  int32_t do_coverage;             /* write @.unique_only_coverage_distribution.tab            */
  int32_t do_errors;               /* write error_rates.tab and base_qual_error_prob.#.tab     */
  const char* output_dir;          /* NULL: compute only, write nothing                        */
  const char* error_rates_file;    /* full path of error_rates.tab (Settings, settings.cpp:1555) */
  /* preprocess_stage (error_count.h:41-52; caller breseq_cmdline.cpp:1969, coverage only): per
   * scored target, the fraction of (position, strand) pairs at positions without redundant reads
   * where no unique read starts (its first query base aligned there) with a stranded aligned end
   * at or beyond required_junction_read_end_min_coordinate(read length) (settings.h:345-354;
   * error_count.cpp:159-166, 187-190, 209-228) */
  int32_t preprocess_stage;
  uint32_t unmatched_end_minimum_read_length;   /* Settings (50)                              */
  double unmatched_end_length_factor;           /* Settings: 1 - require_match_fraction       */
} bsq_error_count_args;

typedef struct bsq_ec_result bsq_ec_result;

/* Histogram of unique-only coverage of coverage group g: hist[c] = positions with c
 * non-deletion, non-redundant reads and no redundant read (error_count.cpp:125-199). */
int bsq_ec_result_histogram(const bsq_ec_result* r, int32_t group, const uint32_t** hist,
                            int64_t* len);
/* Covariate count table (error_count.cpp:854-997); index = read_set + S*ref + 5S*obs + 25S*q. */
int bsq_ec_result_counts(const bsq_ec_result* r, const uint64_t** counts, int64_t* n_bins);
/* log10 probability table after counts_to_log10_prob (error_count.cpp:1005-1026). */
int bsq_ec_result_log10_prob(const bsq_ec_result* r, const double** table, int64_t* n_bins);
/* The same table as identify_mutations sees it: printed to error_rates.tab with 6 significant
 * digits and parsed back with strtod (error_count.cpp:660-690 then 629-654). */
int bsq_ec_result_rates_table(const bsq_ec_result* r, const double** table, int64_t* n_bins);
/* Text of error_rates.tab. */
int bsq_ec_result_rates_text(const bsq_ec_result* r, const char** text, int64_t* len);
/* Preprocess mode: Summary::preprocess_error_count[seq_id].no_pos_hash_per_position_pr of each
 * target (NaN for targets the pass did not visit, or without preprocess_stage). */
int bsq_ec_result_no_pos_hash_pr(const bsq_ec_result* r, const double** per_target, int64_t* n_targets);
void bsq_ec_result_free(bsq_ec_result* r);

/* ------------------------------------------------------------------------------------------ */
/* identify_mutations  (identify_mutations.h:46-60, identify_mutations.cpp:48-88)             */
/* ------------------------------------------------------------------------------------------ */

/* Discordant-pair (DP) candidate-region detection, one group per paired read file set
 * (identify_mutations.cpp:883-908).  Set i is read group i (the @RG index alignments carry). */
typedef struct bsq_dp_settings {
  int32_t n_sets;             /* Settings::read_file_sets.size()                               */
  const uint8_t* paired;      /* 1: cReadFileSet::is_paired() and a paired-mapping-distance
                                 fit exists for the set (Summary, pmdd.find != end)             */
  const double* median;       /* fitted pair distance median; <= 0 disables the set            */
  const double* mad;          /* its median absolute deviation; window = median + 2.42 * mad   */
  const uint32_t* r1_prefix;  /* read-name prefix of the set's R1 file: m_files[0].m_id + 1     */
  const uint32_t* r2_prefix;  /* and of its R2 file: m_files[1].m_id + 1                       */
  int32_t seed;               /* Settings::discordant_pair_seed (default 3, settings.cpp:1340)  */
} bsq_dp_settings;

typedef struct bsq_identify_args {
  const bsq_reference* ref;
  const bsq_alignments* aln;
  bsq_read_files read_files;
  const uint8_t* call_target;              /* per target (NULL = all)                         */
  const double* log10_prob_table;          /* as re-read from error_rates.tab text            */
  int64_t n_bins;
  int32_t read_set_max;
  int32_t quality_max;
  const double* deletion_propagation_cutoff; /* per target; < 0 means the coverage fit failed */
  const double* deletion_seed_cutoff;        /* per target                                   */
  double mutation_cutoff;                  /* consensus score cutoff (10)                    */
  double polymorphism_cutoff;              /* 10 consensus mode, 2 polymorphism mode         */
  double polymorphism_precision_decimal;   /* 1e-6                                           */
  uint32_t polymorphism_precision_places;  /* 3 consensus mode, 8 polymorphism mode          */
  int32_t base_quality_cutoff;             /* Settings::base_quality_cutoff (3)              */
  int32_t skip_missing_coverage_prediction;
  const char* gd_file;                     /* NULL: do not write ra_mc_evidence.gd           */
  const struct bsq_dp_settings* dp;        /* discordant-pair groups; NULL = none             */
  const char* dp_file;                     /* NULL: do not write DP_candidate_regions.csv    */
} bsq_identify_args;

typedef struct bsq_im_result bsq_im_result;

/* ra_mc_evidence.gd text exactly as cGenomeDiff::write would print it (genome_diff.cpp:685). */
int bsq_im_result_gd_text(const bsq_im_result* r, const char** text, int64_t* len);
/* Number of RA / MC / UN entries. */
int bsq_im_result_counts(const bsq_im_result* r, int64_t* n_ra, int64_t* n_mc, int64_t* n_un);
/* DP_candidate_regions.csv text (identify_mutations.cpp:2550-2573); header only when args->dp
 * is NULL or names no paired set. */
int bsq_im_result_dp_text(const bsq_im_result* r, const char** text, int64_t* len);
void bsq_im_result_free(bsq_im_result* r);

/* The discordant-pair accumulator on its own (host; it reads no pileup column): the sliding
 * windows and region state machine of check_discordant_completion (identify_mutations.cpp:
 * 2358-2422), fed at each read's start column as in pileup_callback (:1395-1457), over the
 * targets of call_target (NULL = all) whose deletion_propagation_cutoff is not negative (NULL =
 * none is; pileup_callback returns before the accumulator on those, :1318-1334).  dp->n_sets is
 * also the number of @RG lines read group indices resolve against.  Returns the CSV text;
 * release it with bsq_free_text. */
int bsq_dp_candidate_regions(const bsq_reference* ref, const bsq_alignments* aln,
                             const uint8_t* call_target, const double* deletion_propagation_cutoff,
                             const bsq_dp_settings* dp, char** out_text, int64_t* out_len,
                             bsq_status* status);

/* ------------------------------------------------------------------------------------------ */
/* test_RA_evidence  (identify_mutations.h:63-67, identify_mutations.cpp:687-749)            */
/* ------------------------------------------------------------------------------------------ */

typedef struct bsq_ra_test_settings {
  int32_t polymorphism_prediction;
  double mutation_log10_e_value_cutoff;
  double polymorphism_log10_e_value_cutoff;
  double consensus_frequency_cutoff;
  double polymorphism_frequency_cutoff;
  double consensus_minimum_variant_coverage;
  double consensus_minimum_total_coverage;
  double consensus_minimum_variant_coverage_each_strand;
  double consensus_minimum_total_coverage_each_strand;
  double polymorphism_minimum_variant_coverage;
  double polymorphism_minimum_total_coverage;
  double polymorphism_minimum_variant_coverage_each_strand;
  double polymorphism_minimum_total_coverage_each_strand;
  double polymorphism_fisher_strand_p_value_cutoff;
  double polymorphism_ks_quality_p_value_cutoff;
  int32_t consensus_reject_indel_homopolymer_length;
  int32_t consensus_reject_surrounding_homopolymer_length;
  int32_t polymorphism_reject_indel_homopolymer_length;
  int32_t polymorphism_reject_surrounding_homopolymer_length;
  int32_t polymorphism_no_indels;
} bsq_ra_test_settings;

/* Mode defaults of Settings (settings.cpp:857-950). */
void bsq_ra_test_settings_default(bsq_ra_test_settings* s, int32_t polymorphism_prediction);

/* Classify the RA entries of a GD text (as written by identify_mutations) and return the new
 * GD text.  ref is needed only by the homopolymer filters. */
int bsq_test_ra_evidence(const char* gd_text, const bsq_reference* ref,
                         const bsq_ra_test_settings* s, char** out_text, int64_t* out_len,
                         bsq_status* status);
void bsq_free_text(char* text);

/* breseq's number formatting for evidence fields, formatted_double / to_string(double, precision,
 * scientific) (common.h:845-866, 1542-1563): "NA" for NaN, otherwise %.Nf (scientific 0) or %.Ne
 * (scientific 1).  scientific 2 gives %.Ng, the default ostream form error_rates.tab is written in
 * (error_count.cpp:660-690, N = 6).  Returns the length written (excluding NUL), or -1 if the
 * buffer is too small. */
int32_t bsq_format_double(double v, uint32_t precision, int32_t scientific, char* buf, int32_t buf_len);

/* ------------------------------------------------------------------------------------------ */
/* calculate_trims  (calculate_trims.h:34, calculate_trims.cpp:67-179)                        */
/* ------------------------------------------------------------------------------------------ */

/* Per-position left/right trims of one reference sequence of length len:
 * out[0..len) = left trims, out[len..2len) = right trims (SequenceTrims layout). */
int bsq_sequence_trims(const char* seq, int64_t len, uint8_t* out, bsq_status* status);
/* No counterpart in breseq.  bsq_sequence_trims without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc):
 * the sequence, its repeat extents and the trims are staged in a static device array of the
 * library (32 MiB, held on each GPU it runs on until the process ends).  len <= 8388608 bases,
 * else BSQ_ERR_LIMIT.  Calls that overlap on one GPU share that array and race. */
int bsq_sequence_trims_static(const char* seq, int64_t len, uint8_t* out, bsq_status* status);

/* ------------------------------------------------------------------------------------------ */
/* Coverage distribution fit (coverage_distribution.cpp:345-399)                              */
/* ------------------------------------------------------------------------------------------ */

typedef struct bsq_coverage_fit {
  double nb_fit_size, nb_fit_mu, average, variance, relative_variance;
  double deletion_coverage_propagation_cutoff;
  uint32_t censor_start, censor_end;
} bsq_coverage_fit;

/* fit_coverage_distribution (coverage_distribution.cpp:345-399) of one histogram hist[0..len),
 * rows j >= 1 as read back from its file; sequence_length is the coverage group's summed length
 * (its 0.05/sqrt is the quantile's probability). */
int bsq_fit_coverage_distribution(const uint32_t* hist, int64_t len, double sequence_length,
                                  bsq_coverage_fit* out, bsq_status* status);

/* Summary::unique_coverage[seq_id] as the coverage-distribution stage leaves it
 * (summary.h:116-142, coverage_distribution.cpp:561-588). */
typedef struct bsq_coverage_summary {
  double deletion_coverage_propagation_cutoff;
  double nbinom_size_parameter, nbinom_mean_parameter, nbinom_prob_parameter;
  double nbinom_variance, nbinom_relative_variance;
  double average, variance, relative_variance;
} bsq_coverage_summary;

typedef struct bsq_coverage_stage_args {
  int32_t n_targets;
  const int64_t* target_lengths;      /* ref_seq_info.get_sequence_length of each target           */
  const int32_t* coverage_group;      /* per target: Settings::seq_id_to_coverage_group (NULL: 0)  */
  int32_t n_coverage_groups;          /* Settings::seq_ids_by_coverage_group().size()              */
  const double* const* histograms;    /* per group: n[0..len) as read back from
                                         @.unique_only_coverage_distribution.tab (rows j >= 1)     */
  const int64_t* histogram_lengths;
  int32_t targeted_sequencing;        /* Settings::targeted_sequencing: no fit, summary untouched  */
} bsq_coverage_stage_args;

/* CoverageDistribution::analyze_unique_coverage_distributions (coverage_distribution.h:100-107,
 * coverage_distribution.cpp:510-633): per coverage group, the NB fit of its histogram over the
 * group's summed length, written into every member target's summary.  per_target[n_targets] is
 * filled (all zero for targets the stage leaves alone).  The diagnostic plot is not drawn. */
int bsq_analyze_unique_coverage_distributions(const bsq_coverage_stage_args* args,
                                              bsq_coverage_summary* per_target, bsq_status* status);

/* read_coverage_histogram (coverage_distribution.cpp:34-65) of the text of a
 * @.unique_only_coverage_distribution.tab file: a header line, then "coverage<TAB>count" rows.
 * Writes n[0..max_coverage] into out (cap entries) and the length needed into *len; call with
 * cap 0 to learn it.  An empty table gives length 1 (n[0] = 0). */
int bsq_read_coverage_histogram(const char* text, int64_t text_len, double* out, int64_t cap,
                                int64_t* len, bsq_status* status);

/* ------------------------------------------------------------------------------------------ */
/* Device-resident sessions (MI355X path)                                                     */
/* ------------------------------------------------------------------------------------------ */

typedef struct bsq_session bsq_session;

typedef struct bsq_timing {
  double pack_ms;          /* host CSR packing                                    */
  double h2d_ms;           /* host -> device copies                               */
  double kernel_ms;        /* kernels of the last pass, HIP events on the stream  */
  double d2h_ms;           /* result copies back                                  */
  double host_post_ms;     /* host formatting of results                          */
  double host_pre_ms;      /* host argument processing and table uploads          */
  int64_t aligned_bases;   /* pileup entries (M/=/X bases + D positions)          */
  int64_t candidates;      /* positions sent to the exact FP64 scorer             */
  int64_t n_reads;         /* mapped alignments in the device CSR                 */
  int64_t n_positions;     /* reference columns (incl. one pad per target)        */
  /* per-kernel times of the last pass, HIP events on the session streams:
   * error_count: [0] k_ec_reads (with the difference-array clear)
   *              [5] k_cov_tile_sums + k_cov_tile_scan + k_cov_hist
   * identify:    [1] k_identify_columns  [2] k_insert_slots  [3] k_score_candidates  [4] k_runs
   *              [6] k_fix_columns  [7] number of columns k_fix_columns re-decided (a count)
   *              host work beside the scorer: [0] DP candidate regions (own thread)
   *              [5] MC/UN entries */
  double kernel_detail_ms[8];
  /* bsq_session_pipeline: the coverage-distribution stage on host threads (it runs beside the
   * identify column pass) and how long identify then waited for its cutoffs */
  double fit_ms;
  double fit_wait_ms;
  /* bsq_session_pipeline: wall time of its error_count step and that step's kernels
   * ([0] k_ec_reads, [1] coverage scan and histogram); the fields above then describe identify */
  double pipeline_ec_ms;
  double pipeline_ec_kernel_ms[2];
} bsq_timing;

/* Create a session bound to HIP device `device`.  Fails with BSQ_ERR_DEVICE when no GPU. */
int bsq_session_create(int32_t device, bsq_session** out, bsq_status* status);
/* Pack alignments into the device CSR and copy reference + reads to HBM (once). */
int bsq_session_load(bsq_session* s, const bsq_reference* ref, const bsq_alignments* aln,
                     const bsq_read_files* read_files, bsq_status* status);
/* error_count on the loaded sample.  args->ref / args->aln are ignored (the loaded ones are
 * used). */
int bsq_session_error_count(bsq_session* s, const bsq_error_count_args* args,
                            bsq_ec_result** out, bsq_status* status);
/* identify_mutations on the loaded sample. */
int bsq_session_identify(bsq_session* s, const bsq_identify_args* args, bsq_im_result** out,
                         bsq_status* status);
/* Which writer prints the genome diff of later identify passes (no counterpart in breseq: both
 * write the same text, cGenomeDiff::write, genome_diff.cpp:685-830): on_device 1 (the default)
 * numbers, sorts and prints the entries on the GPU and returns only the text; 0 assembles them on
 * the host from the scorer's results. */
int bsq_session_set_gd_writer(bsq_session* s, int32_t on_device, bsq_status* status);
/* Entries the device writer's lists hold (no counterpart in breseq; default 4096, at most 2^20).
 * A pass whose genome diff has more entries is written again with room for all of them, and the
 * session keeps the larger size; a caller expecting large genome diffs can size it once. */
int bsq_session_set_gd_capacity(bsq_session* s, int32_t entries, bsq_status* status);
/* Diagnostic without a counterpart in breseq: the device writer's number formatter (printf's %.*e
 * when scientific, else %.*f; NaN prints NA) over n values; out holds 512 bytes per value, lens the
 * text lengths (-1 for a precision above 60). */
int bsq_format_doubles_device(const double* v, int64_t n, int32_t precision, int32_t scientific, char* out,
                              int32_t* lens, bsq_status* status);
/* identify_mutations with the candidate screen disabled: every covered column goes through the
 * exact FP64 scorer.  Same output as bsq_session_identify; used to show the screen drops nothing. */
int bsq_session_identify_unscreened(bsq_session* s, const bsq_identify_args* args,
                                    bsq_im_result** out, bsq_status* status);
/* breseq's stages 07-08 on the loaded sample, run back to back as breseq_cmdline.cpp:2293, 2341
 * and 2471 run them:
 *   1. error_count (ec);
 *   2. the coverage-distribution stage on error_count's histograms (cov: lengths, groups and
 *      targeted flag; its histogram fields are ignored), skipped when
 *      im->skip_missing_coverage_prediction is set; then the Settings overrides
 *      deletion_coverage_propagation_cutoff / deletion_coverage_seed_cutoff (0 = unset, < 1 a
 *      fraction of the target's average coverage, >= 1 the value; breseq_cmdline.cpp:2350-2367,
 *      2409-2427);
 *   3. identify_mutations with error_count's table as re-read from error_rates.tab and each
 *      target's cutoffs from step 2 (im's log10 table, covariate maxima and both cutoff arrays are
 *      ignored).
 * The coverage fit runs on host threads while the device runs identify's column pass: only the
 * MC runs need its values, and whether a target's fit failed is known from its average first.
 * per_target[n_targets] receives the stage's summaries (with the overrides applied). */
int bsq_session_pipeline(bsq_session* s, const bsq_error_count_args* ec, const bsq_coverage_stage_args* cov,
                         double deletion_coverage_propagation_cutoff, double deletion_coverage_seed_cutoff,
                         const bsq_identify_args* im, bsq_ec_result** ec_out,
                         bsq_coverage_summary* per_target, bsq_im_result** im_out, bsq_status* status);
/* get_alignment_trims (calculate_trims.cpp:239-285) of every alignment of the loaded sample, the
 * XL / XR tags breseq writes (alignment.cpp:920-946): the larger of the reference's SequenceTrims
 * at the alignment's ends and the read-edge trims of its aligned sequence (edge_trims_for_sequence,
 * :183-230), plus the soft-clipped bases on each side.  xl, xr: n entries, n = the loaded
 * alignments' n_reads, in their order; -1 for records the session did not pack (unmapped). */
int bsq_session_alignment_trims(bsq_session* s, int32_t* xl, int32_t* xr, int64_t n, bsq_status* status);

/* Diagnostic without a counterpart in breseq: how far the column pass's FP32 genotype sums are
 * from the FP64 sums breseq forms (identify_mutations.cpp:1669-1689, in read order), over every
 * column with reads of one identify pass (args as for bsq_session_identify), in log10 units; and
 * whether the certified FP32 decisions plus the exact re-decisions gave every column the predicted /
 * consensus call the exact sums give. */
typedef struct bsq_precision_report {
  int64_t columns;              /* columns with at least one likelihood term                    */
  double max_abs_error;         /* max over columns and genotypes of |FP32 sum - FP64 sum|       */
  double p99_abs_error;         /* its 99th percentile                                          */
  double mean_abs_error;
  double max_rel_error;         /* max |FP32 - FP64| / |FP64|                                    */
  int64_t columns_redecided;    /* columns the exact re-decision kernel recomputed               */
  int64_t decision_mismatches;  /* columns whose final calls differ from the exact ones (0)      */
  int64_t count_mismatches;     /* columns whose FP32 read count differs from the exact (0)      */
} bsq_precision_report;
int bsq_session_precision_audit(bsq_session* s, const bsq_identify_args* args, bsq_precision_report* out,
                                bsq_status* status);
int bsq_session_timing(const bsq_session* s, bsq_timing* out);

/* Launch shape of the column pass of later identify passes (no counterpart in breseq, whose pileup
 * has no launch): `blocks` blocks (0, the default: the grid within 2048 blocks that gives every
 * wavefront the same number of 64-column windows, the fewest that fit; a set count is kept to at
 * most one window per wavefront and rounded down to a multiple of 8 from 8 up) of
 * `waves` wavefronts (4 or 8; 0 is the default, 8), and the likelihood table staged in LDS when it
 * fits (table_in_lds 1, the default) or read from global memory (0).  Results do not depend on it. */
int bsq_session_set_column_launch(bsq_session* s, int32_t blocks, int32_t table_in_lds, int32_t waves,
                                  bsq_status* status);

typedef struct bsq_column_launch {
  int32_t blocks;         /* the setting chosen (0: the default rule)                        */
  int32_t table_in_lds;   /* the setting chosen                                              */
  int32_t grid;           /* blocks the chosen launch runs                                   */
  int32_t n_tried;        /* launches timed                                                  */
  int32_t from_cache;     /* 1 when an earlier tuning of the same sample shape was reused    */
  int32_t waves;          /* wavefronts per block of the chosen launch                       */
  double ms;              /* column pass time of the chosen launch (best of the repeats)     */
  double default_ms;      /* column pass time of the default launch                          */
} bsq_column_launch;

/* Autotuner of the column pass (no counterpart in breseq): runs identify on the loaded sample with
 * `args` under candidate launches, `repeats` times each, keeps the fastest by the pass's own
 * HIP-event time (the default unless another is at least 2% faster) and sets it on the session.
 * The candidates are blocks of 8 and of 4 wavefronts, in grids of 256-2048 blocks and the grids
 * that give every wavefront of a block k windows (k = 1-16), with the table in LDS, then the table
 * in global memory at the best of those.
 * The choice is cached for the process by device, mean aligned length per alignment, mean depth
 * and table size; a later sample of the same shape reuses it without timing anything. */
int bsq_session_tune_columns(bsq_session* s, const bsq_identify_args* args, int32_t repeats, bsq_column_launch* out,
                             bsq_status* status);
void bsq_session_destroy(bsq_session* s);

/* BAM2COV coverage table of one region: coverage_output::table (coverage_output.cpp:190-283) with
 * the rows of pileup_callback (:307-503) over do_pileup's region and downsampling
 * (pileup_base.cpp:118-135, 141-210, 308-360).  Rows: position, reference base, then per strand the
 * unique coverage, the redundant coverage (Sum 1/X1), the redundant read count and the unique reads
 * starting there (or unique / redundant / total with total_only), repeated per read group with
 * per_read_group; then the averages of the region.  *out_text is NUL-terminated, *out_len bytes
 * long; release it with bsq_free_text. */
typedef struct bsq_coverage_table_args {
  int32_t target;             /* target id                                                    */
  int64_t start_1, end_1;     /* region, 1-based inclusive; both 0: the whole target          */
  uint32_t resolution;        /* --resolution (default 600); 0: every position                */
  int32_t total_only;         /* --total-only                                                 */
  int32_t per_read_group;     /* --per-read-group                                             */
  int32_t n_read_groups;      /* @RG lines of the BAM header                                  */
  int32_t csv;                /* --format CSV (comma-separated); else TSV                     */
  int32_t show_average;       /* --show-average: the reference_unique_average_cov line        */
  double reference_average;   /* its value: the target's coverage_average from summary.json   */
} bsq_coverage_table_args;

int bsq_session_coverage_table(bsq_session* s, const bsq_coverage_table_args* args, char** out_text,
                               int64_t* out_len, bsq_status* status);

/* Diagnostic without a counterpart in breseq: checks on device `device` that the scorer's
 * shared-denominator division (one reciprocal, then an FMA correction per quotient) returns the
 * correctly rounded quotient, bit for bit against ordinary division, on n random and adversarial
 * operand pairs.  *mismatches must come back 0; *fast_path counts the pairs that took the
 * shared-reciprocal path. */
int bsq_selftest_division(int32_t device, uint64_t n, uint64_t seed, uint64_t* mismatches,
                          uint64_t* fast_path, bsq_status* status);

/* One-shot conveniences: create a session on device 0, load, run, destroy. */
int bsq_error_count(const bsq_error_count_args* args, bsq_ec_result** out, bsq_status* status);
int bsq_identify_mutations(const bsq_identify_args* args, bsq_im_result** out,
                           bsq_status* status);

/* ------------------------------------------------------------------------------------------ */
/* BAM / FASTA I/O (replaces htslib hts_open/sam_read1/fai_fetch on this path)                */
/* ------------------------------------------------------------------------------------------ */

typedef struct bsq_sample bsq_sample;   /* owns a bsq_reference + bsq_alignments */

int bsq_sample_read(const char* bam_path, const char* fasta_path, bsq_sample** out,
                    bsq_status* status);
int bsq_sample_write_bam(const bsq_sample* s, const char* bam_path, bsq_status* status);
int bsq_sample_write_fasta(const bsq_sample* s, const char* fasta_path, bsq_status* status);
const bsq_reference* bsq_sample_reference(const bsq_sample* s);
const bsq_alignments* bsq_sample_alignments(const bsq_sample* s);
/* Number of @RG lines and their IDs (header order). */
int32_t bsq_sample_n_read_groups(const bsq_sample* s);
const char* bsq_sample_read_group(const bsq_sample* s, int32_t i);
void bsq_sample_free(bsq_sample* s);

/* Synthetic sample generator (seeded, deterministic): random reference, planted variants,
 * reads drawn from the mutated genome and aligned back by construction. */
typedef struct bsq_sim_config {
  uint64_t seed;
  int64_t ref_length;
  double gc;                 /* GC fraction of the reference                        */
  int32_t read_length;
  double coverage;
  int32_t paired;            /* 1: two read files (R1/R2), insert ~ N(mean, sd)      */
  double insert_mean, insert_sd;
  double qual_mean, qual_sd;
  int32_t qual_min, qual_max;
  int32_t n_snps;
  int32_t n_indels;
  int32_t indel_max;
  double variant_freq_min;   /* planted variant frequency range (1,1 = fixed)       */
  double variant_freq_max;
  double redundant_fraction; /* fraction of reads given X1>1                        */
  int32_t mapq;
  /* Discordant pairs: this fraction of fragments (paired only) becomes a pair without the
   * proper-pair flag whose R2 lies at one of n_junctions far sites, in orientation FR, RF or FF. */
  double discordant_fraction;
  int32_t n_junctions;
  /* Missing coverage: n_missing regions of missing_length bp that no fragment covers (a deletion in
   * the sample genome), each reached from both edges by missing_edge_reads repeat-mapped reads
   * (X1 = 2) that extend 10-40 bp into it. */
  int32_t n_missing;
  int32_t missing_length;
  int32_t missing_edge_reads;
} bsq_sim_config;

int bsq_simulate(const bsq_sim_config* cfg, bsq_sample** out, bsq_status* status);

/* SIMULATE-READS (do_simulate_reads, breseq_cmdline.cpp:1150-1330; cSimFastqSequence::simulate_read,
 * simulate_single_ends, simulate_paired_ends, fastq.cpp:1254-1583; cPortableRandom,
 * portable_random.cpp:22-88): byte-identical FASTQ for a given seed.  The option defaults are those
 * of the command (bsq_simreads_defaults).  The 'tiled' mode is not provided. */
typedef struct bsq_simreads_args {
  int32_t mode;                 /* 0 single-end, 1 paired-end                                   */
  int32_t circular;             /* the sequence is circular: fragments may wrap                  */
  uint32_t n_reads;             /* reads (single-end) or pairs; 0: from coverage                 */
  uint32_t coverage;            /* --coverage, used when n_reads is 0                            */
  uint32_t read_length;         /* --length                                                      */
  uint32_t fragment_mean;       /* --mean                                                        */
  uint32_t fragment_stdev;      /* --stdev                                                       */
  uint32_t seed;                /* --seed                                                        */
  uint32_t quality_start, quality_end, quality_stdev, quality_min, quality_max;
  uint32_t indel_error_rate_per_million;
} bsq_simreads_args;
int bsq_simreads_defaults(bsq_simreads_args* args);
/* FASTQ text of the reads (fastq_2: the second mates, paired-end only); release with bsq_free_text. */
int bsq_simulate_reads(const char* seq, int64_t len, const bsq_simreads_args* args, char** fastq_1,
                       int64_t* len_1, char** fastq_2, int64_t* len_2, bsq_status* status);
/* A planted mutation: ref_len reference bases from pos (0-based) replaced by alt (a SNP when both
 * are one base long, an insertion before pos when ref_len is 0, a deletion when alt is empty). */
typedef struct bsq_sim_mutation {
  int64_t pos;
  int32_t ref_len;
  const char* alt;
} bsq_sim_mutation;
/* No counterpart in breseq, which aligns the simulated reads with bowtie2: the reads SIMULATE-READS
 * makes from seq with the mutations applied, aligned to seq by construction (the bases each read was
 * copied from; simulated and planted insertions as I, deletions as D), sorted, with XL/XR as
 * get_alignment_trims writes them, MAPQ mapq and X1 1; a linear sequence of one target `name`. */
int bsq_simulate_reads_sample(const char* name, const char* seq, int64_t len, const bsq_sim_mutation* muts,
                              int32_t n_muts, const bsq_simreads_args* args, int32_t mapq, bsq_sample** out,
                              bsq_status* status);

#ifdef __cplusplus
}
#endif

#endif /* BSQ_H */
