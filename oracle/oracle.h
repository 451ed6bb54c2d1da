/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Single-threaded CPU restatement of breseq's pileup-scoring path.  Only tests/, the smoke check
 * of __graft_entry__ and the CPU-baseline leg of bench.py may load it, and only as the checker.
 * The product (breseq_amd) never links or calls anything here.
 *
 * Parity pins (what fixes each function independently of this restatement) are listed in the
 * header comment of oracle.cpp next to each function.
 */
#ifndef BSQ_ORACLE_H
#define BSQ_ORACLE_H

#include "../include/bsq.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ec_out orc_ec_out;
typedef struct orc_im_out orc_im_out;

/* error_count (error_count.cpp:50-68) -> histograms, counts, log10 table and the text of
 * error_rates.tab / @.unique_only_coverage_distribution.tab / base_qual_error_prob.#.tab */
int orc_error_count(const bsq_error_count_args* args, orc_ec_out** out, bsq_status* status);
int orc_ec_histogram(const orc_ec_out* o, int32_t group, const uint32_t** hist, int64_t* len);
int orc_ec_counts(const orc_ec_out* o, const uint64_t** counts, int64_t* n);
int orc_ec_log10_prob(const orc_ec_out* o, const double** t, int64_t* n);
/* which: 0 = error_rates.tab, 1 = coverage distribution of group `index`,
 *        2 = base_qual_error_prob of read file `index` */
int orc_ec_text(const orc_ec_out* o, int32_t which, int32_t index, const char** text, int64_t* len);
/* preprocess_stage: no_pos_hash_per_position_pr per target (NaN where not visited) */
int orc_ec_no_pos_hash_pr(const orc_ec_out* o, const double** per_target, int64_t* n);
void orc_ec_free(orc_ec_out* o);

/* Parse error_rates.tab text as read_log10_prob_table (error_count.cpp:629-654). */
int orc_read_log10_prob_table(const char* text, double* out, int64_t n_bins);

/* identify_mutations (identify_mutations.cpp:48-88) -> ra_mc_evidence.gd text */
int orc_identify_mutations(const bsq_identify_args* args, orc_im_out** out, bsq_status* status);
int orc_im_gd_text(const orc_im_out* o, const char** text, int64_t* len);
/* DP_candidate_regions.csv text (identify_mutations.cpp:2550-2573) */
int orc_im_dp_text(const orc_im_out* o, const char** text, int64_t* len);
/* Per position (slot 0) diagnostics for error reporting: consensus score, unique cov, ... */
int orc_im_position_scores(const orc_im_out* o, const double** score, int64_t* n);
void orc_im_free(orc_im_out* o);

/* BAM2COV coverage table of one region (coverage_output.cpp:190-503); release with orc_free_text */
int orc_coverage_table(const bsq_reference* ref, const bsq_alignments* aln, const bsq_coverage_table_args* args,
                       char** text, int64_t* len, bsq_status* status);
void orc_free_text(char* text);

/* SequenceTrims (calculate_trims.cpp:67-138) */
int orc_sequence_trims(const char* seq, int64_t len, uint8_t* out);
/* get_alignment_trims (calculate_trims.cpp:183-285): XL / XR of every alignment, -1 when unmapped */
int orc_alignment_trims(const bsq_reference* ref, const bsq_alignments* aln, int32_t* xl, int32_t* xr);

/* fit_allele_frequencies on explicit per-read data (identify_mutations.cpp:3240-3318):
 * r[n*5] (max-normalized likelihoods), lpmax[n], obs[n] (0..4); allowed[5]; tol.
 * Writes f[5]; returns log10 likelihood; *iterations gets the iteration count. */
double orc_fit_alleles(const double* r, const double* lpmax, const int32_t* obs, int64_t n, const int32_t* allowed,
                       double tol, double* f, int32_t* iterations);
/* The scoring of one column's explicit per-read data (r[n*5], lpmax[n], obs[n] as for
 * orc_fit_alleles) for variant allele v: the five-allele fit (f[5]), then out[0] = its log10
 * likelihood, out[1] = variant_presence_score (identify_mutations.cpp:3329-3344, with
 * log10_ref_length), out[2], out[3] = the frequency_lower / frequency_upper of
 * write_RA_frequency_bounds (:3094-3217) at full printed precision. */
int orc_presence_and_bounds(const double* r, const double* lpmax, const int32_t* obs, int64_t n, int32_t v,
                            double tol, double log10_ref_length, double* f, double* out);
/* fill_read_base_likelihoods (identify_mutations.cpp:3359-3384) for one read base: read file rs,
 * quality q, observed base index obs (0..4, reference-strand space), strand +1 forward / -1
 * reverse, mapping quality mapq, against the log10 table (n_bins = S*25*Q).  Writes lp[5], r[5]
 * and *lpmax. */
int orc_read_base_likelihoods(const double* log10_table, int64_t n_bins, int32_t S, int32_t Q, uint32_t rs,
                              uint32_t q, uint32_t obs, int32_t strand, int32_t mapq, double* lp, double* r,
                              double* lpmax);
/* pure_genotype_call (identify_mutations.cpp:3398-3433): returns score, *best gets the index. */
double orc_pure_genotype(const double* lps, uint32_t n, int32_t* best);

/* stats (stats.cpp:2144-2284) */
double orc_fisher_exact_2x2(uint32_t a, uint32_t b, uint32_t c, uint32_t d);
double orc_ks_test_less(const double* x, int64_t nx, const double* y, int64_t ny);

/* Alignment pileup traversal diagnostics: entries of position pos0 of target tid, in order.
 * Writes up to cap (read index, qpos, indel, is_del) quadruples; returns the count. */
int64_t orc_pileup_column(const bsq_alignments* aln, int32_t tid, int64_t pos0, int32_t* out,
                          int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
