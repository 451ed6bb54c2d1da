/*
 * oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A single-threaded CPU restatement of breseq's pileup-scoring path, written from the
 * reference's behaviour.  It is the checker for the MI355X product and the CPU baseline of the
 * benchmark; the product never calls it.  It follows breseq's arithmetic order exactly (same
 * summation order, glibc log10/pow, no FMA contraction: built with -ffp-contract=off), so that
 * its FP64 results are what the reference computes on the same inputs.
 *
 * Function                                   reference                           pinned by (tests/...)
 * ---------------------------------------------------------------------------------------------------
 * resolve_read / pileup columns              htslib resolve_cigar2 via           test_oracle_known_answers.py
 *                                            pileup_base.cpp:141-210,308-359     test_pileup_*: hand-built CIGARs
 * error_count callback, count position       error_count.cpp:125-199, 854-997    test_oracle_known_answers.py
 *                                                                                test_error_count_*: counts by hand
 * counts_to_log10_prob, error_rates.tab      error_count.cpp:660-690,1005-1026   test_error_count_log10_table_is_
 *                                                                                laplace_ratio (closed form, %g)
 * base_qual_error_prob text                  error_count.cpp:697-785             test_oracle_pins.py
 *                                                                                test_base_qual_error_prob_*: by hand
 * identify callback (pos_info, coverage,     identify_mutations.cpp:1309-1700    test_oracle_pins.py
 *   insert-slot loop, slot qualities)                                            test_identify_column_tallies_by_hand,
 *                                                                                test_identify_insert_slots_by_hand
 * fill_read_base_likelihoods                 identify_mutations.cpp:3359-3384    test_oracle_pins.py
 *                                                                                test_read_base_likelihoods_limits_
 *                                                                                and_strand (MAPQ 0 / high, strands)
 * pure_genotype_call                         identify_mutations.cpp:3398-3433    test_pure_genotype_closed_form
 * fit_allele_frequencies (EM)                identify_mutations.cpp:3240-3318    test_em_fixed_point_satisfies_kkt,
 *                                                                                test_em_respects_held_out_allele
 * variant_presence_score, profile bounds     identify_mutations.cpp:3094-3217,   test_oracle_pins.py
 *                                            3329-3344                           test_presence_score_and_bounds_
 *                                                                                closed_form (scipy brentq)
 * fisher_exact_test_2x2                      stats.cpp:2144-2169                 test_fisher_against_scipy;
 *                                                                                test_reference_vectors.py (printed
 *                                                                                fisher_strand_p_value of expected.gd)
 * ks_test_two_sample_less                    stats.cpp:2190-2284                 test_oracle_pins.py
 *                                                                                test_ks_less_ties_brute_force,
 *                                                                                test_ks_less_asymptotic_branch;
 *                                                                                test_ks_less_against_scipy_exact_
 *                                                                                tie_free
 * check_deletion_completion (MC),            identify_mutations.cpp:2117-2164,   test_oracle_pins.py
 *   at_target_end whole-target MC            2262-2344                           test_mc_runs_by_hand,
 *                                                                                test_mc_seed_cutoff_by_hand
 * update_unknown_intervals (UN)              identify_mutations.cpp:2972-3007    test_un_runs_by_hand
 * GD sort / write, number formats            genome_diff.cpp:685-830,            test_reference_vectors.py,
 *                                            common.h:845-866                    test_host.py (printf, expected.gd)
 * SequenceTrims                              calculate_trims.cpp:67-138          test_trims_*: known answers and
 *                                                                                brute force on tiny strings
 * edge trims, get_alignment_trims (XL/XR)    calculate_trims.cpp:183-285         test_oracle_pins.py
 *                                                                                test_alignment_trims_by_hand
 * discordant-pair regions (DP enter step,    identify_mutations.cpp:883-908,     test_dp_regions.py: hand-computed
 *   check_discordant_completion, CSV)        1395-1457, 2358-2422, 2550-2573     windows
 * BAM2COV coverage table (table, rows,       coverage_output.cpp:190-283,        test_coverage_table.py: hand-
 *   region traversal and downsampling)       307-503; pileup_base.cpp:118-360    computed tables; golden tables
 *
 * Not pinned bit-for-bit: lgamma (Cephes lngamma in the reference, glibc lgamma here) -- Fisher /
 * KS p-values agree with the reference's printed 6 significant digits on the expected.gd vectors.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <sstream>
#include <iomanip>
#include <stdexcept>
#include <string>
#include <vector>

using std::string;
using std::vector;

namespace orc {

static const uint32_t UNDEF = 0xFFFFFFFFu;

struct orc_error : std::runtime_error {
  int code;
  orc_error(int c, const string& m) : std::runtime_error(m), code(c) {}
};

static void fail(int code, const string& m) { throw orc_error(code, m); }

/* ---------------------------------------------------------------- base encodings (common.h:202-316) */

static inline bool bam_is_N(uint8_t b) { return b == 0x0f; }

static uint8_t complement_bam(uint8_t b) {
  switch (b) {
    case 0x1: return 0x8;
    case 0x2: return 0x4;
    case 0x4: return 0x2;
    case 0x8: return 0x1;
    case 0xf: return 0xf;
    case '.': return '.';
  }
  fail(BSQ_ERR_DATA, "Unrecognized BAM base");
  return 0;
}

static char complement_char(char c) {
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    case '.': return '.';
    case 'N': return 'N';
  }
  fail(BSQ_ERR_DATA, string("Unrecognized base char ") + c);
  return 0;
}

static uint8_t complement_index(uint8_t i) {
  switch (i) {
    case 0: return 3;
    case 1: return 2;
    case 2: return 1;
    case 3: return 0;
    default: return 4;
  }
}

static int basebam2index(uint8_t b) {
  switch (b) {
    case 0x1: return 0;
    case 0x2: return 1;
    case 0x4: return 2;
    case 0x8: return 3;
    case '.': return 4;
  }
  fail(BSQ_ERR_DATA, "BAM base not allowed");
  return 0;
}

static int basechar2index(char c) {
  switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    case '.': return 4;
    case 'N': return 5;
  }
  fail(BSQ_ERR_DATA, string("Unrecognized base char in reference: ") + c);
  return 0;
}

static const char kIndexChar[6] = {'A', 'C', 'G', 'T', '.', 'N'};

/* ---------------------------------------------------------------- number formatting (common.h:845-866) */

// to_string(double, precision, scientific): the "-0." rewrite in the reference operates on a
// temporary copy and therefore never applies; the 3-digit-exponent trim never fires with glibc.
static string fmt_double(double v, unsigned precision, bool scientific) {
  if (std::isnan(v)) return "NA";
  std::ostringstream os;
  os << (scientific ? std::scientific : std::fixed) << std::setprecision(precision) << v;
  string s = os.str();
  if (scientific && s.size() >= 3 && s[s.size() - 3] == '0') s.erase(s.size() - 3, 1);
  return s;
}

/* ---------------------------------------------------------------- alignment view */

enum { OP_M = 0, OP_I = 1, OP_D = 2, OP_N = 3, OP_S = 4, OP_H = 5, OP_P = 6, OP_EQ = 7, OP_X = 8 };

struct Aln {
  const bsq_alignments* a;
  int64_t n() const { return a->n_reads; }
  uint32_t n_cigar(int64_t r) const { return (uint32_t)(a->cigar_off[r + 1] - a->cigar_off[r]); }
  uint32_t op(int64_t r, uint32_t k) const { return a->cigar[a->cigar_off[r] + k] & 0xf; }
  uint32_t oplen(int64_t r, uint32_t k) const { return a->cigar[a->cigar_off[r] + k] >> 4; }
  uint32_t len(int64_t r) const { return (uint32_t)(a->seq_off[r + 1] - a->seq_off[r]); }
  uint8_t base(int64_t r, int64_t q) const {
    if (q < 0 || q >= (int64_t)len(r)) fail(BSQ_ERR_DATA, "read base index out of range");
    return a->seq[a->seq_off[r] + q];
  }
  uint8_t qual(int64_t r, int64_t q) const {
    if (q < 0 || q >= (int64_t)len(r)) fail(BSQ_ERR_DATA, "read quality index out of range");
    return a->qual[a->seq_off[r] + q];
  }
  bool reversed(int64_t r) const { return (a->flag[r] & 0x10) != 0; }
  uint32_t redundancy(int64_t r) const { return (a->x1 && a->x1[r] >= 0) ? (uint32_t)a->x1[r] : 1u; }
  // bam_endpos: pos + reference length of the CIGAR (ops M, D, N, =, X)
  int64_t endpos(int64_t r) const {
    int64_t e = a->pos[r];
    for (uint32_t k = 0; k < n_cigar(r); k++) {
      uint32_t o = op(r, k);
      if (o == OP_M || o == OP_D || o == OP_N || o == OP_EQ || o == OP_X) e += oplen(r, k);
    }
    return e;
  }
  uint32_t cigar2qlen(int64_t r) const {
    uint32_t l = 0;
    for (uint32_t k = 0; k < n_cigar(r); k++) {
      uint32_t o = op(r, k);
      if (o == OP_M || o == OP_I || o == OP_S || o == OP_EQ || o == OP_X) l += oplen(r, k);
    }
    return l;
  }
  // alignment.cpp:248-289
  uint32_t query_start_1(int64_t r) const {
    uint32_t pos = 1;
    for (uint32_t j = 0; j < n_cigar(r); j++) {
      if (op(r, j) != OP_S) break;
      pos += oplen(r, j);
    }
    return pos;
  }
  uint32_t query_end_1(int64_t r) const {
    int32_t pos = (int32_t)cigar2qlen(r);
    for (int32_t j = (int32_t)n_cigar(r) - 1; j > 0; j--) {
      if (op(r, j) != OP_S) break;
      pos -= oplen(r, j);
    }
    return (uint32_t)pos;
  }
  // alignment.cpp:125-220 with min_qual = 0
  void query_bounds_0(int64_t r, uint32_t& s, uint32_t& e) const {
    uint32_t start = 1, end = cigar2qlen(r);
    uint32_t nc = n_cigar(r);
    for (uint32_t i = 0; i < nc; i++) {
      uint32_t o = op(r, i);
      if (o != OP_S && o != OP_H && o != OP_N) break;
      if (o == OP_S) start += oplen(r, i);
    }
    for (int32_t i = (int32_t)nc - 1; i > 0; --i) {
      uint32_t o = op(r, i);
      if (o != OP_S && o != OP_H && o != OP_N) break;
      if (o == OP_S) end -= oplen(r, i);
    }
    s = start - 1;
    e = end - 1;
  }
  // is_trimmed (alignment.h:389-410)
  bool is_trimmed(int64_t r, int32_t qpos, bool past) const {
    uint32_t q1 = (uint32_t)qpos + 1;
    if (a->xl && a->xl[r] >= 0) {
      if (q1 <= (uint32_t)a->xl[r]) return true;
    }
    if (a->xr && a->xr[r] >= 0) {
      if (len(r) - q1 + 1 <= (uint32_t)a->xr[r]) return true;
      if (past && (len(r) - q1 == (uint32_t)a->xr[r])) return true;
    }
    return false;
  }
};

// Read file index of a record (error_count.h read_file_index + alignment.h:533-591).
static uint32_t read_file_index(const bsq_read_files& rf, const Aln& A, int64_t r) {
  if (rf.n_groups <= 0) return 0;
  uint32_t g = 0;
  if (rf.n_groups > 1) {
    int32_t rg = A.a->rg ? A.a->rg[r] : -1;
    g = (rg < 0 || rg >= rf.n_groups) ? 0u : (uint32_t)rg;
  }
  bool read2 = (A.a->flag[r] & 0x80) != 0;
  uint32_t off = (read2 && rf.count[g] > 1) ? 1 : 0;
  return rf.base[g] + off;
}

/* ---------------------------------------------------------------- pileup traversal
 * Restates what bam_plp_auto hands breseq for each column: every mapped record of the target
 * whose [pos, endpos) covers the column, in file order, with htslib resolve_cigar2 semantics:
 *   M/=/X  : qpos = query index of the base; indel = +len of a following I (summed over I/P
 *            runs), or -len of a following D run (D runs merged), at the op's last base only.
 *   D/N    : is_del = 1, qpos = query index of the next aligned base (state y of the walker).
 * Records with BAM_FUNMAP or without a CIGAR are skipped (bam_plp_push). */

struct Entry { int64_t read; int32_t qpos; int32_t indel; bool is_del; };

struct ResolvedRead {
  int64_t read;
  int64_t pos, end;          // [pos, end)
  vector<int32_t> qpos, indel;
  vector<uint8_t> is_del;
};

static void resolve_read(const Aln& A, int64_t r, ResolvedRead& out) {
  out.read = r;
  out.pos = A.a->pos[r];
  out.end = A.endpos(r);
  int64_t span = out.end - out.pos;
  out.qpos.assign(span, 0);
  out.indel.assign(span, 0);
  out.is_del.assign(span, 0);
  uint32_t nc = A.n_cigar(r);
  // walker state (k, x, y) as in resolve_cigar2
  int64_t x = out.pos;
  int64_t y = 0;
  uint32_t k = 0;
  auto is_ref_op = [](uint32_t o) { return o == OP_M || o == OP_D || o == OP_N || o == OP_EQ || o == OP_X; };
  for (; k < nc; ++k) {
    uint32_t o = A.op(r, k);
    if (is_ref_op(o)) break;
    if (o == OP_I || o == OP_S) y += A.oplen(r, k);
  }
  for (int64_t p = out.pos; p < out.end; p++) {
    // advance to the op containing p
    while (p - x >= (int64_t)A.oplen(r, k)) {
      uint32_t o = A.op(r, k);
      int64_t l = A.oplen(r, k);
      if (o == OP_M || o == OP_EQ || o == OP_X) y += l;
      x += l;
      uint32_t kk = k + 1;
      for (; kk < nc; ++kk) {
        uint32_t o2 = A.op(r, kk);
        if (is_ref_op(o2)) break;
        if (o2 == OP_I || o2 == OP_S) y += A.oplen(r, kk);
      }
      k = kk;
      if (k >= nc) fail(BSQ_ERR_DATA, "CIGAR walk ran past the end");
    }
    uint32_t o = A.op(r, k);
    int64_t l = A.oplen(r, k);
    int32_t indel = 0;
    if (x + l - 1 == p && k + 1 < nc) {
      uint32_t o2 = A.op(r, k + 1);
      int32_t l2 = (int32_t)A.oplen(r, k + 1);
      if (o2 == OP_D && o != OP_D) {
        indel = -l2;
        for (uint32_t kk = k + 2; kk < nc; ++kk) {
          if (A.op(r, kk) == OP_D) indel -= (int32_t)A.oplen(r, kk);
          else break;
        }
      } else if (o2 == OP_I) {
        indel = l2;
        for (uint32_t kk = k + 2; kk < nc; ++kk) {
          uint32_t o3 = A.op(r, kk);
          if (o3 == OP_I) indel += (int32_t)A.oplen(r, kk);
          else if (o3 != OP_P) break;
        }
      } else if (o2 == OP_P && k + 2 < nc) {
        int32_t l3 = 0;
        for (uint32_t kk = k + 2; kk < nc; ++kk) {
          uint32_t o3 = A.op(r, kk);
          if (o3 == OP_I) l3 += (int32_t)A.oplen(r, kk);
          else if (is_ref_op(o3)) break;
        }
        if (l3 > 0) indel = l3;
      }
    }
    int64_t j = p - out.pos;
    out.indel[j] = indel;
    if (o == OP_M || o == OP_EQ || o == OP_X) {
      out.qpos[j] = (int32_t)(y + (p - x));
    } else {
      out.is_del[j] = 1;
      out.qpos[j] = (int32_t)y;
    }
  }
}

// Iterates the columns of one target; calls fn(pos0, entries) for every position 0..L-1.
template <class F>
static void for_each_column(const Aln& A, int32_t tid, int64_t L, F fn) {
  // reads of this target in file order
  vector<int64_t> reads;
  for (int64_t r = 0; r < A.n(); r++) {
    if (A.a->tid[r] != tid) continue;
    if (A.a->flag[r] & 0x4) continue;
    if (A.n_cigar(r) == 0) continue;
    reads.push_back(r);
  }
  for (size_t i = 1; i < reads.size(); i++) {
    if (A.a->pos[reads[i]] < A.a->pos[reads[i - 1]]) fail(BSQ_ERR_DATA, "alignments are not sorted by position");
  }
  std::vector<ResolvedRead> active;   // in file order
  size_t next = 0;
  vector<Entry> col;
  for (int64_t p = 0; p < L; p++) {
    while (next < reads.size() && A.a->pos[reads[next]] <= p) {
      ResolvedRead rr;
      resolve_read(A, reads[next], rr);
      if (rr.end > rr.pos) active.push_back(std::move(rr));
      next++;
    }
    // drop finished reads (keeps order)
    size_t w = 0;
    for (size_t i = 0; i < active.size(); i++) {
      if (active[i].end > p) {
        if (w != i) active[w] = std::move(active[i]);
        w++;
      }
    }
    active.resize(w);
    col.clear();
    for (size_t i = 0; i < active.size(); i++) {
      const ResolvedRead& rr = active[i];
      if (rr.pos > p) continue;
      int64_t j = p - rr.pos;
      col.push_back(Entry{rr.read, rr.qpos[j], rr.indel[j], rr.is_del[j] != 0});
    }
    fn(p, col);
  }
}

// Targets in the order breseq visits them: std::set<string> of seq_ids (pileup_base.cpp:364-385).
static vector<int32_t> target_order(const bsq_reference* ref, const uint8_t* call) {
  vector<int32_t> t;
  for (int32_t i = 0; i < ref->n_targets; i++) {
    if (call && !call[i]) continue;
    t.push_back(i);
  }
  std::sort(t.begin(), t.end(), [&](int32_t a, int32_t b) { return strcmp(ref->names[a], ref->names[b]) < 0; });
  return t;
}

static inline char ref_char(const bsq_reference* ref, int32_t tid, int64_t p) {
  if (p < 0 || p >= ref->lengths[tid]) return '\0';   // fai_fetch strings are NUL terminated
  return ref->seqs[tid][p];
}

/* ---------------------------------------------------------------- error table */

struct ErrorTable {
  int S = 0, Q = 0;
  vector<uint64_t> counts;
  size_t nbins() const { return (size_t)S * 25 * Q; }
  size_t index(uint32_t rs, uint32_t ref, uint32_t obs, uint32_t q) const {
    if (rs >= (uint32_t)S) fail(BSQ_ERR_DATA, "Covariate 'read_set' exceeded its maximum");
    if (q >= (uint32_t)Q) fail(BSQ_ERR_DATA, "Covariate 'quality' exceeded its maximum");
    return rs + (size_t)S * ref + (size_t)5 * S * obs + (size_t)25 * S * q;
  }
};

struct EcOut {
  vector<double> no_pos_hash_pr;   // preprocess mode, per target (NaN when not visited)
  vector<vector<uint32_t>> hist;
  vector<uint64_t> counts;
  vector<double> log10p;
  string rates_text;
  vector<string> cov_text;
  vector<string> bq_text;
};

static void count_alignment_position(const Aln& A, const bsq_reference* ref, int32_t tid, const bsq_read_files& rf,
                                     ErrorTable& T, const Entry& e, int64_t ref_pos) {
  int64_t r = e.read;
  uint32_t reversed = A.reversed(r) ? 1 : 0;
  int32_t q_pos_0 = e.qpos;
  int32_t q_start_0 = (int32_t)A.query_start_1(r) - 1;
  int32_t q_end_0 = (int32_t)A.query_end_1(r) - 1;
  int32_t q_length = (int32_t)A.len(r);
  uint32_t rs = read_file_index(rf, A, r);
  {
    uint8_t obs = A.base(r, q_pos_0);
    char rc = ref_char(ref, tid, ref_pos);
    if (!bam_is_N(obs) && rc != 'N') {
      if (reversed) { obs = complement_bam(obs); rc = complement_char(rc); }
      T.counts[T.index(rs, basechar2index(rc), basebam2index(obs), A.qual(r, q_pos_0))]++;
    }
  }
  if (e.indel == 0) {
    if (q_pos_0 < q_end_0) {
      int32_t mqpos = q_pos_0 + 1 - (int32_t)reversed;
      if (!(mqpos >= 0 && mqpos < q_length)) fail(BSQ_ERR_DATA, "quality index out of range for '..' state");
      uint8_t obs = A.base(r, mqpos);
      char rc = ref_char(ref, tid, ref_pos + 1 - reversed);
      if (!bam_is_N(obs) && rc != 'N') T.counts[T.index(rs, 4, 4, A.qual(r, mqpos))]++;
    }
  } else if (e.indel == -1) {
    int32_t mqpos = q_pos_0 + 1 - (int32_t)reversed;
    if (!(mqpos >= 0 && mqpos < q_length)) fail(BSQ_ERR_DATA, "quality index out of range for 'N.' state");
    uint8_t obs = A.base(r, mqpos);
    char rc = ref_char(ref, tid, ref_pos + 1);
    if (!bam_is_N(obs) && rc != 'N') {
      if (reversed) rc = complement_char(rc);
      T.counts[T.index(rs, basechar2index(rc), 4, A.qual(r, mqpos))]++;
    }
  } else if (e.indel == 1) {
    int32_t mqpos = q_pos_0 + 1;
    if (!(mqpos >= 0 && mqpos < q_length)) fail(BSQ_ERR_DATA, "quality index out of range for '.N' state");
    if (mqpos <= q_end_0 && mqpos >= q_start_0) {
      uint8_t obs = A.base(r, mqpos);
      if (!bam_is_N(obs)) {
        if (reversed) obs = complement_bam(obs);
        T.counts[T.index(rs, 4, basebam2index(obs), A.qual(r, mqpos))]++;
      }
    }
  }
}

static void error_count(const bsq_error_count_args* args, EcOut& out) {
  if (!args || !args->ref || !args->aln) fail(BSQ_ERR_ARG, "missing arguments");
  Aln A{args->aln};
  ErrorTable T;
  T.S = args->read_set_max;
  T.Q = args->quality_max;
  if (T.S <= 0 || T.Q <= 0) fail(BSQ_ERR_ARG, "covariate maxima must be positive");
  T.counts.assign(T.nbins(), 0);
  out.hist.assign(args->n_coverage_groups, vector<uint32_t>());
  out.no_pos_hash_pr.assign(args->ref->n_targets, NAN);
  // required_junction_read_end_min_coordinate (settings.h:345-354)
  auto end_min = [&](uint32_t read_length) -> uint32_t {
    const int32_t m = (int32_t)floor((double)((int32_t)read_length - (int32_t)args->unmatched_end_minimum_read_length) *
                                     args->unmatched_end_length_factor);
    return (m <= 0) ? read_length : read_length - (uint32_t)m;
  };
  // query_stranded_end_1 (alignment.cpp:228-239 over query_bounds_1)
  auto stranded_end_1 = [&](int64_t r) -> uint32_t {
    uint32_t s0, e0;
    A.query_bounds_0(r, s0, e0);
    return A.reversed(r) ? A.len(r) - (s0 + 1) + 1 : e0 + 1;
  };
  for (int32_t tid : target_order(args->ref, args->call_target)) {
    int32_t g = args->coverage_group ? args->coverage_group[tid] : 0;
    if (g < 0 || g >= args->n_coverage_groups) fail(BSQ_ERR_ARG, "coverage group out of range");
    vector<uint32_t>& h = out.hist[g];
    uint64_t found[2] = {0, 0};   // at_target_start (error_count.cpp:209-214)
    for_each_column(A, tid, args->ref->lengths[tid], [&](int64_t p, const vector<Entry>& col) {
      size_t unique_coverage = 0;
      bool has_redundant = false;
      int has_query_start[2] = {0, 0};
      for (const Entry& e : col) {
        if (e.is_del) continue;
        if (A.redundancy(e.read) > 1) { has_redundant = true; continue; }
        ++unique_coverage;
        if (args->preprocess_stage) {
          if (e.qpos + 1 == 1 && stranded_end_1(e.read) >= end_min(A.len(e.read)))
            has_query_start[A.reversed(e.read) ? 1 : 0] = 1;
        }
        if (!args->do_errors) continue;
        count_alignment_position(A, args->ref, tid, args->read_files, T, e, p);
      }
      if (!has_redundant) {
        if (unique_coverage >= h.size()) h.resize(unique_coverage + 1, 0);
        ++h[unique_coverage];
        if (args->preprocess_stage) {
          found[has_query_start[0]]++;
          found[has_query_start[1]]++;
        }
      }
    });
    if (args->preprocess_stage) {   // at_target_end (error_count.cpp:217-229)
      const double total = (double)(found[0] + found[1]);
      out.no_pos_hash_pr[tid] = (total != 0) ? (double)found[0] / total : 1.0;
    }
  }
  out.counts = T.counts;
  // counts_to_log10_prob (error_count.cpp:1005-1026)
  const int S = T.S, Q = T.Q;
  vector<double> sums((size_t)S * 5 * Q, 0.0);
  for (size_t i = 0; i < T.nbins(); i++) {
    size_t rs = i % S, refb = (i / S) % 5, q = (i / (25 * (size_t)S)) % Q;
    sums[rs + (size_t)S * refb + (size_t)5 * S * q] += (double)T.counts[i];
  }
  out.log10p.assign(T.nbins(), 0.0);
  for (size_t i = 0; i < T.nbins(); i++) {
    size_t rs = i % S, refb = (i / S) % 5, q = (i / (25 * (size_t)S)) % Q;
    double c = (double)T.counts[i];
    out.log10p[i] = log10(c + 1.0) - log10(sums[rs + (size_t)S * refb + (size_t)5 * S * q] + 5.0);
  }
  // error_rates.tab text (error_count.cpp:660-690)
  {
    std::ostringstream os;
    os << "read_set=" << S << ",ref_base,obs_base,quality=" << Q << "\n";
    os << "read_set\tref_base\tobs_base\tquality\tlog10_probability\n";
    for (size_t i = 0; i < T.nbins(); i++) {
      size_t rs = i % S, refb = (i / S) % 5, obs = (i / (5 * (size_t)S)) % 5, q = (i / (25 * (size_t)S)) % Q;
      os << rs << "\t" << kIndexChar[refb] << "\t" << kIndexChar[obs] << "\t" << q << "\t" << out.log10p[i] << "\n";
    }
    out.rates_text = os.str();
  }
  // coverage distributions (error_count.cpp:239-253)
  for (size_t g = 0; g < out.hist.size(); g++) {
    std::ostringstream os;
    os << "coverage\tn\n";
    for (size_t j = 1; j < out.hist[g].size(); j++) os << j << "\t" << out.hist[g][j] << "\n";
    out.cov_text.push_back(os.str());
  }
  // base_qual_error_prob.#.tab (error_count.cpp:697-785)
  {
    vector<double> ot((size_t)S * 25 * Q, 0.0);
    for (size_t i = 0; i < T.nbins(); i++) {
      size_t rs = i % S, refb = (i / S) % 5, obs = (i / (5 * (size_t)S)) % 5, q = (i / (25 * (size_t)S)) % Q;
      ot[rs * 25 * Q + refb + 5 * obs + 25 * q] += (double)T.counts[i];
    }
    double running = 0;
    for (int r = 0; r < S; r++) {
      for (size_t k = 0; k < (size_t)25 * Q; k++) {
        size_t i = (size_t)r * 25 * Q + k;
        running += ot[i];
        if (i % 5 == 4) {
          for (size_t j = i - 4; j <= i; j++) {
            if (running > 0) ot[j] /= running;
            if (ot[j] == 0) ot[j] = NAN;
          }
          running = 0;
        }
      }
      std::ostringstream os;
      os << "quality";
      for (int b1 = 0; b1 < 5; b1++)
        for (int b2 = 0; b2 < 5; b2++) os << "\t" << kIndexChar[b1] << kIndexChar[b2];
      os << "\n";
      for (int q = 0; q < Q; q++) {
        os << q;
        for (int b1 = 0; b1 < 5; b1++)
          for (int b2 = 0; b2 < 5; b2++) {
            size_t idx = (size_t)r * 25 * Q + (size_t)q * 25 + b1 * 5 + b2;
            os << "\t";
            if (std::isnan(ot[idx])) os << "NA";
            else os << ot[idx];
          }
        os << "\n";
      }
      out.bq_text.push_back(os.str());
    }
  }
}

/* ---------------------------------------------------------------- statistics */

static double log_choose(double n, double k) { return lgamma(n + 1) - lgamma(k + 1) - lgamma(n - k + 1); }

double fisher_2x2(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {   // stats.cpp:2144-2169
  uint32_t row1 = a + b, row2 = c + d, col1 = a + c, col2 = b + d;
  uint32_t n = row1 + row2;
  uint32_t a_min = (row1 > col2) ? (row1 - col2) : 0;
  uint32_t a_max = std::min(row1, col1);
  double log_denom = log_choose(n, row1);
  double log_p_observed = log_choose(col1, a) + log_choose(col2, row1 - a) - log_denom;
  const double log_rel_err = log(1.0 + 1e-7);
  double total = 0.0;
  for (uint32_t ai = a_min; ai <= a_max; ai++) {
    double lp = log_choose(col1, ai) + log_choose(col2, row1 - ai) - log_denom;
    if (lp <= log_p_observed + log_rel_err) total += exp(lp);
  }
  return std::min(total, 1.0);
}

double ks_less(const vector<double>& x, const vector<double>& y) {   // stats.cpp:2221-2284
  uint32_t nx = (uint32_t)x.size(), ny = (uint32_t)y.size();
  vector<std::pair<double, bool>> comb;
  for (double v : x) comb.push_back({v, true});
  for (double v : y) comb.push_back({v, false});
  std::sort(comb.begin(), comb.end());
  uint32_t cx = 0, cy = 0, bx = 0, by = 0;
  double min_z = 0.0;
  bool have = false;
  vector<bool> boundary(comb.size() + 1, true);
  size_t i = 0;
  while (i < comb.size()) {
    double v = comb[i].first;
    size_t gs = i;
    while (i < comb.size() && comb[i].first == v) {
      if (comb[i].second) cx++; else cy++;
      i++;
    }
    for (size_t k = gs + 1; k < i; k++) boundary[k] = false;
    double z = (double)cx / nx - (double)cy / ny;
    if (!have || z < min_z) { min_z = z; have = true; bx = cx; by = cy; }
  }
  double stat = -min_z;
  if ((double)nx * ny < 10000) {
    uint32_t m = ny, n = nx;
    int64_t thr = (int64_t)by * n - (int64_t)bx * m;
    vector<vector<double>> u(m + 1, vector<double>(n + 1, 0.0));
    for (uint32_t a = 0; a <= m; a++)
      for (uint32_t b = 0; b <= n; b++) {
        double val;
        if (a == 0 && b == 0) val = 1.0;
        else val = (a > 0 ? u[a - 1][b] : 0.0) + (b > 0 ? u[a][b - 1] : 0.0);
        int64_t lvl = (int64_t)a * n - (int64_t)b * m;
        if (boundary[a + b] && lvl >= thr) val = 0.0;
        u[a][b] = val;
      }
    double total = exp(log_choose(m + n, m));
    double p = 1.0 - u[m][n] / total;
    return std::min(std::max(p, 0.0), 1.0);
  }
  double neff = ((double)nx * ny) / (nx + ny);
  return exp(-2.0 * stat * stat * neff);
}

/* ---------------------------------------------------------------- identify_mutations */

struct PData {
  char base;
  uint8_t quality;
  int strand;
  int32_t mq;
  uint32_t rs, obs, q;
  double lp[5], r[5], lpmax;
};

struct Coverage {
  double unique[3] = {0, 0, 0};
  double redundant[3] = {0, 0, 0};
  int raw_redundant[3] = {0, 0, 0};
  int total = 0;
  void sum() {
    unique[1] = unique[0] + unique[2];
    redundant[1] = redundant[0] + redundant[2];
    raw_redundant[1] = raw_redundant[0] + raw_redundant[2];
    total = (int)round(unique[1]) + (int)round(redundant[1]);
  }
  static Coverage nan() {
    Coverage c;
    double v = std::numeric_limits<double>::quiet_NaN();
    for (int i = 0; i < 3; i++) { c.unique[i] = v; c.redundant[i] = v; }
    return c;
  }
};

struct AlleleModel {
  double f[5] = {0, 0, 0, 0, 0};
  double log10_likelihood = 0;
  double sum_w[5] = {0, 0, 0, 0, 0};
  uint32_t n = 0, iterations = 0;
  double reported(int b) const {
    if (n == 0 || b >= 5) return 0.0;
    return (f[b] < 0.5 / (double)n) ? 0.0 : f[b];
  }
  int major() const {
    if (n == 0) return 5;
    int best = 0;
    for (int b = 1; b < 5; b++) if (f[b] > f[best]) best = b;
    return (f[best] > 0.0) ? best : 5;
  }
  int next(int exclude) const {
    if (n == 0) return 5;
    const double thr = 0.5 / (double)n;
    int best = 5;
    for (int b = 0; b < 5; b++) {
      if (b == exclude) continue;
      if (f[b] < thr) continue;
      if (best == 5 || f[b] > f[best]) best = b;
    }
    return best;
  }
  string spectrum(unsigned places) const {
    string s;
    for (int b = 0; b < 5; b++) {
      double fr = reported(b);
      if (fr <= 0.0) continue;
      if (!s.empty()) s += ",";
      s += string(1, kIndexChar[b]) + ":" + fmt_double(fr, places, true);
    }
    return s;
  }
};

typedef std::map<string, string> Fields;

struct GdEntry {
  int type;        // 0 RA, 1 MC, 2 UN
  uint32_t id;
  string seq_id;
  uint32_t pos;    // RA position / MC,UN start
  int32_t ins;     // RA insert position
  uint32_t end;    // MC/UN end
  uint32_t start_range, end_range;
  string ref_base, new_base;
  Fields fields;   // non-spec fields
};

struct Identify {
  const bsq_identify_args* args;
  Aln A;
  vector<double> prob;          // 10^log10 (log10_prob_to_prob)
  double log10_ref_length = 0;
  vector<GdEntry> gd;
  uint32_t next_id = 0;
  vector<double> pos_scores;
  // MC state
  double prop_cutoff = 0, seed_cutoff = 0;
  uint32_t del_start = UNDEF, del_end = UNDEF, red_start = UNDEF, red_end = UNDEF;
  bool reaches_seed = false, red_reached_zero = false;
  Coverage left_outside, left_inside, last_cov;
  uint32_t un_start = UNDEF;

  // Discordant-pair (DP) candidate regions (identify_mutations.h:244-290), per bin = strand*3 + orient
  struct DpRead {
    uint32_t start_pos, end_pos;
    string key;
    bool redundant;
  };
  struct DpGroup {
    double window_width;
    uint32_t r1_prefix, r2_prefix;
    std::deque<DpRead> reads[6];
  };
  struct DpRegion {
    string seq_id;
    uint32_t start, end, max_count;
    char strand;
    string orientation;
    bool redundant;
    string pairs;
  };
  vector<DpGroup> dp_groups;
  vector<int> rg_to_dp;
  int32_t dp_metric[6] = {0, 0, 0, 0, 0, 0};
  uint32_t dp_start[6], dp_max[6] = {0, 0, 0, 0, 0, 0}, dp_red[6] = {0, 0, 0, 0, 0, 0};
  vector<string> dp_desc[6];
  vector<DpRegion> dp_regions;

  explicit Identify(const bsq_identify_args* a) : args(a), A{a->aln} {
    for (int b = 0; b < 6; b++) dp_start[b] = UNDEF;
    // identify_mutations.cpp:893-908
    const bsq_dp_settings* dp = a->dp;
    if (!dp) return;
    rg_to_dp.assign((size_t)dp->n_sets, -1);
    for (int32_t i = 0; i < dp->n_sets; i++) {
      if (!dp->paired || !dp->paired[i]) continue;
      if (dp->median[i] <= 0.0) continue;
      DpGroup g;
      g.window_width = dp->median[i] + 2.42 * dp->mad[i];
      g.r1_prefix = dp->r1_prefix[i];
      g.r2_prefix = dp->r2_prefix[i];
      rg_to_dp[(size_t)i] = (int)dp_groups.size();
      dp_groups.push_back(g);
    }
  }

  // read_group_index_map::index (alignment.h:545-559); the @RG count is the number of read file sets
  uint32_t read_group_index(int64_t r) const {
    const int32_t n = args->dp->n_sets;
    if (n <= 1) return 0;
    const int32_t rg = A.a->rg ? A.a->rg[r] : -1;
    if (rg < 0 || rg >= n) return 0;
    return (uint32_t)rg;
  }

  static string dp_descriptor(const DpRead& r) {   // :1289-1292
    return r.key + "__" + std::to_string(r.start_pos) + "__" + std::to_string(r.end_pos);
  }

  // DP "enter" step for one pileup entry at insert slot 0 (:1399-1457).  Returns false where breseq
  // leaves the per-read loop (`continue`) for this read.
  bool dp_enter(uint32_t position, int64_t r) {
    if (dp_groups.empty()) return true;
    if ((uint32_t)A.a->pos[r] + 1 != position) return true;
    const uint16_t flag = A.a->flag[r];
    const bool is_paired = flag & 0x1, unmapped = flag & 0x4, proper = flag & 0x2;
    if (!(is_paired && !unmapped && !proper && !(flag & (0x100 | 0x800 | 0x8)))) return true;
    const uint32_t rg = read_group_index(r);
    if (rg >= rg_to_dp.size()) return true;
    const int gi = rg_to_dp[rg];
    if (gi < 0) return true;
    const DpGroup& g = dp_groups[(size_t)gi];
    const int s = A.reversed(r) ? 1 : 0;
    const int8_t xp = A.a->xp ? A.a->xp[r] : -1;
    int o;
    if (xp == 0) o = 0;
    else if (xp == 1) o = 1;
    else if (xp == 2) o = 2;
    else return false;
    const int bin = s * 3 + o;
    const bool focal_is_r1 = !(flag & 0x80);
    const string fname(A.a->names + A.a->name_off[r], A.a->names + A.a->name_off[r + 1]);
    const size_t colon = fname.find(':');
    string read1_name, read2_name;
    if (colon != string::npos) {
      const string counter = fname.substr(colon + 1);
      read1_name = std::to_string(g.r1_prefix) + ":" + counter;
      read2_name = std::to_string(g.r2_prefix) + ":" + counter;
    } else {
      read1_name = focal_is_r1 ? fname : "";
      read2_name = focal_is_r1 ? "" : fname;
    }
    const int32_t isize = A.a->tlen ? A.a->tlen[r] : 0;
    const int64_t r1_insert_size = focal_is_r1 ? (int64_t)isize : -(int64_t)isize;
    DpRead dr;
    dr.start_pos = position;
    dr.end_pos = (uint32_t)A.endpos(r);   // reference_end_1 = bam_endpos
    dr.key = read1_name + "__" + read2_name + "__" + std::to_string(r1_insert_size);
    dr.redundant = A.redundancy(r) > 1;
    dp_groups[(size_t)gi].reads[bin].push_back(dr);
    ++dp_metric[bin];
    if (dp_start[bin] != UNDEF) {
      dp_desc[bin].push_back(dp_descriptor(dr));
      if (dr.redundant) ++dp_red[bin];
    }
    return true;
  }

  // check_discordant_completion (:2358-2422)
  void dp_check(int32_t tid, uint32_t position) {
    static const char* kOrientName[3] = {"FR", "RF", "FF"};
    const uint32_t L = (uint32_t)args->ref->lengths[tid];
    for (int bin = 0; bin < 6; bin++) {
      for (DpGroup& g : dp_groups) {
        while (!g.reads[bin].empty() &&
               (static_cast<double>(position) - static_cast<double>(g.reads[bin].front().start_pos) >= g.window_width)) {
          g.reads[bin].pop_front();
          --dp_metric[bin];
        }
      }
      const bool above = (dp_metric[bin] >= args->dp->seed) && (position <= L);
      if (above) {
        if (dp_start[bin] == UNDEF) {
          dp_start[bin] = position;
          dp_max[bin] = (uint32_t)dp_metric[bin];
          dp_desc[bin].clear();
          dp_red[bin] = 0;
          for (DpGroup& g : dp_groups)
            for (const DpRead& r : g.reads[bin]) {
              dp_desc[bin].push_back(dp_descriptor(r));
              if (r.redundant) ++dp_red[bin];
            }
        } else if ((uint32_t)dp_metric[bin] > dp_max[bin]) {
          dp_max[bin] = (uint32_t)dp_metric[bin];
        }
      } else if (dp_start[bin] != UNDEF) {
        DpRegion R;
        R.seq_id = args->ref->names[tid];
        R.start = dp_start[bin];
        R.end = position - 1;
        R.strand = (bin / 3 == 0) ? 'F' : 'R';
        R.orientation = kOrientName[bin % 3];
        R.max_count = dp_max[bin];
        R.redundant = (2 * dp_red[bin] > dp_desc[bin].size());
        for (size_t k = 0; k < dp_desc[bin].size(); k++) {
          if (k) R.pairs += ";";
          R.pairs += dp_desc[bin][k];
        }
        dp_regions.push_back(R);
        dp_start[bin] = UNDEF;
        dp_max[bin] = 0;
        dp_desc[bin].clear();
        dp_red[bin] = 0;
      }
    }
  }

  void dp_reset() {   // at_target_start (:2062-2068)
    for (int bin = 0; bin < 6; bin++) {
      for (DpGroup& g : dp_groups) g.reads[bin].clear();
      dp_metric[bin] = 0;
      dp_start[bin] = UNDEF;
      dp_max[bin] = 0;
      dp_desc[bin].clear();
    }
  }

  // write_dp_candidate_regions (:2550-2573)
  string write_dp() const {
    vector<DpRegion> v = dp_regions;
    std::sort(v.begin(), v.end(), [](const DpRegion& a, const DpRegion& b) {
      if (a.seq_id != b.seq_id) return a.seq_id < b.seq_id;
      if (a.start != b.start) return a.start < b.start;
      if (a.strand != b.strand) return a.strand < b.strand;
      return a.orientation < b.orientation;
    });
    std::ostringstream os;
    os << "seq_id,start,end,strand,orientation,length,max_discordant_count,redundant,discordant_pairs" << "\n";
    for (const DpRegion& r : v)
      os << r.seq_id << "," << r.start << "," << r.end << "," << r.strand << "," << r.orientation << ","
         << (r.end - r.start + 1) << "," << r.max_count << "," << (r.redundant ? 1 : 0) << "," << r.pairs << "\n";
    return os.str();
  }

  void add(GdEntry e) {
    e.id = ++next_id;
    gd.push_back(std::move(e));
  }

  double get_prob(uint32_t rs, uint32_t refb, uint32_t obs, uint32_t q) const {
    const uint32_t S = args->read_set_max, Q = args->quality_max;
    if (rs >= S) fail(BSQ_ERR_DATA, "Covariate 'read_set' exceeded its maximum");
    if (q >= Q) fail(BSQ_ERR_DATA, "Covariate 'quality' exceeded its maximum");
    return prob[rs + S * refb + 5 * S * obs + 25 * S * q];
  }

  void fill_likelihoods(PData& pd) const {   // identify_mutations.cpp:3359-3384
    const double incorrect = pow(10, -static_cast<double>(pd.mq) / 10);
    const double correct = 1 - incorrect;
    const double uniform = 1.0 / 5.0;
    uint32_t obs = pd.obs;
    const bool top = (pd.strand == 1);
    if (!top) obs = complement_index(obs);
    pd.lpmax = -std::numeric_limits<double>::max();
    for (int b = 0; b < 5; b++) {
      uint32_t refb = top ? b : complement_index(b);
      double pr = correct * get_prob(pd.rs, refb, obs, pd.q) + incorrect * uniform;
      if (pr < 0.0) pr = 0.0;
      pd.lp[b] = log10(pr);
      pd.lpmax = std::max(pd.lpmax, pd.lp[b]);
    }
    for (int b = 0; b < 5; b++) pd.r[b] = pow(10, pd.lp[b] - pd.lpmax);
  }

  void pure_genotype_call(const double lps[5], uint32_t n, char& genotype, double& score) const {
    genotype = 'N';
    score = std::numeric_limits<double>::quiet_NaN();
    if (n == 0) return;
    int best = 0;
    for (int b = 1; b < 5; b++) if (lps[b] > lps[best]) best = b;
    genotype = kIndexChar[best];
    double off = -std::numeric_limits<double>::max();
    for (int b = 0; b < 5; b++) if (b != best) off = std::max(off, lps[b]);
    double tot = 0;
    for (int b = 0; b < 5; b++) if (b != best) tot += pow(10, lps[b] - off);
    double lt = log10(tot);
    lt += off;
    score = lps[best] - lt;
  }

  AlleleModel fit(const vector<PData>& pdata, const bool allowed[5]) const {   // :3240-3318
    AlleleModel m;
    m.n = (uint32_t)pdata.size();
    if (m.n == 0) return m;
    double init_total = 0.0;
    uint32_t n_allowed = 0;
    for (int b = 0; b < 5; b++) {
      if (!allowed[b]) continue;
      n_allowed++;
      m.f[b] = 0.5;
    }
    if (n_allowed == 0) return m;
    for (const PData& p : pdata) {
      int obs = basechar2index(p.base);
      if (obs < 5 && allowed[obs]) m.f[obs] += 1.0;
    }
    for (int b = 0; b < 5; b++) init_total += m.f[b];
    for (int b = 0; b < 5; b++) m.f[b] /= init_total;
    const uint32_t kmax = 50;
    const double tol = args->polymorphism_precision_decimal;
    for (m.iterations = 1; m.iterations <= kmax; m.iterations++) {
      double sum_w[5] = {0, 0, 0, 0, 0};
      double ll = 0.0;
      for (const PData& p : pdata) {
        double s = 0.0;
        for (int b = 0; b < 5; b++) if (allowed[b]) s += m.f[b] * p.r[b];
        if (s > 0.0) {
          ll += log10(s) + p.lpmax;
          for (int b = 0; b < 5; b++) {
            if (!allowed[b]) continue;
            sum_w[b] += m.f[b] * p.r[b] / s;
          }
        } else {
          for (int b = 0; b < 5; b++) if (allowed[b]) sum_w[b] += m.f[b];
        }
      }
      double max_delta = 0.0;
      for (int b = 0; b < 5; b++) {
        if (!allowed[b]) continue;
        double fn = sum_w[b] / static_cast<double>(m.n);
        max_delta = std::max(max_delta, fabs(fn - m.f[b]));
        m.f[b] = fn;
      }
      for (int b = 0; b < 5; b++) m.sum_w[b] = sum_w[b];
      m.log10_likelihood = ll;
      if (max_delta < tol) break;
    }
    if (m.iterations > kmax) m.iterations = kmax;
    return m;
  }

  double variant_presence_score(const vector<PData>& pdata, const AlleleModel& full, int v) const {
    if (full.n == 0 || v >= 5) return std::numeric_limits<double>::quiet_NaN();
    bool without[5];
    for (int b = 0; b < 5; b++) without[b] = (b != v);
    AlleleModel nf = fit(pdata, without);
    return (full.log10_likelihood - nf.log10_likelihood) - log10_ref_length;
  }

  double profile(const vector<PData>& pdata, const AlleleModel& full, int v, double f_fixed) const {   // :3094-3143
    if (full.n == 0 || v >= 5) return 0.0;
    double f[5];
    double other = 0.0;
    for (int b = 0; b < 5; b++) if (b != v) other += full.f[b];
    for (int b = 0; b < 5; b++) {
      if (b == v) f[b] = f_fixed;
      else f[b] = (other > 0.0) ? (1.0 - f_fixed) * full.f[b] / other : (1.0 - f_fixed) / static_cast<double>(5 - 1);
    }
    double ll = 0.0;
    for (uint32_t it = 0; it < 50; it++) {
      double sum_w[5] = {0, 0, 0, 0, 0};
      ll = 0.0;
      for (const PData& p : pdata) {
        double s = 0.0;
        for (int b = 0; b < 5; b++) s += f[b] * p.r[b];
        if (s > 0.0) {
          ll += log10(s) + p.lpmax;
          for (int b = 0; b < 5; b++) sum_w[b] += f[b] * p.r[b] / s;
        } else {
          for (int b = 0; b < 5; b++) sum_w[b] += f[b];
        }
      }
      double others = 0.0;
      for (int b = 0; b < 5; b++) if (b != v) others += sum_w[b];
      double max_delta = 0.0;
      for (int b = 0; b < 5; b++) {
        if (b == v) continue;
        double fn = (others > 0.0) ? (1.0 - f_fixed) * sum_w[b] / others : (1.0 - f_fixed) / static_cast<double>(5 - 1);
        max_delta = std::max(max_delta, fabs(fn - f[b]));
        f[b] = fn;
      }
      if (max_delta < args->polymorphism_precision_decimal) break;
    }
    return ll;
  }

  void frequency_bounds(Fields& mut, const vector<PData>& pdata, const AlleleModel& am, int v) const {   // :3175-3217
    const double kDrop = 0.587566;
    const double prec = args->polymorphism_precision_decimal;
    double lower = 0.0, upper = 1.0;
    if (am.n > 0 && v < 5) {
      const double fh = am.f[v];
      const double plmax = profile(pdata, am, v, fh);
      const double target = plmax - kDrop;
      if (profile(pdata, am, v, 0.0) >= target) {
        lower = 0.0;
      } else {
        double lo = 0.0, hi = fh;
        for (uint32_t i = 0; (i < 40) && ((hi - lo) > prec); i++) {
          double mid = 0.5 * (lo + hi);
          if (profile(pdata, am, v, mid) >= target) hi = mid;
          else lo = mid;
        }
        lower = hi;
      }
      if (profile(pdata, am, v, 1.0) >= target) {
        upper = 1.0;
      } else {
        double lo = fh, hi = 1.0;
        for (uint32_t i = 0; (i < 40) && ((hi - lo) > prec); i++) {
          double mid = 0.5 * (lo + hi);
          if (profile(pdata, am, v, mid) >= target) lo = mid;
          else hi = mid;
        }
        upper = lo;
      }
    }
    unsigned places = args->polymorphism_precision_places;
    mut["frequency_lower"] = fmt_double(lower, places, true);
    mut["frequency_upper"] = fmt_double(upper, places, true);
  }

  void check_deletion_completion(const string& seq_id, uint32_t position, const Coverage& c) {   // :2262-2344
    if (position == 1) last_cov = Coverage::nan();
    if (c.unique[1] <= prop_cutoff) {
      if (del_start == UNDEF) {
        del_start = position;
        left_outside = last_cov;
        left_inside = c;
      }
    }
    if (!std::isnan(c.unique[1]) && (c.total <= seed_cutoff)) reaches_seed = true;
    if (del_start != UNDEF && (std::isnan(c.unique[1]) || c.unique[1] > prop_cutoff)) {
      if (reaches_seed) {
        del_end = position - 1;
        if (red_end == UNDEF) red_end = del_end;
        if (red_start == UNDEF) red_start = del_start;
        GdEntry e;
        e.type = 1;
        e.seq_id = seq_id;
        e.pos = del_start;
        e.end = del_end;
        e.start_range = red_start - del_start;
        e.end_range = del_end - red_end;
        e.fields["left_outside_cov"] = fmt_double(left_outside.unique[1], 0, false);
        e.fields["left_inside_cov"] = fmt_double(left_inside.unique[1], 0, false);
        e.fields["right_inside_cov"] = fmt_double(last_cov.unique[1], 0, false);
        e.fields["right_outside_cov"] = fmt_double(c.unique[1], 0, false);
        add(e);
      }
      reaches_seed = false;
      red_reached_zero = false;
      del_start = del_end = red_start = red_end = UNDEF;
    }
    if (del_start != UNDEF) {
      if (c.redundant[1] == 0) {
        red_reached_zero = true;
        red_end = UNDEF;
      } else if (c.redundant[1] > 0) {
        if (!red_reached_zero) red_start = position;
        else if (red_end == UNDEF) red_end = position;
      }
    }
    last_cov = c;
  }

  void update_unknown(const string& seq_id, uint32_t position, bool predicted) {   // :2972-3007
    if (!predicted) {
      if (un_start == UNDEF) un_start = position;
    } else if (un_start != UNDEF) {
      GdEntry e;
      e.type = 2;
      e.seq_id = seq_id;
      e.pos = un_start;
      e.end = position - 1;
      add(e);
      un_start = UNDEF;
    }
  }

  bool position_to_covariates(const Entry& e, int32_t insert_count, PData& pd) const {   // error_count.cpp:1049-1105
    int64_t r = e.read;
    int indel = e.indel < 0 ? 0 : e.indel;
    if (e.is_del) indel = -1;
    uint8_t read_base = (indel < insert_count) ? (uint8_t)'.' : A.base(r, (int64_t)e.qpos + insert_count);
    if (bam_is_N(read_base)) return false;
    uint32_t qb_s, qb_e;
    A.query_bounds_0(r, qb_s, qb_e);
    uint32_t q_pos_0 = (uint32_t)e.qpos;
    uint32_t rev = A.reversed(r) ? 1 : 0;
    if (indel == -1) {
      q_pos_0 += 1 - rev;
      if (bam_is_N(A.base(r, q_pos_0))) return false;
    } else if (insert_count > 0) {
      int32_t max_offset = insert_count;
      if (indel < max_offset) max_offset = indel;
      q_pos_0 += max_offset + 1 - rev;
      if (q_pos_0 > qb_e) return false;
      if (bam_is_N(A.base(r, q_pos_0))) return false;
    }
    pd.obs = basebam2index(read_base);
    pd.q = A.qual(r, q_pos_0);
    pd.rs = read_file_index(args->read_files, A, r);
    return true;
  }

  void callback(int32_t tid, int64_t p0, const vector<Entry>& col) {   // :1309-2022
    const string seq_id = args->ref->names[tid];
    uint32_t position = (uint32_t)(p0 + 1);
    prop_cutoff = args->deletion_propagation_cutoff[tid];
    seed_cutoff = args->deletion_seed_cutoff[tid];
    if (prop_cutoff < 0.0) return;
    int32_t insert_count = -1;
    bool next_exists = true;
    while (next_exists) {
      ++insert_count;
      next_exists = false;
      char ref_base_char = '.';
      if (!insert_count) ref_base_char = ref_char(args->ref, tid, p0);
      uint32_t pos_info[6][3];
      memset(pos_info, 0, sizeof(pos_info));
      Coverage cov;
      vector<PData> pdata;
      double lps[5] = {0, 0, 0, 0, 0};
      for (const Entry& e : col) {
        if (insert_count == 0 && !dp_enter(position, e.read)) continue;
        int64_t r = e.read;
        int indel = e.indel < 0 ? 0 : e.indel;
        if (e.is_del) indel = -1;
        uint8_t read_base = '.';
        bool past = true;
        if (indel >= insert_count) {
          read_base = A.base(r, (int64_t)e.qpos + insert_count);
          past = false;
        }
        if (bam_is_N(read_base)) continue;
        int32_t redundancy = (int32_t)A.redundancy(r);
        int strand = A.reversed(r) ? -1 : 1;
        bool trimmed = A.is_trimmed(r, e.qpos, past);
        if (redundancy == 1) {
          ++cov.unique[1 + strand];
          if (indel > insert_count) next_exists = true;
        } else {
          cov.redundant[1 + strand] += 1.0 / redundancy;
          ++cov.raw_redundant[1 + strand];
        }
        if (redundancy > 1) continue;
        if (trimmed) continue;
        PData pd;
        if (!position_to_covariates(e, insert_count, pd)) continue;
        if ((int32_t)pd.q < args->base_quality_cutoff) continue;
        ++pos_info[basechar2index(kIndexChar[pd.obs])][1 + strand];
        pd.base = kIndexChar[pd.obs];
        pd.quality = (uint8_t)pd.q;
        pd.strand = strand;
        pd.mq = A.a->mapq[r];
        fill_likelihoods(pd);
        pdata.push_back(pd);
        for (int j = 0; j < 5; j++) lps[j] += pdata.back().lp[j];
      }
      cov.sum();
      char best;
      double score;
      pure_genotype_call(lps, (uint32_t)pdata.size(), best, score);
      double cbs = score - log10_ref_length;
      bool base_predicted = (cbs >= args->mutation_cutoff);
      if (insert_count == 0) pos_scores.push_back(cbs);
      int total_cov[3] = {0, 0, 0};
      for (int j = 0; j < 5; j++) {
        double top = pos_info[j][2], bot = pos_info[j][0];
        total_cov[2] += (int)round(top);
        total_cov[0] += (int)round(bot);
      }
      if (insert_count == 0 && !args->skip_missing_coverage_prediction) check_deletion_completion(seq_id, position, cov);
      bool passed_poly = false;
      bool passed_cons = (best != ref_base_char) && (!std::isnan(cbs) && (cbs > 0));
      bool all[5] = {true, true, true, true, true};
      AlleleModel am = fit(pdata, all);
      const int ref_idx = (ref_base_char == 'N') ? 5 : basechar2index(ref_base_char);
      const int major = am.major();
      const int minor = am.next(major);
      const int variant = am.next(ref_idx);
      double variant_score = std::numeric_limits<double>::quiet_NaN();
      if (variant != 5) {
        variant_score = variant_presence_score(pdata, am, variant);
        if (variant_score >= args->polymorphism_cutoff) passed_poly = true;
      }
      if (insert_count == 0) update_unknown(seq_id, position, base_predicted);
      if (passed_cons || passed_poly) {
        const unsigned places = args->polymorphism_precision_places;
        GdEntry e;
        e.type = 0;
        e.seq_id = seq_id;
        e.pos = position;
        e.ins = insert_count;
        e.ref_base = string(1, ref_base_char);
        e.new_base = string(1, kIndexChar[variant]);
        Fields& m = e.fields;
        m["score"] = fmt_double(variant_score, 1, false);
        m["major_base"] = string(1, kIndexChar[major]);
        m["minor_base"] = string(1, kIndexChar[minor]);
        m["major_frequency"] = fmt_double(am.reported(major), places, true);
        m["frequency"] = fmt_double(am.reported(variant), places, true);
        m["allele_frequencies"] = am.spectrum(places);
        frequency_bounds(m, pdata, am, variant);
        // annotate_polymorphism_statistics (:3009-3033)
        {
          const char mj = kIndexChar[major], mn = kIndexChar[minor];
          uint32_t mjt = pos_info[major][2], mjb = pos_info[major][0];
          uint32_t mnt = pos_info[minor][2], mnb = pos_info[minor][0];
          vector<double> majq, minq;
          for (const PData& p : pdata) {
            if (p.base == mj) majq.push_back((double)p.quality);
            if (p.base == mn) minq.push_back((double)p.quality);
          }
          double ks = 1.0;
          if (!majq.empty() && !minq.empty()) ks = ks_less(minq, majq);
          double fs = fisher_2x2(mnt, mnb, mjt, mjb);
          m["ks_quality_p_value"] = fmt_double(ks, 5, true);
          m["fisher_strand_p_value"] = fmt_double(fs, 5, true);
        }
        auto cov_str = [&](int idx) {
          return std::to_string((int32_t)pos_info[idx][2]) + "/" + std::to_string((int32_t)pos_info[idx][0]);
        };
        m["ref_cov"] = cov_str(basechar2index(ref_base_char));
        m["new_cov"] = cov_str(variant);
        m["major_cov"] = cov_str(major);
        m["minor_cov"] = cov_str(minor);
        m["total_cov"] = std::to_string(total_cov[2]) + "/" + std::to_string(total_cov[0]);
        add(e);
      }
    }
    if (!dp_groups.empty()) dp_check(tid, position);
  }

  void run() {
    const bsq_reference* ref = args->ref;
    if (args->n_bins != (int64_t)args->read_set_max * 25 * args->quality_max) fail(BSQ_ERR_ARG, "error table size mismatch");
    prob.resize(args->n_bins);
    for (int64_t i = 0; i < args->n_bins; i++) prob[i] = pow(10, args->log10_prob_table[i]);
    double total_len = 0;
    for (int32_t t = 0; t < ref->n_targets; t++) total_len += (double)ref->lengths[t];
    log10_ref_length = log10(total_len);
    for (int32_t tid : target_order(ref, args->call_target)) {
      del_start = del_end = red_start = red_end = UNDEF;
      un_start = UNDEF;
      const int64_t L = ref->lengths[tid];
      dp_reset();
      for_each_column(A, tid, L, [&](int64_t p, const vector<Entry>& col) { callback(tid, p, col); });
      if (!dp_groups.empty()) dp_check(tid, (uint32_t)L + 1);
      // at_target_end (:2117-2164)
      const string seq_id = ref->names[tid];
      prop_cutoff = args->deletion_propagation_cutoff[tid];
      seed_cutoff = args->deletion_seed_cutoff[tid];
      if (!args->skip_missing_coverage_prediction) check_deletion_completion(seq_id, (uint32_t)L + 1, Coverage::nan());
      update_unknown(seq_id, (uint32_t)L + 1, true);
      if (!args->skip_missing_coverage_prediction && prop_cutoff < 0.0) {
        GdEntry e;
        e.type = 1;
        e.seq_id = seq_id;
        e.pos = 1;
        e.end = (uint32_t)L;
        e.start_range = 0;
        e.end_range = 0;
        e.fields["left_outside_cov"] = "NA";
        e.fields["left_inside_cov"] = fmt_double(0.0, 0, false);
        e.fields["right_inside_cov"] = fmt_double(0.0, 0, false);
        e.fields["right_outside_cov"] = "NA";
        add(e);
      }
    }
  }

  // cGenomeDiff::write (genome_diff.cpp:685-770) with sort (genome_diff_entry.cpp:566-700)
  string write() const {
    vector<const GdEntry*> v;
    for (const GdEntry& e : gd) v.push_back(&e);
    static const int group[3] = {3, 4, 7};
    std::stable_sort(v.begin(), v.end(), [](const GdEntry* a, const GdEntry* b) {
      if (group[a->type] != group[b->type]) return group[a->type] < group[b->type];
      if (a->seq_id != b->seq_id) return a->seq_id < b->seq_id;
      if (a->pos != b->pos) return a->pos < b->pos;
      if (a->type == 0) {
        if (a->ins != b->ins) return a->ins < b->ins;
        if (a->ref_base != b->ref_base) return a->ref_base < b->ref_base;
        if (a->new_base != b->new_base) return a->new_base < b->new_base;
      } else {
        if (a->end != b->end) return a->end < b->end;
        if (a->type == 1) {
          if (a->start_range != b->start_range) return a->start_range < b->start_range;
          if (a->end_range != b->end_range) return a->end_range < b->end_range;
        }
      }
      return a->id < b->id;
    });
    std::ostringstream os;
    os << "#=GENOME_DIFF\t1.0\n";
    for (const GdEntry* e : v) {
      static const char* tn[3] = {"RA", "MC", "UN"};
      os << tn[e->type] << "\t" << e->id << "\t.\t" << e->seq_id << "\t";
      if (e->type == 0) {
        os << e->pos << "\t" << e->ins << "\t" << e->ref_base << "\t" << e->new_base;
      } else if (e->type == 1) {
        os << e->pos << "\t" << e->end << "\t" << e->start_range << "\t" << e->end_range;
      } else {
        os << e->pos << "\t" << e->end;
      }
      for (const auto& kv : e->fields) {
        if (kv.second.empty() || kv.first[0] == '_') continue;
        os << "\t" << kv.first << "=" << kv.second;
      }
      os << "\n";
    }
    return os.str();
  }
};

/* ---------------------------------------------------------------- SequenceTrims (calculate_trims.cpp:44-138) */

static bool repeat_match(const char* s, uint32_t len, uint32_t pos, uint32_t size, uint32_t num) {
  const uint32_t off = size * num;
  for (uint32_t i = 0; i < size; i++) {
    if (pos + i >= len) return false;
    if (pos + i + off >= len) return false;
    if (s[pos + i] != s[pos + i + off]) return false;
  }
  return true;
}

static void sequence_trims(const char* s, uint32_t len, uint8_t* t) {
  memset(t, 0, 2 * (size_t)len);
  for (uint32_t p = 0; p < len; p++) {
    uint32_t mx = 1;
    for (uint32_t size = 1; size <= 18; size++) {
      uint32_t num = 1;
      while (repeat_match(s, len, p, size, num)) num++;
      if (num > 1) mx = std::max(mx, num * size);
    }
    if (mx > 255) mx = 255;
    uint8_t m8 = (uint8_t)mx;
    for (int32_t o = 0; o < m8; o++) {
      if (p + o >= len) continue;
      uint8_t v = (uint8_t)(o + 1);
      t[len + p + o] = std::max(v, t[len + p + o]);
    }
    for (int32_t o = m8; o >= 0; o--) {
      if (p + o >= len) continue;
      uint8_t v = (uint8_t)(m8 - o);
      t[p + o] = std::max(v, t[p + o]);
    }
  }
}

// basebam2char (common.h:263-275): the four bases by their BAM codes, anything else 'N'
static char bam_char(uint8_t b) {
  switch (b) {
    case 1: return 'A';
    case 2: return 'C';
    case 4: return 'G';
    case 8: return 'T';
    default: return 'N';
  }
}

// reverse_complement (fastq.h:271-277 with the table at fastq.cpp:1635-1651): A<->T, C<->G,
// every other character 'N'
static string reverse_complement(const string& s) {
  string r(s.size(), 'N');
  for (size_t i = 0; i < s.size(); i++) {
    const char c = s[s.size() - 1 - i];
    r[i] = c == 'A' ? 'T' : c == 'T' ? 'A' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'N';
  }
  return r;
}

// edge_trims_for_sequence (calculate_trims.cpp:183-230): the longest perfect tandem repeat of unit
// 1..18 starting at the first base (left) and at the first base of the reverse complement (right),
// at least 1 on each side
static void edge_trims(const string& seq, uint32_t& L, uint32_t& R) {
  L = R = 1;
  const string rc = reverse_complement(seq);
  for (int side = 0; side < 2; side++) {
    const string& s = side ? rc : seq;
    uint32_t& t = side ? R : L;
    for (uint32_t size = 1; size <= 18; size++) {
      uint32_t num = 1;
      while (repeat_match(s.data(), (uint32_t)s.size(), 0, size, num)) num++;
      if (num > 1) t = std::max(t, num * size);
    }
  }
}

}  // namespace orc

/* ================================================================ C API */

namespace orc {

// BAM2COV table of one region, restating coverage_output::table (coverage_output.cpp:190-283),
// coverage_output::pileup_callback (:307-503) and the traversal of pileup_base::do_pileup with its
// first-level callback (pileup_base.cpp:118-210, 308-360), step by step.
static string coverage_table(const bsq_reference* ref, const bsq_alignments* aln, const bsq_coverage_table_args* a) {
  const Aln A{aln};
  const int32_t tid = a->target;
  if (tid < 0 || tid >= ref->n_targets) fail(BSQ_ERR_ARG, "target id out of range");
  const int64_t L = ref->lengths[tid];
  uint32_t start = (uint32_t)a->start_1, end = (uint32_t)a->end_1;
  if (start == 0 && end == 0) {
    start = 1;
    end = (uint32_t)L;
  }
  if (start < 1 || (int64_t)end > L || start > end) fail(BSQ_ERR_ARG, "region outside its reference sequence");
  const char* refseq = ref->seqs[tid];
  // coverage_output::table: downsampling step from the resolution
  const uint32_t size = end - start + 1;
  uint32_t downsample = 1;
  if (a->resolution != 0) {
    downsample = static_cast<uint32_t>(floor(static_cast<double>(size) / static_cast<double>(a->resolution)));
    if (downsample < 1) downsample = 1;
  }
  const string d = a->csv ? "," : "\t";
  vector<string> cols;
  if (a->total_only) cols = {"unique_cov", "redundant_cov", "total_cov"};
  else
    cols = {"unique_top_cov", "unique_bot_cov", "redundant_top_cov", "redundant_bot_cov", "raw_redundant_top_cov",
            "raw_redundant_bot_cov", "unique_top_begin", "unique_bot_begin"};
  const size_t n_rg_cols = (size_t)std::max(a->n_read_groups, 1);
  std::ostringstream out;
  out << "position" << d << "ref_base";
  for (size_t c = 0; c < cols.size(); c++) out << d << cols[c];
  if (a->per_read_group)
    for (size_t g = 0; g < n_rg_cols; g++)
      for (size_t c = 0; c < cols.size(); c++) out << d << "RG-" << g << "_" << cols[c];
  out << std::endl;
  double avg_unique = 0, avg_repeat = 0, avg_cov = 0;
  uint32_t num_positions = 0;
  vector<double> rg_unique(n_rg_cols, 0.0), rg_repeat(n_rg_cols, 0.0), rg_cov(n_rg_cols, 0.0);
  // the reads sam_itr_queryi(tid, start-1, end) returns, and the columns they give the pileup
  vector<int64_t> fetched;
  for (int64_t r = 0; r < A.n(); r++) {
    if (A.a->tid[r] != tid || (A.a->flag[r] & 0x4) || A.n_cigar(r) == 0) continue;
    const int64_t p0 = A.a->pos[r], e0 = A.endpos(r);
    if (e0 <= p0) continue;
    if (p0 < (int64_t)end && e0 > (int64_t)start - 1) fetched.push_back(r);
  }
  vector<vector<Entry>> entries((size_t)L + 2);   // by 1-based position, fetched reads in file order
  vector<char> covered((size_t)L + 2, 0);
  for (int64_t r : fetched) {
    ResolvedRead rr;
    resolve_read(A, r, rr);
    for (int64_t p = rr.pos; p < rr.end; p++) {
      const int64_t j = p - rr.pos;
      entries[(size_t)(p + 1)].push_back(Entry{r, rr.qpos[j], rr.indel[j], rr.is_del[j] != 0});
      covered[(size_t)(p + 1)] = 1;
    }
  }
  auto rg_index = [&](int64_t r) -> size_t {   // read_group_index_map::index (alignment.h:545-559)
    if (a->n_read_groups <= 1) return 0;
    const int32_t g = A.a->rg ? A.a->rg[r] : -1;
    return (g < 0 || g >= a->n_read_groups) ? 0 : (size_t)g;
  };
  uint32_t last_position_1 = 0;   // pileup_base::m_last_position_1
  auto handle_position = [&](uint32_t pos_1) {   // pileup_base.cpp:118-135 with clipping on
    if (pos_1 < start) return false;
    if (pos_1 > end) return false;
    if ((pos_1 + start) % downsample != 0) return false;
    return true;
  };
  const size_t ncc = cols.size();
  auto callback = [&](uint32_t pos) {   // coverage_output::pileup_callback
    if (pos == 0) return;
    for (uint32_t i = last_position_1 + 1; i < pos; i++) {
      out << i << d << refseq[i - 1];
      const size_t zero_columns = ncc * (a->per_read_group ? (1 + n_rg_cols) : 1);
      for (size_t c = 0; c < zero_columns; c++) out << d << 0;
      out << std::endl;
      num_positions++;
    }
    const uint8_t ref_base = (uint8_t)refseq[pos - 1];
    uint32_t unique_cov[2] = {0, 0}, raw_redundant_cov[2] = {0, 0}, unique_begin_reads[2] = {0, 0};
    double redundant_cov[2] = {0.0, 0.0};
    const size_t num_rg = a->per_read_group ? n_rg_cols : 0;
    vector<uint32_t> rg_unique_cov(num_rg * 2, 0), rg_raw(num_rg * 2, 0), rg_begin(num_rg * 2, 0);
    vector<double> rg_red(num_rg * 2, 0.0);
    for (const Entry& e : entries[pos]) {
      if (e.is_del) continue;
      const int64_t r = e.read;
      const uint32_t redundancy = A.redundancy(r);
      const uint32_t reversed = A.reversed(r) ? 1 : 0;
      const size_t slot = a->per_read_group ? (rg_index(r) * 2 + reversed) : 0;
      const uint32_t qpos_1 = (uint32_t)e.qpos + 1;
      const bool this_is_first_base = !reversed ? (qpos_1 == 1) : (qpos_1 == A.len(r));
      if (redundancy == 1) {
        unique_cov[reversed]++;
        if (this_is_first_base) unique_begin_reads[reversed]++;
        if (a->per_read_group) {
          rg_unique_cov[slot]++;
          if (this_is_first_base) rg_begin[slot]++;
        }
      } else {
        raw_redundant_cov[reversed]++;
        redundant_cov[reversed] += 1.0 / redundancy;
        if (a->per_read_group) {
          rg_raw[slot]++;
          rg_red[slot] += 1.0 / redundancy;
        }
      }
    }
    num_positions++;
    avg_unique += unique_cov[0] + unique_cov[1];
    avg_repeat += redundant_cov[0] + redundant_cov[1];
    avg_cov += unique_cov[0] + unique_cov[1] + redundant_cov[0] + redundant_cov[1];
    if (a->total_only) {
      out << pos << d << ref_base << d << (unique_cov[0] + unique_cov[1]) << d << (redundant_cov[0] + redundant_cov[1])
          << d << (unique_cov[0] + unique_cov[1] + redundant_cov[0] + redundant_cov[1]);
    } else {
      out << pos << d << ref_base << d << unique_cov[0] << d << unique_cov[1] << d << redundant_cov[0] << d
          << redundant_cov[1] << d << raw_redundant_cov[0] << d << raw_redundant_cov[1] << d << unique_begin_reads[0]
          << d << unique_begin_reads[1];
    }
    for (size_t g = 0; g < num_rg; g++) {
      const size_t t = g * 2, b = g * 2 + 1;
      if (a->total_only) {
        out << d << (rg_unique_cov[t] + rg_unique_cov[b]) << d << (rg_red[t] + rg_red[b]) << d
            << (rg_unique_cov[t] + rg_unique_cov[b] + rg_red[t] + rg_red[b]);
      } else {
        out << d << rg_unique_cov[t] << d << rg_unique_cov[b] << d << rg_red[t] << d << rg_red[b] << d << rg_raw[t] << d
            << rg_raw[b] << d << rg_begin[t] << d << rg_begin[b];
      }
      rg_unique[g] += rg_unique_cov[t] + rg_unique_cov[b];
      rg_repeat[g] += rg_red[t] + rg_red[b];
      rg_cov[g] += rg_unique_cov[t] + rg_unique_cov[b] + rg_red[t] + rg_red[b];
    }
    out << std::endl;
  };
  // first_level_pileup_callback over the pileup's columns (those the fetched reads cover)
  for (uint32_t pos = 1; pos <= (uint32_t)L; pos++) {
    if (!covered[pos]) continue;
    for (uint32_t on_pos_1 = last_position_1 + 1; on_pos_1 <= pos - 1; on_pos_1++) {
      if (handle_position(on_pos_1)) callback(on_pos_1);
      last_position_1 = on_pos_1;
    }
    if (handle_position(pos)) callback(pos);
    last_position_1 = pos;
  }
  // do_pileup: positions after the last one the pileup handled
  for (uint32_t on_pos_1 = last_position_1 + 1; on_pos_1 <= end; on_pos_1++) {
    if (handle_position(on_pos_1)) {
      callback(on_pos_1);
      last_position_1 = on_pos_1;
    }
  }
  if (a->show_average) out << "#" << d << "reference_unique_average_cov" << d << a->reference_average << std::endl;
  out << "#" << d << "region_unique_average_cov" << d << (avg_unique / num_positions) << std::endl;
  out << "#" << d << "region_repeat_average_cov" << d << (avg_repeat / num_positions) << std::endl;
  out << "#" << d << "region_average_cov" << d << (avg_cov / num_positions) << std::endl;
  out << "#" << d << "number_of_positions" << d << num_positions << std::endl;
  if (a->per_read_group) {
    for (size_t g = 0; g < n_rg_cols; g++) {
      const string p = "RG-" + std::to_string(g) + "_";
      out << "#" << d << p << "region_unique_average_cov" << d << (rg_unique[g] / num_positions) << std::endl;
      out << "#" << d << p << "region_repeat_average_cov" << d << (rg_repeat[g] / num_positions) << std::endl;
      out << "#" << d << p << "region_average_cov" << d << (rg_cov[g] / num_positions) << std::endl;
    }
  }
  return out.str();
}

}  // namespace orc

struct orc_ec_out { orc::EcOut o; };
struct orc_im_out { string text; vector<double> scores; string dp_text; };

static void set_status(bsq_status* st, int code, const char* msg) {
  if (!st) return;
  st->code = code;
  snprintf(st->message, sizeof(st->message), "%s", msg ? msg : "");
}

extern "C" {

int orc_error_count(const bsq_error_count_args* args, orc_ec_out** out, bsq_status* st) {
  try {
    orc_ec_out* o = new orc_ec_out();
    try {
      orc::error_count(args, o->o);
    } catch (...) {
      delete o;
      throw;
    }
    *out = o;
    set_status(st, BSQ_OK, "");
    return BSQ_OK;
  } catch (const orc::orc_error& e) {
    set_status(st, e.code, e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_status(st, BSQ_ERR_DATA, e.what());
    return BSQ_ERR_DATA;
  }
}

int orc_ec_histogram(const orc_ec_out* o, int32_t g, const uint32_t** h, int64_t* len) {
  if (!o || g < 0 || g >= (int32_t)o->o.hist.size()) return BSQ_ERR_ARG;
  *h = o->o.hist[g].data();
  *len = (int64_t)o->o.hist[g].size();
  return BSQ_OK;
}

int orc_ec_counts(const orc_ec_out* o, const uint64_t** c, int64_t* n) {
  *c = o->o.counts.data();
  *n = (int64_t)o->o.counts.size();
  return BSQ_OK;
}

int orc_ec_log10_prob(const orc_ec_out* o, const double** t, int64_t* n) {
  *t = o->o.log10p.data();
  *n = (int64_t)o->o.log10p.size();
  return BSQ_OK;
}

int orc_ec_text(const orc_ec_out* o, int32_t which, int32_t index, const char** text, int64_t* len) {
  const string* s = nullptr;
  if (which == 0) s = &o->o.rates_text;
  else if (which == 1 && index >= 0 && index < (int32_t)o->o.cov_text.size()) s = &o->o.cov_text[index];
  else if (which == 2 && index >= 0 && index < (int32_t)o->o.bq_text.size()) s = &o->o.bq_text[index];
  if (!s) return BSQ_ERR_ARG;
  *text = s->c_str();
  *len = (int64_t)s->size();
  return BSQ_OK;
}

int orc_ec_no_pos_hash_pr(const orc_ec_out* o, const double** v, int64_t* n) {
  *v = o->o.no_pos_hash_pr.data();
  *n = (int64_t)o->o.no_pos_hash_pr.size();
  return 0;
}

void orc_ec_free(orc_ec_out* o) { delete o; }

int orc_read_log10_prob_table(const char* text, double* out, int64_t n_bins) {
  // read_log10_prob_table (error_count.cpp:629-654): skip two lines, last tab field of each row
  const char* p = text;
  for (int skip = 0; skip < 2; skip++) {
    p = strchr(p, '\n');
    if (!p) return BSQ_ERR_FORMAT;
    p++;
  }
  for (int64_t i = 0; i < n_bins; i++) {
    const char* e = strchr(p, '\n');
    if (!e) return BSQ_ERR_FORMAT;
    const char* last = p;
    for (const char* q = p; q < e; q++) if (*q == '\t') last = q + 1;
    string field(last, e);
    out[i] = strtod(field.c_str(), NULL);
    p = e + 1;
  }
  return BSQ_OK;
}

int orc_identify_mutations(const bsq_identify_args* args, orc_im_out** out, bsq_status* st) {
  try {
    orc::Identify im(args);
    im.run();
    orc_im_out* o = new orc_im_out();
    o->text = im.write();
    o->dp_text = im.write_dp();
    o->scores = std::move(im.pos_scores);
    *out = o;
    set_status(st, BSQ_OK, "");
    return BSQ_OK;
  } catch (const orc::orc_error& e) {
    set_status(st, e.code, e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_status(st, BSQ_ERR_DATA, e.what());
    return BSQ_ERR_DATA;
  }
}

int orc_im_dp_text(const orc_im_out* o, const char** text, int64_t* len) {
  *text = o->dp_text.c_str();
  *len = (int64_t)o->dp_text.size();
  return BSQ_OK;
}

int orc_im_gd_text(const orc_im_out* o, const char** text, int64_t* len) {
  *text = o->text.c_str();
  *len = (int64_t)o->text.size();
  return BSQ_OK;
}

int orc_im_position_scores(const orc_im_out* o, const double** s, int64_t* n) {
  *s = o->scores.data();
  *n = (int64_t)o->scores.size();
  return BSQ_OK;
}

void orc_im_free(orc_im_out* o) { delete o; }

int orc_coverage_table(const bsq_reference* ref, const bsq_alignments* aln, const bsq_coverage_table_args* args,
                       char** text, int64_t* len, bsq_status* st) {
  try {
    const string t = orc::coverage_table(ref, aln, args);
    char* b = (char*)malloc(t.size() + 1);
    memcpy(b, t.data(), t.size());
    b[t.size()] = 0;
    *text = b;
    *len = (int64_t)t.size();
    set_status(st, BSQ_OK, "");
    return BSQ_OK;
  } catch (const orc::orc_error& e) {
    set_status(st, e.code, e.what());
    return e.code;
  }
}

void orc_free_text(char* t) { free(t); }

int orc_sequence_trims(const char* seq, int64_t len, uint8_t* out) {
  if (len < 0 || len > 0xFFFFFFFFLL) return BSQ_ERR_ARG;
  orc::sequence_trims(seq, (uint32_t)len, out);
  return BSQ_OK;
}

// get_alignment_trims (calculate_trims.cpp:239-285), the XL/XR of every alignment: the larger of
// the reference's SequenceTrims at reference_start_0 / reference_end_0 and the edge trims of
// query_char_sequence (the bases from query_start_0 to query_end_0, alignment.h:190-196), plus
// query_start_0 on the left and read_length - query_end_1 on the right (alignment.cpp:248-289).
// -1 for unmapped records, records without a target or covering no reference base.
int orc_alignment_trims(const bsq_reference* ref, const bsq_alignments* aln, int32_t* xl, int32_t* xr) {
  try {
    const orc::Aln A{aln};
    vector<vector<uint8_t>> trims((size_t)ref->n_targets);
    for (int32_t t = 0; t < ref->n_targets; t++) {
      trims[(size_t)t].resize(2 * (size_t)ref->lengths[t]);
      orc::sequence_trims(ref->seqs[t], (uint32_t)ref->lengths[t], trims[(size_t)t].data());
    }
    for (int64_t r = 0; r < A.n(); r++) {
      xl[r] = xr[r] = -1;
      const int32_t tid = aln->tid[r];
      if ((aln->flag[r] & 0x4) || tid < 0 || A.n_cigar(r) == 0) continue;
      const int64_t start0 = aln->pos[r], end0 = A.endpos(r) - 1;
      if (end0 < start0) continue;
      if (tid >= ref->n_targets || start0 < 0 || end0 >= ref->lengths[tid]) return BSQ_ERR_DATA;
      const uint32_t qs1 = A.query_start_1(r), qe1 = A.query_end_1(r);
      string q;
      for (int64_t p = (int64_t)qs1 - 1; p <= (int64_t)qe1 - 1; p++) q += orc::bam_char(A.base(r, p));
      uint32_t eL, eR;
      orc::edge_trims(q, eL, eR);
      const vector<uint8_t>& T = trims[(size_t)tid];
      const uint32_t gL = T[(size_t)start0], gR = T[(size_t)ref->lengths[tid] + (size_t)end0];
      xl[r] = (int32_t)(std::max(gL, eL) + (qs1 - 1));
      xr[r] = (int32_t)(std::max(gR, eR) + (A.len(r) - qe1));
    }
    return BSQ_OK;
  } catch (const std::exception&) {
    return BSQ_ERR_DATA;
  }
}

double orc_fit_alleles(const double* r, const double* lpmax, const int32_t* obs, int64_t n, const int32_t* allowed,
                       double tol, double* f, int32_t* iterations) {
  bsq_identify_args args;
  memset(&args, 0, sizeof(args));
  args.polymorphism_precision_decimal = tol;
  orc::Identify im(&args);
  vector<orc::PData> pd((size_t)n);
  for (int64_t i = 0; i < n; i++) {
    for (int b = 0; b < 5; b++) pd[i].r[b] = r[i * 5 + b];
    pd[i].lpmax = lpmax[i];
    pd[i].base = orc::kIndexChar[obs[i]];
  }
  bool al[5];
  for (int b = 0; b < 5; b++) al[b] = allowed[b] != 0;
  orc::AlleleModel m = im.fit(pd, al);
  for (int b = 0; b < 5; b++) f[b] = m.f[b];
  *iterations = (int32_t)m.iterations;
  return m.log10_likelihood;
}

int orc_presence_and_bounds(const double* r, const double* lpmax, const int32_t* obs, int64_t n, int32_t v,
                            double tol, double log10_ref_length, double* f, double* out) {
  bsq_identify_args args;
  memset(&args, 0, sizeof(args));
  args.polymorphism_precision_decimal = tol;
  args.polymorphism_precision_places = 17;
  orc::Identify im(&args);
  im.log10_ref_length = log10_ref_length;
  vector<orc::PData> pd((size_t)n);
  for (int64_t i = 0; i < n; i++) {
    for (int b = 0; b < 5; b++) pd[i].r[b] = r[i * 5 + b];
    pd[i].lpmax = lpmax[i];
    pd[i].base = orc::kIndexChar[obs[i]];
  }
  bool al[5] = {true, true, true, true, true};
  orc::AlleleModel m = im.fit(pd, al);
  for (int b = 0; b < 5; b++) f[b] = m.f[b];
  out[0] = m.log10_likelihood;
  out[1] = im.variant_presence_score(pd, m, v);
  orc::Fields fl;
  im.frequency_bounds(fl, pd, m, v);
  out[2] = strtod(fl["frequency_lower"].c_str(), nullptr);
  out[3] = strtod(fl["frequency_upper"].c_str(), nullptr);
  return BSQ_OK;
}

int orc_read_base_likelihoods(const double* log10_table, int64_t n_bins, int32_t S, int32_t Q, uint32_t rs,
                              uint32_t q, uint32_t obs, int32_t strand, int32_t mapq, double* lp, double* r,
                              double* lpmax) {
  try {
    bsq_identify_args args;
    memset(&args, 0, sizeof(args));
    args.log10_prob_table = log10_table;
    args.n_bins = n_bins;
    args.read_set_max = S;
    args.quality_max = Q;
    orc::Identify im(&args);
    im.prob.resize((size_t)n_bins);
    for (int64_t i = 0; i < n_bins; i++) im.prob[(size_t)i] = pow(10, log10_table[i]);
    orc::PData pd;
    pd.rs = rs;
    pd.q = q;
    pd.obs = obs;
    pd.strand = strand;
    pd.mq = mapq;
    im.fill_likelihoods(pd);
    for (int b = 0; b < 5; b++) {
      lp[b] = pd.lp[b];
      r[b] = pd.r[b];
    }
    *lpmax = pd.lpmax;
    return BSQ_OK;
  } catch (const orc::orc_error& e) {
    return e.code;
  }
}

double orc_pure_genotype(const double* lps, uint32_t n, int32_t* best) {
  bsq_identify_args args;
  memset(&args, 0, sizeof(args));
  orc::Identify im(&args);
  char g;
  double score;
  im.pure_genotype_call(lps, n, g, score);
  *best = (g == 'N') ? 5 : (int32_t)(std::string("ACGT.").find(g));
  return score;
}

double orc_fisher_exact_2x2(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return orc::fisher_2x2(a, b, c, d); }

double orc_ks_test_less(const double* x, int64_t nx, const double* y, int64_t ny) {
  vector<double> vx(x, x + nx), vy(y, y + ny);
  return orc::ks_less(vx, vy);
}

int64_t orc_pileup_column(const bsq_alignments* aln, int32_t tid, int64_t pos0, int32_t* out, int64_t cap) {
  try {
    orc::Aln A{aln};
    int64_t n = 0;
    // scan reads overlapping pos0 in file order
    for (int64_t r = 0; r < aln->n_reads; r++) {
      if (aln->tid[r] != tid || (aln->flag[r] & 0x4) || A.n_cigar(r) == 0) continue;
      if (aln->pos[r] > pos0 || A.endpos(r) <= pos0) continue;
      orc::ResolvedRead rr;
      orc::resolve_read(A, r, rr);
      int64_t j = pos0 - rr.pos;
      if (n < cap) {
        out[4 * n + 0] = (int32_t)r;
        out[4 * n + 1] = rr.qpos[j];
        out[4 * n + 2] = rr.indel[j];
        out[4 * n + 3] = rr.is_del[j];
      }
      n++;
    }
    return n;
  } catch (...) {
    return -1;
  }
}

}  // extern "C"
