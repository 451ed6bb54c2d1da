// kernels.h -- parameter blocks of the pileup kernels (passed by value) and the per-column
// screening accumulator.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace bsq {

// Reference base codes per global column (ref_idx[]): A C G T = 0..3, '.' = 4, N = 5, anything
// else (including the pad after each target) = 7.
enum : uint8_t { kRefDot = 4, kRefN = 5, kRefOther = 7 };

// Pass 1, read-parallel: covariate counts and coverage difference arrays.
struct EcParams {
  const DRead* reads;
  const uint32_t* cigar;
  const uint8_t* bases;
  const uint8_t* quals;
  const uint8_t* ref_idx;       // per global column (plus pads), kRef* codes; 16-byte aligned
  const int32_t* pos_target;    // target of each global column, -1 for pads
  const int32_t* target_scored; // 1 if the target is in the call set
  int64_t n_reads;
  int32_t n_pos;
  int32_t read_set_max, quality_max, n_bins;
  int32_t do_errors;
  int32_t* cov_diff;            // [0, n_pos+1): unique reads, [n_pos+1, 2n_pos+2): redundant reads
  unsigned long long* counts;   // n_bins
  uint32_t* err;
  // preprocess mode: 2 bits per column (forward, reverse) set where a qualifying read starts
  uint32_t* start_bits;         // null unless preprocess
  uint32_t end_min_len;         // unmatched_end_minimum_read_length
  double end_factor;            // unmatched_end_length_factor
};

// Pass 1, coverage histogram from the difference arrays (three-phase prefix scan).
struct CovParams {
  const int32_t* cov_diff;      // as EcParams::cov_diff
  const int32_t* pos_target;
  const int32_t* target_group;
  const int32_t* target_scored;
  int32_t n_pos;
  int32_t n_tiles;
  int2* tile_sum;               // per tile: (unique, redundant) sums, then exclusive prefixes
  int32_t n_groups_lds, hist_lds;   // LDS-privatized histogram shape
  int32_t hist_len;                 // global histogram length per group
  unsigned long long* hist;         // n_groups * hist_len
  uint32_t* err;
  const uint32_t* start_bits;       // preprocess mode (EcParams::start_bits), else null
  unsigned long long* start_counts; // per target: (position, strand) pairs without / with a start
};
constexpr int kCovTile = 4096;      // columns per block of the coverage scan (256 threads x 16)
constexpr int kEcChunk = 32;        // consecutive alignments per wavefront work item of k_ec_reads
constexpr int kEcWaves = 4;         // wavefronts per k_ec_reads block (256 threads)

// One row of the strand-folded likelihood table: a read of quality q from read file rs with
// MAPQ slot m observing read-strand base o reads row ((m*S + rs)*Q + q)*5 + o.  Hypotheses are in
// read-strand order; a reverse-strand read maps hypothesis b to t[comp(b)].  FP32: the column pass
// sums these with an explicit error bound and certifies its decisions (exact FP64 values are used
// by the scorer and k_fix_columns).
struct FoldRow {
  float t[5];     // log10 P(o | hypothesis), fill_read_base_likelihoods (identify_mutations.cpp:3359-3384)
  float lpmax;    // max_b t[b]
  float diag;     // t[o]: the read's own observation as the hypothesis
  float pad;
};
static_assert(sizeof(FoldRow) == 32, "FoldRow is two 16-byte LDS words");

// Per-call view of one alignment for the column pass: flat 32-bit fields (two scalar loads, no
// unpacking).  Built by k_prep_reads from the DRead and the call's covariate shape.
struct IdRec {
  int32_t gpos, gend;       // aligned span [gpos, gend) in global columns
  uint32_t seq_off;         // DRead::seq_off (the device byte arrays are below 4 GiB)
  int32_t qoff;             // simple alignments: query position = column + qoff
  uint32_t trim;            // slot 0: trimmed when qpos < (trim & 0xFFFF) or qpos >= (trim >> 16)
                            // (alignment.h:389-410)
  uint32_t row_base;        // ((mslot*S + read file)*Q)*5, 0xFFFFFFFF when the read file is >= S
  uint32_t x1_flags;        // X1 | (kReadReversed | kReadSimple) << 16
  uint32_t pad;
};
static_assert(sizeof(IdRec) == 32, "IdRec is eight dwords");

struct ImParams {
  const DRead* reads;
  const IdRec* idrec;           // per alignment, built for this call
  const uint32_t* cigar;
  const uint8_t* bases;
  const uint8_t* quals;
  const char* ref;
  const int32_t* pos_target;
  const TargetInfo* targets;
  const int32_t* win_rb;
  const int32_t* win_re;
  int64_t n_windows;
  int32_t n_pos;
  int32_t read_set_max, quality_max;
  int32_t base_quality_cutoff;
  const double* tab_lp;         // per key: log10 P(obs | true base b), b = 0..4
  const double* tab_r;          // per key: 10^(lp - lpmax) for b = 0..4, then lpmax
  int32_t tab_keys;             // rows of tab_lp / tab_r
  int32_t n_reads;              // alignments in reads[] (and idrec[])
  const FoldRow* fold;          // strand-folded rows (mslot, read file, quality, read-strand obs);
                                // row fold_rows is all zeros (the row of a read that adds nothing)
  int32_t fold_rows;
  int32_t fold_lds;             // 1: stage the folded table in LDS
  double log10_ref_length;
  double mutation_cutoff;
  double polymorphism_cutoff;
  double precision_decimal;
  double screen_margin;
  int32_t screen_off;           // 1: every covered column goes to the exact scorer
  // per-column outputs
  uint8_t* col_flags;
  int32_t* col_unique;
  int32_t* col_total;
  // candidates
  int2* cand;                   // (global column, insert slot)
  int32_t* n_cand;
  int32_t* score_next;   // work counter of the exact scorer: the next candidate a block takes (zeroed per pass)
  int32_t cand_cap;
  int32_t* ins;                 // columns with insert slots
  int32_t* n_ins;
  int32_t ins_cap;
  int32_t key_cap;              // capacity of a candidate's read list (the most reads of a window)
  int32_t keys_global;          // 1: read lists in key_scratch (key_cap per block), else in LDS
  uint32_t* key_scratch;        // device scratch of the read lists of deep samples
  int32_t rows_cap;             // reads whose table rows the scorer caches in LDS
  RaOut* ra_out;
  // columns whose predicted / consensus decision the column pass could not certify
  int32_t* fix;
  int32_t* n_fix;
  int32_t fix_cap;
  uint32_t* err;
  float* audit;                 // diagnostics: per column 8 floats (FP32 sums, read count), or null
};

// MC / UN interval found by k_runs (kind 0 MC, 1 UN); columns are global.
struct RunRec {
  int32_t kind, s, e;
  int32_t seed;              // MC: seed cutoff reached in the run or at its closing column
  int32_t red_start, red_end;   // MC: redundant sub-range columns, -1 when unset
  int32_t lo, li, ri, ro;    // MC: unique coverage left-outside, left-inside, right-inside, right-outside
};

struct RunParams {
  const int32_t* pos_target;
  const TargetInfo* targets;
  const uint8_t* col_flags;
  const int32_t* col_unique;
  int32_t n_pos;
  int32_t do_mc;
  RunRec* runs;
  int32_t* n_runs;
  int32_t run_cap;
  uint32_t* err;
};

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v);
__global__ void k_repack_reads(const DRead* reads, const int64_t* raw_off, int64_t n_reads, const uint8_t* src_b,
                               const uint8_t* src_q, uint8_t* dst_b, uint8_t* dst_q);
constexpr int kRefIdxFrontPad = 16;   // ref_idx bytes before column 0 (aligned loads may start at -4)
__global__ void k_ec_reads(EcParams P);
__global__ void k_cov_tile_sums(CovParams C);
__global__ void k_cov_tile_scan(CovParams C);
__global__ void k_cov_hist(CovParams C);
constexpr int kColWaves = 8;                 // wavefronts per block of the column kernel (default; 4 also built)
constexpr size_t kFoldLdsBytes = 48 * 1024;  // largest likelihood table staged in LDS (1535 rows)
template <bool kLds, int kWaves>
__global__ void k_identify_columns_t(ImParams P);
__global__ void k_prep_reads(const DRead* reads, int64_t n, int32_t S, int32_t Q, IdRec* out);
template <bool kKeysGlobal>
__global__ void k_fix_columns_t(ImParams P);
__global__ void k_insert_slots(ImParams P);
// BAM2COV per-column counts over a region (coverage_output.cpp:307-503): lane = column, the window's
// reads in file order.
struct CovTabParams {
  const DRead* reads;
  const uint8_t* rg;     // @RG index of each read, 255 when it has none
  const uint32_t* cigar;
  const int32_t* win_rb;
  const int32_t* win_re;
  int32_t h0;            // global column of the first row
  int32_t step;          // columns between rows (the downsampling step)
  int64_t n;             // rows: columns h0 + j * step, j < n
  int32_t rg_filter;     // -1: every read; else only reads of this read group index
  int32_t n_rg;          // @RG lines of the header (read_group_index_map::index resolves against it)
  uint32_t* uniq;        // [2][n]: top, bottom (reversed) strand
  double* red;           // [2][n]: Sum 1/X1 of redundant reads, in read order
  uint32_t* raw;         // [2][n]: redundant reads
  uint32_t* begin;       // [2][n]: unique reads whose first base (in read direction) is here
  uint32_t* err;
};
// Consecutive rows (step 1): one wavefront per 64 columns, lane = column.
__global__ void k_coverage_columns(CovTabParams C);
// Downsampled rows (step > 1): one wavefront per row, lanes over the row's reads.
__global__ void k_coverage_rows(CovTabParams C);

constexpr int kScoreWaves = 4;               // wavefronts per candidate in the exact scorer
constexpr int kScoreGrid = 2048;             // blocks of the exact scorer (each draws candidates from a work counter)
constexpr int kFixGrid = 256;                // blocks of k_fix_columns (at most kScoreGrid)
constexpr size_t kScoreStaticLds = 28 * 1024;   // bound on the scorer's static LDS
constexpr size_t kScoreLdsBudget = 52 * 1024;   // static + dynamic LDS of a scorer block: three per CU
template <bool kKeysGlobal>
__global__ void k_score_candidates_t(ImParams P);
// reference and alignment trims (extra.hip)
__global__ void k_repeat_extent(const char* s, int64_t len, uint8_t* m);
__global__ void k_trim_gather(const uint8_t* m, int64_t len, uint8_t* left_out, uint8_t* right_out);
__global__ void k_alignment_trims(const DRead* reads, int64_t n, const uint32_t* cigar, const uint8_t* bases,
                                  const uint8_t* left, const uint8_t* right, int32_t* xl, int32_t* xr);
__global__ void k_read_starts(EcParams P);
template <bool kKeysGlobal>
__global__ void k_audit_columns_t(ImParams P, const int32_t* cols, int32_t n_cols, double* out, uint32_t* mismatches);
__global__ void k_runs(RunParams P);

}  // namespace bsq
