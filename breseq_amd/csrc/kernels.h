// kernels.h -- parameter blocks of the pileup kernels (passed by value) and the per-column
// screening accumulator.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace bsq {

struct EcParams {
  const DRead* reads;
  const uint32_t* cigar;
  const uint8_t* bases;
  const uint8_t* quals;
  const char* ref;
  const int32_t* pos_target;    // target of each global column, -1 for pads / unscored
  const int32_t* target_group;  // coverage group of each target
  const int32_t* win_rb;        // alignments overlapping window w: [win_rb[w], win_re[w])
  const int32_t* win_re;
  int64_t n_windows;
  int32_t n_pos;
  int32_t read_set_max, quality_max, n_bins;
  int32_t n_groups_lds, hist_lds;   // LDS-privatized histogram shape
  int32_t hist_len;                 // global histogram length per group
  unsigned long long* counts;       // n_bins
  unsigned long long* hist;         // n_groups * hist_len
  uint32_t* err;
};

struct ImParams {
  const DRead* reads;
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
  uint8_t* cand_cons;           // consensus test passed (computed from the summed likelihoods)
  int32_t* n_cand;
  int32_t cand_cap;
  int32_t* ins;                 // columns with insert slots
  int32_t* n_ins;
  int32_t ins_cap;
  int32_t key_cap;              // LDS capacity of the scorer's read list
  RaOut* ra_out;
  uint32_t* err;
};

struct RunParams {
  const int32_t* pos_target;
  const TargetInfo* targets;
  const uint8_t* col_flags;
  const int32_t* col_unique;
  int32_t n_pos;
  int32_t do_mc;
  int2* events;
  int32_t* n_events;
  int32_t event_cap;
  const int2* runs;
  int4* run_out;
};

// Screening sums of one column (FP32 is enough: the screen compares an upper bound with a margin).
struct ColumnAcc {
  uint32_t n;
  uint32_t cnt[5];
  float SL;        // Sum_i max_b log10 P_i(b)
  float S[25];     // S[o*5+b] = Sum over reads observing o of log10 P_i(b)
  __device__ void reset() {
    n = 0;
    SL = 0.0f;
    for (int i = 0; i < 5; i++) cnt[i] = 0;
    for (int i = 0; i < 25; i++) S[i] = 0.0f;
  }
  __device__ void add(int obs, bool top, const double l[5], float lpmax) {
    (void)top;
    n++;
    SL += lpmax;
#pragma unroll
    for (int o = 0; o < 5; o++) {
      if (o == obs) {
        cnt[o]++;
#pragma unroll
        for (int b = 0; b < 5; b++) S[o * 5 + b] += (float)l[b];
      }
    }
  }
};

__global__ void k_error_count(EcParams P);
__global__ void k_identify_columns(ImParams P);
__global__ void k_insert_slots(ImParams P);
__global__ void k_score_candidates(ImParams P, int cand_begin, int cand_end);
__global__ void k_runs(RunParams P);
__global__ void k_mc_detail(RunParams P, int n_runs);

}  // namespace bsq
