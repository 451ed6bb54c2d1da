// dp_regions.h -- discordant-pair (DP) candidate regions: the read selection done when a sample
// is packed, and the sliding-window region detector run at identify time.  Host side only.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bsq.h"

namespace bsq {

// One read that enters a DP window: paired, mapped, not a proper pair, not secondary,
// supplementary or mate-unmapped, with an XP orientation (identify_mutations.cpp:1399-1421).
struct DpRead {
  int32_t tid;
  uint32_t start_1;        // reference_start_1: the column at which it enters
  uint32_t end_1;          // reference_end_1 (bam_endpos)
  int32_t rg;              // @RG index the record carries (-1: none)
  uint8_t bin;             // strand * 3 + orientation; kDpNoOrientation: no usable XP tag
  bool read1;              // !(flag & BAM_FREAD2)
  bool redundant;          // X1 > 1
  bool has_colon;          // the read name holds a ':'
  int32_t tlen;            // BAM TLEN
  std::string name_part;   // text after the first ':' of the name, or the whole name
};

// A read that passes the flag tests but carries no XP of FR / RF / FF.  breseq skips such a
// read entirely at its start column once DP groups exist (the `continue` at
// identify_mutations.cpp:1411 and :1415 leaves the per-read pileup loop), which its own BAMs
// never contain (mark_pair_info always writes XP, resolve_alignments.cpp:1241-1272); the region
// detector refuses them with BSQ_ERR_DATA instead of changing the column tallies.
constexpr uint8_t kDpNoOrientation = 0xFF;

// Selects the DP reads in file order.  Reads that never enter a pileup column (unmapped, no
// CIGAR, zero reference span) are skipped.  Needs a->xp; with a->xp == NULL nothing is selected.
void dp_select(const bsq_alignments* a, int32_t n_targets, std::vector<DpRead>& out);

// DP_candidate_regions.csv for the selected reads over the targets with call[t] != 0 (call ==
// NULL: all) and prop[t] >= 0 (prop == NULL: all), sorted as write_dp_candidate_regions sorts
// (identify_mutations.cpp:2550-2573).
std::string dp_regions_csv(int32_t n_targets, const char* const* names, const int64_t* lengths, const uint8_t* call,
                           const double* prop, const bsq_dp_settings* dp, const std::vector<DpRead>& reads);

}  // namespace bsq
