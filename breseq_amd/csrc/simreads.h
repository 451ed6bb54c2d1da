// simreads.h -- SIMULATE-READS restated (simreads.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bsq.h"

namespace bsq {

// One simulated read with its alignment by construction, forward-strand oriented: seq / qual as a
// BAM record holds them, ops the CIGAR (M / I / D runs), left0 the 0-based leftmost aligned base.
struct SimReadOut {
  std::string name, seq, qual;
  std::vector<std::pair<char, uint32_t>> ops;
  int64_t left0 = 0, ref_span = 0;
  int strand = 1;
  int mate = 0;   // 0 single-end, 1 / 2 the mates of a pair
};

// Reads of `ref` as SIMULATE-READS makes them; each output is optional (nullptr to skip).
void simulate_reads(const std::string& ref, const bsq_simreads_args& a, std::vector<SimReadOut>* reads,
                    std::string* fastq_1, std::string* fastq_2);

}  // namespace bsq
