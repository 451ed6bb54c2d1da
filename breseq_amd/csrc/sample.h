// sample.h -- owning container for one sequencing sample (reference + sorted alignments), the
// structures behind bsq_sample.  Host side only.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bsq.h"

namespace bsq {

struct Sample {
  // reference
  std::vector<std::string> names;
  std::vector<std::string> seqs;
  std::vector<int64_t> lengths;
  std::vector<const char*> name_ptrs, seq_ptrs;
  // alignments (structure of arrays, file order)
  std::vector<int32_t> tid, pos;
  std::vector<uint16_t> flag;
  std::vector<uint8_t> mapq;
  std::vector<int64_t> cigar_off, seq_off;
  std::vector<uint32_t> cigar;
  std::vector<uint8_t> seq, qual;
  std::vector<int32_t> x1, xl, xr, rg;
  std::vector<std::string> read_names;   // optional (BAM round trip)
  std::vector<int32_t> mate_tid, mate_pos, tlen;
  std::vector<int8_t> xp;                // XP:Z pair orientation code (0 FR, 1 RF, 2 FF, -1 none)
  std::vector<int64_t> name_off;         // CSR copy of read_names for the C view
  std::string name_chars;
  // read groups (@RG ID / LB) and read files
  std::vector<std::string> rg_ids, rg_libs;
  std::vector<uint32_t> rf_base, rf_count;
  std::vector<std::string> rf_names;
  std::vector<const char*> rf_name_ptrs;

  bsq_reference ref_view{};
  bsq_alignments aln_view{};

  void finalize();   // refresh the C views after the vectors changed
  int64_t n_reads() const { return (int64_t)pos.size(); }
};

}  // namespace bsq

struct bsq_sample {
  bsq::Sample s;
};
