// evidence.h -- host-side assembly of breseq's output files from device results: the error-rate
// and coverage tables of error_count, the RA/MC/UN genome diff of identify_mutations, and the
// test_RA_evidence classification that reads that genome diff back.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bsq.h"
#include "engine.h"

namespace bsq {

// counts_to_log10_prob (error_count.cpp:1005-1026)
void counts_to_log10_prob(int S, int Q, const std::vector<uint64_t>& counts, std::vector<double>& out);
// the entries of qualities [q0, q1) only, into out (the whole table's storage); S <= 256
void counts_to_log10_prob_range(int S, int Q, const std::vector<uint64_t>& counts, double* out, int q0, int q1);
// write_log10_prob_table (error_count.cpp:660-690)
std::string error_rates_text(int S, int Q, const std::vector<double>& log10p);
// print_coverage (error_count.cpp:239-253)
std::string coverage_text(const std::vector<uint32_t>& hist);
// write_base_qual_only_prob_table (error_count.cpp:697-785), one text per read file
std::vector<std::string> base_qual_texts(int S, int Q, const std::vector<uint64_t>& counts);
// read_log10_prob_table (error_count.cpp:629-654)
bool parse_error_rates(const std::string& text, std::vector<double>& out, int& S, int& Q);

struct GdEntry {
  int type = 0;          // 0 RA, 1 MC, 2 UN
  uint32_t id = 0;
  std::string seq_id;
  uint32_t pos = 0;      // RA position, MC/UN start
  int32_t ins = 0;       // RA insert position
  uint32_t end = 0;      // MC/UN end
  uint32_t start_range = 0, end_range = 0;
  std::string ref_base, new_base;
  std::vector<std::pair<const char*, std::string>> fields;   // other key=value pairs (static names)
  std::string fields_text;   // or those pairs already written as "\tkey=value..." (RA entries)
  // traversal order key (target rank, emission column, kind) used to number entries as add() does
  int64_t order_target = 0, order_pos = 0, order_sub = 0;
};

// Fields of an RA entry from the exact scorer's result (identify_mutations.cpp:1836-1910).
void ra_entry_from_result(const RaOut& r, const std::string& seq_id, uint32_t position,
                          uint32_t precision_places, GdEntry& e);

// Number entries in traversal order, then sort and print as cGenomeDiff::write does.
std::string write_gd(std::vector<GdEntry>& entries);

// test_RA_evidence over GD text (identify_mutations.cpp:522-749).
std::string test_ra_evidence_text(const std::string& gd_text, const bsq_reference* ref,
                                  const bsq_ra_test_settings& s);

}  // namespace bsq
