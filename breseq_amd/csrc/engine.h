// engine.h -- device data layout of the pileup-scoring engine, shared by the HIP kernels and the
// host code that packs samples and formats results.
#pragma once

#include <cstdint>

namespace bsq {

// One mapped alignment in HBM (40 bytes).  Positions are global: target t occupies
// [tgt_off[t], tgt_off[t] + L_t) and is followed by one NUL pad byte in the reference array,
// which plays the part of the terminator of the fai_fetch string (pileup_base.cpp:36-44).
struct DRead {
  int64_t seq_off;    // first base of the read in the device bases[] / quals[] (8-byte aligned)
  int32_t gpos;       // global position of the first aligned reference base
  int32_t gend;       // one past the last aligned reference base (bam_endpos)
  uint32_t cig_off;   // first CIGAR word in cigar[]
  uint16_t l_qseq;    // read length
  uint8_t n_cigar;    // CIGAR ops (<= 255)
  uint8_t flags;      // kReadReversed | kReadSimple
  uint16_t y0;        // simple reads: query index of the first aligned base
  uint16_t qs0;       // query_start_0 (alignment.cpp:248-267)
  uint16_t qe0;       // query_end_0   (alignment.cpp:270-289)
  uint16_t qbe0;      // end of query_bounds_0 (alignment.cpp:125-220, min_qual = 0)
  uint16_t x1;        // X1 redundancy, 1 when absent (alignment.cpp:52-57)
  int16_t xl;         // XL trim tag, -1 when absent
  int16_t xr;         // XR trim tag, -1 when absent
  uint8_t rfile;      // read file index: the read_set covariate
  uint8_t mslot;      // index of this read's MAPQ in the session's MAPQ list
};
static_assert(sizeof(DRead) == 40, "DRead must stay 40 bytes");

enum : uint8_t { kReadReversed = 1, kReadSimple = 2 };

// Per-position summary written by the identify column kernel (slot 0).
enum : uint8_t {
  kPosPredicted = 1,   // consensus score >= mutation cutoff (UN calling, :1684)
  kPosRedundant = 2,   // redundant coverage > 0 (MC redundant sub-range, :2329-2341)
  kPosInsert = 4,      // a unique read carries an insertion after this base (:1598-1599)
  kPosCandidate = 8,   // sent to the exact FP64 scorer
  kPosConsensus = 16,  // best genotype != reference and score > 0 (:1789)
  kPosSeed = 32,       // total coverage <= seed cutoff (:2280)
  kPosValid = 128      // position belongs to a target that is scored
};

// Result of the exact scorer for one (position, insert slot).
struct RaOut {
  int32_t gpos;
  int32_t slot;
  uint8_t emitted;
  uint8_t ref_idx;       // 0..4, 5 = N
  uint8_t major, minor, variant;
  uint8_t pad[3];
  uint32_t n;
  double f[5];
  double variant_score;
  double lower, upper;
  double ks_p, fisher_p;
  uint32_t cov[5][2];    // pos_info[base][top, bottom]
  int32_t total_top, total_bot;
};

struct TargetInfo {
  int32_t off;           // global offset
  int32_t len;
  int32_t group;         // coverage group
  int32_t scored;        // in call_mutations_seq_id_set
  double prop_cutoff;
  double seed_cutoff;
};

}  // namespace bsq
