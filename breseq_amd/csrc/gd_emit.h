// gd_emit.h -- device writer of ra_mc_evidence.gd (gd_emit.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.h"
#include "kernels.h"

namespace bsq {

constexpr int kGdSlot = 1024;            // bytes of one entry's line slot
constexpr int kGdMaxEntries = 1 << 20;   // entries one pass may write on the device
constexpr int kGdFmtStride = 512;        // bytes per value of k_format_doubles
constexpr int kGdPieces = 16;            // pieces of one line, one per lane of a 16-lane group
constexpr int kGdPiece = 480;            // bytes of one piece (a printed number with its key fits)
constexpr int kGdRankTile = 64;          // keys per LDS tile of k_gd_rank

// One genome-diff entry: numbering key (k0..k2: target in traversal order, emission column, kind)
// and output key (o0..o2: evidence type, seq_id order, position, the type's fields, ID).
struct GdEnt {
  uint64_t k0, k1, k2;
  uint64_t o0, o1, o2;
  int32_t type;   // 0 RA, 1 MC, 2 UN, 3 MC of a whole target (failed coverage fit)
  int32_t src;    // index into its own list: candidate, MC/UN run, or failed target
  int32_t t;      // target
  uint32_t id;
};

struct GdParams {
  // the lists as the identify pass left them: counts read on the device (clamped to the list
  // capacities), so the writer is queued behind the scorer without a host round trip
  const RaOut* ra;
  const int32_t* n_cand;
  int32_t cand_cap;
  const RunRec* runs;
  const int32_t* n_runs;
  int32_t run_cap;
  const int32_t* whole;   // targets whose coverage fit failed (whole-target MC), in traversal order
  int32_t n_whole;
  const int32_t* pos_target;
  int32_t n_pos;
  const TargetInfo* targets;
  int32_t n_targets;
  const int32_t* rank;      // target -> its place in the traversal order
  const int32_t* lexrank;   // target -> its place in seq_id byte order
  const char* names;        // seq_ids, concatenated
  const int32_t* name_off;  // n_targets + 1 offsets into names
  int32_t places;           // polymorphism_precision_places
  GdEnt* ents;
  int32_t* n_ents;
  int32_t ent_cap;
  int32_t* n_type;          // entries written per type: RA, MC, UN
  uint32_t* err;            // 1 an index out of range, 2 entry list full, 4 a line beyond its slot,
                            // 8 a number the formatter refused
  char* slots;              // ent_cap * kGdSlot
  int32_t* rank_id;         // ent_cap numbering ranks, zeroed by k_gd_collect
  int32_t* rank_out;        // ent_cap output ranks, zeroed by k_gd_collect
  int32_t* lens;            // ent_cap
};

__global__ void k_gd_collect(GdParams P);
constexpr int kGdRankGridX = 16, kGdRankGridY = 64;   // fixed grid of k_gd_rank (grid-stride over entries x tiles)
__global__ void k_gd_rank(const GdEnt* ents, const int32_t* n_ents, int cap, int which, int32_t* rank);
__global__ void k_gd_number(GdEnt* ents, const int32_t* n_ents, int cap, const int32_t* rank);
__global__ void k_gd_lines(GdParams P, const int32_t* out_rank);
__global__ void k_gd_offsets(const int32_t* lens, const int32_t* n_ents, int cap, int64_t* offsets, int64_t* total);
__global__ void k_gd_gather(const char* slots, const int32_t* lens, const int64_t* offsets, const int32_t* n_ents,
                            int cap, char* text);
__global__ void k_format_doubles(const double* v, int64_t n, int32_t precision, int32_t scientific, char* out,
                                 int32_t* lens);

}  // namespace bsq
