// dp_regions.cpp -- discordant-pair (DP) candidate regions (identify_mutations.cpp:883-908,
// 1395-1457, 2358-2422, 2550-2573).
//
// breseq keeps, per (focal strand x pair orientation) bin and per paired read group, a deque of
// the discordant reads whose start lies within the group's window width of the current column,
// and re-examines every bin at every reference column.  The window count only changes at a
// read's start column (it enters) or at the first column at or beyond start + width (it leaves),
// and "past the end of the target" only at L+1, so the detector here visits just those event
// columns (plus column 1), in column order, with the same per-column steps: enter reads that
// start here (appending to an open region), drop reads that left, then open / widen / close the
// bin's region.  Between event columns none of the state the region output depends on changes.
// Cost: O(discordant reads) instead of O(6 L).
#include "dp_regions.h"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstring>
#include <deque>

#include "host_util.h"

namespace bsq {

static constexpr int kDpBins = 6;
static const char* const kOrientName[3] = {"FR", "RF", "FF"};

void dp_select(const bsq_alignments* a, int32_t n_targets, std::vector<DpRead>& out) {
  out.clear();
  if (!a || !a->xp) return;
  for (int64_t r = 0; r < a->n_reads; r++) {
    const uint16_t f = a->flag[r];
    // is_paired, !unmapped, !proper_pair, not secondary / supplementary / mate unmapped
    if (!(f & 0x1) || (f & (0x2 | 0x4 | 0x8 | 0x100 | 0x800))) continue;
    const int8_t xp = a->xp[r];
    const int32_t tid = a->tid[r];
    if (tid < 0 || tid >= n_targets) continue;
    const int64_t c0 = a->cigar_off[r], c1 = a->cigar_off[r + 1];
    int64_t span = 0;
    for (int64_t k = c0; k < c1; k++) {
      const uint32_t op = a->cigar[k] & 15;
      if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) span += a->cigar[k] >> 4;
    }
    if (span == 0) continue;   // never in a pileup column
    DpRead d;
    d.tid = tid;
    d.start_1 = (uint32_t)a->pos[r] + 1;
    d.end_1 = (uint32_t)(a->pos[r] + span);
    d.rg = a->rg ? a->rg[r] : -1;
    d.bin = (xp < 0 || xp > 2) ? kDpNoOrientation : (uint8_t)(((f & 0x10) ? 3 : 0) + xp);
    d.read1 = !(f & 0x80);
    d.redundant = a->x1 && a->x1[r] > 1;
    d.tlen = a->tlen ? a->tlen[r] : 0;
    const char* nm = a->names ? a->names + a->name_off[r] : "";
    const size_t nl = a->names ? (size_t)(a->name_off[r + 1] - a->name_off[r]) : 0;
    const void* colon = nl ? memchr(nm, ':', nl) : nullptr;
    d.has_colon = colon != nullptr;
    d.name_part = colon ? std::string((const char*)colon + 1, nm + nl) : std::string(nm, nl);
    out.push_back(std::move(d));
  }
}

namespace {

struct Group {
  double width;
  uint32_t r1_prefix, r2_prefix;
};

struct Region {
  std::string seq_id;
  uint32_t start, end, max_count;
  char strand;
  const char* orientation;
  bool redundant;
  std::string pairs;
};

void append_u(std::string& s, uint64_t v) {
  char b[24];
  s.append(b, std::to_chars(b, b + sizeof(b), v).ptr);
}

void append_i(std::string& s, int64_t v) {
  char b[24];
  s.append(b, std::to_chars(b, b + sizeof(b), v).ptr);
}

// Appends <read1>__<read2>__<r1_insert_size>__<start>__<end> (dp_descriptor,
// identify_mutations.cpp:1289-1292).
void append_descriptor(std::string& s, const DpRead& r, const Group& g) {
  if (r.has_colon) {
    append_u(s, g.r1_prefix);
    s += ':';
    s += r.name_part;
    s += "__";
    append_u(s, g.r2_prefix);
    s += ':';
    s += r.name_part;
  } else {
    if (r.read1) s += r.name_part;
    s += "__";
    if (!r.read1) s += r.name_part;
  }
  s += "__";
  append_i(s, r.read1 ? (int64_t)r.tlen : -(int64_t)r.tlen);
  s += "__";
  append_u(s, r.start_1);
  s += "__";
  append_u(s, r.end_1);
}

// First column at which a read entering at s has left a window of width w: the smallest p >= s
// with (double)p - (double)s >= w; UINT64_MAX when it never leaves (w NaN or +inf).
uint64_t leave_column(uint32_t s, double w) {
  if (!(w > 0)) return (w <= 0) ? s : UINT64_MAX;
  if (std::isinf(w)) return UINT64_MAX;
  const double c = std::ceil(w);
  if (c > 1e15) return UINT64_MAX;
  return (uint64_t)s + (uint64_t)c;
}

}  // namespace

std::string dp_regions_csv(int32_t n_targets, const char* const* names, const int64_t* lengths, const uint8_t* call,
                           const double* prop, const bsq_dp_settings* dp, const std::vector<DpRead>& reads) {
  std::string out = "seq_id,start,end,strand,orientation,length,max_discordant_count,redundant,discordant_pairs\n";
  // groups in read-file-set order (identify_mutations.cpp:893-908)
  std::vector<Group> groups;
  std::vector<int32_t> group_of_set;
  if (dp) {
    if (dp->n_sets < 0) throw BsqError(BSQ_ERR_ARG, "negative number of read file sets");
    group_of_set.assign((size_t)dp->n_sets, -1);
    for (int32_t i = 0; i < dp->n_sets; i++) {
      if (!dp->paired || !dp->paired[i]) continue;
      if (dp->median[i] <= 0.0) continue;
      Group g;
      g.width = dp->median[i] + 2.42 * dp->mad[i];
      g.r1_prefix = dp->r1_prefix[i];
      g.r2_prefix = dp->r2_prefix[i];
      group_of_set[(size_t)i] = (int32_t)groups.size();
      groups.push_back(g);
    }
  }
  if (groups.empty()) return out;
  const int32_t n_sets = dp->n_sets;
  const int32_t seed = dp->seed;
  // group of a record (read_group_index_map::index, then the set -> group map)
  auto group_of = [&](const DpRead& r) -> int32_t {
    const int32_t rg = (n_sets <= 1 || r.rg < 0 || r.rg >= n_sets) ? 0 : r.rg;
    return rg < n_sets ? group_of_set[(size_t)rg] : -1;
  };

  std::vector<Region> regions;
  // targets the pileup visits with the accumulator on: called, and a coverage fit that worked
  auto scored = [&](int32_t t) { return (!call || call[t]) && (!prop || !(prop[t] < 0.0)); };
  // reads of each (target, bin), file order
  std::vector<std::vector<std::vector<int32_t>>> by_tb((size_t)n_targets, std::vector<std::vector<int32_t>>(kDpBins));
  for (size_t i = 0; i < reads.size(); i++) {
    const DpRead& r = reads[i];
    if (r.tid < 0 || r.tid >= n_targets) continue;
    if (!scored(r.tid)) continue;
    if (group_of(r) < 0) continue;
    if (r.bin == kDpNoOrientation)
      throw BsqError(BSQ_ERR_DATA, "paired, not properly paired alignment without an XP:Z orientation of FR/RF/FF "
                                   "in a paired read group (breseq drops it from its start column's tallies)");
    by_tb[(size_t)r.tid][r.bin].push_back((int32_t)i);
  }
  struct Ev {
    uint64_t col;
    int32_t kind;    // 0 enter, 1 leave (enters of a column come first, in file order)
    int32_t idx;
  };
  std::vector<Ev> ev;
  std::vector<std::deque<int32_t>> window(groups.size());
  std::vector<int32_t> desc;   // the open region's reads, formatted when it closes
  for (int32_t t = 0; t < n_targets; t++) {
    if (!scored(t)) continue;
    const uint64_t L = (uint64_t)lengths[t];
    for (int bin = 0; bin < kDpBins; bin++) {
      const std::vector<int32_t>& rs = by_tb[(size_t)t][(size_t)bin];
      if (rs.empty() && seed > 0) continue;   // the count stays 0 < seed: nothing opens
      ev.clear();
      for (int32_t i : rs) {
        const DpRead& r = reads[(size_t)i];
        if (r.start_1 > L) throw BsqError(BSQ_ERR_FORMAT, "discordant read starts past its reference");
        ev.push_back({r.start_1, 0, i});
        const uint64_t lc = leave_column(r.start_1, groups[(size_t)group_of(r)].width);
        if (lc <= L + 1) ev.push_back({lc, 1, i});
      }
      std::stable_sort(ev.begin(), ev.end(), [](const Ev& x, const Ev& y) {
        if (x.col != y.col) return x.col < y.col;
        return x.kind < y.kind;
      });
      for (auto& w : window) w.clear();
      int64_t metric = 0;
      bool open = false;
      uint32_t r_start = 0, r_max = 0, r_redundant = 0;
      desc.clear();
      size_t e = 0;
      uint64_t col = 1;
      for (;;) {
        // enter: reads starting at this column, in file order (pileup order)
        while (e < ev.size() && ev[e].col == col && ev[e].kind == 0) {
          const DpRead& r = reads[(size_t)ev[e].idx];
          const int32_t g = group_of(r);
          window[(size_t)g].push_back(ev[e].idx);
          metric++;
          if (open) {
            desc.push_back(ev[e].idx);
            if (r.redundant) r_redundant++;
          }
          e++;
        }
        // leave: reads whose window has passed, from the front of their group's window
        while (e < ev.size() && ev[e].col == col && ev[e].kind == 1) {
          const DpRead& r = reads[(size_t)ev[e].idx];
          std::deque<int32_t>& w = window[(size_t)group_of(r)];
          if (w.empty() || w.front() != ev[e].idx) throw BsqError(BSQ_ERR_FORMAT, "discordant reads are not sorted by position");
          w.pop_front();
          metric--;
          e++;
        }
        const bool above = metric >= seed && col <= L;
        if (above) {
          if (!open) {
            open = true;
            r_start = (uint32_t)col;
            r_max = (uint32_t)metric;
            desc.clear();
            r_redundant = 0;
            for (size_t g = 0; g < groups.size(); g++)
              for (int32_t i : window[g]) {
                desc.push_back(i);
                if (reads[(size_t)i].redundant) r_redundant++;
              }
          } else if ((uint32_t)metric > r_max) {
            r_max = (uint32_t)metric;
          }
        } else if (open) {
          Region R;
          R.seq_id = names[t];
          R.start = r_start;
          R.end = (uint32_t)(col - 1);
          R.strand = (bin / 3 == 0) ? 'F' : 'R';
          R.orientation = kOrientName[bin % 3];
          R.max_count = r_max;
          R.redundant = 2 * (uint64_t)r_redundant > desc.size();
          R.pairs.reserve(desc.size() * 48);
          for (size_t k = 0; k < desc.size(); k++) {
            if (k) R.pairs += ';';
            const DpRead& dr = reads[(size_t)desc[k]];
            append_descriptor(R.pairs, dr, groups[(size_t)group_of(dr)]);
          }
          regions.push_back(std::move(R));
          open = false;
          r_max = 0;
          r_redundant = 0;
          desc.clear();
        }
        if (col > L) break;
        col = (e < ev.size()) ? std::min<uint64_t>(ev[e].col, L + 1) : L + 1;
      }
    }
  }
  std::sort(regions.begin(), regions.end(), [](const Region& x, const Region& y) {
    if (x.seq_id != y.seq_id) return x.seq_id < y.seq_id;
    if (x.start != y.start) return x.start < y.start;
    if (x.strand != y.strand) return x.strand < y.strand;
    return strcmp(x.orientation, y.orientation) < 0;
  });
  for (const Region& R : regions) {
    out += R.seq_id;
    out += ',';
    append_u(out, R.start);
    out += ',';
    append_u(out, R.end);
    out += ',';
    out += R.strand;
    out += ',';
    out += R.orientation;
    out += ',';
    append_u(out, (uint32_t)(R.end - R.start + 1));
    out += ',';
    append_u(out, R.max_count);
    out += ',';
    out += R.redundant ? '1' : '0';
    out += ',';
    out += R.pairs;
    out += '\n';
  }
  return out;
}

}  // namespace bsq

using namespace bsq;

extern "C" int bsq_dp_candidate_regions(const bsq_reference* ref, const bsq_alignments* aln, const uint8_t* call_target,
                                        const double* deletion_propagation_cutoff, const bsq_dp_settings* dp,
                                        char** out_text, int64_t* out_len, bsq_status* status) {
  try {
    if (!ref || !aln || !out_text || !out_len) throw BsqError(BSQ_ERR_ARG, "null argument");
    std::vector<DpRead> reads;
    dp_select(aln, ref->n_targets, reads);
    const std::string s = dp_regions_csv(ref->n_targets, ref->names, ref->lengths, call_target,
                                         deletion_propagation_cutoff, dp, reads);
    char* buf = (char*)malloc(s.size() + 1);
    if (!buf) throw BsqError(BSQ_ERR_LIMIT, "out of host memory");
    memcpy(buf, s.data(), s.size());
    buf[s.size()] = '\0';
    *out_text = buf;
    *out_len = (int64_t)s.size();
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    set_status(status, e.code, e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_status(status, BSQ_ERR_ARG, e.what());
    return BSQ_ERR_ARG;
  }
}
