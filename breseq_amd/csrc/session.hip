// session.hip -- device-resident sessions: packs a sample's alignments once into the device CSR,
// runs the pileup kernels, and assembles breseq's output files from their results.
//
// Entry points (include/bsq.h) replace:
//   bsq_session_error_count -> error_count()            error_count.cpp:50-68
//   bsq_session_identify    -> identify_mutations()     identify_mutations.cpp:48-88
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <charconv>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstring>
#include <fstream>
#include <future>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "dp_regions.h"
#include "engine.h"
#include "evidence.h"
#include "covfit.h"
#include "host_util.h"
#include "dec_format.h"
#include "gd_emit.h"
#include "kernels.h"
#include "sample.h"

namespace bsq {

#define HIP_OK(expr)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) throw BsqError(BSQ_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) count = 1;
    HIP_OK(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
};

// Pinned host memory, so that a hipMemcpyAsync from it returns at once; it is rewritten only after
// the stream that read it has been synchronized.
template <class T>
struct PinnedBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) count = 1;
    HIP_OK(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault));
    n = count;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  ~PinnedBuf() { release(); }
};

static const int kWave = 64;

struct Session {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;   // side stream: work that runs beside the exact scorer
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evk[8] = {};
  hipEvent_t evc = nullptr;   // just before the column pass (its time excludes the table uploads)
  bsq_timing timing{};
  bool loaded = false;
  bool audit_on = false;              // the next identify pass records its FP32 column sums
  DBuf<float> d_audit;                // 8 floats per column while audit_on
  ImParams last_P{};                  // parameters of the last identify pass (the audit reuses them)
  // packed metadata (host)
  int n_targets = 0;
  std::vector<std::string> names;
  std::vector<int64_t> lengths;
  std::vector<int32_t> tgt_off;
  int32_t n_pos = 0;        // global columns incl. pads
  int64_t n_windows = 0;
  int32_t max_window_reads = 0;
  int64_t n_dreads = 0;
  int64_t aligned_bases = 0;
  std::vector<int32_t> mapq_values;   // mslot -> MAPQ
  int32_t n_read_files = 1;
  std::vector<int32_t> h_pos_target;
  std::string h_ref;                  // padded reference (for host-side formatting)
  std::vector<DpRead> dp_reads;       // discordant-pair candidates, file order
  // smallest capacities of identify's result lists, raised when a pass overflows one (the pass
  // is then run again with room for every entry)
  struct {
    int32_t cand = 0, ins = 0, runs = 0, fix = 0;
  } id_caps;
  // pinned sources and destinations of the identify pass's small copies, so that queuing them never
  // waits for the device (a pageable copy does)
  PinnedBuf<TargetInfo> h_targets0;    // per-target parameters as the pass starts
  PinnedBuf<TargetInfo> h_targets;     // per-target parameters with the fitted cutoffs
  PinnedBuf<double> h_tab_lp;          // the likelihood tables of the column pass
  PinnedBuf<FoldRow> h_fold;
  PinnedBuf<int32_t> h_counters;       // [0..7] the pass's counters, [8] MC/UN runs found
  std::vector<int32_t> h_gpos;        // global first column of each packed read (sorted)
  std::vector<int64_t> h_src;         // input alignment index of each packed read
  int64_t n_input_reads = 0;          // alignments the sample was loaded from (packed or not)
  std::vector<int32_t> h_gend_max;    // running maximum of the packed reads' gend, in read order
  // device buffers
  DBuf<DRead> d_reads;
  DBuf<uint32_t> d_cigar;
  DBuf<uint8_t> d_bases, d_quals;
  DBuf<char> d_ref;
  DBuf<uint8_t> d_ref_idx;        // kRef* code per global column
  DBuf<int32_t> d_cov_diff;       // pass-1 coverage difference arrays (unique, redundant)
  DBuf<int2> d_tile_sum;          // pass-1 coverage scan tile sums
  DBuf<int32_t> d_pos_target, d_win_rb, d_win_re, d_target_group, d_target_scored;
  DBuf<uint8_t> d_raw_b, d_raw_q;     // the caller's bases and qualities, as uploaded
  DBuf<int64_t> d_raw_off;
  DBuf<FoldRow> d_fold;
  DBuf<RunRec> d_runrec;
  DBuf<TargetInfo> d_targets;
  DBuf<unsigned long long> d_counts, d_hist;
  DBuf<uint32_t> d_err;
  DBuf<double> d_tab_lp, d_tab_r;
  PinnedBuf<double> h_tab_r;       // the scorer's ratio table, uploaded while the column pass runs
  PinnedBuf<unsigned long long> h_ec_counts, h_ec_hist;   // error_count's tables, copied back
  PinnedBuf<uint32_t> h_ec_err;
  int32_t idrec_sm = -1, idrec_q = -1;   // covariate maxima of an alignment view already queued
  DBuf<uint32_t> d_start_bits;        // preprocess mode: read-start bits, 2 per column
  DBuf<unsigned long long> d_start_counts;   // preprocess mode: per target (no start, start)
  DBuf<uint32_t> d_key_scratch;      // read lists of the scorer and k_fix_columns for deep samples
  DBuf<uint8_t> d_col_flags;
  DBuf<int32_t> d_col_unique, d_col_total;
  DBuf<int2> d_cand;
  DBuf<int32_t> d_counters;   // [0] n_cand, [1] n_ins, [2] n_runs, [3] error bits, [4] n_fix
  DBuf<int32_t> d_fix;
  DBuf<IdRec> d_idrec;
  DBuf<uint32_t> d_ct_u, d_ct_raw, d_ct_beg;   // BAM2COV column counts (coverage table)
  DBuf<double> d_ct_red;
  DBuf<int32_t> d_ct_misc;                    // error bits of the coverage-table kernels
  DBuf<uint8_t> d_rg;                         // @RG index of each read (255: none)
  DBuf<int32_t> d_ins;
  DBuf<RaOut> d_ra;
  // launch of the column pass (bsq_session_set_column_launch, bsq_session_tune_columns)
  int32_t col_blocks = 0;   // 0: one window per wavefront, at most 2048 blocks
  int32_t col_lds = 1;      // 1: the likelihood table staged in LDS when it fits; 0: read from global memory
  int32_t col_waves = kColWaves;   // wavefronts per block: 8 or 4
  // device writer of the genome diff (gd_emit.hip)
  bool gd_on_device = true;
  int gd_cap = 4096;                       // entries the writer's lists hold (grows when a pass needs more)
  size_t gd_text_hint = (size_t)64 << 10;  // text bytes copied back with the writer's counters
  PinnedBuf<int32_t> h_gd_cnt, h_gd_meta;
  PinnedBuf<char> h_gd_text, h_gd_names;
  DBuf<GdEnt> d_gd_ents;
  DBuf<int32_t> d_gd_rank_id, d_gd_rank_out, d_gd_lens, d_gd_meta, d_gd_cnt;
  DBuf<int64_t> d_gd_off;
  DBuf<char> d_gd_slots, d_gd_names, d_gd_text;
  DBuf<int2> d_events, d_runs;
  DBuf<int4> d_run_out;

  ~Session() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (evc) (void)hipEventDestroy(evc);
    for (hipEvent_t e : evk)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    if (stream2) (void)hipStreamDestroy(stream2);
  }
};

static uint32_t read_file_of(const bsq_read_files* rf, const bsq_alignments* a, int64_t r) {
  if (!rf || rf->n_groups <= 0) return 0;
  uint32_t g = 0;
  if (rf->n_groups > 1) {
    int32_t rg = a->rg ? a->rg[r] : -1;
    g = (rg < 0 || rg >= rf->n_groups) ? 0u : (uint32_t)rg;
  }
  const bool read2 = (a->flag[r] & 0x80) != 0;
  return rf->base[g] + ((read2 && rf->count[g] > 1) ? 1u : 0u);
}

static bool ref_consuming(uint32_t op) { return op == 0 || op == 2 || op == 3 || op == 7 || op == 8; }

// Pinned staging buffers for host-to-device copies of large pageable arrays, shared by the sessions
// of a device: two 32 MiB buffers per set, host threads fill one while the copy engine drains the
// other.  A set is taken for one upload and put back afterwards.
struct StagingSet {
  int device = -1;
  uint8_t* buf[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
};
constexpr size_t kStageBytes = (size_t)32 << 20;
// Sets per device: more uploads at once than this share the PCIe link without gaining bandwidth,
// so a further upload waits for a set to come back instead of pinning more memory.
constexpr int kStageSetsPerDevice = 3;
static std::mutex g_stage_mu;
static std::condition_variable g_stage_cv;
static std::vector<StagingSet*> g_stage_free;
static std::map<int, int> g_stage_count;   // sets created per device

static StagingSet* take_staging(int device) {
  {
    std::unique_lock<std::mutex> g(g_stage_mu);
    for (;;) {
      for (size_t i = 0; i < g_stage_free.size(); i++)
        if (g_stage_free[i]->device == device) {
          StagingSet* s = g_stage_free[i];
          g_stage_free.erase(g_stage_free.begin() + (std::ptrdiff_t)i);
          return s;
        }
      if (g_stage_count[device] < kStageSetsPerDevice) {
        g_stage_count[device]++;
        break;
      }
      g_stage_cv.wait(g);
    }
  }
  StagingSet* s = new StagingSet();
  s->device = device;
  try {
    for (int b = 0; b < 2; b++) {
      HIP_OK(hipHostMalloc((void**)&s->buf[b], kStageBytes, hipHostMallocDefault));
      HIP_OK(hipEventCreateWithFlags(&s->done[b], hipEventDisableTiming));
    }
  } catch (...) {
    for (int b = 0; b < 2; b++) {
      if (s->buf[b]) (void)hipHostFree(s->buf[b]);
      if (s->done[b]) (void)hipEventDestroy(s->done[b]);
    }
    delete s;
    {
      std::lock_guard<std::mutex> g(g_stage_mu);
      g_stage_count[device]--;
    }
    g_stage_cv.notify_all();
    throw;
  }
  return s;
}

static void put_staging(StagingSet* s) {
  {
    std::lock_guard<std::mutex> g(g_stage_mu);
    g_stage_free.push_back(s);
  }
  g_stage_cv.notify_all();
}

// dst (device) <- src (pageable host), in 32 MiB chunks through a staging set; the copies are
// queued on S.stream, and the set is returned once the last one has completed.  Small arrays go
// straight to hipMemcpyAsync.
static void upload_staged(Session& S, void* dst, const void* src, size_t bytes) {
  if (bytes < (kStageBytes >> 2)) {
    HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S.stream));
    return;
  }
  StagingSet* st = take_staging(S.device);
  try {
    size_t i = 0;
    for (size_t off = 0; off < bytes; off += kStageBytes, i++) {
      const int b = (int)(i & 1);
      const size_t n = std::min(kStageBytes, bytes - off);
      if (i >= 2) HIP_OK(hipEventSynchronize(st->done[b]));   // the copy that last read this buffer
      const int parts = 8;
      pool_run(parts, [&](int k) {
        const size_t p0 = n * (size_t)k / parts, p1 = n * (size_t)(k + 1) / parts;
        memcpy(st->buf[b] + p0, (const uint8_t*)src + off + p0, p1 - p0);
      });
      HIP_OK(hipMemcpyAsync((uint8_t*)dst + off, st->buf[b], n, hipMemcpyHostToDevice, S.stream));
      HIP_OK(hipEventRecord(st->done[b], S.stream));
    }
    for (int b = 0; b < 2; b++) HIP_OK(hipEventSynchronize(st->done[b]));
  } catch (...) {
    (void)hipStreamSynchronize(S.stream);
    put_staging(st);
    throw;
  }
  put_staging(st);
}

// Pack the alignments into the device CSR (DRead records + window index) and upload.
static void session_load(Session& S, const bsq_reference* ref, const bsq_alignments* a, const bsq_read_files* rf) {
  const double t0 = now_ms();
  S.idrec_sm = S.idrec_q = -1;   // no alignment view of the new sample yet
  if (!ref || !a) throw BsqError(BSQ_ERR_ARG, "null reference or alignments");
  S.n_targets = ref->n_targets;
  S.names.clear();
  S.lengths.clear();
  S.tgt_off.clear();
  int64_t off = 0;
  for (int t = 0; t < ref->n_targets; t++) {
    S.names.push_back(ref->names[t]);
    S.lengths.push_back(ref->lengths[t]);
    S.tgt_off.push_back((int32_t)off);
    off += ref->lengths[t] + 1;
  }
  if (off >= ((int64_t)1 << 31) - 4096) throw BsqError(BSQ_ERR_LIMIT, "total reference length exceeds 2^31");
  S.n_pos = (int32_t)off;
  S.h_ref.assign((size_t)off + kWave, '\0');
  S.h_pos_target.assign((size_t)off, -1);
  for (int t = 0; t < ref->n_targets; t++) {
    memcpy(&S.h_ref[S.tgt_off[t]], ref->seqs[t], (size_t)ref->lengths[t]);
    std::fill(S.h_pos_target.begin() + S.tgt_off[t], S.h_pos_target.begin() + S.tgt_off[t] + ref->lengths[t], t);
  }
  // reads, packed in two parallel passes over the input: (1) each alignment's record and whether it
  // is packed (or the first invariant it breaks); (2) after prefix sums over the packed ones, the
  // byte and CIGAR offsets and the copies.  Errors are those of the first offending alignment in
  // input order, as a sequential walk would report them.
  const int64_t n_in = a->n_reads;
  std::vector<uint8_t> keep((size_t)n_in, 0);
  struct Fail {
    int64_t r = -1;
    int code = 0;
    std::string msg;
  };
  const int n_chunks = (int)std::min<int64_t>(256, std::max<int64_t>(1, n_in / 4096));
  auto chunk_begin = [&](int c) { return n_in * c / n_chunks; };
  std::vector<Fail> fails(n_chunks);
  std::vector<int64_t> c_kept(n_chunks, 0), c_bytes(n_chunks, 0), c_cig(n_chunks, 0), c_nab(n_chunks, 0);
  std::vector<int32_t> c_maxspan(n_chunks, 1);
  std::vector<uint32_t> c_maxrf(n_chunks, 0);
  std::vector<std::array<uint8_t, 256>> c_mq(n_chunks);
  pool_run(n_chunks, [&](int c) {
    Fail& f = fails[(size_t)c];
    std::array<uint8_t, 256>& mq = c_mq[(size_t)c];
    mq.fill(0);
    auto fail = [&](int64_t r, int code, const char* msg) {
      f.r = r;
      f.code = code;
      f.msg = msg;
    };
    // the chunk's sums are kept in locals and stored once: neighbouring chunks' slots share cache
    // lines, and updating them per alignment from many threads would bounce those lines
    int64_t kept = 0, bytes = 0, ncig = 0, nab_c = 0;
    int32_t mspan = 1;
    uint32_t mrf = 0;
    for (int64_t r = chunk_begin(c); r < chunk_begin(c + 1) && f.r < 0; r++) {
      if (a->flag[r] & 0x4) continue;
      const int32_t tid = a->tid[r];
      if (tid < 0) continue;
      if (tid >= ref->n_targets) {
        fail(r, BSQ_ERR_FORMAT, "alignment target id out of range");
        break;
      }
      const int64_t c0 = a->cigar_off[r], c1 = a->cigar_off[r + 1];
      const int64_t nc = c1 - c0;
      if (nc == 0) continue;
      if (nc > 255) {
        fail(r, BSQ_ERR_LIMIT, "alignment with more than 255 CIGAR operations");
        break;
      }
      const int64_t lq = a->seq_off[r + 1] - a->seq_off[r];
      if (lq > 65535) {
        fail(r, BSQ_ERR_LIMIT, "read longer than 65535 bases");
        break;
      }
      int64_t span = 0, qlen = 0;
      for (int64_t k = c0; k < c1; k++) {
        const uint32_t op = a->cigar[k] & 15, len = a->cigar[k] >> 4;
        if (ref_consuming(op)) span += len;
        if (op == 0 || op == 1 || op == 4 || op == 7 || op == 8) qlen += len;
      }
      if (span == 0) continue;   // covers no reference base: not in any pileup, not packed
      if (a->pos[r] < 0 || a->pos[r] + span > ref->lengths[tid]) {
        fail(r, BSQ_ERR_FORMAT, "alignment extends past its reference");
        break;
      }
      if (qlen != lq) {
        fail(r, BSQ_ERR_FORMAT, "CIGAR query length differs from SEQ length");
        break;
      }
      const int32_t x1 = (a->x1 && a->x1[r] >= 0) ? a->x1[r] : 1;
      if (x1 > 65535) {
        fail(r, BSQ_ERR_LIMIT, "X1 redundancy above 65535");
        break;
      }
      const uint32_t rfi = read_file_of(rf, a, r);
      if (rfi > 255) {
        fail(r, BSQ_ERR_LIMIT, "more than 256 read files");
        break;
      }
      if (a->rg && a->rg[r] >= 255) {
        fail(r, BSQ_ERR_LIMIT, "more than 255 read groups");
        break;
      }
      mq[a->mapq[r]] = 1;
      keep[(size_t)r] = 1;
      kept++;
      bytes += (lq + 7) & ~(int64_t)7;
      ncig += nc;
      nab_c += span;
      mspan = std::max<int32_t>(mspan, (int32_t)span);
      mrf = std::max(mrf, rfi);
    }
    c_kept[(size_t)c] = kept;
    c_bytes[(size_t)c] = bytes;
    c_cig[(size_t)c] = ncig;
    c_nab[(size_t)c] = nab_c;
    c_maxspan[(size_t)c] = mspan;
    c_maxrf[(size_t)c] = mrf;
  });
  std::vector<int64_t> o_kept(n_chunks + 1, 0), o_bytes(n_chunks + 1, 0), o_cig(n_chunks + 1, 0);
  for (int c = 0; c < n_chunks; c++) {
    o_kept[c + 1] = o_kept[c] + c_kept[c];
    o_bytes[c + 1] = o_bytes[c] + c_bytes[c];
    o_cig[c + 1] = o_cig[c] + c_cig[c];
  }
  int64_t fail_r = -1;
  Fail first;
  for (int c = 0; c < n_chunks; c++)
    if (fails[c].r >= 0) {
      first = fails[c];
      fail_r = fails[c].r;
      break;
    }
  const int64_t n_pack = o_kept[n_chunks];
  // the packed arrays are written (and so first touched) by the threads that fill them
  std::unique_ptr<DRead[]> dr(new DRead[(size_t)std::max<int64_t>(1, n_pack)]);
  std::vector<uint8_t> h_rg((size_t)n_pack);   // @RG index of each packed read (255: none)
  std::vector<int64_t> raw_off((size_t)n_pack), src((size_t)n_pack);
  std::vector<uint32_t> cig((size_t)o_cig[n_chunks]);
  pool_run(n_chunks, [&](int c) {
    int64_t k = o_kept[c], bytes = o_bytes[c], co = o_cig[c];
    for (int64_t r = chunk_begin(c); r < chunk_begin(c + 1); r++) {
      if (!keep[(size_t)r]) continue;
      const int64_t c0 = a->cigar_off[r], c1 = a->cigar_off[r + 1];
      const int64_t lq = a->seq_off[r + 1] - a->seq_off[r];
      int64_t span = 0, qlen = 0;
      int n_ref_ops = 0, n_ins = 0;
      uint32_t first_ref_op = 0;
      int64_t y0 = 0;
      bool seen_ref = false;
      for (int64_t q = c0; q < c1; q++) {
        const uint32_t op = a->cigar[q] & 15, len = a->cigar[q] >> 4;
        cig[(size_t)(co + (q - c0))] = a->cigar[q];
        if (ref_consuming(op)) {
          span += len;
          if (!seen_ref) first_ref_op = op;
          seen_ref = true;
          n_ref_ops++;
        }
        if (op == 0 || op == 1 || op == 4 || op == 7 || op == 8) qlen += len;
        if (op == 1 || op == 6) n_ins++;
        if (!seen_ref && (op == 1 || op == 4)) y0 += len;
      }
      DRead d;
      memset(&d, 0, sizeof(d));
      const int64_t gp = (int64_t)S.tgt_off[a->tid[r]] + a->pos[r];
      d.seq_off = bytes;
      d.cig_off = (uint32_t)co;
      d.gpos = (int32_t)gp;
      d.gend = (int32_t)(gp + span);
      d.l_qseq = (uint16_t)lq;
      d.n_cigar = (uint8_t)(c1 - c0);
      const bool simple = (n_ref_ops == 1) && (n_ins == 0) && (first_ref_op == 0 || first_ref_op == 7 || first_ref_op == 8);
      d.flags = (uint8_t)(((a->flag[r] & 0x10) ? kReadReversed : 0) | (simple ? kReadSimple : 0));
      d.y0 = (uint16_t)y0;
      // query_start_1 / query_end_1 / query_bounds_1 (alignment.cpp:125-289)
      {
        uint32_t qs = 1;
        for (int64_t q = c0; q < c1; q++) {
          if ((a->cigar[q] & 15) != 4) break;
          qs += a->cigar[q] >> 4;
        }
        int32_t qe = (int32_t)qlen;
        for (int64_t q = c1 - 1; q > c0; q--) {
          if ((a->cigar[q] & 15) != 4) break;
          qe -= (int32_t)(a->cigar[q] >> 4);
        }
        int32_t be = (int32_t)qlen;
        for (int64_t q = c1 - 1; q > c0; q--) {
          const uint32_t op = a->cigar[q] & 15;
          if (op != 4 && op != 5 && op != 3) break;
          if (op == 4) be -= (int32_t)(a->cigar[q] >> 4);
        }
        d.qs0 = (uint16_t)(qs - 1);
        d.qe0 = (uint16_t)(qe - 1);
        d.qbe0 = (uint16_t)(be - 1);
      }
      d.x1 = (uint16_t)((a->x1 && a->x1[r] >= 0) ? a->x1[r] : 1);
      d.xl = (int16_t)((a->xl && a->xl[r] >= 0) ? std::min(a->xl[r], 32767) : -1);
      d.xr = (int16_t)((a->xr && a->xr[r] >= 0) ? std::min(a->xr[r], 32767) : -1);
      d.rfile = (uint8_t)read_file_of(rf, a, r);
      d.mslot = a->mapq[r];   // remapped below
      bytes += (lq + 7) & ~(int64_t)7;
      co += c1 - c0;
      raw_off[(size_t)k] = a->seq_off[r];
      src[(size_t)k] = r;
      h_rg[(size_t)k] = (a->rg && a->rg[r] >= 0) ? (uint8_t)a->rg[r] : (uint8_t)255;
      dr[(size_t)k] = d;
      k++;
    }
  });
  // coordinate order over consecutive packed alignments; the sequential walk reports whichever
  // comes first in input order: an alignment that starts before its packed predecessor, or the
  // first alignment that breaks another invariant
  {
    std::vector<int64_t> unsorted_at(n_chunks, -1);
    pool_run(n_chunks, [&](int c) {
      for (int64_t k = std::max<int64_t>(1, o_kept[c]); k < o_kept[c + 1]; k++)
        if (dr[(size_t)k].gpos < dr[(size_t)k - 1].gpos) {
          unsorted_at[(size_t)c] = src[(size_t)k];
          break;
        }
    });
    int64_t sort_r = -1;
    for (int c = 0; c < n_chunks; c++)
      if (unsorted_at[c] >= 0) {
        sort_r = unsorted_at[c];
        break;
      }
    if (sort_r >= 0 && (fail_r < 0 || sort_r < fail_r))
      throw BsqError(BSQ_ERR_FORMAT, "alignments are not sorted by coordinate");
    if (fail_r >= 0) throw BsqError(first.code, first.msg);
  }
  const int64_t dev_bytes = o_bytes[n_chunks];
  int32_t maxspan = 1;
  int64_t nab = 0;
  uint32_t max_rfile = 0;
  std::array<uint8_t, 256> mq_seen{};
  for (int c = 0; c < n_chunks; c++) {
    maxspan = std::max(maxspan, c_maxspan[c]);
    nab += c_nab[c];
    max_rfile = std::max(max_rfile, c_maxrf[c]);
    for (int v = 0; v < 256; v++) mq_seen[v] |= c_mq[c][v];
  }
  // MAPQ slots in ascending MAPQ order
  S.mapq_values.clear();
  {
    std::array<uint8_t, 256> slot_of{};
    uint8_t k = 0;
    for (int v = 0; v < 256; v++)
      if (mq_seen[v]) {
        slot_of[v] = k++;
        S.mapq_values.push_back((uint8_t)v);
      }
    pool_run(n_chunks, [&](int c) {
      for (int64_t i = o_kept[c]; i < o_kept[c + 1]; i++) dr[(size_t)i].mslot = slot_of[dr[(size_t)i].mslot];
    });
  }
  S.n_read_files = (int32_t)max_rfile + 1;
  S.n_dreads = (int64_t)(size_t)n_pack;
  S.h_src = std::move(src);
  S.n_input_reads = a->n_reads;
  S.h_gpos.resize((size_t)n_pack);
  S.h_gend_max.resize((size_t)n_pack);
  for (size_t i = 0; i < (size_t)n_pack; i++) {
    S.h_gpos[i] = dr[i].gpos;
    S.h_gend_max[i] = std::max(dr[i].gend, i ? S.h_gend_max[i - 1] : (int32_t)-1);
  }
  dp_select(a, ref->n_targets, S.dp_reads);
  if (S.n_dreads >= ((int64_t)1 << 31)) throw BsqError(BSQ_ERR_LIMIT, "more than 2^31 alignments");
  S.aligned_bases = nab;
  // window index: alignments overlapping [64w, 64w+64)
  S.n_windows = ((int64_t)S.n_pos + kWave - 1) / kWave;
  std::vector<int32_t> rb(S.n_windows), re(S.n_windows);
  {
    size_t lo = 0, hi = 0;
    int32_t mx = 0;
    for (int64_t w = 0; w < S.n_windows; w++) {
      const int64_t w0 = w * kWave, w1 = w0 + kWave;
      while (lo < (size_t)n_pack && (int64_t)dr[lo].gpos < w0 - maxspan) lo++;
      while (hi < (size_t)n_pack && (int64_t)dr[hi].gpos < w1) hi++;
      rb[w] = (int32_t)lo;
      re[w] = (int32_t)hi;
      mx = std::max<int32_t>(mx, (int32_t)(hi - lo));
    }
    S.max_window_reads = mx;
  }
  const double t1 = now_ms();
  // upload
  const double t2 = now_ms();
  S.d_reads.alloc((size_t)n_pack);
  HIP_OK(hipMemcpyAsync(S.d_reads.p, dr.get(), (size_t)n_pack * sizeof(DRead), hipMemcpyHostToDevice, S.stream));
  S.d_rg.alloc(h_rg.size());
  HIP_OK(hipMemcpyAsync(S.d_rg.p, h_rg.data(), h_rg.size(), hipMemcpyHostToDevice, S.stream));
  S.d_cigar.alloc(cig.size());
  HIP_OK(hipMemcpyAsync(S.d_cigar.p, cig.data(), cig.size() * sizeof(uint32_t), hipMemcpyHostToDevice, S.stream));
  // bases and qualities: uploaded as given, then repacked on the device to the aligned layout
  const size_t nbases = (size_t)a->seq_off[a->n_reads];
  // the column pass addresses a read's bytes with a 32-bit offset
  if (dev_bytes + 16 >= ((int64_t)1 << 32)) throw BsqError(BSQ_ERR_LIMIT, "more than 4 GiB of read bases");
  S.d_bases.alloc((size_t)dev_bytes + 16);
  S.d_quals.alloc((size_t)dev_bytes + 16);
  HIP_OK(hipMemsetAsync(S.d_bases.p, 0, (size_t)dev_bytes + 16, S.stream));
  HIP_OK(hipMemsetAsync(S.d_quals.p, 0, (size_t)dev_bytes + 16, S.stream));
  if (nbases && n_pack > 0) {
    // the uploaded arrays stay with the session for its next sample (no allocation per load)
    S.d_raw_b.alloc(nbases);
    S.d_raw_q.alloc(nbases);
    S.d_raw_off.alloc(raw_off.size());
    upload_staged(S, S.d_raw_b.p, a->seq, nbases);
    upload_staged(S, S.d_raw_q.p, a->qual, nbases);
    HIP_OK(hipMemcpyAsync(S.d_raw_off.p, raw_off.data(), raw_off.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                          S.stream));
    hipLaunchKernelGGL(k_repack_reads, dim3(2048), dim3(256), 0, S.stream, S.d_reads.p, S.d_raw_off.p, (int64_t)n_pack,
                       S.d_raw_b.p, S.d_raw_q.p, S.d_bases.p, S.d_quals.p);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(S.stream));
  }
  S.d_ref.alloc(S.h_ref.size());
  HIP_OK(hipMemcpyAsync(S.d_ref.p, S.h_ref.data(), S.h_ref.size(), hipMemcpyHostToDevice, S.stream));
  {
    std::vector<uint8_t> ri(S.h_ref.size());
    const int nch = (int)std::min<size_t>(64, std::max<size_t>(1, ri.size() >> 16));
    pool_run(nch, [&](int c) {
      const size_t b = ri.size() * (size_t)c / nch, e = ri.size() * (size_t)(c + 1) / nch;
      for (size_t i = b; i < e; i++) {
        switch (S.h_ref[i]) {
          case 'A': ri[i] = 0; break;
          case 'C': ri[i] = 1; break;
          case 'G': ri[i] = 2; break;
          case 'T': ri[i] = 3; break;
          case '.': ri[i] = kRefDot; break;
          case 'N': ri[i] = kRefN; break;
          default: ri[i] = kRefOther; break;
        }
      }
    });
    S.d_ref_idx.alloc(ri.size() + kRefIdxFrontPad);
    HIP_OK(hipMemsetAsync(S.d_ref_idx.p, kRefOther, ri.size() + kRefIdxFrontPad, S.stream));
    HIP_OK(hipMemcpyAsync(S.d_ref_idx.p + kRefIdxFrontPad, ri.data(), ri.size(), hipMemcpyHostToDevice, S.stream));
    HIP_OK(hipStreamSynchronize(S.stream));
  }
  S.d_pos_target.alloc(S.h_pos_target.size());
  HIP_OK(hipMemcpyAsync(S.d_pos_target.p, S.h_pos_target.data(), S.h_pos_target.size() * sizeof(int32_t), hipMemcpyHostToDevice, S.stream));
  S.d_win_rb.alloc(rb.size());
  S.d_win_re.alloc(re.size());
  HIP_OK(hipMemcpyAsync(S.d_win_rb.p, rb.data(), rb.size() * sizeof(int32_t), hipMemcpyHostToDevice, S.stream));
  HIP_OK(hipMemcpyAsync(S.d_win_re.p, re.data(), re.size() * sizeof(int32_t), hipMemcpyHostToDevice, S.stream));
  S.d_err.alloc(1);
  S.d_counters.alloc(8);
  HIP_OK(hipStreamSynchronize(S.stream));
  S.timing.pack_ms = t1 - t0;
  S.timing.h2d_ms = now_ms() - t2;
  S.timing.aligned_bases = S.aligned_bases;
  S.timing.n_reads = S.n_dreads;
  S.timing.n_positions = S.n_pos;
  S.loaded = true;
}

static std::vector<int32_t> target_order(const Session& S, const uint8_t* call) {
  std::vector<int32_t> t;
  for (int i = 0; i < S.n_targets; i++)
    if (!call || call[i]) t.push_back(i);
  std::sort(t.begin(), t.end(), [&](int32_t a, int32_t b) { return S.names[a] < S.names[b]; });
  return t;
}

static void raise_err_bits(uint32_t err) {
  if (!err) return;
  std::string m = "alignment data violates a breseq invariant:";
  if (err & 1) m += " read base not A/C/G/T/N;";
  if (err & 2) m += " reference base not A/C/G/T/N;";
  if (err & 4) m += " quality above the 'quality' covariate maximum;";
  if (err & 8) m += " read file index above the 'read_set' covariate maximum;";
  if (err & 16) m += " quality index outside the read;";
  if (err & 32) m += " CIGAR does not cover its alignment span;";
  if (err & 64) m += " device buffer capacity exceeded;";
  if (err & 128) throw BsqError(BSQ_ERR_DEVICE, "a device list held an index outside its buffer (skipped)");
  throw BsqError(BSQ_ERR_DATA, m);
}

static void check_err(Session& S) {
  uint32_t err = 0;
  HIP_OK(hipMemcpyAsync(&err, S.d_err.p, sizeof(uint32_t), hipMemcpyDeviceToHost, S.stream));
  HIP_OK(hipStreamSynchronize(S.stream));
  raise_err_bits(err);
}


}  // namespace bsq

using namespace bsq;

struct bsq_session {
  Session s;
};

struct bsq_ec_result {
  std::vector<double> no_pos_hash_pr;   // preprocess mode, per target (NaN when not visited)
  std::vector<std::vector<uint32_t>> hist;
  std::vector<uint64_t> counts;
  std::vector<double> log10p;
  std::string rates_text;
  std::vector<double> rates_table;
};

struct bsq_im_result {
  std::string text;
  std::string dp_text;
  int64_t n_ra = 0, n_mc = 0, n_un = 0;
};

namespace bsq {

// defer_rates_text: leave out.rates_text to the caller, which prints it off the critical path
// (only when nothing is written to files here).
// queue_idrec: also queue identify's alignment view for these covariate maxima behind the tables'
// copies (the pipeline's identify pass follows), so that it is built while the host works.
static void run_error_count(Session& S, const bsq_error_count_args* args, bsq_ec_result& out,
                            bool defer_rates_text = false, bool queue_idrec = false) {
  if (!S.loaded) throw BsqError(BSQ_ERR_ARG, "session has no sample loaded");
  S.timing.fit_ms = S.timing.fit_wait_ms = S.timing.pipeline_ec_ms = 0.0;
  S.timing.pipeline_ec_kernel_ms[0] = S.timing.pipeline_ec_kernel_ms[1] = 0.0;
  const double th0 = now_ms();
  const int Sm = args->read_set_max, Q = args->quality_max;
  if (Sm <= 0 || Q <= 0 || Sm > 255 || Q > 256) throw BsqError(BSQ_ERR_ARG, "covariate maxima out of range");
  const int n_bins = Sm * 25 * Q;
  const int G = std::max(1, args->n_coverage_groups);
  // scored targets: the call set only
  std::vector<int32_t> group(S.n_targets, 0), scored(S.n_targets, 0);
  for (int t = 0; t < S.n_targets; t++) {
    const bool call = !args->call_target || args->call_target[t];
    int g = args->coverage_group ? args->coverage_group[t] : 0;
    if (call && (g < 0 || g >= G)) throw BsqError(BSQ_ERR_ARG, "coverage group out of range");
    group[t] = std::max(0, std::min(g, G - 1));
    scored[t] = call ? 1 : 0;
  }
  S.d_target_group.alloc(group.size());
  S.d_target_scored.alloc(scored.size());
  HIP_OK(hipMemcpyAsync(S.d_target_group.p, group.data(), group.size() * sizeof(int32_t), hipMemcpyHostToDevice, S.stream));
  HIP_OK(hipMemcpyAsync(S.d_target_scored.p, scored.data(), scored.size() * sizeof(int32_t), hipMemcpyHostToDevice, S.stream));
  const int hist_len = S.max_window_reads + 2;
  S.d_counts.alloc(n_bins);
  S.d_hist.alloc((size_t)G * hist_len);
  HIP_OK(hipMemsetAsync(S.d_counts.p, 0, n_bins * sizeof(unsigned long long), S.stream));
  HIP_OK(hipMemsetAsync(S.d_hist.p, 0, (size_t)G * hist_len * sizeof(unsigned long long), S.stream));
  HIP_OK(hipMemsetAsync(S.d_err.p, 0, sizeof(uint32_t), S.stream));
  const int64_t n_diff = 2 * ((int64_t)S.n_pos + 1);
  S.d_cov_diff.alloc((size_t)n_diff);
  const int n_tiles = (int)(((int64_t)S.n_pos + 1 + kCovTile - 1) / kCovTile);
  S.d_tile_sum.alloc((size_t)n_tiles);
  EcParams P;
  P.reads = S.d_reads.p;
  P.cigar = S.d_cigar.p;
  P.bases = S.d_bases.p;
  P.quals = S.d_quals.p;
  P.ref_idx = S.d_ref_idx.p + kRefIdxFrontPad;
  P.pos_target = S.d_pos_target.p;
  P.target_scored = S.d_target_scored.p;
  P.n_reads = S.n_dreads;
  P.n_pos = S.n_pos;
  P.read_set_max = Sm;
  P.quality_max = Q;
  P.n_bins = n_bins;
  P.do_errors = args->do_errors ? 1 : 0;
  P.cov_diff = S.d_cov_diff.p;
  P.counts = S.d_counts.p;
  P.err = S.d_err.p;
  CovParams C;
  C.cov_diff = S.d_cov_diff.p;
  C.pos_target = S.d_pos_target.p;
  C.target_group = S.d_target_group.p;
  C.target_scored = S.d_target_scored.p;
  C.n_pos = S.n_pos;
  C.n_tiles = n_tiles;
  C.tile_sum = S.d_tile_sum.p;
  C.n_groups_lds = std::min(G, 4);
  C.hist_lds = 1024;
  C.hist_len = hist_len;
  C.hist = S.d_hist.p;
  C.err = S.d_err.p;
  const bool pre = args->preprocess_stage != 0;
  P.start_bits = nullptr;
  P.end_min_len = args->unmatched_end_minimum_read_length;
  P.end_factor = args->unmatched_end_length_factor;
  C.start_bits = nullptr;
  C.start_counts = nullptr;
  if (pre) {
    const size_t words = ((size_t)S.n_pos + 15) / 16;
    S.d_start_bits.alloc(words);
    S.d_start_counts.alloc((size_t)2 * S.n_targets);
    HIP_OK(hipMemsetAsync(S.d_start_bits.p, 0, words * sizeof(uint32_t), S.stream));
    HIP_OK(hipMemsetAsync(S.d_start_counts.p, 0, (size_t)2 * S.n_targets * sizeof(unsigned long long), S.stream));
    P.start_bits = S.d_start_bits.p;
    C.start_bits = S.d_start_bits.p;
    C.start_counts = S.d_start_counts.p;
  }
  const size_t lds_table = (size_t)n_bins * sizeof(uint32_t);
  const size_t lds_hist = (size_t)C.n_groups_lds * C.hist_lds * sizeof(uint32_t);
  if (lds_table > 64 * 1024) throw BsqError(BSQ_ERR_LIMIT, "covariate table too large for LDS");
  // Two work items per wavefront, in the grid that splits the items evenly over the wavefronts
  // (rounded up to a multiple of the 8 XCDs, so that no wavefront takes a third item while the
  // others idle).  Measured on the MI355X, balanced grids of 1-4 items per wavefront all ran about
  // 20% faster than 2,048 blocks of 11-12 items per wavefront; the per-block flush of the LDS
  // table does not limit it.
  const int64_t chunks = (S.n_dreads + kEcChunk - 1) / kEcChunk;
  int64_t ec_blocks = std::max<int64_t>(1, (chunks + 2 * kEcWaves - 1) / (2 * kEcWaves));
  if (ec_blocks >= 8) ec_blocks = (ec_blocks + 7) / 8 * 8;
  if (ec_blocks > INT32_MAX) throw BsqError(BSQ_ERR_LIMIT, "too many reads for one error_count launch");
  const int ec_grid = (int)ec_blocks;
  S.timing.host_pre_ms = now_ms() - th0;
  HIP_OK(hipEventRecord(S.ev0, S.stream));
  hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, S.stream, S.d_cov_diff.p, n_diff, 0);
  hipLaunchKernelGGL(k_ec_reads, dim3(ec_grid), dim3(256), lds_table, S.stream, P);
  HIP_OK(hipGetLastError());
  if (pre) hipLaunchKernelGGL(k_read_starts, dim3((unsigned)((S.n_dreads + 255) / 256)), dim3(256), 0, S.stream, P);
  HIP_OK(hipEventRecord(S.evk[0], S.stream));
  hipLaunchKernelGGL(k_cov_tile_sums, dim3(n_tiles), dim3(256), 0, S.stream, C);
  hipLaunchKernelGGL(k_cov_tile_scan, dim3(1), dim3(1024), 0, S.stream, C);
  hipLaunchKernelGGL(k_cov_hist, dim3(n_tiles), dim3(256), lds_hist, S.stream, C);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(S.ev1, S.stream));
  // the tables come back with the kernels' own completion: pinned copies queued behind them
  S.h_ec_counts.alloc((size_t)n_bins);
  S.h_ec_hist.alloc((size_t)G * hist_len);
  HIP_OK(hipMemcpyAsync(S.h_ec_counts.p, S.d_counts.p, n_bins * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        S.stream));
  HIP_OK(hipMemcpyAsync(S.h_ec_hist.p, S.d_hist.p, (size_t)G * hist_len * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, S.stream));
  S.h_ec_err.alloc(1);
  HIP_OK(hipMemcpyAsync(S.h_ec_err.p, S.d_err.p, sizeof(uint32_t), hipMemcpyDeviceToHost, S.stream));
  HIP_OK(hipEventRecord(S.evk[7], S.stream));
  S.idrec_sm = S.idrec_q = -1;
  if (queue_idrec) {
    S.d_idrec.alloc((size_t)std::max<int64_t>(1, S.n_dreads));
    hipLaunchKernelGGL(k_prep_reads, dim3(1024), dim3(256), 0, S.stream, S.d_reads.p, S.n_dreads, Sm, Q, S.d_idrec.p);
    HIP_OK(hipGetLastError());
    S.idrec_sm = Sm;
    S.idrec_q = Q;
  }
  const double t0 = now_ms();
  HIP_OK(hipEventSynchronize(S.evk[7]));
  S.timing.d2h_ms = now_ms() - t0;
  float ms = 0, ms_reads = 0;
  HIP_OK(hipEventElapsedTime(&ms, S.ev0, S.ev1));
  HIP_OK(hipEventElapsedTime(&ms_reads, S.ev0, S.evk[0]));
  S.timing.kernel_ms = ms;
  for (double& d : S.timing.kernel_detail_ms) d = 0.0;
  S.timing.kernel_detail_ms[0] = ms_reads;
  S.timing.kernel_detail_ms[5] = ms - ms_reads;
  raise_err_bits(*S.h_ec_err.p);
  const unsigned long long* counts = S.h_ec_counts.p;
  const unsigned long long* hist = S.h_ec_hist.p;
  const double t1 = now_ms();
  out.counts.assign(counts, counts + n_bins);
  out.hist.assign(G, std::vector<uint32_t>());
  for (int g = 0; g < G; g++) {
    int64_t last = -1;
    for (int c = 0; c < hist_len; c++)
      if (hist[(size_t)g * hist_len + c]) last = c;
    // error_count resizes its table to the largest coverage seen, so only groups that were visited
    // have entries
    out.hist[g].assign((size_t)(last + 1), 0);
    for (int64_t c = 0; c <= last; c++) out.hist[g][c] = (uint32_t)hist[(size_t)g * hist_len + c];
  }
  out.no_pos_hash_pr.assign(S.n_targets, std::numeric_limits<double>::quiet_NaN());
  if (pre) {
    std::vector<unsigned long long> sc((size_t)2 * S.n_targets);
    HIP_OK(hipMemcpy(sc.data(), S.d_start_counts.p, sc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int t = 0; t < S.n_targets; t++) {
      if (!scored[t]) continue;
      const double total = (double)(sc[2 * t] + sc[2 * t + 1]);
      out.no_pos_hash_pr[t] = (total != 0) ? (double)sc[2 * t] / total : 1.0;
    }
  }
  // the Laplace table and the table identify_mutations reads back from error_rates.tab (each value
  // as strtod returns it from its %g text, computed directly), split over host threads by quality
  out.log10p.resize((size_t)n_bins);
  out.rates_table.resize((size_t)n_bins);
  {
    const int tasks = std::min(8, Q);
    pool_run(tasks, [&](int k) {
      const int q0 = Q * k / tasks, q1 = Q * (k + 1) / tasks;
      counts_to_log10_prob_range(Sm, Q, out.counts, out.log10p.data(), q0, q1);
      for (size_t i = (size_t)25 * Sm * q0; i < (size_t)25 * Sm * q1; i++) out.rates_table[i] = g6_round_trip(out.log10p[i]);
    });
  }
  if (!defer_rates_text || args->output_dir) out.rates_text = error_rates_text(Sm, Q, out.log10p);
  if (args->output_dir) {
    const std::string dir = args->output_dir;
    if (args->do_coverage) {
      for (int g = 0; g < G; g++) {
        std::ofstream f(dir + "/" + std::to_string(g) + ".unique_only_coverage_distribution.tab");
        if (!f) throw BsqError(BSQ_ERR_IO, "could not write coverage distribution in " + dir);
        f << coverage_text(out.hist[g]);
      }
    }
    if (args->do_errors) {
      const std::string rates = args->error_rates_file ? std::string(args->error_rates_file) : dir + "/error_rates.tab";
      std::ofstream f(rates);
      if (!f) throw BsqError(BSQ_ERR_IO, "could not write " + rates);
      f << out.rates_text;
      std::vector<std::string> bq = base_qual_texts(Sm, Q, out.counts);
      for (int r = 0; r < Sm; r++) {
        std::string name = (r < args->read_files.n_files && args->read_files.names) ? args->read_files.names[r] : std::to_string(r);
        std::ofstream fb(dir + "/base_qual_error_prob." + name + ".tab");
        if (!fb) throw BsqError(BSQ_ERR_IO, "could not write base_qual_error_prob in " + dir);
        fb << bq[r];
      }
    }
  }
  S.timing.host_post_ms = now_ms() - t1;
}

// fill_read_base_likelihoods (identify_mutations.cpp:3359-3384) for every (MAPQ, read file,
// quality, observed base, strand): the per-read cache depends on nothing else, so it is computed
// once per key on the host with the same libm calls and looked up by the kernels.
//   lp / rr: keyed by (mslot, read file, quality, top-strand obs, strand) -- the exact scorer
//   fold: keyed by (mslot, read file, quality, read-strand obs), hypotheses in read-strand
//                order; a bottom-strand read reads the row of the complemented observation with
//                its hypotheses complemented (the reference complements both before the lookup).
// A bottom-strand key looks up the same probabilities as the top-strand key of the complemented
// observation, with the hypotheses complemented, so its values are that key's values permuted:
// only the top-strand keys are computed.  The column pass needs lp and fold; rr (the scorer's
// ratios) is built by the second call while the column pass runs.
struct LikelihoodTables {
  std::vector<double> lp, rr;
  std::vector<FoldRow> fold;
  size_t keys = 0;
};

static inline int comp_class(int i) { return i < 4 ? 3 - i : 4; }

static void build_column_tables(const Session& S, const bsq_identify_args* a, LikelihoodTables& L) {
  const int Sm = a->read_set_max, Q = a->quality_max;
  const size_t nm = S.mapq_values.size();
  L.keys = nm * Sm * Q * 10;
  L.lp.assign(L.keys * 5, 0.0);
  // one extra all-zero row: the row a filtered read reads
  L.fold.assign(nm * Sm * Q * 5 + 1, FoldRow{{0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, 0.0f, 0.0f, 0.0f});
  // the bins and rows of one quality depend on nothing else (a thread pool's wake-up costs more
  // than the few thousand logarithms this takes, so the qualities run in order)
  std::vector<double> prob(a->n_bins);
  auto qualities = [&](int q0, int q1) {
    for (int64_t i = (int64_t)25 * Sm * q0; i < (int64_t)25 * Sm * q1; i++) prob[i] = pow(10, a->log10_prob_table[i]);
    for (size_t m = 0; m < nm; m++) {
      const double incorrect = pow(10, -static_cast<double>(S.mapq_values[m]) / 10);
      const double correct = 1 - incorrect;
      const double uniform = 1.0 / 5.0;
      for (int rs = 0; rs < Sm; rs++)
        for (int q = q0; q < q1; q++) {
          const size_t kq = ((m * Sm + rs) * Q + q) * 10;
          for (int obs = 0; obs < 5; obs++) {
            double l[5];
            double mx = -std::numeric_limits<double>::max();
            for (int b = 0; b < 5; b++) {
              double pr = correct * prob[rs + Sm * b + 5 * Sm * obs + 25 * Sm * q] + incorrect * uniform;
              if (pr < 0.0) pr = 0.0;
              l[b] = log10(pr);
              mx = std::max(mx, l[b]);
            }
            const size_t top = kq + obs * 2, bot = kq + comp_class(obs) * 2 + 1;
            for (int b = 0; b < 5; b++) {
              L.lp[top * 5 + b] = l[b];
              L.lp[bot * 5 + comp_class(b)] = l[b];
            }
            // read-strand row == top-strand row for a top-strand read
            const size_t row = ((m * Sm + rs) * Q + q) * 5 + obs;
            for (int b = 0; b < 5; b++) L.fold[row].t[b] = (float)l[b];
            L.fold[row].lpmax = (float)mx;
            L.fold[row].diag = (float)l[obs];
          }
        }
    }
  };
  qualities(0, Q);
}

static void build_scorer_ratios(LikelihoodTables& L) {
  L.rr.assign(L.keys * 6, 0.0);
  for (size_t k = 0; k < L.keys; k += 2) {   // top-strand keys; the bottom key of comp(obs) follows
    const double* l = &L.lp[k * 5];
    const size_t obs = (k % 10) / 2;
    const size_t bot = k - obs * 2 + comp_class((int)obs) * 2 + 1;
    double mx = -std::numeric_limits<double>::max();
    for (int b = 0; b < 5; b++) mx = std::max(mx, l[b]);
    for (int b = 0; b < 5; b++) {
      const double r = pow(10, l[b] - mx);
      L.rr[k * 6 + b] = r;
      L.rr[bot * 6 + comp_class(b)] = r;
    }
    L.rr[k * 6 + 5] = mx;
    L.rr[bot * 6 + 5] = mx;
  }
}

// One identify pass.  Returns false, with the session's minimum capacities raised, when one of the
// device result lists (candidates, insert columns, MC/UN runs, columns to re-decide) was too small;
// the caller runs the pass again.
// When `late` is given, a->deletion_propagation_cutoff holds only each target's sign (whether the
// coverage fit failed: the column pass needs no more); the values come from `late`, which is
// waited for just before the MC runs are found, so a coverage fit running on host threads
// overlaps the column pass.
using LateCutoffs = std::shared_future<std::vector<double>>;

// The genome diff of one identify pass written on the device (gd_emit.hip): the scorer's RA results,
// the MC / UN runs and one MC per target in `whole` (failed coverage fit) become numbered, sorted
// lines; only the text comes back.  The kernels read the list counts on the device, so the whole
// writer is queued behind the scorer and the MC/UN runs, with no host round trip; its counters and
// a first stretch of the text come back by pinned copies queued behind it.  rank: each target's
// place in the traversal order.
struct GdInputs {
  const int32_t* n_cand;   // device counters of the identify pass
  int32_t cand_cap;
  const int32_t* n_runs;
  int32_t run_cap;
  std::vector<int32_t> whole;
  std::vector<int64_t> rank;
  uint32_t places;
};

struct GdQueued {
  int cap = 0;          // entries the writer's lists hold
  size_t prefix = 0;    // text bytes copied back with the counters
};

// Blocks of the column pass.  A set count is kept to at most one window per wavefront and rounded
// down to a multiple of the 8 XCDs from 8 blocks up (the windows of one block then sit on one XCD's
// L2).  The default takes the fewest windows per wavefront k that keeps the grid within 2048
// blocks, in the grid that gives every wavefront exactly k windows (the last block aside), rounded
// up to a multiple of 8: a wavefront with one window more than the others holds its block.
static int64_t column_blocks(const Session& S) {
  const int64_t w = S.col_waves;
  const int64_t full = std::max<int64_t>(1, (S.n_windows + w - 1) / w);
  if (S.col_blocks > 0) {
    int64_t g = std::min<int64_t>(full, (int64_t)S.col_blocks);
    if (g >= 8) g &= ~(int64_t)7;
    return g;
  }
  const int64_t k = std::max<int64_t>(1, (S.n_windows + w * 2048 - 1) / (w * 2048));
  int64_t g = std::max<int64_t>(1, (S.n_windows + w * k - 1) / (w * k));
  if (g >= 8) g = (g + 7) / 8 * 8;
  return g;
}

static GdQueued queue_gd_on_device(Session& S, const GdInputs& in) {
  if (in.places > (uint32_t)dec::kMaxPrecision) throw BsqError(BSQ_ERR_ARG, "polymorphism_precision_places above 60");
  const int T = S.n_targets;
  GdQueued q;
  q.cap = std::max(1, std::min(S.gd_cap, kGdMaxEntries));
  S.d_gd_ents.alloc((size_t)q.cap);
  S.d_gd_rank_id.alloc((size_t)q.cap);
  S.d_gd_rank_out.alloc((size_t)q.cap);
  S.d_gd_lens.alloc((size_t)q.cap);
  S.d_gd_off.alloc((size_t)q.cap + 1);
  S.d_gd_slots.alloc((size_t)q.cap * kGdSlot);
  S.d_gd_text.alloc((size_t)q.cap * kGdSlot);
  // per-pass tables: traversal rank, seq_id byte-order rank, name offsets, failed targets
  std::vector<int32_t> order_lex(T);
  for (int t = 0; t < T; t++) order_lex[t] = t;
  std::sort(order_lex.begin(), order_lex.end(), [&](int x, int y) { return S.names[x] < S.names[y]; });
  const size_t n_meta = (size_t)3 * T + 1 + in.whole.size() + 1;
  S.h_gd_meta.alloc(n_meta);
  int32_t* m_rank = S.h_gd_meta.p;
  int32_t* m_lex = m_rank + T;
  int32_t* m_noff = m_lex + T;
  int32_t* m_whole = m_noff + T + 1;
  std::string names;
  for (int t = 0; t < T; t++) {
    m_rank[t] = (int32_t)std::max<int64_t>(0, in.rank[t]);
    m_noff[t] = (int32_t)names.size();
    names += S.names[t];
  }
  m_noff[T] = (int32_t)names.size();
  for (int k = 0; k < T; k++) m_lex[order_lex[k]] = k;
  for (size_t k = 0; k < in.whole.size(); k++) m_whole[k] = in.whole[k];
  m_whole[in.whole.size()] = 0;
  S.h_gd_names.alloc(std::max<size_t>(1, names.size()));
  std::copy(names.begin(), names.end(), S.h_gd_names.p);
  S.d_gd_meta.alloc(n_meta);
  S.d_gd_names.alloc(std::max<size_t>(1, names.size()));
  S.d_gd_cnt.alloc(8);
  HIP_OK(hipMemcpyAsync(S.d_gd_meta.p, S.h_gd_meta.p, n_meta * sizeof(int32_t), hipMemcpyHostToDevice, S.stream));
  if (!names.empty())
    HIP_OK(hipMemcpyAsync(S.d_gd_names.p, S.h_gd_names.p, names.size(), hipMemcpyHostToDevice, S.stream));
  HIP_OK(hipMemsetAsync(S.d_gd_cnt.p, 0, 8 * sizeof(int32_t), S.stream));
  GdParams G;
  G.ra = S.d_ra.p;
  G.n_cand = in.n_cand;
  G.cand_cap = in.cand_cap;
  G.runs = S.d_runrec.p;
  G.n_runs = in.n_runs;
  G.run_cap = in.run_cap;
  G.whole = S.d_gd_meta.p + (m_whole - S.h_gd_meta.p);
  G.n_whole = (int32_t)in.whole.size();
  G.pos_target = S.d_pos_target.p;
  G.n_pos = (int32_t)S.n_pos;
  G.targets = S.d_targets.p;
  G.n_targets = T;
  G.rank = S.d_gd_meta.p;
  G.lexrank = S.d_gd_meta.p + T;
  G.names = S.d_gd_names.p;
  G.name_off = S.d_gd_meta.p + 2 * T;
  G.places = (int32_t)in.places;
  G.ents = S.d_gd_ents.p;
  G.n_ents = S.d_gd_cnt.p + 0;
  G.ent_cap = q.cap;
  G.n_type = S.d_gd_cnt.p + 1;
  G.err = (uint32_t*)(S.d_gd_cnt.p + 4);
  G.slots = S.d_gd_slots.p;
  G.rank_id = S.d_gd_rank_id.p;
  G.rank_out = S.d_gd_rank_out.p;
  G.lens = S.d_gd_lens.p;
  int64_t* d_total = (int64_t*)(S.d_gd_cnt.p + 6);
  const int cap = q.cap;
  const dim3 rank_grid(kGdRankGridX, kGdRankGridY);
  hipLaunchKernelGGL(k_gd_collect, dim3(256), dim3(256), 0, S.stream, G);
  hipLaunchKernelGGL(k_gd_rank, rank_grid, dim3(256), 0, S.stream, S.d_gd_ents.p, G.n_ents, cap, 0, S.d_gd_rank_id.p);
  hipLaunchKernelGGL(k_gd_number, dim3(64), dim3(256), 0, S.stream, S.d_gd_ents.p, G.n_ents, cap, S.d_gd_rank_id.p);
  hipLaunchKernelGGL(k_gd_rank, rank_grid, dim3(256), 0, S.stream, S.d_gd_ents.p, G.n_ents, cap, 1, S.d_gd_rank_out.p);
  hipLaunchKernelGGL(k_gd_lines, dim3((unsigned)std::min(2048, (cap + 3) / 4)), dim3(64), 0, S.stream, G, S.d_gd_rank_out.p);
  hipLaunchKernelGGL(k_gd_offsets, dim3(1), dim3(1024), 0, S.stream, S.d_gd_lens.p, G.n_ents, cap, S.d_gd_off.p, d_total);
  hipLaunchKernelGGL(k_gd_gather, dim3(256), dim3(256), 0, S.stream, S.d_gd_slots.p, S.d_gd_lens.p, S.d_gd_off.p, G.n_ents,
                     cap, S.d_gd_text.p);
  HIP_OK(hipGetLastError());
  S.h_gd_cnt.alloc(8);
  HIP_OK(hipMemcpyAsync(S.h_gd_cnt.p, S.d_gd_cnt.p, 8 * sizeof(int32_t), hipMemcpyDeviceToHost, S.stream));
  q.prefix = std::min((size_t)q.cap * kGdSlot, S.gd_text_hint);
  S.h_gd_text.alloc(q.prefix);
  HIP_OK(hipMemcpyAsync(S.h_gd_text.p, S.d_gd_text.p, q.prefix, hipMemcpyDeviceToHost, S.stream));
  return q;
}

// After the stream has been synchronized: the writer's checks and the text.  A writer whose entry
// list overflowed is queued again with room for every entry, after the stream is idle.
static std::string finish_gd_on_device(Session& S, const GdInputs& in, GdQueued q, int64_t& n_ra, int64_t& n_mc,
                                       int64_t& n_un) {
  for (int attempt = 0;; attempt++) {
    const int32_t* cnt = S.h_gd_cnt.p;
    const uint32_t gerr = (uint32_t)cnt[4];
    const int n = cnt[0];
    if ((gerr & 2u) || n > q.cap) {
      if (n > kGdMaxEntries || attempt > 0) throw BsqError(BSQ_ERR_LIMIT, "more genome diff entries than the device writer holds");
      S.gd_cap = std::min<int64_t>(kGdMaxEntries, (int64_t)n + n / 4 + 1024);
      q = queue_gd_on_device(S, in);
      HIP_OK(hipStreamSynchronize(S.stream));
      continue;
    }
    if (gerr & 1u) throw BsqError(BSQ_ERR_DEVICE, "genome diff writer: an index outside its list");
    if (gerr & 4u) throw BsqError(BSQ_ERR_LIMIT, "genome diff writer: an entry longer than its line slot");
    if (gerr & 8u) throw BsqError(BSQ_ERR_ARG, "genome diff writer: a number precision it cannot print");
    int64_t text_len;
    std::memcpy(&text_len, &cnt[6], sizeof(text_len));
    n_ra = cnt[1];
    n_mc = cnt[2];
    n_un = cnt[3];
    std::string text = "#=GENOME_DIFF\t1.0\n";
    const size_t head = text.size();
    text.resize(head + (size_t)text_len);
    if (n > 0 && text_len > 0) {
      if ((size_t)text_len <= q.prefix) {
        std::memcpy(&text[head], S.h_gd_text.p, (size_t)text_len);
      } else {
        HIP_OK(hipMemcpyAsync(&text[head], S.d_gd_text.p, (size_t)text_len, hipMemcpyDeviceToHost, S.stream));
        HIP_OK(hipStreamSynchronize(S.stream));
      }
    }
    // the next pass copies back twice this text's length with its counters
    S.gd_text_hint = std::max<size_t>((size_t)64 << 10, 2 * (size_t)text_len);
    return text;
  }
}

// The files of an identify pass: the genome diff text and the DP candidate regions.
static void finish_identify_outputs(Session& S, const bsq_identify_args* a, bsq_im_result& out,
                                    std::future<std::string>& dp_text, const double& dp_ms, double run_host_ms) {
  if (a->gd_file) {
    std::ofstream f(a->gd_file);
    if (!f) throw BsqError(BSQ_ERR_IO, std::string("could not write ") + a->gd_file);
    f << out.text;
  }
  out.dp_text = dp_text.get();
  S.timing.kernel_detail_ms[0] = dp_ms;
  if (a->dp_file) {
    std::ofstream f(a->dp_file);
    if (!f) throw BsqError(BSQ_ERR_IO, std::string("could not write ") + a->dp_file);
    f << out.dp_text;
  }
  S.timing.kernel_detail_ms[5] = run_host_ms;
}

static bool run_identify_once(Session& S, const bsq_identify_args* a, bsq_im_result& out, int screen_off,
                              const LateCutoffs* late) {
  if (!S.loaded) throw BsqError(BSQ_ERR_ARG, "session has no sample loaded");
  out.n_ra = out.n_mc = out.n_un = 0;
  S.timing.fit_wait_ms = 0.0;
  const int Sm = a->read_set_max, Q = a->quality_max;
  if (Sm <= 0 || Q <= 0 || Sm > 255 || Q > 256) throw BsqError(BSQ_ERR_ARG, "covariate maxima out of range");
  if (a->n_bins != (int64_t)Sm * 25 * Q) throw BsqError(BSQ_ERR_ARG, "error table size does not match its covariates");
  if (!a->deletion_propagation_cutoff || !a->deletion_seed_cutoff) throw BsqError(BSQ_ERR_ARG, "missing per-target cutoffs");
  const double th0 = now_ms();
  // the DP candidate regions need nothing from the device: their own thread, from the start
  std::vector<const char*> tn;
  for (const std::string& n : S.names) tn.push_back(n.c_str());
  double dp_ms = 0.0;
  std::future<std::string> dp_text = run_async([&] {
    const double d0 = now_ms();
    std::string t = dp_regions_csv(S.n_targets, tn.data(), S.lengths.data(), a->call_target,
                                   a->deletion_propagation_cutoff, a->dp, S.dp_reads);
    dp_ms = now_ms() - d0;
    return t;
  });
  WaitOnExit<std::future<std::string>> dp_wait{dp_text};   // it reads this frame's locals
  // the per-call alignment view depends only on the sample and the covariate maxima: it is built on
  // the device while the host prepares the tables
  HIP_OK(hipEventRecord(S.ev0, S.stream));
  if (S.idrec_sm != Sm || S.idrec_q != Q) {   // not already queued by the pipeline's error_count
    S.d_idrec.alloc((size_t)std::max<int64_t>(1, S.n_dreads));
    hipLaunchKernelGGL(k_prep_reads, dim3(1024), dim3(256), 0, S.stream, S.d_reads.p, S.n_dreads, Sm, Q, S.d_idrec.p);
    HIP_OK(hipGetLastError());
  }
  S.idrec_sm = S.idrec_q = -1;   // a later pass builds its own
  // targets
  std::vector<TargetInfo> T(S.n_targets);
  for (int t = 0; t < S.n_targets; t++) {
    T[t].off = S.tgt_off[t];
    T[t].len = (int32_t)S.lengths[t];
    T[t].group = 0;
    T[t].scored = (!a->call_target || a->call_target[t]) ? 1 : 0;
    T[t].prop_cutoff = a->deletion_propagation_cutoff[t];
    T[t].seed_cutoff = a->deletion_seed_cutoff[t];
    if (T[t].scored && !a->skip_missing_coverage_prediction && T[t].prop_cutoff >= 0.0 && T[t].seed_cutoff > T[t].prop_cutoff)
      throw BsqError(BSQ_ERR_ARG, "deletion seed cutoff above the propagation cutoff is not supported");
  }
  S.d_targets.alloc(T.size());
  S.h_targets0.alloc(T.size());
  std::copy(T.begin(), T.end(), S.h_targets0.p);
  HIP_OK(hipMemcpyAsync(S.d_targets.p, S.h_targets0.p, T.size() * sizeof(TargetInfo), hipMemcpyHostToDevice, S.stream));
  LikelihoodTables L;
  build_column_tables(S, a, L);
  const std::vector<FoldRow>& fold = L.fold;
  S.d_tab_lp.alloc(L.lp.size());
  S.d_tab_r.alloc(L.keys * 6);
  S.d_fold.alloc(fold.size());
  S.h_tab_lp.alloc(L.lp.size());
  S.h_fold.alloc(fold.size());
  std::copy(L.lp.begin(), L.lp.end(), S.h_tab_lp.p);
  std::copy(fold.begin(), fold.end(), S.h_fold.p);
  HIP_OK(hipMemcpyAsync(S.d_tab_lp.p, S.h_tab_lp.p, L.lp.size() * sizeof(double), hipMemcpyHostToDevice, S.stream));
  HIP_OK(hipMemcpyAsync(S.d_fold.p, S.h_fold.p, fold.size() * sizeof(FoldRow), hipMemcpyHostToDevice, S.stream));
  double total_len = 0;
  for (int t = 0; t < S.n_targets; t++) total_len += (double)S.lengths[t];
  S.d_col_flags.alloc(S.n_pos);
  S.d_col_unique.alloc(S.n_pos);
  S.d_col_total.alloc(S.n_pos);
  const int cand_cap = std::max<int>(screen_off ? S.n_pos + 1024 : std::max<int>(1 << 16, S.n_pos / 64), S.id_caps.cand);
  const int ins_cap = std::max<int>(std::max<int>(1 << 14, S.n_pos / 256), S.id_caps.ins);
  const int run_cap = std::max<int>(std::max<int>(1 << 16, S.n_pos / 16), S.id_caps.runs);
  S.d_cand.alloc(cand_cap);
  S.d_ins.alloc(ins_cap);
  S.d_ra.alloc(cand_cap);
  S.d_runrec.alloc(run_cap);
  const int fix_cap = std::max<int>(std::max<int>(1 << 16, S.n_pos / 256), S.id_caps.fix);
  S.d_fix.alloc(fix_cap);
  ImParams P;
  P.reads = S.d_reads.p;
  P.idrec = S.d_idrec.p;
  P.cigar = S.d_cigar.p;
  P.bases = S.d_bases.p;
  P.quals = S.d_quals.p;
  P.ref = S.d_ref.p;
  P.pos_target = S.d_pos_target.p;
  P.targets = S.d_targets.p;
  P.win_rb = S.d_win_rb.p;
  P.win_re = S.d_win_re.p;
  P.n_windows = S.n_windows;
  P.n_pos = S.n_pos;
  P.read_set_max = Sm;
  P.quality_max = Q;
  P.base_quality_cutoff = a->base_quality_cutoff;
  P.tab_lp = S.d_tab_lp.p;
  P.tab_r = S.d_tab_r.p;
  P.tab_keys = (int32_t)L.keys;
  P.n_reads = (int32_t)S.n_dreads;
  P.fold = S.d_fold.p;
  P.fold_rows = (int32_t)fold.size() - 1;
  const size_t fold_bytes = fold.size() * sizeof(FoldRow);
  P.fold_lds = (fold_bytes <= kFoldLdsBytes && S.col_lds) ? 1 : 0;
  P.log10_ref_length = log10(total_len);
  P.mutation_cutoff = a->mutation_cutoff;
  P.polymorphism_cutoff = a->polymorphism_cutoff;
  P.precision_decimal = a->polymorphism_precision_decimal;
  P.screen_margin = 0.1;
  P.screen_off = screen_off;
  P.col_flags = S.d_col_flags.p;
  P.col_unique = S.d_col_unique.p;
  P.col_total = S.d_col_total.p;
  P.cand = S.d_cand.p;
  P.n_cand = S.d_counters.p + 0;
  P.score_next = S.d_counters.p + 5;
  P.cand_cap = cand_cap;
  P.ins = S.d_ins.p;
  P.n_ins = S.d_counters.p + 1;
  P.ins_cap = ins_cap;
  P.key_cap = std::max(64, S.max_window_reads);
  // a candidate's read list (4 B per read of its window) stays in LDS while it takes at most half
  // of the scorer's dynamic LDS; deeper samples keep it in a device scratch list per block
  const size_t score_dyn_budget = kScoreLdsBudget - kScoreStaticLds;
  const size_t key_list_bytes = ((size_t)P.key_cap * sizeof(uint32_t) + 15) & ~(size_t)15;
  P.keys_global = (key_list_bytes > score_dyn_budget / 2) ? 1 : 0;
  const size_t key_bytes = P.keys_global ? 0 : key_list_bytes;
  if (P.keys_global) {
    S.d_key_scratch.alloc((size_t)kScoreGrid * (size_t)P.key_cap);
    P.key_scratch = S.d_key_scratch.p;
  } else {
    P.key_scratch = nullptr;
  }
  P.rows_cap = (int32_t)((score_dyn_budget - key_bytes) / (6 * sizeof(double)));
  const size_t score_lds = key_bytes + (size_t)P.rows_cap * 6 * sizeof(double);
  P.ra_out = S.d_ra.p;
  P.err = (uint32_t*)(S.d_counters.p + 3);
  P.fix = S.d_fix.p;
  P.n_fix = S.d_counters.p + 4;
  P.fix_cap = fix_cap;
  RunParams R;
  R.pos_target = S.d_pos_target.p;
  R.targets = S.d_targets.p;
  R.col_flags = S.d_col_flags.p;
  R.col_unique = S.d_col_unique.p;
  R.n_pos = S.n_pos;
  R.do_mc = a->skip_missing_coverage_prediction ? 0 : 1;
  R.runs = S.d_runrec.p;
  R.n_runs = S.d_counters.p + 2;
  R.run_cap = run_cap;
  R.err = P.err;
  P.audit = nullptr;
  if (S.audit_on) {
    S.d_audit.alloc((size_t)8 * S.n_pos);
    HIP_OK(hipMemsetAsync(S.d_audit.p, 0, (size_t)8 * S.n_pos * sizeof(float), S.stream));
    P.audit = S.d_audit.p;
  }
  S.last_P = P;
  HIP_OK(hipMemsetAsync(S.d_counters.p, 0, 8 * sizeof(int32_t), S.stream));
  const double th1 = now_ms();
  // the whole pass is queued without a host round trip: later kernels read the counts on device
  {
    const int64_t g = column_blocks(S);
    HIP_OK(hipEventRecord(S.evc, S.stream));
    const dim3 blk(S.col_waves * 64);
    if (S.col_waves == 4) {
      if (P.fold_lds) hipLaunchKernelGGL((k_identify_columns_t<true, 4>), dim3((int)g), blk, fold_bytes, S.stream, P);
      else hipLaunchKernelGGL((k_identify_columns_t<false, 4>), dim3((int)g), blk, 0, S.stream, P);
    } else {
      if (P.fold_lds) hipLaunchKernelGGL((k_identify_columns_t<true, 8>), dim3((int)g), blk, fold_bytes, S.stream, P);
      else hipLaunchKernelGGL((k_identify_columns_t<false, 8>), dim3((int)g), blk, 0, S.stream, P);
    }
  }
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(S.evk[0], S.stream));
  if (P.keys_global)
    hipLaunchKernelGGL(k_fix_columns_t<true>, dim3(kFixGrid), dim3(64), 0, S.stream, P);
  else
    hipLaunchKernelGGL(k_fix_columns_t<false>, dim3(kFixGrid), dim3(64), key_bytes, S.stream, P);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(S.evk[3], S.stream));
  // the scorer's ratio table, built while the column pass runs
  build_scorer_ratios(L);
  S.h_tab_r.alloc(L.rr.size());
  std::copy(L.rr.begin(), L.rr.end(), S.h_tab_r.p);
  HIP_OK(hipMemcpyAsync(S.d_tab_r.p, S.h_tab_r.p, L.rr.size() * sizeof(double), hipMemcpyHostToDevice, S.stream));
  // the insert slots and the scorer need no coverage cutoff: queued before the host waits for the fit
  hipLaunchKernelGGL(k_insert_slots, dim3(1024), dim3(64), 0, S.stream, P);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(S.evk[1], S.stream));
  if (P.keys_global)
    hipLaunchKernelGGL(k_score_candidates_t<true>, dim3(kScoreGrid), dim3(kScoreWaves * 64), score_lds, S.stream, P);
  else
    hipLaunchKernelGGL(k_score_candidates_t<false>, dim3(kScoreGrid), dim3(kScoreWaves * 64), score_lds, S.stream, P);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(S.ev1, S.stream));
  // MC/UN runs read only the final column flags and coverage and the fitted cutoffs, so they run on
  // the side stream beside the insert slots and the scorer
  HIP_OK(hipStreamWaitEvent(S.stream2, S.evk[3], 0));
  std::vector<double> prop_final(a->deletion_propagation_cutoff, a->deletion_propagation_cutoff + S.n_targets);
  if (late) {
    const double w0 = now_ms();
    const std::vector<double>& c = late->get();
    S.timing.fit_wait_ms = now_ms() - w0;
    for (int t = 0; t < S.n_targets; t++) {
      if ((c[t] < 0.0) != (prop_final[t] < 0.0)) throw BsqError(BSQ_ERR_ARG, "coverage fit changed a cutoff's sign");
      prop_final[t] = c[t];
      T[t].prop_cutoff = c[t];
    }
    S.h_targets.alloc(T.size());
    std::copy(T.begin(), T.end(), S.h_targets.p);
    HIP_OK(hipMemcpyAsync(S.d_targets.p, S.h_targets.p, T.size() * sizeof(TargetInfo), hipMemcpyHostToDevice,
                          S.stream2));
  }
  HIP_OK(hipEventRecord(S.evk[4], S.stream2));
  hipLaunchKernelGGL(k_runs, dim3((S.n_pos + 255) / 256), dim3(256), 0, S.stream2, R);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(S.evk[5], S.stream2));
  S.h_counters.alloc(9);
  int32_t* counters = S.h_counters.p;
  HIP_OK(hipMemcpyAsync(counters + 8, S.d_counters.p + 2, sizeof(int32_t), hipMemcpyDeviceToHost, S.stream2));
  std::vector<int32_t> order = target_order(S, a->call_target);
  std::vector<int64_t> rank(S.n_targets, -1);
  for (size_t i = 0; i < order.size(); i++) rank[order[i]] = (int64_t)i;
  GdInputs gd_in;
  GdQueued gd_q;
  if (S.gd_on_device) {
    // the writer reads the MC/UN runs too: the stream waits for them
    HIP_OK(hipStreamWaitEvent(S.stream, S.evk[5], 0));
    gd_in.n_cand = S.d_counters.p + 0;
    gd_in.cand_cap = cand_cap;
    gd_in.n_runs = S.d_counters.p + 2;
    gd_in.run_cap = run_cap;
    if (!a->skip_missing_coverage_prediction)
      for (int t : order)
        if (prop_final[t] < 0.0) gd_in.whole.push_back(t);
    gd_in.rank = rank;
    gd_in.places = a->polymorphism_precision_places;
    gd_q = queue_gd_on_device(S, gd_in);
  }
  HIP_OK(hipMemcpyAsync(counters, S.d_counters.p, 8 * sizeof(int32_t), hipMemcpyDeviceToHost, S.stream));
  std::vector<GdEntry> run_entries;
  HIP_OK(hipStreamSynchronize(S.stream2));
  const int32_t n_runs_seen = counters[8];
  const double tr0 = now_ms();
  if (!S.gd_on_device && n_runs_seen <= run_cap) {   // past the cap the call fails below, with the other limit checks
    std::vector<RunRec> runs(n_runs_seen);
    if (n_runs_seen)
      HIP_OK(hipMemcpyAsync(runs.data(), S.d_runrec.p, n_runs_seen * sizeof(RunRec), hipMemcpyDeviceToHost, S.stream2));
    HIP_OK(hipStreamSynchronize(S.stream2));
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (const RunRec& rr_ : runs) {
      const int t = S.h_pos_target[rr_.s];
      const int32_t base = S.tgt_off[t];
      GdEntry e;
      e.seq_id = S.names[t];
      e.pos = (uint32_t)(rr_.s - base + 1);
      e.end = (uint32_t)(rr_.e - base + 1);
      e.order_target = rank[t];
      e.order_pos = e.end + 1;
      if (rr_.kind == 0) {
        if (!rr_.seed) continue;   // seed never reached
        const uint32_t red_start = (rr_.red_start >= 0) ? (uint32_t)(rr_.red_start - base + 1) : e.pos;
        const uint32_t red_end = (rr_.red_end >= 0) ? (uint32_t)(rr_.red_end - base + 1) : e.end;
        e.type = 1;
        e.start_range = red_start - e.pos;
        e.end_range = e.end - red_end;
        e.fields = {{"left_inside_cov", fmt_double((double)rr_.li, 0, false)},
                    {"left_outside_cov", fmt_double(rr_.lo < 0 ? nan : (double)rr_.lo, 0, false)},
                    {"right_inside_cov", fmt_double((double)rr_.ri, 0, false)},
                    {"right_outside_cov", fmt_double(rr_.ro < 0 ? nan : (double)rr_.ro, 0, false)}};
        e.order_sub = 0;
        out.n_mc++;
      } else {
        e.type = 2;
        e.order_sub = 1;
        out.n_un++;
      }
      run_entries.push_back(std::move(e));
    }
  }
  const double run_host_ms = now_ms() - tr0;
  HIP_OK(hipStreamSynchronize(S.stream));
  float ms = 0;
  HIP_OK(hipEventElapsedTime(&ms, S.ev0, S.ev1));
  S.timing.kernel_ms = ms;
  for (double& d : S.timing.kernel_detail_ms) d = 0.0;
  {
    float k0 = 0, k1 = 0, k2 = 0, k3 = 0, kf = 0;
    HIP_OK(hipEventElapsedTime(&k0, S.evc, S.evk[0]));
    HIP_OK(hipEventElapsedTime(&kf, S.evk[0], S.evk[3]));
    HIP_OK(hipEventElapsedTime(&k1, S.evk[3], S.evk[1]));
    HIP_OK(hipEventElapsedTime(&k2, S.evk[1], S.ev1));
    HIP_OK(hipEventElapsedTime(&k3, S.evk[4], S.evk[5]));
    S.timing.kernel_detail_ms[1] = k0;
    S.timing.kernel_detail_ms[2] = k1;
    S.timing.kernel_detail_ms[3] = k2;
    S.timing.kernel_detail_ms[4] = k3;
    S.timing.kernel_detail_ms[6] = kf;
    S.timing.kernel_detail_ms[7] = (double)counters[4];
  }
  const uint32_t errbits = (uint32_t)counters[3];
  // the list counters count every append, also those past the end: a list that overflowed gets
  // room for all of them (and a quarter more) and the pass is repeated
  bool grow = false;
  auto need = [&](int32_t got, int cap, int32_t& minimum) {
    if (got <= cap) return;
    const int64_t want = (int64_t)got + got / 4 + 1024;
    if (want > (int64_t)(1 << 30)) throw BsqError(BSQ_ERR_LIMIT, "identify result list beyond 2^30 entries");
    minimum = (int32_t)want;
    grow = true;
  };
  need(counters[0], cand_cap, S.id_caps.cand);
  need(counters[1], ins_cap, S.id_caps.ins);
  need(counters[2], run_cap, S.id_caps.runs);
  need(counters[4], fix_cap, S.id_caps.fix);
  if (grow) return false;
  raise_err_bits(errbits);
  const int n_cand = counters[0];
  S.timing.candidates = n_cand;
  if (S.gd_on_device) {
    const double t1 = now_ms();
    int64_t n_ra = 0, n_mc = 0, n_un = 0;
    out.text = finish_gd_on_device(S, gd_in, gd_q, n_ra, n_mc, n_un);
    out.n_ra = n_ra;
    out.n_mc = n_mc;
    out.n_un = n_un;
    S.timing.d2h_ms = 0.0;
    finish_identify_outputs(S, a, out, dp_text, dp_ms, run_host_ms);
    S.timing.host_post_ms = now_ms() - t1;
    S.timing.host_pre_ms = th1 - th0;
    return true;
  }
  const double t0 = now_ms();
  std::vector<RaOut> ra(n_cand);
  if (n_cand) HIP_OK(hipMemcpyAsync(ra.data(), S.d_ra.p, n_cand * sizeof(RaOut), hipMemcpyDeviceToHost, S.stream));
  HIP_OK(hipStreamSynchronize(S.stream));
  S.timing.d2h_ms = now_ms() - t0;
  const double t1 = now_ms();
  // assemble entries: RA, then the MC/UN runs, then whole-target MC of failed fits
  std::vector<GdEntry> entries;
  entries.reserve((size_t)n_cand + run_entries.size() + order.size());
  for (const RaOut& r : ra) {
    if (!r.emitted) continue;
    const int t = S.h_pos_target[r.gpos];
    GdEntry e;
    const uint32_t position = (uint32_t)(r.gpos - S.tgt_off[t] + 1);
    ra_entry_from_result(r, S.names[t], position, a->polymorphism_precision_places, e);
    e.order_target = rank[t];
    e.order_pos = position;
    e.order_sub = 2 + r.slot;
    entries.push_back(std::move(e));
    out.n_ra++;
  }
  for (GdEntry& e : run_entries) entries.push_back(std::move(e));
  if (!a->skip_missing_coverage_prediction) {
    for (int t : order) {
      if (prop_final[t] >= 0.0) continue;
      GdEntry e;
      e.type = 1;
      e.seq_id = S.names[t];
      e.pos = 1;
      e.end = (uint32_t)S.lengths[t];
      e.fields = {{"left_inside_cov", fmt_double(0.0, 0, false)},
                  {"left_outside_cov", "NA"},
                  {"right_inside_cov", fmt_double(0.0, 0, false)},
                  {"right_outside_cov", "NA"}};
      e.order_target = rank[t];
      e.order_pos = S.lengths[t] + 2;
      e.order_sub = 0;
      entries.push_back(std::move(e));
      out.n_mc++;
    }
  }
  out.text = write_gd(entries);
  finish_identify_outputs(S, a, out, dp_text, dp_ms, run_host_ms);
  S.timing.host_post_ms = now_ms() - t1;
  S.timing.host_pre_ms = th1 - th0;
  return true;
}

static void run_identify(Session& S, const bsq_identify_args* a, bsq_im_result& out, int screen_off,
                         const LateCutoffs* late = nullptr) {
  for (int attempt = 0; attempt < 3; attempt++)
    if (run_identify_once(S, a, out, screen_off, late)) return;
  throw BsqError(BSQ_ERR_LIMIT, "identify result lists could not be sized for this sample");
}

// Precision audit of the column pass: one identify pass that keeps its FP32 genotype sums, then the
// exact FP64 sums of every column with reads (k_audit_columns), compared per column and genotype,
// and the final predicted / consensus flags compared with the exact calls.
static void run_precision_audit(Session& S, const bsq_identify_args* a, bsq_precision_report& rep) {
  bsq_im_result out;
  S.audit_on = true;
  try {
    run_identify(S, a, out, 0);
  } catch (...) {
    S.audit_on = false;
    throw;
  }
  S.audit_on = false;
  const int64_t redecided = (int64_t)S.timing.kernel_detail_ms[7];
  std::vector<float> f32((size_t)8 * S.n_pos);
  HIP_OK(hipMemcpy(f32.data(), S.d_audit.p, f32.size() * sizeof(float), hipMemcpyDeviceToHost));
  std::vector<int32_t> cols;
  for (int64_t p = 0; p < S.n_pos; p++) {
    uint32_t n;
    std::memcpy(&n, &f32[(size_t)8 * p + 5], sizeof n);
    if (n > 0) cols.push_back((int32_t)p);
  }
  rep = bsq_precision_report{};
  rep.columns = (int64_t)cols.size();
  rep.columns_redecided = redecided;
  if (cols.empty()) return;
  DBuf<int32_t> d_cols;
  DBuf<double> d_out;
  DBuf<uint32_t> d_mis;
  d_cols.alloc(cols.size());
  d_out.alloc((size_t)6 * cols.size());
  d_mis.alloc(1);
  HIP_OK(hipMemcpy(d_cols.p, cols.data(), cols.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_OK(hipMemset(d_mis.p, 0, sizeof(uint32_t)));
  HIP_OK(hipMemset(S.d_counters.p + 3, 0, sizeof(int32_t)));
  ImParams P = S.last_P;
  const size_t key_bytes = P.keys_global ? 0 : (((size_t)P.key_cap * sizeof(uint32_t) + 15) & ~(size_t)15);
  if (P.keys_global)
    hipLaunchKernelGGL(k_audit_columns_t<true>, dim3(kFixGrid), dim3(64), 0, S.stream, P, d_cols.p,
                       (int32_t)cols.size(), d_out.p, d_mis.p);
  else
    hipLaunchKernelGGL(k_audit_columns_t<false>, dim3(kFixGrid), dim3(64), key_bytes, S.stream, P, d_cols.p,
                       (int32_t)cols.size(), d_out.p, d_mis.p);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(S.stream));
  int32_t errbits = 0;
  HIP_OK(hipMemcpy(&errbits, S.d_counters.p + 3, sizeof(int32_t), hipMemcpyDeviceToHost));
  raise_err_bits((uint32_t)errbits);
  std::vector<double> x64((size_t)6 * cols.size());
  uint32_t mis = 0;
  HIP_OK(hipMemcpy(x64.data(), d_out.p, x64.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(&mis, d_mis.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
  std::vector<double> abs_err;
  abs_err.reserve((size_t)5 * cols.size());
  double sum = 0.0, mx = 0.0, mx_rel = 0.0;
  for (size_t i = 0; i < cols.size(); i++) {
    uint32_t n32;
    std::memcpy(&n32, &f32[(size_t)8 * cols[i] + 5], sizeof n32);
    if ((double)n32 != x64[6 * i + 5]) rep.count_mismatches++;
    for (int b = 0; b < 5; b++) {
      const double e = std::fabs((double)f32[(size_t)8 * cols[i] + b] - x64[6 * i + b]);
      abs_err.push_back(e);
      sum += e;
      mx = std::max(mx, e);
      if (x64[6 * i + b] != 0.0) mx_rel = std::max(mx_rel, e / std::fabs(x64[6 * i + b]));
    }
  }
  std::sort(abs_err.begin(), abs_err.end());
  rep.max_abs_error = mx;
  rep.p99_abs_error = abs_err[(size_t)((abs_err.size() - 1) * 0.99)];
  rep.mean_abs_error = sum / (double)abs_err.size();
  rep.max_rel_error = mx_rel;
  rep.decision_mismatches = mis;
}

// BAM2COV table of one region (coverage_output::table, coverage_output.cpp:190-283; rows of
// pileup_callback, :307-503).  The counts come from k_coverage_columns (consecutive rows) or
// k_coverage_rows (downsampled rows), computed only for the positions that print counts: one pass
// for all reads, one more per read group with per_read_group.  Which positions get a row follows
// the reference's traversal (pileup_base.cpp:118-210, 308-360):
//   * a position is handled when it lies in the region and (position + start) is a multiple of the
//     downsampling step floor(size / resolution) (at least 1);
//   * up to the last column any read overlapping the region covers, positions are visited one by
//     one, so each handled position prints its own row;
//   * after it, only handled positions advance the traversal, so a handled position also prints
//     zero rows for the unhandled positions since the previous one (from position 1 when no read
//     overlaps the region); those rows count in the averages' denominator only.
// So the table is: the handled positions up to that last column, with counts, then every position
// after it up to the last handled one, as zero rows.
static std::string coverage_table(Session& S, const bsq_coverage_table_args* a) {
  if (!a) throw BsqError(BSQ_ERR_ARG, "null argument");
  if (a->target < 0 || a->target >= S.n_targets) throw BsqError(BSQ_ERR_ARG, "target id out of range");
  const int64_t L = S.lengths[(size_t)a->target];
  int64_t start = a->start_1, end = a->end_1;
  if (start == 0 && end == 0) {
    start = 1;
    end = L;
  }
  if (start < 1 || end > L || start > end) throw BsqError(BSQ_ERR_ARG, "region outside its reference sequence");
  const int64_t off = S.tgt_off[(size_t)a->target];
  const int n_rg = std::max(1, a->n_read_groups);
  const int passes = a->per_read_group ? 1 + n_rg : 1;
  int64_t ds = 1;
  if (a->resolution != 0) {
    ds = (int64_t)std::floor((double)(uint32_t)(end - start + 1) / (double)a->resolution);
    if (ds < 1) ds = 1;
  }
  // the last column covered by a read that overlaps the region (sam_itr_queryi(start-1, end)):
  // reads are sorted by first column, so it is the running maximum of gend over the reads that
  // begin at or before the region's last column, when that maximum reaches into the region
  int64_t last_pile = 0;   // 1-based; 0: no read overlaps the region
  {
    const int32_t g0 = (int32_t)(off + start - 1), g1 = (int32_t)(off + end - 1);
    const size_t hi = (size_t)(std::upper_bound(S.h_gpos.begin(), S.h_gpos.end(), g1) - S.h_gpos.begin());
    if (hi > 0 && S.h_gend_max[hi - 1] > g0) last_pile = (int64_t)S.h_gend_max[hi - 1] - off;
  }
  // handled positions: (position + start) a multiple of the step
  const int64_t rem = (2 * start) % ds;
  const int64_t first = start + (rem ? ds - rem : 0);
  const int64_t row_end = std::min(end, last_pile);
  const int64_t n = row_end >= first ? (row_end - first) / ds + 1 : 0;   // rows with counts
  const int64_t tail_end = end >= first ? first + (end - first) / ds * ds : 0;
  const int64_t z0 = std::max<int64_t>(last_pile, 0) + 1;                // zero rows z0..tail_end
  const int64_t nz = tail_end >= z0 ? tail_end - z0 + 1 : 0;
  // counts come back to the host for every pass: 24 bytes per row and pass
  if ((double)n * passes > (double)(1 << 28))
    throw BsqError(BSQ_ERR_LIMIT, "coverage table too large: split the region or lower the resolution");
  std::vector<std::vector<uint32_t>> hu(passes), hraw(passes), hbeg(passes);
  std::vector<std::vector<double>> hred(passes);
  if (n > 0) {
    S.d_ct_u.alloc((size_t)(2 * n));
    S.d_ct_raw.alloc((size_t)(2 * n));
    S.d_ct_beg.alloc((size_t)(2 * n));
    S.d_ct_red.alloc((size_t)(2 * n));
    S.d_ct_misc.alloc(1);
    HIP_OK(hipMemsetAsync(S.d_ct_misc.p, 0, sizeof(int32_t), S.stream));
    for (int k = 0; k < passes; k++) {
      CovTabParams C;
      C.reads = S.d_reads.p;
      C.rg = S.d_rg.p;
      C.cigar = S.d_cigar.p;
      C.win_rb = S.d_win_rb.p;
      C.win_re = S.d_win_re.p;
      C.h0 = (int32_t)(off + first - 1);
      C.step = (int32_t)ds;
      C.n = n;
      C.rg_filter = k - 1;
      C.n_rg = a->n_read_groups;
      C.uniq = S.d_ct_u.p;
      C.red = S.d_ct_red.p;
      C.raw = S.d_ct_raw.p;
      C.begin = S.d_ct_beg.p;
      C.err = (uint32_t*)S.d_ct_misc.p;
      if (ds == 1) {
        const int64_t nwin = (C.h0 + n - 1) / 64 - C.h0 / 64 + 1;
        hipLaunchKernelGGL(k_coverage_columns, dim3((unsigned)std::min<int64_t>((nwin + 3) / 4, 4096)), dim3(256), 0,
                           S.stream, C);
      } else {
        hipLaunchKernelGGL(k_coverage_rows, dim3((unsigned)std::min<int64_t>((n + 3) / 4, 4096)), dim3(256), 0,
                           S.stream, C);
      }
      HIP_OK(hipGetLastError());
      hu[k].resize((size_t)(2 * n));
      hraw[k].resize((size_t)(2 * n));
      hbeg[k].resize((size_t)(2 * n));
      hred[k].resize((size_t)(2 * n));
      HIP_OK(hipMemcpyAsync(hu[k].data(), S.d_ct_u.p, (size_t)(2 * n) * 4, hipMemcpyDeviceToHost, S.stream));
      HIP_OK(hipMemcpyAsync(hraw[k].data(), S.d_ct_raw.p, (size_t)(2 * n) * 4, hipMemcpyDeviceToHost, S.stream));
      HIP_OK(hipMemcpyAsync(hbeg[k].data(), S.d_ct_beg.p, (size_t)(2 * n) * 4, hipMemcpyDeviceToHost, S.stream));
      HIP_OK(hipMemcpyAsync(hred[k].data(), S.d_ct_red.p, (size_t)(2 * n) * 8, hipMemcpyDeviceToHost, S.stream));
    }
    int32_t got = 0;
    HIP_OK(hipMemcpyAsync(&got, S.d_ct_misc.p, sizeof(got), hipMemcpyDeviceToHost, S.stream));
    HIP_OK(hipStreamSynchronize(S.stream));
    raise_err_bits((uint32_t)got);
  }
  const char d = a->csv ? ',' : '\t';
  static const char* const kCols8[8] = {"unique_top_cov", "unique_bot_cov", "redundant_top_cov", "redundant_bot_cov",
                                        "raw_redundant_top_cov", "raw_redundant_bot_cov", "unique_top_begin",
                                        "unique_bot_begin"};
  static const char* const kCols3[3] = {"unique_cov", "redundant_cov", "total_cov"};
  const int ncols = a->total_only ? 3 : 8;
  const char* const* cols = a->total_only ? kCols3 : kCols8;
  auto put_u = [](std::string& o, uint64_t v) {
    char b[24];
    o.append(b, std::to_chars(b, b + sizeof(b), v).ptr);
  };
  std::string out;
  out += "position";
  out += d;
  out += "ref_base";
  for (int c = 0; c < ncols; c++) {
    out += d;
    out += cols[c];
  }
  for (int g = 0; a->per_read_group && g < n_rg; g++)
    for (int c = 0; c < ncols; c++) {
      out += d;
      out += "RG-";
      put_u(out, (uint64_t)g);
      out += '_';
      out += cols[c];
    }
  out += '\n';
  const char* ref = S.h_ref.data() + off;   // reference bases of the target
  // averages: per-row sums in row order, as the reference accumulates them; zero rows add nothing
  double avg_u = 0.0, avg_r = 0.0, avg_t = 0.0;
  std::vector<double> rg_u((size_t)n_rg, 0.0), rg_r((size_t)n_rg, 0.0), rg_t((size_t)n_rg, 0.0);
  for (int k = 0; k < passes; k++) {
    double su = 0.0, sr = 0.0, st = 0.0;
    for (int64_t j = 0; j < n; j++) {
      const uint32_t u = hu[k][(size_t)j] + hu[k][(size_t)(n + j)];
      const double r0 = hred[k][(size_t)j], r1 = hred[k][(size_t)(n + j)];
      su += (double)u;
      sr += r0 + r1;
      st += (double)u + r0 + r1;
    }
    if (k == 0) {
      avg_u = su;
      avg_r = sr;
      avg_t = st;
    } else {
      rg_u[(size_t)(k - 1)] = su;
      rg_r[(size_t)(k - 1)] = sr;
      rg_t[(size_t)(k - 1)] = st;
    }
  }
  // rows: n with counts at first + j*ds, then zero rows z0..tail_end; formatted in parallel chunks
  const int64_t n_rows = n + nz;
  const int threads = (int)std::min<int64_t>(host_threads(), std::max<int64_t>(1, n_rows / 4096));
  std::vector<std::string> parts((size_t)threads);
  parallel_ranges(n_rows, threads, [&](int t, int64_t i0, int64_t i1) {
    std::string& o = parts[(size_t)t];
    o.reserve((size_t)(i1 - i0) * (size_t)(12 + 4 * ncols * passes));
    for (int64_t i = i0; i < i1; i++) {
      if (i >= n) {
        const int64_t pos = z0 + (i - n);
        put_u(o, (uint64_t)pos);
        o += d;
        o += ref[pos - 1];
        for (int c = 0; c < ncols * passes; c++) {
          o += d;
          o += '0';
        }
        o += '\n';
        continue;
      }
      const int64_t pos = first + i * ds;
      const size_t jt = (size_t)i, jb = (size_t)(n + i);
      put_u(o, (uint64_t)pos);
      o += d;
      o += ref[pos - 1];
      for (int k = 0; k < passes; k++) {
        const uint32_t u0 = hu[k][jt], u1 = hu[k][jb];
        const double r0 = hred[k][jt], r1 = hred[k][jb];
        if (a->total_only) {
          // the reference's expression order: (u0 + u1) unsigned, then + r0, + r1
          o += d;
          put_u(o, (uint64_t)(uint32_t)(u0 + u1));
          o += d;
          o += fmt_g(r0 + r1);
          o += d;
          o += fmt_g((double)(uint32_t)(u0 + u1) + r0 + r1);
        } else {
          o += d;
          put_u(o, u0);
          o += d;
          put_u(o, u1);
          o += d;
          o += fmt_g(r0);
          o += d;
          o += fmt_g(r1);
          o += d;
          put_u(o, hraw[k][jt]);
          o += d;
          put_u(o, hraw[k][jb]);
          o += d;
          put_u(o, hbeg[k][jt]);
          o += d;
          put_u(o, hbeg[k][jb]);
        }
      }
      o += '\n';
    }
  });
  size_t total = out.size() + 512 + 160 * (size_t)n_rg;
  for (const std::string& p : parts) total += p.size();
  out.reserve(total);
  for (const std::string& p : parts) out += p;
  const uint32_t n_positions = (uint32_t)n_rows;
  auto foot = [&](const std::string& name, const std::string& value) {
    out += '#';
    out += d;
    out += name;
    out += d;
    out += value;
    out += '\n';
  };
  if (a->show_average) foot("reference_unique_average_cov", fmt_g(a->reference_average));
  foot("region_unique_average_cov", fmt_g(avg_u / (double)n_positions));
  foot("region_repeat_average_cov", fmt_g(avg_r / (double)n_positions));
  foot("region_average_cov", fmt_g(avg_t / (double)n_positions));
  {
    std::string v;
    put_u(v, (uint64_t)n_positions);
    foot("number_of_positions", v);
  }
  for (int g = 0; a->per_read_group && g < n_rg; g++) {
    const std::string p = "RG-" + std::to_string(g) + "_";
    foot(p + "region_unique_average_cov", fmt_g(rg_u[(size_t)g] / (double)n_positions));
    foot(p + "region_repeat_average_cov", fmt_g(rg_r[(size_t)g] / (double)n_positions));
    foot(p + "region_average_cov", fmt_g(rg_t[(size_t)g] / (double)n_positions));
  }
  return out;
}

}  // namespace bsq

/* ------------------------------------------------------------------------------------------ */
/* C ABI                                                                                       */

#define BSQ_TRY try {
#define BSQ_CATCH(status)                          \
  }                                                \
  catch (const BsqError& e) {                      \
    set_status(status, e.code, e.what());          \
    return e.code;                                 \
  }                                                \
  catch (const std::bad_alloc&) {                  \
    set_status(status, BSQ_ERR_LIMIT, "out of host memory"); \
    return BSQ_ERR_LIMIT;                          \
  }                                                \
  catch (const std::exception& e) {                \
    set_status(status, BSQ_ERR_ARG, e.what());     \
    return BSQ_ERR_ARG;                            \
  }

extern "C" {

int bsq_session_create(int32_t device, bsq_session** out, bsq_status* status) {
  BSQ_TRY
  if (!out) throw BsqError(BSQ_ERR_ARG, "null output pointer");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw BsqError(BSQ_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= n) throw BsqError(BSQ_ERR_DEVICE, "device index out of range");
  std::unique_ptr<bsq_session> s(new bsq_session());
  s->s.device = device;
  HIP_OK(hipSetDevice(device));
  HIP_OK(hipStreamCreateWithFlags(&s->s.stream, hipStreamNonBlocking));
  HIP_OK(hipStreamCreateWithFlags(&s->s.stream2, hipStreamNonBlocking));
  HIP_OK(hipEventCreate(&s->s.ev0));
  HIP_OK(hipEventCreate(&s->s.ev1));
  HIP_OK(hipEventCreate(&s->s.evc));
  for (hipEvent_t& e : s->s.evk) HIP_OK(hipEventCreate(&e));
  *out = s.release();
  set_status(status, BSQ_OK, "");
  return BSQ_OK;
  BSQ_CATCH(status)
}

int bsq_session_load(bsq_session* s, const bsq_reference* ref, const bsq_alignments* aln, const bsq_read_files* rf,
                     bsq_status* status) {
  BSQ_TRY
  if (!s) throw BsqError(BSQ_ERR_ARG, "null session");
  HIP_OK(hipSetDevice(s->s.device));
  session_load(s->s, ref, aln, rf);
  set_status(status, BSQ_OK, "");
  return BSQ_OK;
  BSQ_CATCH(status)
}

int bsq_session_error_count(bsq_session* s, const bsq_error_count_args* args, bsq_ec_result** out, bsq_status* status) {
  BSQ_TRY
  if (!s || !args || !out) throw BsqError(BSQ_ERR_ARG, "null argument");
  HIP_OK(hipSetDevice(s->s.device));
  std::unique_ptr<bsq_ec_result> r(new bsq_ec_result());
  run_error_count(s->s, args, *r);
  *out = r.release();
  set_status(status, BSQ_OK, "");
  return BSQ_OK;
  BSQ_CATCH(status)
}

int bsq_session_identify(bsq_session* s, const bsq_identify_args* args, bsq_im_result** out, bsq_status* status) {
  BSQ_TRY
  if (!s || !args || !out) throw BsqError(BSQ_ERR_ARG, "null argument");
  HIP_OK(hipSetDevice(s->s.device));
  std::unique_ptr<bsq_im_result> r(new bsq_im_result());
  run_identify(s->s, args, *r, 0);
  *out = r.release();
  set_status(status, BSQ_OK, "");
  return BSQ_OK;
  BSQ_CATCH(status)
}

// Same as bsq_session_identify, but every covered column is scored exactly (the screen is off).
// Used by the tests to show that the screen never drops an entry.
int bsq_session_identify_unscreened(bsq_session* s, const bsq_identify_args* args, bsq_im_result** out,
                                    bsq_status* status) {
  BSQ_TRY
  if (!s || !args || !out) throw BsqError(BSQ_ERR_ARG, "null argument");
  HIP_OK(hipSetDevice(s->s.device));
  std::unique_ptr<bsq_im_result> r(new bsq_im_result());
  run_identify(s->s, args, *r, 1);
  *out = r.release();
  set_status(status, BSQ_OK, "");
  return BSQ_OK;
  BSQ_CATCH(status)
}

int bsq_session_coverage_table(bsq_session* s, const bsq_coverage_table_args* args, char** out_text,
                               int64_t* out_len, bsq_status* status) {
  BSQ_TRY
  if (!s || !args || !out_text || !out_len) throw BsqError(BSQ_ERR_ARG, "null argument");
  if (!s->s.loaded) throw BsqError(BSQ_ERR_ARG, "no sample loaded");
  HIP_OK(hipSetDevice(s->s.device));
  const std::string t = coverage_table(s->s, args);
  char* buf = (char*)malloc(t.size() + 1);
  if (!buf) throw BsqError(BSQ_ERR_LIMIT, "out of host memory");
  memcpy(buf, t.data(), t.size());
  buf[t.size()] = '\0';
  *out_text = buf;
  *out_len = (int64_t)t.size();
  set_status(status, BSQ_OK, "");
  return BSQ_OK;
  BSQ_CATCH(status)
}

int bsq_session_pipeline(bsq_session* s, const bsq_error_count_args* ec, const bsq_coverage_stage_args* cov,
                         double prop_option, double seed_option, const bsq_identify_args* im, bsq_ec_result** ec_out,
                         bsq_coverage_summary* per_target, bsq_im_result** im_out, bsq_status* status) {
  std::unique_ptr<bsq_ec_result> e(new bsq_ec_result);
  std::unique_ptr<bsq_im_result> r(new bsq_im_result);
  try {
    if (!s || !ec || !cov || !im || !ec_out || !per_target || !im_out) throw BsqError(BSQ_ERR_ARG, "null argument");
    Session& S = s->s;
    if (!S.loaded) throw BsqError(BSQ_ERR_ARG, "session has no sample loaded");
    if (cov->n_targets != S.n_targets || !cov->target_lengths) throw BsqError(BSQ_ERR_ARG, "coverage stage targets do not match the sample");
    const double e0 = now_ms();
    const bool text_later = !ec->output_dir;
    run_error_count(S, ec, *e, text_later, true);
    const double ec_wall = now_ms() - e0;
    const double ec_k0 = S.timing.kernel_detail_ms[0], ec_k1 = S.timing.kernel_detail_ms[5];
    const int NT = S.n_targets;
    const int G = (int)e->hist.size();
    // the stage reads the histograms as its file holds them (rows j >= 1 as doubles)
    auto H = std::make_shared<std::vector<std::vector<double>>>(G);
    for (int g = 0; g < G; g++) (*H)[g].assign(e->hist[g].begin(), e->hist[g].end());
    std::vector<double> avg(NT, 0.0);
    std::vector<int32_t> grp(NT, 0);
    for (int t = 0; t < NT; t++) {
      grp[t] = cov->coverage_group ? cov->coverage_group[t] : 0;
      if (grp[t] < 0 || grp[t] >= std::max(G, 1)) throw BsqError(BSQ_ERR_ARG, "coverage group of a target out of range");
      if (grp[t] < G) avg[t] = coverage_histogram_mean((*H)[grp[t]].data(), (int64_t)(*H)[grp[t]].size());
    }
    const bool skip = im->skip_missing_coverage_prediction != 0;
    const bool fit = !skip && !cov->targeted_sequencing;
    // the overrides after the stage; with neither the fitted cutoff is negative only for a failed
    // fit, which needs an average of at most 3
    auto apply_options = [=](std::vector<bsq_coverage_summary>& sm) {
      for (int t = 0; t < NT; t++) {
        if (!skip && prop_option) {
          sm[t].deletion_coverage_propagation_cutoff = (prop_option < 1) ? sm[t].average * prop_option : prop_option;
        }
      }
    };
    // the seed override reads the stage's average (0 when the stage fitted nothing)
    std::vector<double> seed(NT, 0.0), sign(NT, 1.0);
    for (int t = 0; t < NT; t++) {
      if (!skip && seed_option) seed[t] = (seed_option < 1) ? (fit ? avg[t] : 0.0) * seed_option : seed_option;
      if (!fit) sign[t] = 0.0;   // the summary keeps its default cutoff of 0
      if (fit && !prop_option && !(avg[t] > 3)) sign[t] = -2.0;   // decided by the fit
    }
    auto summaries = std::make_shared<std::vector<bsq_coverage_summary>>(NT);
    auto args = std::make_shared<bsq_coverage_stage_args>(*cov);
    auto hp = std::make_shared<std::vector<const double*>>(G);
    auto hl = std::make_shared<std::vector<int64_t>>(G);
    for (int g = 0; g < G; g++) {
      (*hp)[g] = (*H)[g].data();
      (*hl)[g] = (int64_t)(*H)[g].size();
    }
    args->histograms = hp->data();
    args->histogram_lengths = hl->data();
    args->n_coverage_groups = G;
    double* fit_ms = &S.timing.fit_ms;
    S.timing.fit_ms = 0.0;
    // the error-rate text is printed on the fit's thread after the fit, while the column pass runs
    bsq_ec_result* ep = e.get();
    const int ec_sm = ec->read_set_max, ec_q = ec->quality_max;
    LateCutoffs late = run_async([=]() {
      const double f0 = now_ms();
      if (fit) coverage_stage(args.get(), summaries->data(), host_threads());
      (void)H;
      apply_options(*summaries);
      std::vector<double> c(NT);
      for (int t = 0; t < NT; t++) c[t] = (*summaries)[t].deletion_coverage_propagation_cutoff;
      *fit_ms = now_ms() - f0;
      if (text_later) ep->rates_text = error_rates_text(ec_sm, ec_q, ep->log10p);
      return c;
    }).share();
    WaitOnExit<LateCutoffs> late_wait{late};   // it writes the error_count result
    // targets whose sign the fit decides wait for it before the column pass
    bool wait_first = false;
    for (int t = 0; t < NT; t++) wait_first = wait_first || sign[t] < -1.0;
    std::vector<double> prov(NT);
    if (wait_first) {
      const std::vector<double>& c = late.get();
      for (int t = 0; t < NT; t++) prov[t] = c[t];
    } else {
      for (int t = 0; t < NT; t++) prov[t] = sign[t];
    }
    bsq_identify_args a2 = *im;
    a2.log10_prob_table = e->rates_table.data();
    a2.n_bins = (int64_t)e->rates_table.size();
    a2.read_set_max = ec->read_set_max;
    a2.quality_max = ec->quality_max;
    a2.deletion_propagation_cutoff = prov.data();
    a2.deletion_seed_cutoff = seed.data();
    run_identify(S, &a2, *r, 0, &late);
    late.wait();
    for (int t = 0; t < NT; t++) per_target[t] = (*summaries)[t];
    S.timing.pipeline_ec_ms = ec_wall;
    S.timing.pipeline_ec_kernel_ms[0] = ec_k0;
    S.timing.pipeline_ec_kernel_ms[1] = ec_k1;
    *ec_out = e.release();
    *im_out = r.release();
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& ex) {
    set_status(status, ex.code, ex.what());
    return ex.code;
  }
}

int bsq_session_precision_audit(bsq_session* s, const bsq_identify_args* args, bsq_precision_report* out,
                                bsq_status* status) {
  try {
    if (!s || !args || !out) throw BsqError(BSQ_ERR_ARG, "null argument");
    run_precision_audit(s->s, args, *out);
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    s->s.audit_on = false;
    set_status(status, e.code, e.what());
    return e.code;
  }
}

int bsq_session_alignment_trims(bsq_session* s, int32_t* xl, int32_t* xr, int64_t n, bsq_status* status) {
  try {
    if (!s || (n > 0 && (!xl || !xr))) throw BsqError(BSQ_ERR_ARG, "null argument");
    Session& S = s->s;
    if (!S.loaded) throw BsqError(BSQ_ERR_ARG, "session has no sample loaded");
    if (n != S.n_input_reads) throw BsqError(BSQ_ERR_ARG, "output arrays must hold one entry per loaded alignment");
    // the SequenceTrims of every target at its global columns: left row, then right row
    DBuf<uint8_t> d_m, d_trims;
    DBuf<int32_t> d_xl, d_xr;
    d_m.alloc((size_t)S.n_pos);
    d_trims.alloc((size_t)2 * S.n_pos);
    for (int t = 0; t < S.n_targets; t++) {
      const int64_t len = S.lengths[t], off = S.tgt_off[t];
      if (len == 0) continue;
      const unsigned blocks = (unsigned)((len + 255) / 256);
      hipLaunchKernelGGL(k_repeat_extent, dim3(blocks), dim3(256), 0, S.stream, S.d_ref.p + off, len, d_m.p + off);
      hipLaunchKernelGGL(k_trim_gather, dim3(blocks), dim3(256), 0, S.stream, d_m.p + off, len, d_trims.p + off,
                         d_trims.p + S.n_pos + off);
    }
    const int64_t nd = S.n_dreads;
    d_xl.alloc((size_t)std::max<int64_t>(1, nd));
    d_xr.alloc((size_t)std::max<int64_t>(1, nd));
    if (nd > 0)
      hipLaunchKernelGGL(k_alignment_trims, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, S.stream, S.d_reads.p, nd,
                         S.d_cigar.p, S.d_bases.p, d_trims.p, d_trims.p + S.n_pos, d_xl.p, d_xr.p);
    HIP_OK(hipGetLastError());
    std::vector<int32_t> hl((size_t)nd), hr((size_t)nd);
    if (nd > 0) {
      HIP_OK(hipMemcpyAsync(hl.data(), d_xl.p, (size_t)nd * sizeof(int32_t), hipMemcpyDeviceToHost, S.stream));
      HIP_OK(hipMemcpyAsync(hr.data(), d_xr.p, (size_t)nd * sizeof(int32_t), hipMemcpyDeviceToHost, S.stream));
    }
    HIP_OK(hipStreamSynchronize(S.stream));
    for (int64_t i = 0; i < n; i++) xl[i] = xr[i] = -1;
    for (int64_t i = 0; i < nd; i++) {
      xl[S.h_src[(size_t)i]] = hl[(size_t)i];
      xr[S.h_src[(size_t)i]] = hr[(size_t)i];
    }
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    set_status(status, e.code, e.what());
    return e.code;
  }
}

int bsq_session_set_gd_writer(bsq_session* s, int32_t on_device, bsq_status* status) {
  if (!s) {
    set_status(status, BSQ_ERR_ARG, "null session");
    return BSQ_ERR_ARG;
  }
  s->s.gd_on_device = on_device != 0;
  set_status(status, BSQ_OK, "");
  return BSQ_OK;
}

int bsq_session_set_gd_capacity(bsq_session* s, int32_t entries, bsq_status* status) {
  if (!s || entries < 1 || entries > kGdMaxEntries) {
    set_status(status, BSQ_ERR_ARG, "null session, or a writer capacity outside 1..2^20");
    return BSQ_ERR_ARG;
  }
  s->s.gd_cap = entries;
  set_status(status, BSQ_OK, "");
  return BSQ_OK;
}

int bsq_session_set_column_launch(bsq_session* s, int32_t blocks, int32_t table_in_lds, int32_t waves,
                                  bsq_status* status) {
  if (!s || blocks < 0 || (table_in_lds != 0 && table_in_lds != 1) || (waves != 0 && waves != 4 && waves != 8)) {
    set_status(status, BSQ_ERR_ARG, "null session, negative blocks, table_in_lds not 0 or 1, or waves not 0, 4 or 8");
    return BSQ_ERR_ARG;
  }
  s->s.col_blocks = blocks;
  s->s.col_lds = table_in_lds;
  s->s.col_waves = waves ? waves : kColWaves;
  set_status(status, BSQ_OK, "");
  return BSQ_OK;
}

}  // extern "C"

namespace {
// Launches chosen by bsq_session_tune_columns, shared by the sessions of the process: a sample
// of the same shape on the same device reuses the choice.  The shape is the mean aligned length
// per alignment (25-base steps), the mean depth (half octaves) and the likelihood-table size
// (read files x qualities), which between them set the work per window and the LDS footprint.
struct ColTuneKey {
  int device, read_len, depth, rows;
  bool operator<(const ColTuneKey& o) const {
    if (device != o.device) return device < o.device;
    if (read_len != o.read_len) return read_len < o.read_len;
    if (depth != o.depth) return depth < o.depth;
    return rows < o.rows;
  }
};
std::mutex g_col_tune_mu;
std::map<ColTuneKey, bsq_column_launch> g_col_tune;
}  // namespace

extern "C" {

int bsq_session_tune_columns(bsq_session* s, const bsq_identify_args* args, int32_t repeats, bsq_column_launch* out,
                             bsq_status* status) {
  BSQ_TRY
  if (!s || !args || !out) throw BsqError(BSQ_ERR_ARG, "null argument");
  if (repeats < 1 || repeats > 100) throw BsqError(BSQ_ERR_ARG, "repeats outside 1..100");
  Session& S = s->s;
  if (!S.loaded) throw BsqError(BSQ_ERR_ARG, "no sample loaded");
  HIP_OK(hipSetDevice(S.device));
  const double per_read = (double)S.aligned_bases / (double)std::max<int64_t>(1, S.n_dreads);
  const double depth = (double)S.aligned_bases / (double)std::max<int32_t>(1, S.n_pos);
  const ColTuneKey key{S.device, (int)std::lround(per_read / 25.0), (int)std::floor(2.0 * std::log2(std::max(1.0, depth))),
                       args->read_set_max * args->quality_max};
  {
    std::lock_guard<std::mutex> lk(g_col_tune_mu);
    auto it = g_col_tune.find(key);
    if (it != g_col_tune.end()) {
      *out = it->second;
      out->from_cache = 1;
      S.col_blocks = out->blocks;
      S.col_lds = out->table_in_lds;
      S.col_waves = out->waves;
      set_status(status, BSQ_OK, "");
      return BSQ_OK;
    }
  }
  const int32_t keep_blocks = S.col_blocks, keep_lds = S.col_lds, keep_waves = S.col_waves;
  bsq_column_launch best{};
  try {
    // the best of `repeats` identify passes, by the column pass's own HIP-event time
    auto time_launch = [&](int32_t blocks, int32_t lds, int32_t waves) {
      S.col_blocks = blocks;
      S.col_lds = lds;
      S.col_waves = waves;
      double t = INFINITY;
      for (int r = 0; r < repeats; r++) {
        bsq_im_result res;
        run_identify(S, args, res, 0);
        t = std::min(t, S.timing.kernel_detail_ms[1]);
      }
      return t;
    };
    {
      bsq_im_result warm;   // first pass of the sample: uploads and table builds
      S.col_blocks = 0;
      S.col_lds = 1;
      S.col_waves = kColWaves;
      run_identify(S, args, warm, 0);
    }
    const bool fits = S.last_P.fold_lds != 0;
    const double default_ms = time_launch(0, 1, kColWaves);
    best.blocks = 0;
    best.table_in_lds = 1;
    best.waves = kColWaves;
    const int64_t g_default = column_blocks(S);
    best.grid = (int32_t)g_default;
    best.ms = default_ms;
    best.default_ms = default_ms;
    best.n_tried = 1;
    // another launch replaces the default only when it is at least 2% faster
    auto consider = [&](int64_t g, int32_t lds, int32_t waves) {
      const double t = time_launch((int32_t)g, lds, waves);
      best.n_tried++;
      const bool is_default = best.blocks == 0 && best.table_in_lds == 1 && best.waves == kColWaves;
      if (t < (is_default ? 0.98 * best.ms : best.ms)) {
        best.blocks = (int32_t)g;
        best.table_in_lds = lds;
        best.waves = waves;
        best.grid = (int32_t)g;
        best.ms = t;
      }
    };
    // Candidate grids for blocks of 8 and of 4 wavefronts: fixed sizes, and the grids that give
    // every wavefront of a block the same number k of windows (the last block aside), which leaves
    // no wavefront idle at a block's end.
    for (int32_t waves : {8, 4}) {
      S.col_waves = waves;
      std::vector<int64_t> grids;
      auto add_grid = [&](int64_t b) {
        S.col_blocks = (int32_t)std::min<int64_t>(std::max<int64_t>(b, 1), INT32_MAX);
        const int64_t g = column_blocks(S);
        if (!(waves == kColWaves && g == g_default) && std::find(grids.begin(), grids.end(), g) == grids.end())
          grids.push_back(g);
      };
      for (int64_t b : {(int64_t)256, (int64_t)512, (int64_t)1024, (int64_t)2048}) add_grid(b);
      for (int64_t k : {1, 2, 3, 4, 5, 6, 8, 12, 16}) {
        const int64_t g = (S.n_windows + waves * k - 1) / (waves * k);
        add_grid((g + 7) / 8 * 8);
      }
      for (int64_t g : grids) consider(g, fits ? 1 : 0, waves);
    }
    // the table in global memory at the best launch (the same launch when the table does not fit)
    if (fits) consider(best.grid, 0, best.waves);
  } catch (...) {
    S.col_blocks = keep_blocks;
    S.col_lds = keep_lds;
    S.col_waves = keep_waves;
    throw;
  }
  S.col_blocks = best.blocks;
  S.col_lds = best.table_in_lds;
  S.col_waves = best.waves;
  best.from_cache = 0;
  {
    std::lock_guard<std::mutex> lk(g_col_tune_mu);
    g_col_tune[key] = best;
  }
  *out = best;
  set_status(status, BSQ_OK, "");
  return BSQ_OK;
  BSQ_CATCH(status)
}

int bsq_session_timing(const bsq_session* s, bsq_timing* out) {
  if (!s || !out) return BSQ_ERR_ARG;
  *out = s->s.timing;
  return BSQ_OK;
}

void bsq_session_destroy(bsq_session* s) {
  if (!s) return;
  (void)hipSetDevice(s->s.device);
  delete s;
}

int bsq_error_count(const bsq_error_count_args* args, bsq_ec_result** out, bsq_status* status) {
  bsq_session* s = nullptr;
  int rc = bsq_session_create(0, &s, status);
  if (rc) return rc;
  rc = bsq_session_load(s, args->ref, args->aln, &args->read_files, status);
  if (!rc) rc = bsq_session_error_count(s, args, out, status);
  bsq_session_destroy(s);
  return rc;
}

int bsq_identify_mutations(const bsq_identify_args* args, bsq_im_result** out, bsq_status* status) {
  bsq_session* s = nullptr;
  int rc = bsq_session_create(0, &s, status);
  if (rc) return rc;
  rc = bsq_session_load(s, args->ref, args->aln, &args->read_files, status);
  if (!rc) rc = bsq_session_identify(s, args, out, status);
  bsq_session_destroy(s);
  return rc;
}

int bsq_ec_result_histogram(const bsq_ec_result* r, int32_t group, const uint32_t** hist, int64_t* len) {
  if (!r || group < 0 || group >= (int32_t)r->hist.size()) return BSQ_ERR_ARG;
  *hist = r->hist[group].data();
  *len = (int64_t)r->hist[group].size();
  return BSQ_OK;
}

int bsq_ec_result_counts(const bsq_ec_result* r, const uint64_t** counts, int64_t* n) {
  if (!r) return BSQ_ERR_ARG;
  *counts = r->counts.data();
  *n = (int64_t)r->counts.size();
  return BSQ_OK;
}

int bsq_ec_result_log10_prob(const bsq_ec_result* r, const double** t, int64_t* n) {
  if (!r) return BSQ_ERR_ARG;
  *t = r->log10p.data();
  *n = (int64_t)r->log10p.size();
  return BSQ_OK;
}

int bsq_ec_result_rates_table(const bsq_ec_result* r, const double** t, int64_t* n) {
  if (!r) return BSQ_ERR_ARG;
  *t = r->rates_table.data();
  *n = (int64_t)r->rates_table.size();
  return BSQ_OK;
}

int bsq_ec_result_rates_text(const bsq_ec_result* r, const char** text, int64_t* len) {
  if (!r) return BSQ_ERR_ARG;
  *text = r->rates_text.c_str();
  *len = (int64_t)r->rates_text.size();
  return BSQ_OK;
}

int bsq_ec_result_no_pos_hash_pr(const bsq_ec_result* r, const double** v, int64_t* n) {
  if (!r || !v || !n) return BSQ_ERR_ARG;
  *v = r->no_pos_hash_pr.data();
  *n = (int64_t)r->no_pos_hash_pr.size();
  return BSQ_OK;
}

void bsq_ec_result_free(bsq_ec_result* r) { delete r; }

int bsq_im_result_gd_text(const bsq_im_result* r, const char** text, int64_t* len) {
  if (!r) return BSQ_ERR_ARG;
  *text = r->text.c_str();
  *len = (int64_t)r->text.size();
  return BSQ_OK;
}

int bsq_im_result_dp_text(const bsq_im_result* r, const char** text, int64_t* len) {
  if (!r) return BSQ_ERR_ARG;
  *text = r->dp_text.c_str();
  *len = (int64_t)r->dp_text.size();
  return BSQ_OK;
}

int bsq_im_result_counts(const bsq_im_result* r, int64_t* n_ra, int64_t* n_mc, int64_t* n_un) {
  if (!r) return BSQ_ERR_ARG;
  *n_ra = r->n_ra;
  *n_mc = r->n_mc;
  *n_un = r->n_un;
  return BSQ_OK;
}

void bsq_im_result_free(bsq_im_result* r) { delete r; }

void bsq_ra_test_settings_default(bsq_ra_test_settings* s, int32_t poly) {
  memset(s, 0, sizeof(*s));
  s->polymorphism_prediction = poly;
  s->mutation_log10_e_value_cutoff = 10;
  s->consensus_frequency_cutoff = poly ? 0.95 : 0.50;
  s->polymorphism_log10_e_value_cutoff = poly ? 2 : 10;
  s->polymorphism_frequency_cutoff = poly ? 0.05 : 0.10;
  s->polymorphism_minimum_variant_coverage_each_strand = 2;
  s->polymorphism_fisher_strand_p_value_cutoff = 0.05;
  s->polymorphism_ks_quality_p_value_cutoff = 0;
}

int bsq_test_ra_evidence(const char* gd_text, const bsq_reference* ref, const bsq_ra_test_settings* s, char** out_text,
                         int64_t* out_len, bsq_status* status) {
  BSQ_TRY
  if (!gd_text || !s || !out_text) throw BsqError(BSQ_ERR_ARG, "null argument");
  std::string r = test_ra_evidence_text(gd_text, ref, *s);
  char* buf = (char*)malloc(r.size() + 1);
  if (!buf) throw std::bad_alloc();
  memcpy(buf, r.data(), r.size());
  buf[r.size()] = '\0';
  *out_text = buf;
  if (out_len) *out_len = (int64_t)r.size();
  set_status(status, BSQ_OK, "");
  return BSQ_OK;
  BSQ_CATCH(status)
}

void bsq_free_text(char* text) { free(text); }

}  // extern "C"
