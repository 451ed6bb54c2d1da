// identify_columns.hip -- pass 2, insert slot 0 of every reference column, on CDNA4 (gfx950):
// coverage tallies, the per-genotype log10 likelihood sums, the pure-genotype call and a
// conservative screen that routes columns to the exact FP64 scorer
// (identify_mutations_pileup::pileup_callback, identify_mutations.cpp:1561-1821).
//
// One 64-wide wavefront owns 64 consecutive columns (lane = column) and walks the alignments that
// overlap them in file order.  Every lane sees the same alignment in the same iteration, so the
// 40-byte record is a scalar load and the bases and qualities of consecutive columns are
// consecutive bytes of that alignment (one coalesced load each).
//
// The likelihood sums are FP32 (one 32-byte table row per read, two LDS reads); the column's
// predicted and consensus decisions are certified against an explicit bound on the difference to
// the reference's FP64 sums, and the rare columns within that bound of a threshold are decided by
// k_fix_columns from exact FP64 sums in read order.  Coverage counts are integers and the 1/X1
// redundancy sums stay FP64 in read order, so those outputs are the reference's exactly.
//
// Per alignment the lane work is branch-light: filtered lanes (no entry, N base, trimmed, low
// quality) read the all-zero table row instead of branching; the strand permutation of the
// hypotheses is a wave-uniform branch; redundant alignments take a wave-uniform path that only
// adds 1/X1.  The next alignment's record and bytes are loaded while the current one is added.
// Within each batch of 64 alignments, the unique single-run ones go first through a lean loop per
// strand (class, trimming and quality checks as table lookups and range tests), and the others
// follow in file order, so the FP64 1/X1 sums keep the read order.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.h"
#include "kernels.h"

namespace bsq {

namespace {

constexpr int kWaveCols = 64;

__device__ __forceinline__ uint32_t rdlane(uint32_t v, int j) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, j);
}

// BAM 4-bit base code -> A0 C1 G2 T3; every other code -> 15.
__device__ __forceinline__ uint32_t base_idx4(uint32_t b) {
  const uint32_t i = (uint32_t)(0xFFFFFFF3FFF2F10FULL >> ((b & 15u) * 4u)) & 15u;
  return b > 15u ? 15u : i;
}

// Screening sums of one column (FP32).  Only observations other than the reference base update the
// per-class sums; the reference class follows from n.
struct Screen {
  uint32_t n;          // reads with a likelihood term
  uint32_t cnt[5];     // per observed class, reference class excluded
  float SL;            // Sum_i max_b log10 P_i(b)
  float Dall;          // Sum_i log10 P_i(observed class)
  float G[5];          // per observed class v: Sum (log10 P_i(reference) - log10 P_i(v))
  float Gabs;          // Sum |terms of G| (rounding bound)
};

// What one alignment contributes to one lane's column at insert slot 0.
struct Entry {
  bool cover;          // the alignment has a pileup entry at this column
  bool is_del;
  int qpos;            // clamped into the read when !cover
  int ondel;           // -1 deletion, else max(indel, 0)
  uint32_t b, q;       // base code and quality at qpos (b = '.' for deletions)
  uint32_t qd;         // deletions: quality of the next read base in read direction
  bool del_ok;         // deletions: that base exists and is not N
  bool del_oob;        // deletions: that base lies outside the read (an error if the entry is used)
  bool trimmed;        // is_trimmed at slot 0 (alignment.h:389-410)
};

template <bool kLds>
struct Table;

template <>
struct Table<true> {
  const FoldRow* rows;   // in LDS
  __device__ __forceinline__ const FoldRow& operator[](int i) const {
    typedef __attribute__((address_space(3))) const FoldRow LdsRow;
    return *(const FoldRow*)&((LdsRow*)rows)[i];
  }
};

template <>
struct Table<false> {
  const FoldRow* __restrict__ rows;   // in HBM (L2-resident)
  __device__ __forceinline__ const FoldRow& operator[](int i) const { return rows[i]; }
};

}  // namespace

// Screen (identify_mutations.cpp:3240-3344 bounded from above): true unless no allele v can reach
// the presence-score cutoff.
//   * The full model's log likelihood is at most Sum_i max_b log10 P_i(b) = SL.
//   * The null model (v held out) starts EM from Laplace-smoothed counts f0 and EM never lowers the
//     likelihood, so its value is at least the likelihood at f0, which per read is at least
//     log10 f0(c) + log10 P_i(c) for any held-in class c: the read's own class, or the reference
//     class for reads that observe v.  Summed: Sum_{b!=v} cnt_b log10 f0_b + cnt_v log10 f0_ref
//     + Dall + G[v].
// The FP32 rounding of every sum is bounded explicitly and added to the margin.  Columns whose
// reference base is N go to the exact scorer whenever they have reads.
__device__ __forceinline__ bool screen_passes(const Screen& A, int ref_idx, double log10_ref_len, double cutoff,
                                              double margin) {
  if (A.n == 0) return false;
  if (ref_idx < 0 || ref_idx > 4) return true;
  // class counts; the reference class is what the other classes leave of n
  float cnt[5], Lb[5];
  uint32_t others = 0;
#pragma unroll
  for (int b = 0; b < 5; b++) others += (b == ref_idx) ? 0u : A.cnt[b];
  float sumL = 0.0f, magL = 0.0f, Lref = 0.0f;
#pragma unroll
  for (int b = 0; b < 5; b++) {
    cnt[b] = (float)((b == ref_idx) ? A.n - others : A.cnt[b]);
    // log10 f0_b = log10(0.5 + cnt_b) - log10(init_total)
    Lb[b] = log10f(0.5f + cnt[b]);
    sumL += cnt[b] * Lb[b];
    magL += cnt[b] * fabsf(Lb[b]);
    if (b == ref_idx) Lref = Lb[b];
  }
  const float n = (float)A.n;
  const float u = 5.9604645e-08f;   // 2^-24
  const float mag0 = fabsf(A.SL) + fabsf(A.Dall) + A.Gabs + magL;
#pragma unroll
  for (int v = 0; v < 5; v++) {
    if (v == ref_idx) continue;
    const float Li = log10f(2.0f + (n - cnt[v]));
    // Sum_{b!=v} cnt_b log10 f0_b + cnt_v log10 f0_ref, regrouped
    const float lb = A.Dall + A.G[v] + (sumL - cnt[v] * Lb[v]) + cnt[v] * Lref - n * Li;
    const float mag = mag0 + n * fabsf(Li) + cnt[v] * (fabsf(Lb[v]) + fabsf(Lref));
    // every FP32 term carries at most (n + 16) u of relative error through its sums, the
    // conversions from FP64 and log10f; doubled for safety
    const double bound = 2.0 * (double)(n + 16.0f) * (double)u * (double)mag;
    const double ub = (double)A.SL - (double)lb - log10_ref_len;
    if (ub >= cutoff - margin - bound) return true;
  }
  return false;
}

// Entry of alignment R at column p for the general (non-simple) CIGAR path; loads issued here.
__device__ __forceinline__ void prepare_general(const ImParams& P, const DRead& R, int p, Entry& e, uint32_t& err) {
  int qpos = 0, indel = 0;
  bool is_del = false;
  e.cover = entry_at(R, P.cigar, p, qpos, indel, is_del, err);
  e.is_del = e.cover && is_del;
  e.ondel = is_del ? -1 : (indel < 0 ? 0 : indel);
  e.qpos = e.cover ? qpos : 0;
  const uint8_t* sq = P.bases + R.seq_off;
  const uint8_t* ql = P.quals + R.seq_off;
  e.b = '.';
  e.q = 0;
  e.qd = 0;
  e.del_ok = false;
  e.del_oob = false;
  e.trimmed = is_trimmed(R, e.qpos, e.ondel < 0);
  if (e.cover && !is_del) {
    e.b = sq[e.qpos];
    e.q = ql[e.qpos];
  } else if (e.is_del) {
    // deletion: quality of the next read base in read direction (error_count.cpp:1075-1080)
    const int rev = (R.flags & kReadReversed) ? 1 : 0;
    const int qq = qpos + 1 - rev;
    if (qq < 0 || qq >= (int)R.l_qseq) {
      // reported only if the entry is used: breseq skips redundant and trimmed reads before it
      // looks at this base (identify_mutations.cpp:1591-1626)
      e.del_oob = true;
    } else if (sq[qq] != 15) {
      e.qd = ql[qq];
      e.del_ok = true;
    }
  }
}

// The per-call alignment view of the column pass (thread per alignment).
__global__ void __launch_bounds__(256) k_prep_reads(const DRead* reads, int64_t n, int32_t S, int32_t Q, IdRec* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const DRead R = reads[i];
    IdRec c;
    c.gpos = R.gpos;
    c.gend = R.gend;
    c.seq_off = (uint32_t)(uint64_t)R.seq_off;
    c.qoff = (int32_t)R.y0 - R.gpos;
    // q1 = qpos + 1 <= XL  <=>  qpos < XL;  l_qseq - q1 + 1 <= XR  <=>  qpos >= l_qseq - XR.
    // Query positions are below 65535, so 0xFFFF stands for "no right trim".
    const uint32_t lo = (R.xl >= 0) ? (uint32_t)R.xl : 0u;
    const uint32_t hi = (R.xr >= 0) ? (uint32_t)max(0, (int32_t)R.l_qseq - (int32_t)R.xr) : 0xFFFFu;
    c.trim = lo | (hi << 16);
    c.row_base = ((int)R.rfile < S) ? (((uint32_t)R.mslot * (uint32_t)S + R.rfile) * (uint32_t)Q) * 5u : 0xffffffffu;
    c.x1_flags = (uint32_t)R.x1 | ((uint32_t)R.flags << 16);
    c.pad = 0;
    out[i] = c;
  }
}

// Certified pure_genotype_call (identify_mutations.cpp:3398-3433) from FP32 sums.
//   * Each FP32 sum lps[b] adds n terms of one sign, each rounded once from FP64, in n - 1 rounded
//     additions: with u = 2^-24 and g = (n + 2) u it is within g / (1 - g) |lps[b]| of the exact sum
//     (the recursive-summation bound gamma_n = n u / (1 - n u), with two more u for the term
//     roundings and the reference's own FP64 summation error).  E is the largest such bound; a
//     column so deep that g reaches 1/2 is never certified.
//   * The reference's best is certain when the top sum leads the next by more than 2E.
//   * Its score lps[best] - log10 Sum_{b != best} 10^lps[b] moves by at most 2E when every sum moves
//     by at most E; here each 10^d (d <= 0) is evaluated from (float)d, adding at most 3e-6 (terms
//     below 10^-38 of the total underflow, which changes the total by less than that).
// A decision within 2E + kScoreSlack of its threshold is left to k_fix_columns, which recomputes
// the column's sums exactly in read order.
constexpr double kScoreSlack = 1e-4;

__device__ __forceinline__ double genotype_score_approx(const float lps[5], int best) {
  float off = -3.4028235e38f;
#pragma unroll
  for (int b = 0; b < 5; b++)
    if (b != best) off = fmaxf(off, lps[b]);
  float tot = 0.0f;
#pragma unroll
  for (int b = 0; b < 5; b++)
    if (b != best) tot += exp10f((float)((double)lps[b] - (double)off));
  return ((double)lps[best] - (double)off) - (double)log10f(tot);
}

__device__ __forceinline__ void finish_column(const ImParams& P, int p, int t, int ref_idx, const float lps[5],
                                              double red_bot, double red_top, int uniq, bool insert_slot,
                                              const Screen& A, uint32_t& err) {
  if (P.audit) {
    // diagnostics: the column's FP32 sums and read count, for the precision audit
    float4* a = (float4*)(P.audit + (size_t)8 * p);
    a[0] = make_float4(lps[0], lps[1], lps[2], lps[3]);
    a[1] = make_float4(lps[4], __uint_as_float(A.n), 0.0f, 0.0f);
  }
  // position_coverage::sum (identify_mutations.h:115-120)
  const double r1 = red_bot + red_top;
  const int total = (int)round((double)uniq) + (int)round(r1);
  if (ref_idx < 0) err |= kErrRefChar;
  uint8_t fl = kPosValid;
  bool cons = false, cons_unsure = false;
  if (A.n > 0) {
    int best = 0;
    float E = 0.0f;
#pragma unroll
    for (int b = 0; b < 5; b++) {
      if (lps[b] > lps[best]) best = b;
      E = fmaxf(E, fabsf(lps[b]));
    }
    const double g = ((double)A.n + 2.0) * 5.9604644775390625e-08;
    const double err_sum = (g < 0.5) ? g / (1.0 - g) * (double)E * 1.0001 : INFINITY;
    float second = -3.4028235e38f;
#pragma unroll
    for (int b = 0; b < 5; b++)
      if (b != best) second = fmaxf(second, lps[b]);
    const bool best_unsure = ((double)lps[best] - (double)second) <= 2.0 * err_sum;
    const double margin = 2.0 * err_sum + kScoreSlack;
    const double cbs = genotype_score_approx(lps, best) - P.log10_ref_length;
    const bool pred_unsure = best_unsure || fabs(cbs - P.mutation_cutoff) <= margin;
    if (cbs >= P.mutation_cutoff) fl |= kPosPredicted;
    if (best != ref_idx) cons = cbs > 0.0;
    cons_unsure = best_unsure || (best != ref_idx && fabs(cbs) <= margin);
    if (cons) fl |= kPosConsensus;
    if (pred_unsure || cons_unsure) {
      const int k = atomicAdd(P.n_fix, 1);
      if (k < P.fix_cap) P.fix[k] = p;
      else err |= kErrCapacity;
    }
  }
  if (r1 > 0.0) fl |= kPosRedundant;
  if (insert_slot) fl |= kPosInsert;
  if ((double)total <= P.targets[t].seed_cutoff) fl |= kPosSeed;
  // an uncertain consensus call goes to the exact scorer, which makes the call itself
  const bool cand = cons || cons_unsure || (P.screen_off && A.n > 0) ||
                    screen_passes(A, ref_idx, P.log10_ref_length, P.polymorphism_cutoff, P.screen_margin);
  if (cand) {
    fl |= kPosCandidate;
    const int slot = atomicAdd(P.n_cand, 1);
    if (slot < P.cand_cap) P.cand[slot] = make_int2(p, 0);
    else err |= kErrCapacity;
  }
  if (insert_slot) {
    const int slot = atomicAdd(P.n_ins, 1);
    if (slot < P.ins_cap) P.ins[slot] = p;
    else err |= kErrCapacity;
  }
  P.col_flags[p] = fl;
  P.col_unique[p] = uniq;
  P.col_total[p] = total;
}

// The exact FP64 genotype sums of column p (slot 0), in read order, by one wavefront: its pdata
// gathered by order-preserving compaction into keys[], then five lanes add the FP64 likelihoods in
// that order (identify_mutations.cpp:1561-1657, 1669-1689).  Every lane returns the sums in lps[]
// and the number of reads.  sums[5] is LDS scratch of the wavefront.
__device__ __forceinline__ int exact_column_sums(const ImParams& P, int p, uint32_t* keys, double* sums, double lps[5],
                                                 uint32_t& err) {
  const int lane = threadIdx.x & 63;
  const int S = P.read_set_max, Q = P.quality_max;
  const int64_t w = p / kWaveCols;
  const int rb = P.win_rb[w], re = P.win_re[w];
  if (rb < 0 || re < rb || re > P.n_reads) {
    err |= kErrIndex;
    for (int b = 0; b < 5; b++) lps[b] = 0.0;
    return 0;
  }
  int n = 0;
  for (int r0 = rb; r0 < re; r0 += 64) {
    const int r = r0 + lane;
    bool use = false;
    uint32_t key = 0;
    int obs = 0, strand_idx = 0;
    if (r < re) use = pdata_key(P.reads[r], P.cigar, P.bases, P.quals, p, 0, S, Q, P.base_quality_cutoff, key, obs,
                                strand_idx, err);
    if (use && key >= (uint32_t)P.tab_keys) {
      err |= kErrIndex;
      key = 0;
    }
    const unsigned long long mask = __ballot(use);
    if (use) {
      const int idx = n + __popcll(mask & ((1ull << lane) - 1ull));
      if (idx < P.key_cap) keys[idx] = key;
      else err |= kErrCapacity;
    }
    n += __popcll(mask);
  }
  n = min(n, P.key_cap);
  __syncthreads();
  if (lane < 5) {
    double acc = 0.0;
    for (int j = 0; j < n; j++) acc += P.tab_lp[(size_t)keys[j] * 5 + lane];
    sums[lane] = acc;
  }
  __syncthreads();
  for (int b = 0; b < 5; b++) lps[b] = sums[b];
  __syncthreads();
  return n;
}

// The read list of a column's wavefront: in LDS, or (kKeysGlobal, deep samples) this block's slice
// of the device scratch list.  A compile-time choice, so that the list's accesses keep their address
// space.
template <bool kKeysGlobal>
__device__ __forceinline__ uint32_t* column_key_list(const ImParams& P, uint32_t* lds) {
  if constexpr (kKeysGlobal) return P.key_scratch + (size_t)blockIdx.x * (size_t)P.key_cap;
  else return lds;
}

// Exact pure-genotype decisions for the columns the column kernel could not certify: one wavefront
// per column, the reference's call made from its exact sums.
template <bool kKeysGlobal>
__global__ void __launch_bounds__(64) k_fix_columns_t(ImParams P) {
  extern __shared__ uint32_t fix_lds_keys[];   // P.key_cap, unless the list is in device scratch
  uint32_t* fix_keys = column_key_list<kKeysGlobal>(P, fix_lds_keys);
  __shared__ double fix_lps[5];
  const int lane = threadIdx.x;
  const int n_fix = min(*P.n_fix, P.fix_cap);
  uint32_t err = 0;
  for (int i = blockIdx.x; i < n_fix; i += gridDim.x) {
    const int p = P.fix[i];
    if (p < 0 || p >= P.n_pos || p / kWaveCols >= P.n_windows) {
      err |= kErrIndex;
      continue;
    }
    double lps[5];
    const int n = exact_column_sums(P, p, fix_keys, fix_lps, lps, err);
    int best;
    double score;
    pure_genotype(lps, (uint32_t)n, best, score);
    const double cbs = score - P.log10_ref_length;
    const int ref_idx = char_index(P.ref[p]);
    if (lane == 0) {
      uint8_t fl = P.col_flags[p] & (uint8_t)~(kPosPredicted | kPosConsensus);
      if (cbs >= P.mutation_cutoff) fl |= kPosPredicted;
      if ((best != ref_idx) && !isnan(cbs) && (cbs > 0.0)) fl |= kPosConsensus;
      P.col_flags[p] = fl;
    }
    __syncthreads();
  }
  if (err) atomicOr(P.err, err);
}

// Precision audit (diagnostics, after an identify pass that wrote P.audit): for every listed
// column, the exact FP64 sums beside the column pass's FP32 ones, and whether the final predicted
// and consensus flags equal the calls the exact sums make.  out[6*i .. 6*i+4] = FP64 sums,
// out[6*i+5] = reads; *mismatches counts columns whose flags differ.
template <bool kKeysGlobal>
__global__ void __launch_bounds__(64) k_audit_columns_t(ImParams P, const int32_t* cols, int32_t n_cols, double* out,
                                                        uint32_t* mismatches) {
  extern __shared__ uint32_t audit_lds_keys[];
  uint32_t* keys = column_key_list<kKeysGlobal>(P, audit_lds_keys);
  __shared__ double sums[5];
  const int lane = threadIdx.x;
  uint32_t err = 0;
  for (int i = blockIdx.x; i < n_cols; i += gridDim.x) {
    const int p = cols[i];
    if (p < 0 || p >= P.n_pos || p / kWaveCols >= P.n_windows) {
      err |= kErrIndex;
      continue;
    }
    double lps[5];
    const int n = exact_column_sums(P, p, keys, sums, lps, err);
    int best;
    double score;
    pure_genotype(lps, (uint32_t)n, best, score);
    const double cbs = score - P.log10_ref_length;
    const int ref_idx = char_index(P.ref[p]);
    if (lane == 0) {
      uint8_t want = 0;
      if (cbs >= P.mutation_cutoff) want |= kPosPredicted;
      if ((best != ref_idx) && !isnan(cbs) && (cbs > 0.0)) want |= kPosConsensus;
      const uint8_t have = P.col_flags[p] & (kPosPredicted | kPosConsensus);
      if (have != want) atomicAdd(mismatches, 1u);
      for (int b = 0; b < 5; b++) out[(size_t)6 * i + b] = lps[b];
      out[(size_t)6 * i + 5] = (double)n;
    }
    __syncthreads();
  }
  if (err) atomicOr(P.err, err);
}

template <bool kLds, int kWaves>
__global__ void __launch_bounds__(kWaves * 64) k_identify_columns_t(ImParams P) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Table<kLds> tab;
  if constexpr (kLds) {
    FoldRow* lds_rows = (FoldRow*)smem_raw;
    const int n4 = (P.fold_rows + 1) * (int)(sizeof(FoldRow) / 16);
    const uint4* src = (const uint4*)P.fold;
    uint4* dst = (uint4*)lds_rows;
    // the all-zero last row goes first, so that "no term" is row 0
    constexpr int kRow4 = (int)(sizeof(FoldRow) / 16);
    for (int i = threadIdx.x; i < n4; i += blockDim.x) dst[(i + kRow4) % n4] = src[i];
    __syncthreads();
    tab.rows = lds_rows;
  } else {
    tab.rows = P.fold;
  }
  // LDS rows are shifted by one (the zero row first); rows in HBM keep the zero row last
  constexpr uint32_t kRowOff = kLds ? 1u : 0u;
  const int zero_row = kLds ? 0 : P.fold_rows;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t chunk = (P.n_windows + gridDim.x - 1) / gridDim.x;
  const int64_t w_begin = (int64_t)xcd_major_block() * chunk;
  const int64_t w_end = min<int64_t>(w_begin + chunk, P.n_windows);
  uint32_t err = 0;
  const int S = P.read_set_max, Q = P.quality_max;
  const int qcut = P.base_quality_cutoff;
  // counted qualities: qcut <= q < Q, i.e. (q - qcut) < qspan as unsigned
  const uint32_t qspan = (Q > qcut) ? (uint32_t)(Q - qcut) : 0u;

  for (int64_t w = w_begin + wave; w < w_end; w += kWaves) {
    const int p = (int)(w * kWaveCols + lane);
    const int t = (p < P.n_pos) ? P.pos_target[p] : -1;
    const bool lane_ok = (t >= 0) && P.targets[t].scored && (P.targets[t].prop_cutoff >= 0.0);
    const int ref_idx = lane_ok ? char_index(P.ref[p]) : 0;
    // hypothesis index of the reference base for a forward / reverse read
    const int e_top = (ref_idx >= 0 && ref_idx < 5) ? ref_idx : 0;
    const int e_bot = (e_top < 4) ? 3 - e_top : e_top;
    // the reference base as a read-strand class (7 when it is not A, C, G or T)
    const uint32_t ref_rs_top = (ref_idx >= 0 && ref_idx < 4) ? (uint32_t)ref_idx : 7u;
    const uint32_t ref_rs_bot = (ref_idx >= 0 && ref_idx < 4) ? (uint32_t)(ref_idx ^ 3) : 7u;
    const int rb = __builtin_amdgcn_readfirstlane(P.win_rb[w]);
    const int re = __builtin_amdgcn_readfirstlane(P.win_re[w]);
    float lps0 = 0.0f, lps1 = 0.0f, lps2 = 0.0f, lps3 = 0.0f, lps4 = 0.0f;
    double red_bot = 0.0, red_top = 0.0;
    int uniq = 0;
    bool insert_slot = false;
    Screen A;
    A.n = 0;
    A.SL = 0.0f;
    A.Dall = 0.0f;
    A.Gabs = 0.0f;
#pragma unroll
    for (int i = 0; i < 5; i++) {
      A.cnt[i] = 0;
      A.G[i] = 0.0f;
    }

    // The window's alignment records come in batches of 64 (lane i holds alignment r0 + i) and
    // their fields are read out with v_readlane, so the loop never waits on a scalar load (whose
    // counter it would share with the LDS table reads).  The bytes of alignment j + 1 are requested
    // before alignment j is added.
    for (int r0 = rb; r0 < re; r0 += 64) {
      const int m = min(64, re - r0);
      uint4 ra = make_uint4(0u, 0u, 0u, 0u), rc = make_uint4(0u, 0u, 0u, 0u);
      if (lane < m) {
        const uint4* src = (const uint4*)(P.idrec + r0 + lane);
        ra = src[0];
        rc = src[1];
      }
      // ra = {gpos, gend, seq_off, qoff}, rc = {trim, row_base, x1_flags, pad}
      // Per-alignment view for the fast path (single-run alignments with X1 == 1 and a read file
      // in the table): span length | untrimmed length << 16, the byte offset of the first aligned
      // base, the first untrimmed column, and row_base << 2 | fast << 1.  The span of
      // a single-run alignment and the untrimmed length are below 65536.
      const uint32_t tr_lo = rc.x & 0xFFFFu, tr_hi = rc.x >> 16;
      const uint32_t f_len = (ra.y - ra.x) | ((tr_hi > tr_lo ? tr_hi - tr_lo : 0u) << 16);
      const uint32_t f_at = ra.z + ra.w + ra.x;
      const uint32_t f_clo = tr_lo - ra.w;
      const bool f_fast = (rc.z & 0xFFFFu) == 1u && (rc.z & (kReadSimple << 16)) && rc.y < (1u << 29);
      const uint32_t f_row = f_fast ? (((rc.y + kRowOff) << 2) | 2u) : 0u;
      // bytes of a single-run alignment at this lane's column (its last aligned base when the
      // column is outside it), and the column's offset into the span
      uint32_t nb = 0u, nq = 0u, nd = 0u;
      auto load_bytes = [&](int j) {
        const uint32_t gpos = rdlane(ra.x, j);
        const uint32_t len = rdlane(f_len, j) & 0xFFFFu;
        nd = (uint32_t)p - gpos;
        const uint32_t at = rdlane(f_at, j) + min(nd, len - 1u);
        nb = P.bases[at];
        nq = P.quals[at];
      };
      // The unique single-run alignments of the batch go first (forward ones, then reverse
      // ones), the others after them in file order: only the others add to the FP64 sums of
      // 1/X1 (which keep the reference's read order); the FP32 sums are certified for any order
      // and the counts are order-free.
      const unsigned long long live = (m == 64) ? ~0ull : ((1ull << m) - 1ull);
      const bool f_rev = (rc.z >> 16) & kReadReversed;
      unsigned long long sm = live & ~__ballot(f_fast);
      // one loop per strand, so that the strand is a compile-time constant
      auto fast_loop = [&](auto rev_c, unsigned long long fm) {
        constexpr bool rev = decltype(rev_c)::value;
        // read-strand class per BAM base code (A1 C2 G4 T8), two bits each
        constexpr uint32_t kLut = rev ? 0x0000012Cu : 0x00030210u;
        const uint32_t ref_rs = rev ? ref_rs_bot : ref_rs_top;
        const int e_ref = rev ? e_bot : e_top;
        if (!fm) return;
        struct Bytes {
          uint32_t b, q, d;
        };
        // bytes of alignment j at this lane's column (its last aligned base when the column is
        // outside it) and the column's offset into the span
        auto fetch = [&](int j) {
          const uint32_t gpos = rdlane(ra.x, j);
          const uint32_t len = rdlane(f_len, j) & 0xFFFFu;
          Bytes x;
          x.d = (uint32_t)p - gpos;
          const uint32_t at = rdlane(f_at, j) + min(x.d, len - 1u);
          x.b = P.bases[at];
          x.q = P.quals[at];
          return x;
        };
        auto next = [&]() {
          if (!fm) return -1;
          const int j = __builtin_ctzll(fm);
          fm &= fm - 1ull;
          return j;
        };
        auto add = [&](int j, const Bytes& x) {
          const uint32_t bq = x.b, qq = x.q, dj = x.d;
          const uint32_t fr = rdlane(f_row, j);
          const uint32_t fl = rdlane(f_len, j);
          const uint32_t clo = rdlane(f_clo, j);
          // the conditions combine with & (not &&): evaluated in full on every lane, they are
          // selects and mask operations, where short-circuit forms become branches around the
          // loaded bytes
          const bool valid = lane_ok & (dj < (fl & 0xFFFFu)) & (bq != 15u);
          uniq += valid ? 1 : 0;
          const bool cand = valid & (((uint32_t)p - clo) < (fl >> 16));
          const uint32_t bqc = min(bq, 16u);
          const bool acgt = ((0x116u >> bqc) & 1u) != 0u;
          const uint32_t o_rs = __builtin_amdgcn_ubfe(kLut, 2u * bqc, 2u);
          const bool ok = cand & acgt & ((qq - (uint32_t)qcut) < qspan);
          if (cand & !ok) {
            // not counted: quality below the cutoff, or an error (the reference asserts)
            // (the base-code test restated, so that the common path keeps one mask of it)
            if (__popc(bq) != 1 || bq > 8u) err |= kErrBaseCode;
            else if ((int)qq >= qcut && qq >= (uint32_t)Q) err |= kErrQuality;
          }
          const int row = ok ? (int)((fr >> 2) + qq * 5u + o_rs) : zero_row;
          const FoldRow& T = tab[row];
          const float t0 = T.t[0], t1 = T.t[1], t2 = T.t[2], t3 = T.t[3], t4 = T.t[4];
          if (!rev) {
            lps0 += t0;
            lps1 += t1;
            lps2 += t2;
            lps3 += t3;
          } else {
            lps0 += t3;
            lps1 += t2;
            lps2 += t1;
            lps3 += t0;
          }
          lps4 += t4;
          const float dg = T.diag;
          A.SL += T.lpmax;
          A.Dall += dg;
          A.n += ok ? 1u : 0u;
          if (ok & (o_rs != ref_rs)) {
            // one branch for the rare non-reference observation; inside it every class is
            // updated by a select (adding 0 leaves a sum unchanged), not by a branch per class
            const uint32_t obs = rev ? (o_rs ^ 3u) : o_rs;
            const float e = T.t[e_ref];
            const float g = e - dg;
            A.Gabs += fabsf(e) + fabsf(dg);
#pragma unroll
            for (int o = 0; o < 4; o++) {
              const bool hit = obs == (uint32_t)o;
              A.G[o] += hit ? g : 0.0f;
              A.cnt[o] += hit ? 1u : 0u;
            }
          }
        };
        // unrolled by two so that the bytes in flight alternate between two register sets
        int j = next();
        Bytes xa = fetch(j), xb;
        for (;;) {
          const int j1 = next();
          if (j1 >= 0) xb = fetch(j1);
          add(j, xa);
          if (j1 < 0) break;
          j = next();
          if (j >= 0) xa = fetch(j);
          add(j1, xb);
          if (j < 0) break;
        }
      };
      fast_loop(std::integral_constant<bool, false>{}, __ballot(f_fast && !f_rev) & live);
      fast_loop(std::integral_constant<bool, true>{}, __ballot(f_fast && f_rev) & live);
      while (sm) {
        const int j = __builtin_ctzll(sm);
        sm &= sm - 1ull;
        if (rdlane(rc.z, j) & (kReadSimple << 16)) load_bytes(j);
        const uint32_t dj = nd;
        uint32_t bq = nb, qq = nq;
        const uint32_t xf = rdlane(rc.z, j);
        const uint32_t x1 = xf & 0xFFFFu, rflags = xf >> 16;
        const uint32_t row_base = rdlane(rc.y, j);
        // entry bits: 1 cover, 2 deletion, 4 deletion quality usable, 8 trimmed, 16 insertion follows,
        // 32 deletion quality base outside the read
        uint32_t eb;
        uint32_t qd = 0u;
        if (rflags & kReadSimple) {
          const int gpos = (int)rdlane(ra.x, j), gend = (int)rdlane(ra.y, j);
          const int qpos = p + (int)rdlane(ra.w, j);
          const uint32_t tr = rdlane(rc.x, j);
          const bool cover = lane_ok && dj < (uint32_t)(gend - gpos);
          const bool trimmed = qpos < (int)(tr & 0xFFFFu) || qpos >= (int)(tr >> 16);
          eb = (cover ? 1u : 0u) | (trimmed ? 8u : 0u);
        } else {
          Entry E;
          prepare_general(P, load_read(P.reads, r0 + j), p, E, err);
          eb = ((E.cover && lane_ok) ? 1u : 0u) | (E.is_del ? 2u : 0u) | (E.del_ok ? 4u : 0u) |
               (E.trimmed ? 8u : 0u) | (E.ondel > 0 ? 16u : 0u) | (E.del_oob ? 32u : 0u);
          bq = E.b;
          qq = E.q;
          qd = E.qd;
        }
        const bool is_del = (eb & 2u) != 0u;
        const bool top = !(rflags & kReadReversed);
        const uint32_t base = is_del ? (uint32_t)'.' : bq;
        const bool valid = (eb & 1u) && base != 15u;
        if (x1 != 1u) {
          // coverage of a read with X1 != 1 is 1/X1 (identify_mutations.cpp:1606-1614); a redundant
          // read (X1 > 1) adds nothing else
          const double inv = 1.0 / (double)x1;
          if (top) red_top += valid ? inv : 0.0;
          else red_bot += valid ? inv : 0.0;
          if (x1 > 1u) continue;
        } else {
          uniq += valid ? 1 : 0;
          insert_slot = insert_slot || (valid && (eb & 16u));
        }
        // observed class (reference strand) and quality (error_count.cpp:1049-1105), as selects;
        // a base code other than A C G T N, a read file beyond the table or a quality beyond its
        // maximum is reported (the reference asserts)
        const uint32_t o4 = base_idx4(bq);
        const uint32_t obs = is_del ? 4u : o4;
        const uint32_t q = is_del ? qd : qq;
        // a deletion whose quality base lies outside the read, on an untrimmed unique read
        if (valid && !(eb & 8u) && (eb & 32u)) err |= kErrQueryIndex;
        const bool cand = valid && !(eb & 8u) && (!is_del || (eb & 4u));
        const bool bad_base = cand && !is_del && o4 > 3u;
        bool ok = cand && !bad_base && q >= (uint32_t)qcut;
        const bool bad_rs = ok && row_base == 0xffffffffu;
        const bool bad_q = ok && !bad_rs && q >= (uint32_t)Q;
        err |= (bad_base ? kErrBaseCode : 0u) | (bad_rs ? kErrReadSet : 0u) | (bad_q ? kErrQuality : 0u);
        ok = ok && !bad_rs && !bad_q;
        const uint32_t o_rs = (obs < 4u && !top) ? (obs ^ 3u) : obs;
        const int row = ok ? (int)(row_base + kRowOff + q * 5u + o_rs) : zero_row;
        const FoldRow& T = tab[row];
        const float t0 = T.t[0], t1 = T.t[1], t2 = T.t[2], t3 = T.t[3], t4 = T.t[4];
        if (top) {
          lps0 += t0;
          lps1 += t1;
          lps2 += t2;
          lps3 += t3;
        } else {
          lps0 += t3;
          lps1 += t2;
          lps2 += t1;
          lps3 += t0;
        }
        lps4 += t4;
        const float dg = T.diag;
        A.SL += T.lpmax;
        A.Dall += dg;
        A.n += ok ? 1u : 0u;
        if (ok && (int)obs != ref_idx) {
          const float e = T.t[top ? e_top : e_bot];
          const float g = e - dg;
          A.Gabs += fabsf(e) + fabsf(dg);
#pragma unroll
          for (int o = 0; o < 5; o++) {
            const bool hit = (int)obs == o;
            A.G[o] += hit ? g : 0.0f;
            A.cnt[o] += hit ? 1u : 0u;
          }
        }
      }
    }
    if (!lane_ok) {
      if (p < P.n_pos) {
        P.col_flags[p] = 0;
        P.col_unique[p] = 0;
        P.col_total[p] = 0;
      }
    } else {
      const float lps[5] = {lps0, lps1, lps2, lps3, lps4};
      finish_column(P, p, t, ref_idx, lps, red_bot, red_top, uniq, insert_slot, A, err);
    }
  }
  if (err) atomicOr(P.err, err);
}

template __global__ void k_fix_columns_t<false>(ImParams P);
template __global__ void k_fix_columns_t<true>(ImParams P);
template __global__ void k_audit_columns_t<false>(ImParams P, const int32_t*, int32_t, double*, uint32_t*);
template __global__ void k_audit_columns_t<true>(ImParams P, const int32_t*, int32_t, double*, uint32_t*);
template __global__ void k_identify_columns_t<true, 8>(ImParams P);
template __global__ void k_identify_columns_t<false, 8>(ImParams P);
template __global__ void k_identify_columns_t<true, 4>(ImParams P);
template __global__ void k_identify_columns_t<false, 4>(ImParams P);

}  // namespace bsq
