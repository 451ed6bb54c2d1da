// kernels.hip -- CDNA4 (gfx950) kernels of the pileup-scoring engine.
//
//  k_insert_slots     enumerates insert slots >= 1 of columns whose unique reads carry insertions
//  k_score_candidates wave per candidate: allele-frequency EM, presence score, profile-likelihood
//                     bounds, Fisher strand and KS quality tests (identify_mutations.cpp:3009-3318)
//  k_runs             MC and UN interval detection (identify_mutations.cpp:2262-2344, 2972-3007)
//
// Column kernels map one 64-wide wavefront onto 64 consecutive reference columns and walk the
// alignments overlapping that window in file order: every lane sees the same alignment in the
// same iteration (scalar loads of the 40-byte record), and the bases/qualities of consecutive
// columns are consecutive bytes of that alignment, so the loads coalesce.  Each lane accumulates
// its column's sums sequentially in file order, which is the reference's summation order.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace bsq {

#define WAVE 64

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

/* ------------------------------------------------------------------------------------------ */
/* Insert slots >= 1                                                                           */

// One wavefront per column whose unique reads carry insertions enumerates the insert slots the
// reference visits (identify_mutations.cpp:1359-1361, 1598-1599): slot 1 always, and slot s+1 while
// some unique read with a longer insertion has a base other than N at slot s.  Lanes take the
// window's alignments.  Every visited slot goes to the exact scorer, which makes the consensus and
// polymorphism calls itself.
__global__ void __launch_bounds__(64) k_insert_slots(ImParams P) {
  const int n_ins = min(*P.n_ins, P.ins_cap);
  const int lane = threadIdx.x;
  uint32_t err = 0;
  for (int i = blockIdx.x; i < n_ins; i += gridDim.x) {
    const int p = P.ins[i];
    const int64_t w = p / WAVE;
    const int rb = P.win_rb[w], re = P.win_re[w];
    for (int slot = 1; slot < 65536; slot++) {
      if (lane == 0) {
        const int c = atomicAdd(P.n_cand, 1);
        if (c < P.cand_cap) P.cand[c] = make_int2(p, slot);
        else err |= kErrCapacity;
      }
      bool next = false;
      for (int r0 = rb; r0 < re; r0 += WAVE) {
        const int r = r0 + lane;
        if (r >= re) continue;
        const DRead R = P.reads[r];
        if (R.x1 != 1) continue;
        int qpos, indel;
        bool is_del;
        if (!entry_at(R, P.cigar, p, qpos, indel, is_del, err) || is_del || indel <= slot) continue;
        if (P.bases[R.seq_off + qpos + slot] != 15) next = true;
      }
      if (!__any(next)) break;
    }
  }
  if (err) atomicOr(P.err, err);
}

/* ------------------------------------------------------------------------------------------ */
/* Exact FP64 scorer: one wavefront per candidate (position, slot)                             */

struct WaveScratch {
  double term[6][WAVE];   // per-read terms of one chunk: log10 s_i + lpmax_i, w_i(b)
  double sums[6];
  double f[5];
  int flag;
};

// One EM-type pass over the candidate's reads with frequencies f (uniform across the wave).
// Per-read terms are computed in parallel; the six sums are accumulated in read order by six
// lanes, which is the reference's order, so the results are bit-identical to a sequential loop.
__device__ void em_pass(const uint32_t* keys, int n, const double* __restrict__ tab_r, const double f[5],
                        const bool allowed[5], WaveScratch& W, double sums[6]) {
  const int lane = threadIdx.x % WAVE;
  double acc = 0.0;
  for (int c0 = 0; c0 < n; c0 += WAVE) {
    const int i = c0 + lane;
    if (i < n) {
      const double* tr = tab_r + (size_t)keys[i] * 6;
      double rr[5];
      for (int b = 0; b < 5; b++) rr[b] = tr[b];
      const double lpmax = tr[5];
      double s = 0.0;
      for (int b = 0; b < 5; b++)
        if (allowed[b]) s += f[b] * rr[b];
      if (s > 0.0) {
        W.term[0][lane] = log10(s) + lpmax;
        for (int b = 0; b < 5; b++) W.term[1 + b][lane] = allowed[b] ? f[b] * rr[b] / s : 0.0;
      } else {
        W.term[0][lane] = 0.0;
        for (int b = 0; b < 5; b++) W.term[1 + b][lane] = allowed[b] ? f[b] : 0.0;
      }
    }
    __syncthreads();
    if (lane < 6) {
      const int m = min(WAVE, n - c0);
      for (int j = 0; j < m; j++) acc += W.term[lane][j];
    }
    __syncthreads();
  }
  if (lane < 6) W.sums[lane] = acc;
  __syncthreads();
  for (int k = 0; k < 6; k++) sums[k] = W.sums[k];
  __syncthreads();
}

struct AlleleFit {
  double f[5];
  double ll;
  uint32_t n;
};

// fit_allele_frequencies (identify_mutations.cpp:3240-3318)
__device__ AlleleFit fit_alleles_counts(const uint32_t* keys, const uint32_t obs_of[5], int n,
                                        const double* __restrict__ tab_r, const bool allowed[5], double tol,
                                        WaveScratch& W) {
  AlleleFit m;
  for (int b = 0; b < 5; b++) m.f[b] = 0.0;
  m.ll = 0.0;
  m.n = (uint32_t)n;
  if (n == 0) return m;
  int n_allowed = 0;
  for (int b = 0; b < 5; b++)
    if (allowed[b]) {
      n_allowed++;
      m.f[b] = 0.5;
    }
  if (n_allowed == 0) return m;
  // m.f[obs] += 1.0 once per read in the reference: exact for integer counts below 2^53
  for (int b = 0; b < 5; b++)
    if (allowed[b]) m.f[b] += (double)obs_of[b];
  double init_total = 0.0;
  for (int b = 0; b < 5; b++) init_total += m.f[b];
  for (int b = 0; b < 5; b++) m.f[b] /= init_total;
  double sums[6];
  for (int it = 1; it <= 50; it++) {
    em_pass(keys, n, tab_r, m.f, allowed, W, sums);
    double max_delta = 0.0;
    for (int b = 0; b < 5; b++) {
      if (!allowed[b]) continue;
      const double fn = sums[1 + b] / (double)n;
      max_delta = fmax(max_delta, fabs(fn - m.f[b]));
      m.f[b] = fn;
    }
    m.ll = sums[0];
    if (max_delta < tol) break;
  }
  return m;
}

// profile_log10_likelihood (identify_mutations.cpp:3094-3143)
__device__ double profile_ll(const uint32_t* keys, int n, const double* __restrict__ tab_r, const AlleleFit& full,
                             int v, double f_fixed, double tol, WaveScratch& W) {
  double f[5];
  double other = 0.0;
  for (int b = 0; b < 5; b++)
    if (b != v) other += full.f[b];
  for (int b = 0; b < 5; b++) {
    if (b == v) f[b] = f_fixed;
    else f[b] = (other > 0.0) ? (1.0 - f_fixed) * full.f[b] / other : (1.0 - f_fixed) / 4.0;
  }
  const bool all[5] = {true, true, true, true, true};
  double ll = 0.0;
  double sums[6];
  for (int it = 0; it < 50; it++) {
    em_pass(keys, n, tab_r, f, all, W, sums);
    ll = sums[0];
    double others = 0.0;
    for (int b = 0; b < 5; b++)
      if (b != v) others += sums[1 + b];
    double max_delta = 0.0;
    for (int b = 0; b < 5; b++) {
      if (b == v) continue;
      const double fn = (others > 0.0) ? (1.0 - f_fixed) * sums[1 + b] / others : (1.0 - f_fixed) / 4.0;
      max_delta = fmax(max_delta, fabs(fn - f[b]));
      f[b] = fn;
    }
    if (max_delta < tol) break;
  }
  return ll;
}

__device__ __forceinline__ double log_choose(double n, double k) { return lgamma(n + 1) - lgamma(k + 1) - lgamma(n - k + 1); }

// fisher_exact_test_2x2 (stats.cpp:2144-2169)
__device__ double fisher_2x2(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  const uint32_t row1 = a + b, row2 = c + d, col1 = a + c, col2 = b + d;
  const uint32_t n = row1 + row2;
  const uint32_t a_min = (row1 > col2) ? (row1 - col2) : 0;
  const uint32_t a_max = min(row1, col1);
  const double log_denom = log_choose(n, row1);
  const double lpo = log_choose(col1, a) + log_choose(col2, row1 - a) - log_denom;
  const double lre = log(1.0 + 1e-7);
  double total = 0.0;
  for (uint32_t ai = a_min; ai <= a_max; ai++) {
    const double lp = log_choose(col1, ai) + log_choose(col2, row1 - ai) - log_denom;
    if (lp <= lpo + lre) total += exp(lp);
  }
  return fmin(total, 1.0);
}

// ks_test_two_sample_less (stats.cpp:2190-2284) from per-quality histograms of x (minor) and y
// (major).  Pooled ties are whole quality values, so the sorted pooled sample is determined by the
// two histograms.
__device__ double ks_less_hist(const uint16_t* hx, const uint16_t* hy, uint32_t nx, uint32_t ny, double* row) {
  uint32_t cx = 0, cy = 0, bx = 0, by = 0;
  double min_z = 0.0;
  bool have = false;
  for (int v = 0; v < 128; v++) {
    if (hx[v] == 0 && hy[v] == 0) continue;
    cx += hx[v];
    cy += hy[v];
    const double z = (double)cx / nx - (double)cy / ny;
    if (!have || z < min_z) {
      min_z = z;
      have = true;
      bx = cx;
      by = cy;
    }
  }
  const double stat = -min_z;
  if ((double)nx * ny < 10000) {
    const uint32_t m = ny, n = nx;
    const int64_t thr = (int64_t)by * n - (int64_t)bx * m;
    // boundary[k]: k is 0 or a cumulative pooled count at the end of a tie group
    auto boundary = [&](uint32_t k) -> bool {
      if (k == 0) return true;
      uint32_t cum = 0;
      for (int v = 0; v < 128; v++) {
        const uint32_t g = (uint32_t)hx[v] + hy[v];
        if (!g) continue;
        if (k <= cum) return false;
        cum += g;
        if (k == cum) return true;
        if (k < cum) return false;
      }
      return true;
    };
    // u[i][j], i over m (y), j over n (x); keep one row over the smaller dimension
    if (m <= n) {
      // iterate j outer, i inner: row[i] = u[i][j]
      for (uint32_t j = 0; j <= n; j++) {
        for (uint32_t i = 0; i <= m; i++) {
          double val;
          if (i == 0 && j == 0) val = 1.0;
          else val = (i > 0 ? row[i - 1] : 0.0) + (j > 0 ? row[i] : 0.0);
          const int64_t lvl = (int64_t)i * n - (int64_t)j * m;
          if (lvl >= thr && boundary(i + j)) val = 0.0;
          row[i] = val;
        }
      }
      const double total = exp(log_choose(m + n, m));
      const double pv = 1.0 - row[m] / total;
      return fmin(fmax(pv, 0.0), 1.0);
    } else {
      for (uint32_t i = 0; i <= m; i++) {
        for (uint32_t j = 0; j <= n; j++) {
          double val;
          if (i == 0 && j == 0) val = 1.0;
          else val = (i > 0 ? row[j] : 0.0) + (j > 0 ? row[j - 1] : 0.0);
          const int64_t lvl = (int64_t)i * n - (int64_t)j * m;
          if (lvl >= thr && boundary(i + j)) val = 0.0;
          row[j] = val;
        }
      }
      const double total = exp(log_choose(m + n, m));
      const double pv = 1.0 - row[n] / total;
      return fmin(fmax(pv, 0.0), 1.0);
    }
  }
  const double neff = ((double)nx * ny) / (nx + ny);
  return exp(-2.0 * stat * stat * neff);
}

__global__ void __launch_bounds__(64) k_score_candidates(ImParams P) {
  extern __shared__ uint32_t keys[];   // P.key_cap entries
  __shared__ WaveScratch W;
  __shared__ uint32_t cov[6][3];
  __shared__ uint16_t qh[2][128];
  __shared__ int n_keys;
  __shared__ double ks_row[128];
  const int lane = threadIdx.x;
  const int S = P.read_set_max, Q = P.quality_max;
  const int n_cand = min(*P.n_cand, P.cand_cap);
  for (int ci = blockIdx.x; ci < n_cand; ci += gridDim.x) {
    const int2 cd = P.cand[ci];
    const int p = cd.x, slot = cd.y;
    const int64_t w = p / WAVE;
    const int rb = P.win_rb[w], re = P.win_re[w];
    uint32_t err = 0;
    if (lane < 18) cov[lane / 3][lane % 3] = 0;
    if (lane == 0) n_keys = 0;
    __syncthreads();
    // gather pdata keys in file order (wave-parallel, order-preserving compaction)
    for (int r0 = rb; r0 < re; r0 += WAVE) {
      const int r = r0 + lane;
      bool use = false;
      uint32_t key = 0;
      int obs = 0, strand_idx = 0;
      if (r < re) {
        const DRead R = P.reads[r];
        int qpos, indel;
        bool is_del;
        if (entry_at(R, P.cigar, p, qpos, indel, is_del, err)) {
          const int ondel = is_del ? -1 : (indel < 0 ? 0 : indel);
          uint8_t read_base = '.';
          bool past = true;
          if (ondel >= slot) {
            read_base = P.bases[R.seq_off + qpos + slot];
            past = false;
          }
          if (read_base != 15 && R.x1 <= 1 && !is_trimmed(R, qpos, past)) {   // X1 > 1: redundant
            int q;
            if (slot_observation(R, P.bases, P.quals, qpos, ondel, slot, read_base, obs, q, err) &&
                q >= P.base_quality_cutoff && (int)R.rfile < S && q < Q) {
              const bool top = !(R.flags & kReadReversed);
              key = ((((uint32_t)R.mslot * S + R.rfile) * Q + q) * 10u) + obs * 2u + (top ? 0u : 1u);
              strand_idx = top ? 2 : 0;
              use = true;
            }
          }
        }
      }
      const unsigned long long mask = __ballot(use);
      const int before = __popcll(mask & ((1ull << lane) - 1ull));
      const int base = n_keys;
      if (use) {
        const int idx = base + before;
        if (idx < P.key_cap) keys[idx] = key;
        else err |= kErrCapacity;
        atomicAdd(&cov[obs][strand_idx], 1u);
      }
      __syncthreads();
      if (lane == 0) n_keys = base + __popcll(mask);
      __syncthreads();
    }
    const int n = min(n_keys, P.key_cap);
    uint32_t obs_count[5];
    for (int b = 0; b < 5; b++) obs_count[b] = cov[b][0] + cov[b][2];
    const double tol = P.precision_decimal;
    const char refc = (slot == 0) ? P.ref[p] : '.';
    const int ref_idx = char_index(refc) < 0 ? 5 : char_index(refc);
    // pure_genotype_call on the summed likelihoods of the same reads, in read order
    // (identify_mutations.cpp:1669-1689, 3398-3433): five lanes each add one hypothesis' terms.
    if (lane < 5) {
      double acc = 0.0;
      for (int i = 0; i < n; i++) acc += P.tab_lp[(size_t)keys[i] * 5 + lane];
      W.sums[lane] = acc;
    }
    __syncthreads();
    double lps[5];
    for (int b = 0; b < 5; b++) lps[b] = W.sums[b];
    __syncthreads();
    int best;
    double gscore;
    pure_genotype(lps, (uint32_t)n, best, gscore);
    const double cbs = gscore - P.log10_ref_length;
    const bool cons = (best != ref_idx) && !isnan(cbs) && (cbs > 0.0);
    // full fit
    const bool all[5] = {true, true, true, true, true};
    const AlleleFit full = fit_alleles_counts(keys, obs_count, n, P.tab_r, all, tol, W);
    // allele_model::major_index / next_index (identify_mutations.cpp:3061-3086)
    auto major_index = [&](const AlleleFit& m) -> int {
      if (m.n == 0) return 5;
      int best = 0;
      for (int b = 1; b < 5; b++)
        if (m.f[b] > m.f[best]) best = b;
      return (m.f[best] > 0.0) ? best : 5;
    };
    auto next_index = [&](const AlleleFit& m, int excl) -> int {
      if (m.n == 0) return 5;
      const double thr = 0.5 / (double)m.n;
      int best = 5;
      for (int b = 0; b < 5; b++) {
        if (b == excl) continue;
        if (m.f[b] < thr) continue;
        if (best == 5 || m.f[b] > m.f[best]) best = b;
      }
      return best;
    };
    const int major = major_index(full);
    const int minor = next_index(full, major);
    const int variant = next_index(full, ref_idx);
    double vscore = __longlong_as_double(0x7ff8000000000000LL);
    bool poly = false;
    if (variant != 5 && full.n > 0) {
      bool without[5];
      for (int b = 0; b < 5; b++) without[b] = (b != variant);
      AlleleFit nf = fit_alleles_counts(keys, obs_count, n, P.tab_r, without, tol, W);
      vscore = (full.ll - nf.ll) - P.log10_ref_length;
      poly = vscore >= P.polymorphism_cutoff;
    }
    const bool emitted = cons || poly;
    RaOut out;
    out.gpos = p;
    out.slot = slot;
    out.emitted = emitted ? 1 : 0;
    out.ref_idx = (uint8_t)ref_idx;
    out.major = (uint8_t)major;
    out.minor = (uint8_t)minor;
    out.variant = (uint8_t)variant;
    out.n = (uint32_t)n;
    for (int b = 0; b < 5; b++) out.f[b] = full.f[b];
    out.variant_score = vscore;
    out.lower = 0.0;
    out.upper = 1.0;
    out.ks_p = 1.0;
    out.fisher_p = 1.0;
    for (int b = 0; b < 5; b++) {
      out.cov[b][0] = cov[b][2];
      out.cov[b][1] = cov[b][0];
    }
    int tt = 0, tb = 0;
    for (int b = 0; b < 5; b++) {
      tt += (int)cov[b][2];
      tb += (int)cov[b][0];
    }
    out.total_top = tt;
    out.total_bot = tb;
    if (emitted) {
      // write_RA_frequency_bounds (identify_mutations.cpp:3175-3217)
      if (full.n > 0 && variant < 5) {
        const double kDrop = 0.587566;
        const double fh = full.f[variant];
        const double plmax = profile_ll(keys, n, P.tab_r, full, variant, fh, tol, W);
        const double target = plmax - kDrop;
        if (profile_ll(keys, n, P.tab_r, full, variant, 0.0, tol, W) >= target) {
          out.lower = 0.0;
        } else {
          double lo = 0.0, hi = fh;
          for (int i = 0; (i < 40) && ((hi - lo) > tol); i++) {
            const double mid = 0.5 * (lo + hi);
            if (profile_ll(keys, n, P.tab_r, full, variant, mid, tol, W) >= target) hi = mid;
            else lo = mid;
          }
          out.lower = hi;
        }
        if (profile_ll(keys, n, P.tab_r, full, variant, 1.0, tol, W) >= target) {
          out.upper = 1.0;
        } else {
          double lo = fh, hi = 1.0;
          for (int i = 0; (i < 40) && ((hi - lo) > tol); i++) {
            const double mid = 0.5 * (lo + hi);
            if (profile_ll(keys, n, P.tab_r, full, variant, mid, tol, W) >= target) lo = mid;
            else hi = mid;
          }
          out.upper = lo;
        }
      }
      // annotate_polymorphism_statistics (identify_mutations.cpp:3009-3033)
      for (int i = lane; i < 256; i += WAVE) (&qh[0][0])[i] = 0;
      __syncthreads();
      const int mj = major, mn = minor;
      if (lane == 0) {
        uint32_t nx = 0, ny = 0;
        for (int i = 0; i < n; i++) {
          const uint32_t k = keys[i];
          const int ob = (int)((k % 10u) / 2u);
          const int q = (int)((k / 10u) % (uint32_t)Q);
          if (ob == mj && q < 128) {
            qh[1][q]++;
            ny++;
          }
          if (ob == mn && q < 128) {
            qh[0][q]++;
            nx++;
          }
        }
        double ks = 1.0;
        if (nx > 0 && ny > 0) ks = ks_less_hist(qh[0], qh[1], nx, ny, ks_row);
        const uint32_t mjt = (mj < 5) ? cov[mj][2] : 0, mjb = (mj < 5) ? cov[mj][0] : 0;
        const uint32_t mnt = (mn < 5) ? cov[mn][2] : 0, mnb = (mn < 5) ? cov[mn][0] : 0;
        out.ks_p = ks;
        out.fisher_p = fisher_2x2(mnt, mnb, mjt, mjb);
      }
    }
    if (lane == 0) P.ra_out[ci] = out;
    if (err) atomicOr(P.err, err);
    __syncthreads();
  }
}

/* ------------------------------------------------------------------------------------------ */
/* MC / UN runs                                                                                */

// One thread per column; the thread at the first column of a run walks it to its end, so every
// run is found and described by exactly one thread:
//   MC: maximal runs with unique coverage <= the propagation cutoff; the seed is reached if any
//       column of the run, or the column closing it, has total coverage <= the seed cutoff
//       (identify_mutations.cpp:2270-2281); the redundant sub-range follows :2326-2343.
//   UN: maximal runs of columns without a predicted base (:2972-3007).
__global__ void __launch_bounds__(256) k_runs(RunParams P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P.n_pos) return;
  const int t = P.pos_target[p];
  if (t < 0) return;
  const TargetInfo T = P.targets[t];
  if (!T.scored || T.prop_cutoff < 0.0) return;
  const int first = T.off, last = T.off + T.len - 1;
  auto in_mc = [&](int q) -> bool { return (double)P.col_unique[q] <= T.prop_cutoff; };
  auto in_un = [&](int q) -> bool { return !(P.col_flags[q] & kPosPredicted); };
  if (P.do_mc && in_mc(p) && (p == first || !in_mc(p - 1))) {
    bool seed = false, reached_zero = false;
    int red_start = -1, red_end = -1;
    int q = p;
    for (; q <= last && in_mc(q); q++) {
      const uint8_t fl = P.col_flags[q];
      if (fl & kPosSeed) seed = true;
      if (!(fl & kPosRedundant)) {
        reached_zero = true;
        red_end = -1;
      } else if (!reached_zero) {
        red_start = q;
      } else if (red_end < 0) {
        red_end = q;
      }
    }
    const int e = q - 1;
    if (e < last && (P.col_flags[e + 1] & kPosSeed)) seed = true;
    const int k = atomicAdd(P.n_runs, 1);
    if (k < P.run_cap) {
      RunRec R;
      R.kind = 0;
      R.s = p;
      R.e = e;
      R.seed = seed ? 1 : 0;
      R.red_start = red_start;
      R.red_end = red_end;
      R.lo = (p > first) ? P.col_unique[p - 1] : -1;
      R.li = P.col_unique[p];
      R.ri = P.col_unique[e];
      R.ro = (e < last) ? P.col_unique[e + 1] : -1;
      P.runs[k] = R;
    } else {
      atomicOr(P.err, kErrCapacity);
    }
  }
  if (in_un(p) && (p == first || !in_un(p - 1))) {
    int q = p;
    while (q <= last && in_un(q)) q++;
    const int k = atomicAdd(P.n_runs, 1);
    if (k < P.run_cap) {
      RunRec R;
      R.kind = 1;
      R.s = p;
      R.e = q - 1;
      R.seed = 0;
      R.red_start = R.red_end = -1;
      R.lo = R.li = R.ri = R.ro = 0;
      P.runs[k] = R;
    } else {
      atomicOr(P.err, kErrCapacity);
    }
  }
}

}  // namespace bsq
