// score_candidates.hip -- the exact FP64 scorer of candidate (column, insert slot) pairs on CDNA4
// (gfx950): allele-frequency EM, variant presence score, profile-likelihood frequency bounds,
// Fisher strand and KS quality tests (identify_mutations.cpp:1669-1910, 3009-3318; stats.cpp
// 2144-2284).
//
// One 4-wavefront block per candidate.  Every sum the reference forms over its per-read data
// (pdata) is accumulated in read order, so results are bit-identical to a sequential loop:
//   * EM passes: per-read responsibilities are computed one read per lane from a per-candidate LDS
//     cache of the reads' table rows; five lanes then add the five term columns in read order.  The
//     log likelihood, which the reference only reports from the last pass of each loop, is formed
//     once per loop from that pass's frequencies.
//   * The candidate's EM work is a chain of ~1000 passes (null fit, profile likelihoods,
//     bisections).  It runs as phases of one loop; in each phase every wavefront has at most one EM
//     task:
//       phase 0: full fit | pure-genotype sums
//       phase 1: null fit | profile at the ML frequency | profile at 0 | profile at 1, then
//                Fisher and KS
//       phase 2+: bisection rounds for both bounds together.  Three wavefronts on one bound
//                 evaluate the midpoint and both possible next midpoints, which is exactly two
//                 sequential bisection steps per round; when both bounds still run, the upper
//                 one advances one step per round on the fourth wavefront.
//   * Fisher terms are evaluated across lanes and added in order by one lane; the exact KS lattice
//     is filled one anti-diagonal at a time (each cell is one addition of two neighbours, so the
//     order of the cells does not change any value).
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace bsq {

namespace {

constexpr int kSW = kScoreWaves;
static_assert(kSW == 4, "phase roles are laid out for four wavefronts");
constexpr int kTermCols = 64;    // reads per EM chunk (one per lane)
constexpr int kTermRow = 65;     // LDS row stride of a term column: the summing lanes hit different banks
constexpr int kKsDiag = 100;     // exact KS runs when m*n < 10000, so min(m, n) <= 99

// result slots of the phase tasks
enum { kResFull = 0, kResNull = 1, kResPlmax = 2, kResP0 = 3, kResP1 = 4, kResBisect = 5 };   // + side*3 + k

struct ScoreShared {
  double term[kSW][5][kTermRow];
  double full_f[5];
  double lps[5];
  double res[kResBisect + 6];
  double lo[2], hi[2];
  double target, lower, upper, ks_p, fisher_p;
  double ks_a[kKsDiag + 1], ks_b[kKsDiag + 1];
  double fisher[64];
  uint32_t ks_bnd[(10000 + 2 + 31) / 32 + 1];
  uint32_t qh[2][128];
  uint32_t cov[6][3];
  uint32_t wcount[kSW];
  int32_t n_keys;
  int32_t iter[2];
  int32_t active[2];
  int32_t emitted;
  int32_t done;
};

static_assert(sizeof(ScoreShared) <= kScoreStaticLds, "scorer static LDS above its declared bound");

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The candidate's reads: per read the table row r[0..4] = 10^(lp_b - lpmax), r[5] = lpmax, cached
// in LDS when it fits, else read from the table in HBM.
struct Rows {
  const double* lds;            // n x 6, or nullptr
  const uint32_t* keys;
  const double* __restrict__ tab_r;
  __device__ __forceinline__ void get(int i, double r[6]) const {
    if (lds) {
      typedef __attribute__((address_space(3))) const double LdsD;
      const LdsD* p = (const LdsD*)lds + (size_t)i * 6;
#pragma unroll
      for (int k = 0; k < 6; k++) r[k] = p[k];
    } else {
      const double* p = tab_r + (size_t)keys[i] * 6;
#pragma unroll
      for (int k = 0; k < 6; k++) r[k] = p[k];
    }
  }
};

// acc + t[0] + t[1] + ... + t[m-1], added left to right (one lane).  The LDS reads run 8 terms
// ahead of the additions so the dependent chain does not wait on LDS latency.
__device__ __forceinline__ double sum_in_order(double acc, const double* t, int m) {
  int j = 0;
  if (m >= 8) {
    double c[8];
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = t[k];
    for (; j + 16 <= m; j += 8) {
      double nx[8];
#pragma unroll
      for (int k = 0; k < 8; k++) nx[k] = t[j + 8 + k];
#pragma unroll
      for (int k = 0; k < 8; k++) acc += c[k];
#pragma unroll
      for (int k = 0; k < 8; k++) c[k] = nx[k];
    }
#pragma unroll
    for (int k = 0; k < 8; k++) acc += c[k];
    j += 8;
  }
  for (; j < m; j++) acc += t[j];
  return acc;
}

// One EM pass with frequencies f (wave-uniform).  with_ll == false: sums[1+b] = Sum_i w_i(b), the
// responsibilities the update and the stopping test use.  with_ll == true: sums[0] = Sum_i
// (log10 s_i + lpmax_i) over reads with s_i > 0, the log likelihood the reference returns.  Every
// sum is in read order (identify_mutations.cpp:3268-3300).  The log likelihood is only needed for
// the last pass of an EM loop, so the loops compute it once, after they stop, from the
// frequencies of that pass: the same terms in the same order as the reference's last pass.
__device__ __forceinline__ void em_pass(const Rows& R, int n, const double f[5], uint32_t allowed, double* term,
                                        bool with_ll, double sums[6]) {
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int c0 = 0; c0 < n; c0 += kTermCols) {
    const int i = c0 + lane;
    if (i < n) {
      double r[6];
      R.get(i, r);
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < 5; b++)
        if (allowed & (1u << b)) s += f[b] * r[b];
      if (with_ll) {
        term[lane] = (s > 0.0) ? log10(s) + r[5] : 0.0;
      } else if (s > 0.0) {
#pragma unroll
        for (int b = 0; b < 5; b++) term[b * kTermRow + lane] = (allowed & (1u << b)) ? f[b] * r[b] / s : 0.0;
      } else {
#pragma unroll
        for (int b = 0; b < 5; b++) term[b * kTermRow + lane] = (allowed & (1u << b)) ? f[b] : 0.0;
      }
    }
    wave_sync();
    if (lane < (with_ll ? 1 : 5)) acc = sum_in_order(acc, term + lane * kTermRow, min(kTermCols, n - c0));
    wave_sync();
  }
  if (with_ll) {
    sums[0] = readlane_d(acc, 0);
  } else {
#pragma unroll
    for (int b = 0; b < 5; b++) sums[1 + b] = readlane_d(acc, b);
  }
}

// The EM loops of the reference, one wavefront:
//   v < 0:  fit_allele_frequencies (identify_mutations.cpp:3240-3318) over the alleles in
//           `allowed`, started from the Laplace-smoothed counts obs_of; returns the last pass's
//           log10 likelihood and leaves the frequencies in f.
//   v >= 0: profile_log10_likelihood (:3094-3143): f[v] pinned to f_fixed, the other alleles
//           started from full_f rescaled; returns the last pass's log10 likelihood.
__device__ __noinline__ double run_em(const Rows& R, int n, const uint32_t* obs_of, uint32_t allowed,
                                      const double* full_f, int v, double f_fixed, double tol, double* term,
                                      double* f_out) {
  double f[5];
  if (v < 0) {
#pragma unroll
    for (int b = 0; b < 5; b++) f[b] = 0.0;
    if (n == 0 || allowed == 0) {
      if (f_out)
        for (int b = 0; b < 5; b++) f_out[b] = 0.0;
      return 0.0;
    }
#pragma unroll
    for (int b = 0; b < 5; b++)
      if (allowed & (1u << b)) f[b] = 0.5 + (double)obs_of[b];   // 0.5 + one per observing read
    double init_total = 0.0;
#pragma unroll
    for (int b = 0; b < 5; b++) init_total += f[b];
#pragma unroll
    for (int b = 0; b < 5; b++) f[b] /= init_total;
  } else {
    double other = 0.0;
#pragma unroll
    for (int b = 0; b < 5; b++)
      if (b != v) other += full_f[b];
#pragma unroll
    for (int b = 0; b < 5; b++) {
      if (b == v) f[b] = f_fixed;
      else f[b] = (other > 0.0) ? (1.0 - f_fixed) * full_f[b] / other : (1.0 - f_fixed) / 4.0;
    }
    allowed = 31u;
  }
  double sums[6];
  double f_last[5];
  const int first = (v < 0) ? 1 : 0;
#pragma unroll 1
  for (int it = first; it < first + 50; it++) {
    em_pass(R, n, f, allowed, term, false, sums);
#pragma unroll
    for (int b = 0; b < 5; b++) f_last[b] = f[b];
    double max_delta = 0.0;
    if (v < 0) {
#pragma unroll
      for (int b = 0; b < 5; b++) {
        if (!(allowed & (1u << b))) continue;
        const double fn = sums[1 + b] / (double)n;
        max_delta = fmax(max_delta, fabs(fn - f[b]));
        f[b] = fn;
      }
    } else {
      double others = 0.0;
#pragma unroll
      for (int b = 0; b < 5; b++)
        if (b != v) others += sums[1 + b];
#pragma unroll
      for (int b = 0; b < 5; b++) {
        if (b == v) continue;
        const double fn = (others > 0.0) ? (1.0 - f_fixed) * sums[1 + b] / others : (1.0 - f_fixed) / 4.0;
        max_delta = fmax(max_delta, fabs(fn - f[b]));
        f[b] = fn;
      }
    }
    if (max_delta < tol) break;
  }
  // log10 likelihood of the last pass
  em_pass(R, n, f_last, allowed, term, true, sums);
  const double ll = sums[0];
  if (f_out)
    for (int b = 0; b < 5; b++) f_out[b] = f[b];
  return ll;
}

__device__ __noinline__ double log_choose(double n, double k) { return lgamma(n + 1) - lgamma(k + 1) - lgamma(n - k + 1); }

// fisher_exact_test_2x2 (stats.cpp:2144-2169): terms across lanes, summed in order by lane 0.
__device__ __noinline__ double fisher_2x2_w(uint32_t a, uint32_t b, uint32_t c, uint32_t d, double* buf) {
  const int lane = threadIdx.x & 63;
  const uint32_t row1 = a + b, row2 = c + d, col1 = a + c, col2 = b + d;
  const uint32_t n = row1 + row2;
  const uint32_t a_min = (row1 > col2) ? (row1 - col2) : 0;
  const uint32_t a_max = min(row1, col1);
  const double log_denom = log_choose(n, row1);
  const double lpo = log_choose(col1, a) + log_choose(col2, row1 - a) - log_denom;
  const double lre = log(1.0 + 1e-7);
  double total = 0.0;
  for (uint32_t base = a_min; base <= a_max; base += 64) {
    const uint32_t ai = base + (uint32_t)lane;
    double t = 0.0;
    if (ai <= a_max) {
      const double lp = log_choose(col1, ai) + log_choose(col2, row1 - ai) - log_denom;
      if (lp <= lpo + lre) t = exp(lp);
    }
    buf[lane] = t;
    wave_sync();
    if (lane == 0) {
      const uint32_t m = min<uint32_t>(64u, a_max - base + 1u);
      for (uint32_t j = 0; j < m; j++) total += buf[j];
    }
    wave_sync();
  }
  return fmin(readlane_d(total, 0), 1.0);
}

// ks_test_two_sample_less (stats.cpp:2190-2284) from per-quality histograms of x (minor allele) and
// y (major allele); pooled ties are whole quality values.
__device__ __noinline__ double ks_less_w(ScoreShared& SS, uint32_t nx, uint32_t ny) {
  const int lane = threadIdx.x & 63;
  const uint32_t* hx = SS.qh[0];
  const uint32_t* hy = SS.qh[1];
  // statistic and the counts at its minimum (lane-redundant: 128 steps)
  uint32_t cx = 0, cy = 0, bx = 0, by = 0;
  double min_z = 0.0;
  bool have = false;
#pragma unroll 1
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
  if ((double)nx * ny >= 10000) {
    const double neff = ((double)nx * ny) / (nx + ny);
    return exp(-2.0 * stat * stat * neff);
  }
  const uint32_t m = ny, n = nx;   // cells (i, j), i <= m over y, j <= n over x
  const int64_t thr = (int64_t)by * n - (int64_t)bx * m;
  const uint32_t D = m + n;
  // boundary(d): d is 0 or the pooled count at the end of a tie group
  const uint32_t nw = D / 32 + 1;
  for (uint32_t k = lane; k < nw; k += 64) SS.ks_bnd[k] = 0;
  wave_sync();
  if (lane == 0) {
    SS.ks_bnd[0] |= 1u;
    uint32_t cum = 0;
#pragma unroll 1
    for (int v = 0; v < 128; v++) {
      const uint32_t g = hx[v] + hy[v];
      if (!g) continue;
      cum += g;
      SS.ks_bnd[cum >> 5] |= 1u << (cum & 31);
    }
  }
  wave_sync();
  // the diagonal index runs over the smaller dimension
  const bool by_i = (m <= n);
  const uint32_t small = by_i ? m : n, large = by_i ? n : m;
  double* A = SS.ks_a;   // diagonal d-1
  double* B = SS.ks_b;   // diagonal d
  for (uint32_t d = 0; d <= D; d++) {
    const bool bnd = (SS.ks_bnd[d >> 5] >> (d & 31)) & 1u;
    const uint32_t s_lo = (d > large) ? d - large : 0u;
    const uint32_t s_hi = min(small, d);
    for (uint32_t s = s_lo + lane; s <= s_hi; s += 64) {
      const uint32_t i = by_i ? s : d - s;
      const uint32_t j = by_i ? d - s : s;
      double val;
      if (d == 0) {
        val = 1.0;
      } else {
        // u[i-1][j] + u[i][j-1], both on diagonal d-1
        const double up = (i > 0) ? A[by_i ? s - 1 : s] : 0.0;
        const double left = (j > 0) ? A[by_i ? s : s - 1] : 0.0;
        val = up + left;
      }
      const int64_t lvl = (int64_t)i * n - (int64_t)j * m;
      if (lvl >= thr && bnd) val = 0.0;
      B[s] = val;
    }
    wave_sync();
    double* t = A;
    A = B;
    B = t;
  }
  const double umn = A[by_i ? m : n];
  const double total = exp(log_choose(m + n, m));
  const double pv = 1.0 - umn / total;
  return fmin(fmax(pv, 0.0), 1.0);
}

// pure_genotype_call kept out of line: its FP64 pow/log10 would otherwise set the register
// footprint of the whole kernel.
__device__ __noinline__ void pure_genotype_call(const double* lps, uint32_t n, int* best, double* score) {
  double l[5];
  for (int b = 0; b < 5; b++) l[b] = lps[b];
  pure_genotype(l, n, *best, *score);
}

// allele_model::major_index / next_index (identify_mutations.cpp:3061-3086)
__device__ __forceinline__ int major_index(const double f[5], uint32_t n) {
  if (n == 0) return 5;
  int best = 0;
  for (int b = 1; b < 5; b++)
    if (f[b] > f[best]) best = b;
  return (f[best] > 0.0) ? best : 5;
}

__device__ __forceinline__ int next_index(const double f[5], uint32_t n, int excl) {
  if (n == 0) return 5;
  const double thr = 0.5 / (double)n;
  int best = 5;
  for (int b = 0; b < 5; b++) {
    if (b == excl) continue;
    if (f[b] < thr) continue;
    if (best == 5 || f[b] > f[best]) best = b;
  }
  return best;
}

}  // namespace

__global__ void __launch_bounds__(kSW * 64) k_score_candidates(ImParams P) {
  extern __shared__ __align__(16) unsigned char dyn[];
  uint32_t* keys = (uint32_t*)dyn;
  double* rows = (double*)(dyn + (((size_t)P.key_cap * sizeof(uint32_t) + 15) & ~(size_t)15));
  __shared__ ScoreShared SS;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  double* term = &SS.term[wave][0][0];
  const int S = P.read_set_max, Q = P.quality_max;
  const double tol = P.precision_decimal;
  const double kDrop = 0.587566;   // kProfileLikelihoodLog10Drop (identify_mutations.cpp:3173)
  const int n_cand = min(*P.n_cand, P.cand_cap);
  uint32_t err = 0;
  for (int ci = blockIdx.x; ci < n_cand; ci += gridDim.x) {
    const int2 cd = P.cand[ci];
    const int p = cd.x, slot = cd.y;
    const int64_t w = p / 64;
    const int rb = P.win_rb[w], re = P.win_re[w];
    if (tid < 18) SS.cov[tid / 3][tid % 3] = 0;
    if (tid == 0) SS.n_keys = 0;
    __syncthreads();
    // ---- pdata in file order (identify_mutations.cpp:1561-1657): order-preserving compaction
    for (int r0 = rb; r0 < re; r0 += kSW * 64) {
      const int r = r0 + tid;
      bool use = false;
      uint32_t key = 0;
      int obs = 0, strand_idx = 0;
      if (r < re) {
        const DRead R = P.reads[r];
        use = pdata_key(R, P.cigar, P.bases, P.quals, p, slot, S, Q, P.base_quality_cutoff, key, obs, strand_idx, err);
      }
      const unsigned long long mask = __ballot(use);
      if (lane == 0) SS.wcount[wave] = (uint32_t)__popcll(mask);
      __syncthreads();
      int base = SS.n_keys;
      for (int k = 0; k < wave; k++) base += (int)SS.wcount[k];
      if (use) {
        const int idx = base + __popcll(mask & ((1ull << lane) - 1ull));
        if (idx < P.key_cap) keys[idx] = key;
        else err |= kErrCapacity;
        atomicAdd(&SS.cov[obs][strand_idx], 1u);
      }
      __syncthreads();
      if (tid == 0) {
        int tot = 0;
        for (int k = 0; k < kSW; k++) tot += (int)SS.wcount[k];
        SS.n_keys += tot;
      }
      __syncthreads();
    }
    const int n = min(SS.n_keys, P.key_cap);
    uint32_t obs_count[5];
    for (int b = 0; b < 5; b++) obs_count[b] = SS.cov[b][0] + SS.cov[b][2];
    // ---- table rows of the reads, cached in LDS
    const bool cached = n <= P.rows_cap;
    if (cached)
      for (int i = tid; i < n * 6; i += kSW * 64) rows[i] = P.tab_r[(size_t)keys[i / 6] * 6 + i % 6];
    if (tid == 0) {
      SS.done = 0;
      SS.emitted = 0;
      SS.active[0] = SS.active[1] = 0;
      SS.lower = 0.0;
      SS.upper = 1.0;
      SS.ks_p = 1.0;
      SS.fisher_p = 1.0;
    }
    __syncthreads();
    Rows RR;
    RR.lds = cached ? rows : nullptr;
    RR.keys = keys;
    RR.tab_r = P.tab_r;
    const char refc = (slot == 0) ? P.ref[p] : '.';
    const int ref_idx = char_index(refc) < 0 ? 5 : char_index(refc);
    int major = 5, minor = 5, variant = 5;
    bool cons = false;
    // ---- phases: one EM task per wavefront and phase, a single EM call site
    for (int ph = 0;; ph++) {
      int mode = 0;   // 0 none, 1 fit, 2 profile
      uint32_t allowed = 31u;
      double x = 0.0;
      int res_slot = -1;
      if (ph == 0) {
        if (wave == 0) {
          mode = 1;
          res_slot = kResFull;
        }
      } else if (ph == 1) {
        if (variant != 5 && n > 0) {
          if (wave == 0) {
            mode = 1;
            allowed = 31u & ~(1u << variant);
            res_slot = kResNull;
          } else if (wave <= 3) {
            mode = 2;
            x = (wave == 1) ? SS.full_f[variant] : (wave == 2) ? 0.0 : 1.0;
            res_slot = (wave == 1) ? kResPlmax : (wave == 2) ? kResP0 : kResP1;
          }
        }
      } else {
        // both bounds running: the lower one on wavefronts 0-2, the upper one on wavefront 3;
        // one running: wavefronts 0-2 on it
        const bool a0 = SS.active[0] != 0, a1 = SS.active[1] != 0;
        int side = -1, k = 0;
        if (a0 && a1) {
          side = (wave < 3) ? 0 : 1;
          k = (wave < 3) ? wave : 0;
        } else if ((a0 || a1) && wave < 3) {
          side = a0 ? 0 : 1;
          k = wave;
        }
        if (side >= 0) {
          const double lo = SS.lo[side], hi = SS.hi[side];
          const double mid = 0.5 * (lo + hi);
          x = (k == 0) ? mid : (k == 1) ? 0.5 * (lo + mid) : 0.5 * (mid + hi);
          mode = 2;
          res_slot = kResBisect + side * 3 + k;
        }
      }
      if (mode != 0) {
        double f[5];
        const double ll = run_em(RR, n, obs_count, allowed, SS.full_f, (mode == 2) ? variant : -1, x, tol, term,
                                 (ph == 0) ? f : nullptr);
        if (lane == 0) {
          SS.res[res_slot] = ll;
          if (ph == 0)
            for (int b = 0; b < 5; b++) SS.full_f[b] = f[b];
        }
      }
      if (ph == 0 && wave == 1) {
        // pure_genotype_call on the summed likelihoods of the same reads, in read order
        // (identify_mutations.cpp:1669-1689, 3398-3433)
        double acc = 0.0;
        for (int c0 = 0; c0 < n; c0 += kTermCols) {
          if (c0 + lane < n) {
            const double* lp = P.tab_lp + (size_t)keys[c0 + lane] * 5;
#pragma unroll
            for (int b = 0; b < 5; b++) term[b * kTermRow + lane] = lp[b];
          }
          wave_sync();
          if (lane < 5) acc = sum_in_order(acc, term + lane * kTermRow, min(kTermCols, n - c0));
          wave_sync();
        }
        if (lane < 5) SS.lps[lane] = acc;
      }
      if (ph == 1 && wave == 3) {
        // annotate_polymorphism_statistics (identify_mutations.cpp:3009-3033); used if emitted
        for (int i = lane; i < 256; i += 64) (&SS.qh[0][0])[i] = 0;
        wave_sync();
        uint32_t nx = 0, ny = 0;
        for (int i = lane; i < n; i += 64) {
          const uint32_t k = keys[i];
          const int ob = (int)((k % 10u) / 2u);
          const int q = (int)((k / 10u) % (uint32_t)Q);
          if (ob == major && q < 128) {
            atomicAdd(&SS.qh[1][q], 1u);
            ny++;
          }
          if (ob == minor && q < 128) {
            atomicAdd(&SS.qh[0][q], 1u);
            nx++;
          }
        }
        for (int o = 32; o > 0; o >>= 1) {
          nx += __shfl_xor(nx, o, 64);
          ny += __shfl_xor(ny, o, 64);
        }
        wave_sync();
        const double ks = (nx > 0 && ny > 0) ? ks_less_w(SS, nx, ny) : 1.0;
        const uint32_t mjt = (major < 5) ? SS.cov[major][2] : 0, mjb = (major < 5) ? SS.cov[major][0] : 0;
        const uint32_t mnt = (minor < 5) ? SS.cov[minor][2] : 0, mnb = (minor < 5) ? SS.cov[minor][0] : 0;
        const double fs = fisher_2x2_w(mnt, mnb, mjt, mjb, SS.fisher);
        if (lane == 0) {
          SS.ks_p = ks;
          SS.fisher_p = fs;
        }
      }
      __syncthreads();
      if (ph == 0) {
        // every thread derives the same calls from the shared sums
        double full_f[5], lps[5];
        for (int b = 0; b < 5; b++) {
          full_f[b] = SS.full_f[b];
          lps[b] = SS.lps[b];
        }
        int best;
        double gscore;
        pure_genotype_call(lps, (uint32_t)n, &best, &gscore);
        const double cbs = gscore - P.log10_ref_length;
        cons = (best != ref_idx) && !isnan(cbs) && (cbs > 0.0);
        major = major_index(full_f, (uint32_t)n);
        minor = next_index(full_f, (uint32_t)n, major);
        variant = next_index(full_f, (uint32_t)n, ref_idx);
        continue;   // phase 1 reads the derived calls; no shared state changes here
      }
      if (tid == 0) {
        if (ph == 1) {
          bool poly = false;
          if (variant != 5 && n > 0) poly = ((SS.res[kResFull] - SS.res[kResNull]) - P.log10_ref_length) >= P.polymorphism_cutoff;
          SS.emitted = (cons || poly) ? 1 : 0;
          // write_RA_frequency_bounds (identify_mutations.cpp:3175-3217)
          if (SS.emitted && variant < 5 && n > 0) {
            const double fh = SS.full_f[variant];
            const double target = SS.res[kResPlmax] - kDrop;
            SS.target = target;
            if (SS.res[kResP0] < target) {
              SS.lo[0] = 0.0;
              SS.hi[0] = fh;
              SS.iter[0] = 0;
              SS.active[0] = (fh - 0.0) > tol ? 1 : 0;
              SS.lower = fh;
            }
            if (SS.res[kResP1] < target) {
              SS.lo[1] = fh;
              SS.hi[1] = 1.0;
              SS.iter[1] = 0;
              SS.active[1] = (1.0 - fh) > tol ? 1 : 0;
              SS.upper = fh;
            }
          }
        } else {
          // each running bisection advances two steps from three evaluations (midpoint and both
          // possible next midpoints), or one step from one evaluation
          const bool both = SS.active[0] && SS.active[1];
          for (int side = 0; side < 2; side++) {
            if (!SS.active[side]) continue;
            const bool three = !(both && side == 1);
            const double lo = SS.lo[side], hi = SS.hi[side];
            const double mid = 0.5 * (lo + hi);
            const double* r3 = &SS.res[kResBisect + side * 3];
            double l = lo, h = hi;
            int i = SS.iter[side];
            // lower bound: a profile >= target moves hi down; upper bound: moves lo up
            const bool ge = r3[0] >= SS.target;
            const bool left = (side == 0) ? ge : !ge;   // the interval kept is [lo, mid]
            if (left) h = mid;
            else l = mid;
            i++;
            if (three && i < 40 && (h - l) > tol) {
              const double m2 = 0.5 * (l + h);   // == the left or right next midpoint evaluated
              const bool ge2 = (left ? r3[1] : r3[2]) >= SS.target;
              const bool left2 = (side == 0) ? ge2 : !ge2;
              if (left2) h = m2;
              else l = m2;
              i++;
            }
            SS.lo[side] = l;
            SS.hi[side] = h;
            SS.iter[side] = i;
            SS.active[side] = (i < 40 && (h - l) > tol) ? 1 : 0;
            if (side == 0) SS.lower = h;
            else SS.upper = l;
          }
        }
        SS.done = (SS.active[0] || SS.active[1]) ? 0 : 1;
      }
      __syncthreads();
      if (SS.done) break;
    }
    if (tid == 0) {
      RaOut out;
      const bool emitted = SS.emitted != 0;
      out.gpos = p;
      out.slot = slot;
      out.emitted = emitted ? 1 : 0;
      out.ref_idx = (uint8_t)ref_idx;
      out.major = (uint8_t)major;
      out.minor = (uint8_t)minor;
      out.variant = (uint8_t)variant;
      out.pad[0] = out.pad[1] = out.pad[2] = 0;
      out.n = (uint32_t)n;
      for (int b = 0; b < 5; b++) out.f[b] = SS.full_f[b];
      out.variant_score = (variant != 5 && n > 0) ? (SS.res[kResFull] - SS.res[kResNull]) - P.log10_ref_length
                                                  : __longlong_as_double(0x7ff8000000000000LL);
      out.lower = emitted ? SS.lower : 0.0;
      out.upper = emitted ? SS.upper : 1.0;
      out.ks_p = emitted ? SS.ks_p : 1.0;
      out.fisher_p = emitted ? SS.fisher_p : 1.0;
      int tt = 0, tb = 0;
      for (int b = 0; b < 5; b++) {
        out.cov[b][0] = SS.cov[b][2];
        out.cov[b][1] = SS.cov[b][0];
        tt += (int)SS.cov[b][2];
        tb += (int)SS.cov[b][0];
      }
      out.total_top = tt;
      out.total_bot = tb;
      P.ra_out[ci] = out;
    }
    __syncthreads();
  }
  if (err) atomicOr(P.err, err);
}

}  // namespace bsq
