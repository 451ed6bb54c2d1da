// score_candidates.hip -- the exact FP64 scorer of candidate (column, insert slot) pairs on CDNA4
// (gfx950): allele-frequency EM, variant presence score, profile-likelihood frequency bounds,
// Fisher strand and KS quality tests (identify_mutations.cpp:1669-1910, 3009-3318; stats.cpp
// 2144-2284).
//
// One 4-wavefront block per candidate.  Every sum the reference forms over its per-read data
// (pdata) is accumulated in read order, so results are bit-identical to a sequential loop:
//   * EM passes: per-read terms (log10 s_i + lpmax_i and the five responsibilities) are computed
//     two reads per lane from a per-candidate LDS cache of the reads' table rows; six lanes then
//     add the six term columns in read order.
//   * The candidate's EM work is a chain of ~1000 passes (null fit, profile likelihoods, bisection);
//     it is spread over the four wavefronts:
//       stage 1: null fit | profile at the ML frequency | profile at 0 | profile at 1
//       stage 2: bisection for the lower, then the upper bound; three wavefronts evaluate the
//                midpoint and both possible next midpoints, which is exactly two sequential
//                bisection steps per round.
//   * Fisher terms are evaluated across lanes and added in order by one lane; the exact KS lattice
//     is filled one anti-diagonal at a time (each cell is one addition of two neighbours, so the
//     order of the cells does not change any value).
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace bsq {

namespace {

constexpr int kSW = kScoreWaves;
constexpr int kTermCols = 128;   // reads per EM chunk (two per lane)
constexpr int kKsDiag = 100;     // exact KS runs when m*n < 10000, so min(m, n) <= 99

struct ScoreShared {
  double term[kSW][6][kTermCols];
  double full_f[5];
  double full_ll;
  double res[kSW];
  double lps[5];
  double lo, hi;
  double ks_a[kKsDiag + 1], ks_b[kKsDiag + 1];
  double fisher[64];
  uint32_t ks_bnd[(10000 + 2 + 31) / 32 + 1];
  uint32_t qh[2][128];
  uint32_t cov[6][3];
  uint32_t wcount[kSW];
  int32_t n_keys;
  int32_t iter;
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

// One EM pass with frequencies f (wave-uniform): sums[0] = Sum_i (log10 s_i + lpmax_i) over reads
// with s_i > 0, sums[1+b] = Sum_i w_i(b); every sum in read order (identify_mutations.cpp:3268-3300).
__device__ __forceinline__ void em_pass(const Rows& R, int n, const double f[5], uint32_t allowed, double* term,
                                        double sums[6]) {
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int c0 = 0; c0 < n; c0 += kTermCols) {
#pragma unroll 1
    for (int k = 0; k < 2; k++) {
      const int j = lane + 64 * k;
      const int i = c0 + j;
      if (i < n) {
        double r[6];
        R.get(i, r);
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < 5; b++)
          if (allowed & (1u << b)) s += f[b] * r[b];
        if (s > 0.0) {
          term[j] = log10(s) + r[5];
#pragma unroll
          for (int b = 0; b < 5; b++) term[(1 + b) * kTermCols + j] = (allowed & (1u << b)) ? f[b] * r[b] / s : 0.0;
        } else {
          term[j] = 0.0;
#pragma unroll
          for (int b = 0; b < 5; b++) term[(1 + b) * kTermCols + j] = (allowed & (1u << b)) ? f[b] : 0.0;
        }
      }
    }
    wave_sync();
    if (lane < 6) acc = sum_in_order(acc, term + lane * kTermCols, min(kTermCols, n - c0));
    wave_sync();
  }
#pragma unroll
  for (int k = 0; k < 6; k++) sums[k] = readlane_d(acc, k);
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
  double ll = 0.0;
  double sums[6];
  const int first = (v < 0) ? 1 : 0;
#pragma unroll 1
  for (int it = first; it < first + 50; it++) {
    em_pass(R, n, f, allowed, term, sums);
    ll = sums[0];
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
    // ---- pdata in file order (identify_mutations.cpp:1561-1657): order-preserving compaction over
    //      256 alignments per round
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
      if (tid == 0) SS.n_keys += (int)(SS.wcount[0] + SS.wcount[1] + SS.wcount[2] + SS.wcount[3]);
      __syncthreads();
    }
    const int n = min(SS.n_keys, P.key_cap);
    uint32_t obs_count[5];
    for (int b = 0; b < 5; b++) obs_count[b] = SS.cov[b][0] + SS.cov[b][2];
    // ---- table rows of the reads, cached in LDS
    const bool cached = n <= P.rows_cap;
    if (cached)
      for (int i = tid; i < n * 6; i += kSW * 64) rows[i] = P.tab_r[(size_t)keys[i / 6] * 6 + i % 6];
    __syncthreads();
    Rows RR;
    RR.lds = cached ? rows : nullptr;
    RR.keys = keys;
    RR.tab_r = P.tab_r;
    const char refc = (slot == 0) ? P.ref[p] : '.';
    const int ref_idx = char_index(refc) < 0 ? 5 : char_index(refc);
    // ---- full fit (wave 0) | pure-genotype sums (wave 1, lanes 0..4)
    if (wave == 0) {
      double f[5];
      const double ll = run_em(RR, n, obs_count, 31u, nullptr, -1, 0.0, tol, term, f);
      if (lane == 0) {
        for (int b = 0; b < 5; b++) SS.full_f[b] = f[b];
        SS.full_ll = ll;
      }
    } else if (wave == 1) {
      // pure_genotype_call on the summed likelihoods of the same reads, in read order
      // (identify_mutations.cpp:1669-1689, 3398-3433): 128 reads' terms staged in LDS, then five
      // lanes add them in order
      double acc = 0.0;
      for (int c0 = 0; c0 < n; c0 += kTermCols) {
        for (int j = lane; j < kTermCols && c0 + j < n; j += 64) {
          const double* lp = P.tab_lp + (size_t)keys[c0 + j] * 5;
#pragma unroll
          for (int b = 0; b < 5; b++) term[b * kTermCols + j] = lp[b];
        }
        wave_sync();
        if (lane < 5) acc = sum_in_order(acc, term + lane * kTermCols, min(kTermCols, n - c0));
        wave_sync();
      }
      if (lane < 5) SS.lps[lane] = acc;
    }
    __syncthreads();
    double full_f[5], lps[5];
    for (int b = 0; b < 5; b++) {
      full_f[b] = SS.full_f[b];
      lps[b] = SS.lps[b];
    }
    const double full_ll = SS.full_ll;
    int best;
    double gscore;
    pure_genotype(lps, (uint32_t)n, best, gscore);
    const double cbs = gscore - P.log10_ref_length;
    const bool cons = (best != ref_idx) && !isnan(cbs) && (cbs > 0.0);
    const int major = major_index(full_f, (uint32_t)n);
    const int minor = next_index(full_f, (uint32_t)n, major);
    const int variant = next_index(full_f, (uint32_t)n, ref_idx);
    const bool bounds_apply = (n > 0) && (variant < 5);
    // ---- stage 1: null fit | profile at f_ML | profile at 0 | profile at 1
    if (wave == 0) {
      if (variant != 5 && n > 0) {
        const double ll = run_em(RR, n, obs_count, 31u & ~(1u << variant), nullptr, -1, 0.0, tol, term, nullptr);
        if (lane == 0) SS.res[0] = ll;
      }
    } else if (bounds_apply) {
      const double x = (wave == 1) ? full_f[variant] : (wave == 2) ? 0.0 : 1.0;
      const double ll = run_em(RR, n, nullptr, 31u, SS.full_f, variant, x, tol, term, nullptr);
      if (lane == 0) SS.res[wave] = ll;
    }
    __syncthreads();
    double vscore = __longlong_as_double(0x7ff8000000000000LL);
    bool poly = false;
    if (variant != 5 && n > 0) {
      vscore = (full_ll - SS.res[0]) - P.log10_ref_length;
      poly = vscore >= P.polymorphism_cutoff;
    }
    const bool emitted = cons || poly;
    double lower = 0.0, upper = 1.0;
    // ---- stage 2: write_RA_frequency_bounds (identify_mutations.cpp:3175-3217)
    if (emitted && bounds_apply) {
      const double kDrop = 0.587566;
      const double fh = full_f[variant];
      const double target = SS.res[1] - kDrop;
      const bool lower_zero = SS.res[2] >= target;
      const bool upper_one = SS.res[3] >= target;
      for (int side = 0; side < 2; side++) {
        if (side == 0 && lower_zero) continue;
        if (side == 1 && upper_one) continue;
        __syncthreads();
        if (tid == 0) {
          SS.lo = (side == 0) ? 0.0 : fh;
          SS.hi = (side == 0) ? fh : 1.0;
          SS.iter = 0;
        }
        __syncthreads();
        for (;;) {
          const double lo = SS.lo, hi = SS.hi;
          const int it = SS.iter;
          if (!(it < 40 && (hi - lo) > tol)) break;
          const double mid = 0.5 * (lo + hi);
          const double midL = 0.5 * (lo + mid), midR = 0.5 * (mid + hi);
          if (wave > 0) {
            const double x = (wave == 1) ? mid : (wave == 2) ? midL : midR;
            const double ll = run_em(RR, n, nullptr, 31u, SS.full_f, variant, x, tol, term, nullptr);
            if (lane == 0) SS.res[wave] = ll;
          }
          __syncthreads();
          if (tid == 0) {
            // two steps of the sequential bisection from the three evaluations
            double l = lo, h = hi;
            int i = it;
            // lower bound: profile >= target moves hi down; upper bound: moves lo up
            const bool ge = SS.res[1] >= target;
            const bool left = (side == 0) ? ge : !ge;   // the interval kept is [lo, mid]
            if (left) h = mid;
            else l = mid;
            i++;
            if (i < 40 && (h - l) > tol) {
              const double m2 = 0.5 * (l + h);   // == midL when left, midR otherwise
              const bool ge2 = (left ? SS.res[2] : SS.res[3]) >= target;
              const bool left2 = (side == 0) ? ge2 : !ge2;
              if (left2) h = m2;
              else l = m2;
              i++;
            }
            SS.lo = l;
            SS.hi = h;
            SS.iter = i;
          }
          __syncthreads();
        }
        if (side == 0) lower = SS.hi;
        else upper = SS.lo;
      }
    }
    // ---- annotate_polymorphism_statistics (identify_mutations.cpp:3009-3033), wave 0
    double ks_p = 1.0, fisher_p = 1.0;
    if (emitted && wave == 0) {
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
      if (nx > 0 && ny > 0) ks_p = ks_less_w(SS, nx, ny);
      const uint32_t mjt = (major < 5) ? SS.cov[major][2] : 0, mjb = (major < 5) ? SS.cov[major][0] : 0;
      const uint32_t mnt = (minor < 5) ? SS.cov[minor][2] : 0, mnb = (minor < 5) ? SS.cov[minor][0] : 0;
      fisher_p = fisher_2x2_w(mnt, mnb, mjt, mjb, SS.fisher);
    }
    if (tid == 0) {
      RaOut out;
      out.gpos = p;
      out.slot = slot;
      out.emitted = emitted ? 1 : 0;
      out.ref_idx = (uint8_t)ref_idx;
      out.major = (uint8_t)major;
      out.minor = (uint8_t)minor;
      out.variant = (uint8_t)variant;
      out.pad[0] = out.pad[1] = out.pad[2] = 0;
      out.n = (uint32_t)n;
      for (int b = 0; b < 5; b++) out.f[b] = full_f[b];
      out.variant_score = vscore;
      out.lower = lower;
      out.upper = upper;
      out.ks_p = ks_p;
      out.fisher_p = fisher_p;
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
