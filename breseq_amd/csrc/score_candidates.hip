// score_candidates.hip -- the exact FP64 scorer of candidate (column, insert slot) pairs on CDNA4
// (gfx950): allele-frequency EM, variant presence score, profile-likelihood frequency bounds,
// Fisher strand and KS quality tests (identify_mutations.cpp:1669-1910, 3009-3318; stats.cpp
// 2144-2284).
//
// One 4-wavefront block per candidate.  Every sum the reference forms over its per-read data
// (pdata) is accumulated in read order, so results are bit-identical to a sequential loop:
//   * EM: a wavefront runs up to two EM loops ("tasks") in lockstep.  Per pass, each lane takes
//     one read and computes that read's responsibilities for every running task from a
//     per-candidate LDS cache of the reads' table rows; then lane 5t+b adds term column b of task
//     t in read order.  The in-order additions are the pass's critical path and its largest
//     instruction count, and one pass of them serves both tasks.  The log likelihood, which
//     the reference reports only from the last pass of a loop, is formed once per loop from that
//     pass's frequencies: after its last update a task makes one more pass in "likelihood" mode.
//   * The candidate's EM work is a chain of ~1000 passes (null fit, profile likelihoods,
//     bisections), run as phases:
//       phase 0: full fit | pure-genotype sums
//       phase 1: null fit + a speculative bisection midpoint | profile at the ML frequency and
//                at 1, then Fisher and KS | profile at 0 + a midpoint | two more midpoints
//       phase 2+: bisection rounds.  Each round evaluates the midpoints a sequential bisection
//                 would visit along one predicted path: eight levels with one bound open, four
//                 per bound with both open (two midpoints per wavefront).  The path follows the
//                 crossing point of a model of the profile likelihood (below); the round then
//                 keeps every level up to the first one where the real decision leaves the path.
//                 A bound whose prediction goes wrong within its first two levels switches to
//                 full trees (seven midpoints = three levels, or three = two levels).  The first
//                 four path levels (two per bound with both open) run in phase 1.  Only
//                 midpoints the sequential bisection visits are used, so the bounds are its own.
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
constexpr int kTW = 2;           // EM tasks per wavefront
constexpr int kTermCols = 64;    // reads per EM chunk (one per lane)
constexpr int kTermRow = 65;     // LDS row stride of a term column: the summing lanes hit different banks
constexpr int kKsDiag = 100;     // exact KS runs when m*n < 10000, so min(m, n) <= 99
constexpr int kBisectPts = 8;    // midpoints evaluated per bound and round (one per task slot)

// result slots of the phase tasks
enum { kResFull = 0, kResNull = 1, kResPlmax = 2, kResP0 = 3, kResP1 = 4, kResBisect = 5 };   // + side*8 + node
constexpr int kNumRes = kResBisect + 2 * kBisectPts;

// EM task states
enum { kTaskIdle = 0, kTaskIter = 1, kTaskLast = 2 };

// The EM tasks of one wavefront (LDS).
struct WaveTasks {
  double term[kTW][5][kTermRow];   // per task: five term columns of the current chunk
  double f[kTW][5];                // current frequencies
  double fl[kTW][5];               // frequencies of the last pass (the likelihood pass uses them)
  double sum[kTW][5];              // the pass's in-order sums
  double delta[kTW][5];            // |new - old| frequency per allele of the last update
  double x[kTW];                   // profile: pinned frequency
  int32_t mode[kTW];               // 1 fit, 2 profile
  uint32_t allowed[kTW];
  int32_t v[kTW];                  // profile: pinned allele
  int32_t state[kTW];
  int32_t it[kTW];
  int32_t res[kTW];                // result slot
};

struct ScoreShared {
  WaveTasks wt[kSW];
  double full_f[5];
  double lps[5];
  double res[kNumRes];
  double pts[2][kBisectPts];       // bisection midpoints of the round, per bound
  double vlo[2], vhi[2];           // profile values at the ends of each bound's interval
  double lo[2], hi[2];
  double target, lower, upper, ks_p, fisher_p;
  double ks_a[kKsDiag + 1], ks_b[kKsDiag + 1];
  double fisher[64];
  uint32_t ks_bnd[(10000 + 2 + 31) / 32 + 1];
  uint32_t qh[2][128];
  uint32_t cov[6][3];
  uint32_t obs[5];                 // reads per observed class (both strands)
  uint32_t wcount[kSW];
  int32_t n_keys;
  int32_t iter[2];
  int32_t active[2];
  int32_t depth[2];                // bisection levels evaluated this round, per bound
  int32_t path[2];                 // this round's midpoints follow one predicted path (else: a tree)
  int32_t tree_only[2];            // a predicted path went wrong early: trees from then on
  int32_t spec[2];                 // first path levels per bound evaluated beside phase 1
  int32_t emitted;
  int32_t done;
};

// LDS views.  Helpers that are not inlined reach shared memory through these, so their accesses
// compile to LDS instructions rather than generic (flat) ones, whose waits also cover memory loads.
typedef __attribute__((address_space(3))) double LdsDouble;
typedef __attribute__((address_space(3))) uint32_t LdsU32;
typedef __attribute__((address_space(3))) WaveTasks LdsTasks;
typedef __attribute__((address_space(3))) ScoreShared LdsShared;

static_assert(sizeof(ScoreShared) <= kScoreStaticLds, "scorer static LDS above its declared bound");

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double readfirstlane_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
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

// acc + t[0] + t[1] + ... + t[m-1], added left to right (one lane).  LDS latency (~120 cycles)
// is 25 dependent FP64 additions, so the reads run three groups of eight (24 terms) ahead of
// the additions.
__device__ __forceinline__ double sum_in_order(double acc, const LdsDouble* t, int m) {
  int j = 0;
  if (m >= 32) {
    double c0[8], c1[8], c2[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      c0[k] = t[k];
      c1[k] = t[8 + k];
      c2[k] = t[16 + k];
    }
    for (; j + 32 <= m; j += 8) {
      double nx[8];
#pragma unroll
      for (int k = 0; k < 8; k++) nx[k] = t[j + 24 + k];
#pragma unroll
      for (int k = 0; k < 8; k++) acc += c0[k];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        c0[k] = c1[k];
        c1[k] = c2[k];
        c2[k] = nx[k];
      }
    }
#pragma unroll
    for (int k = 0; k < 8; k++) acc += c0[k];
#pragma unroll
    for (int k = 0; k < 8; k++) acc += c1[k];
#pragma unroll
    for (int k = 0; k < 8; k++) acc += c2[k];
    j += 24;
  }
  for (; j < m; j++) acc += t[j];
  return acc;
}

// The same for exactly K terms (K a multiple of 8), unrolled in groups of eight, each group's LDS
// reads issued two groups before it is added.
template <int K>
__device__ __forceinline__ double sum_in_order_k(double acc, const LdsDouble* t) {
  static_assert(K % 8 == 0, "groups of eight");
  constexpr int G = K / 8;
  double g0[8], g1[8], g2[8];
#pragma unroll
  for (int k = 0; k < 8; k++) g0[k] = t[k];
  if (G > 1) {
#pragma unroll
    for (int k = 0; k < 8; k++) g1[k] = t[8 + k];
  }
#pragma unroll
  for (int grp = 0; grp < G; grp++) {
    double* cur = (grp % 3 == 0) ? g0 : (grp % 3 == 1) ? g1 : g2;
    double* nxt = (grp % 3 == 0) ? g2 : (grp % 3 == 1) ? g0 : g1;
    if (grp + 2 < G) {
#pragma unroll
      for (int k = 0; k < 8; k++) nxt[k] = t[(grp + 2) * 8 + k];
    }
#pragma unroll
    for (int k = 0; k < 8; k++) acc += cur[k];
  }
  return acc;
}

// A chunk's column of m terms, padded in LDS with -0.0 up to the next multiple of 16: adding -0.0
// leaves every double unchanged (+0.0 included), so the result is the in-order sum of the m terms.
__device__ __forceinline__ double sum_chunk(double acc, const LdsDouble* t, int m) {
  switch ((m + 15) >> 4) {
    case 1: return sum_in_order_k<16>(acc, t);
    case 2: return sum_in_order_k<32>(acc, t);
    case 3: return sum_in_order_k<48>(acc, t);
    default: return sum_in_order_k<64>(acc, t);
  }
}

// Starts EM task slot t of a wavefront (one lane): the starting point of the reference's loops.
//   mode 1: fit_allele_frequencies (identify_mutations.cpp:3240-3318) over the alleles in
//           `allowed`, from the Laplace-smoothed counts obs_of.
//   mode 2: profile_log10_likelihood (:3094-3143): f[v] pinned to x, the other alleles started
//           from full_f rescaled.
// A fit with no reads or no allowed allele has likelihood 0 and frequencies 0 without a pass.
__device__ __noinline__ void task_start(LdsTasks* W, int t, int mode, uint32_t allowed, int v, double x, int res,
                           const LdsU32* obs_of, const LdsDouble* full_f, int n, LdsDouble* res_out) {
  double f[5];
  W->mode[t] = mode;
  W->v[t] = v;
  W->x[t] = x;
  W->res[t] = res;
  W->it[t] = 0;
  if (mode == 1) {
#pragma unroll
    for (int b = 0; b < 5; b++) f[b] = 0.0;
    if (n == 0 || allowed == 0) {
#pragma unroll
      for (int b = 0; b < 5; b++) W->f[t][b] = 0.0;
      W->state[t] = kTaskIdle;
      res_out[res] = 0.0;
      return;
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
      if (b == v) f[b] = x;
      else f[b] = (other > 0.0) ? (1.0 - x) * full_f[b] / other : (1.0 - x) / 4.0;
    }
    allowed = 31u;
  }
  W->allowed[t] = allowed;
#pragma unroll
  for (int b = 0; b < 5; b++) W->f[t][b] = f[b];
  W->state[t] = kTaskIter;
}

// The likelihood pass of task t for one chunk of reads: term column 0 gets log10 s_i + lpmax_i
// (0 for s_i = 0), s_i = Sum_{allowed b} f_b r_i(b) in allele order.  Kept out of line: the
// inlined log10 would set the register budget of every EM pass, and runs once per EM loop.
__device__ __noinline__ void likelihood_terms(const Rows R, int n, int c0, LdsTasks* W, int t, double f0, double f1,
                                              double f2, double f3, double f4, uint32_t al) {
  const int lane = threadIdx.x & 63;
  const int i = c0 + lane;
  if (i >= n) return;
  double r[6];
  R.get(i, r);
  const double f[5] = {f0, f1, f2, f3, f4};
  double s = 0.0;
#pragma unroll
  for (int b = 0; b < 5; b++) {
    const double x = f[b] * r[b];
    if (al & (1u << b)) s += x;
  }
  W->term[t][0][lane] = (s > 0.0) ? log10(s) + r[5] : 0.0;
}

// Runs the started tasks of a wavefront to completion; each task's log10 likelihood goes to
// res_out[its slot] and its final frequencies stay in W->f.  Per pass, for every running task:
//   kTaskIter: sums[b] = Sum_i w_i(b), w_i(b) = f_b r_i(b) / s_i, s_i = Sum_{allowed b} f_b r_i(b)
//   kTaskLast: sums[0] = Sum_i (log10 s_i + lpmax_i) over reads with s_i > 0
// every sum in read order (identify_mutations.cpp:3268-3300); then the reference's loop body
// (:3268-3318, 3094-3143) updates the frequencies: 50 passes at most, stop when the largest change
// is below tol, then one likelihood pass from that last pass's frequencies.
__device__ __noinline__ void run_tasks(const Rows R, int n, LdsTasks* W, double tol, LdsDouble* res_out) {
  const int lane = threadIdx.x & 63;
  const int my_t = min(lane / 5, kTW - 1), my_b = lane % 5;
  const bool sum_lane = lane < 5 * kTW;
  const LdsDouble* my_col = &W->term[my_t][my_b][0];
  for (;;) {
    // the pass's task states and frequencies, read once (wave-uniform)
    int st[kTW];
    uint32_t al[kTW];
    double f[kTW][5];
    bool any = false;
#pragma unroll
    for (int t = 0; t < kTW; t++) {
      st[t] = __builtin_amdgcn_readfirstlane(W->state[t]);
      al[t] = (uint32_t)__builtin_amdgcn_readfirstlane((int)W->allowed[t]);
      const LdsDouble* src = (st[t] == kTaskLast) ? W->fl[t] : W->f[t];
#pragma unroll
      for (int b = 0; b < 5; b++) f[t][b] = readfirstlane_d(src[b]);
      any = any || (st[t] != kTaskIdle);
    }
    if (!any) break;
    double acc = 0.0;
    for (int c0 = 0; c0 < n; c0 += kTermCols) {
      const int i = c0 + lane;
#pragma unroll
      for (int t = 0; t < kTW; t++)
        if (st[t] == kTaskLast) likelihood_terms(R, n, c0, W, t, f[t][0], f[t][1], f[t][2], f[t][3], f[t][4], al[t]);
      if (i >= n) {
        // padding of the chunk's last partial group of 16
#pragma unroll
        for (int t = 0; t < kTW; t++) {
          LdsDouble* tm = &W->term[t][0][0];
#pragma unroll
          for (int b = 0; b < 5; b++) tm[b * kTermRow + lane] = -0.0;
        }
      } else {
        double r[6];
        R.get(i, r);
#pragma unroll
        for (int t = 0; t < kTW; t++) {
          if (st[t] != kTaskIter) continue;
          double x[5];
          double s = 0.0;
#pragma unroll
          for (int b = 0; b < 5; b++) {
            x[b] = f[t][b] * r[b];
            if (al[t] & (1u << b)) s += x[b];
          }
          LdsDouble* tm = &W->term[t][0][0];
          double q[5];
          if (s > 0.0) {
            // f_b r_b / s correctly rounded from one reciprocal (the SharedDiv step, with the
            // out-of-range fallback taken per lane only when needed)
            const double y = 1.0 / s;
            bool slow = !(s >= 0x1p-500 && s <= 0x1p500);
#pragma unroll
            for (int b = 0; b < 5; b++) {
              const double q0 = x[b] * y;
              q[b] = fma(fma(-s, q0, x[b]), y, q0);
              slow = slow || !(x[b] == 0.0 || (x[b] >= 0x1p-500 && x[b] <= 0x1p500));
            }
            if (slow) {
#pragma unroll
              for (int b = 0; b < 5; b++) q[b] = x[b] / s;
            }
          } else {
#pragma unroll
            for (int b = 0; b < 5; b++) q[b] = f[t][b];
          }
#pragma unroll
          for (int b = 0; b < 5; b++) tm[b * kTermRow + lane] = (al[t] & (1u << b)) ? q[b] : 0.0;
        }
      }
      wave_sync();
      // lanes of idle tasks add stale columns; their sums are not used
      if (sum_lane) acc = sum_chunk(acc, my_col, min(kTermCols, n - c0));
      wave_sync();
    }
    if (sum_lane) W->sum[my_t][my_b] = acc;
    wave_sync();
    // update, lane 5t+b for allele b of task t
    if (sum_lane) {
      const int mst = W->state[my_t];
      if (mst == kTaskLast) {
        if (my_b == 0) {
          res_out[W->res[my_t]] = acc;
          W->state[my_t] = kTaskIdle;
        }
      } else if (mst == kTaskIter) {
        const double fo = W->f[my_t][my_b];
        const double mine = W->sum[my_t][my_b];
        double fn = fo;
        bool upd = false;
        if (W->mode[my_t] == 1) {
          upd = (W->allowed[my_t] >> my_b) & 1u;
          // one correctly rounded division: a reciprocal hoisted out of the pass loop would be
          // spilled to scratch under this function's register budget
          if (upd) fn = mine / (double)n;
        } else {
          const int v = W->v[my_t];
          const double x = W->x[my_t];
          double others = 0.0;
#pragma unroll
          for (int b = 0; b < 5; b++)
            if (b != v) others += W->sum[my_t][b];
          upd = my_b != v;
          if (upd) fn = (others > 0.0) ? SharedDiv(others)((1.0 - x) * mine) : (1.0 - x) / 4.0;
        }
        W->fl[my_t][my_b] = fo;
        W->f[my_t][my_b] = fn;
        W->delta[my_t][my_b] = upd ? fabs(fn - fo) : 0.0;
      }
    }
    wave_sync();
    if (lane < kTW && W->state[lane] == kTaskIter) {
      double max_delta = 0.0;
#pragma unroll
      for (int b = 0; b < 5; b++) max_delta = fmax(max_delta, W->delta[lane][b]);
      const int it = W->it[lane] + 1;
      W->it[lane] = it;
      if (max_delta < tol || it >= 50) W->state[lane] = kTaskLast;
    }
    wave_sync();
  }
}

__device__ __noinline__ double log_choose(double n, double k) { return lgamma(n + 1) - lgamma(k + 1) - lgamma(n - k + 1); }

// fisher_exact_test_2x2 (stats.cpp:2144-2169): terms across lanes, summed in order by lane 0.
__device__ __noinline__ double fisher_2x2_w(uint32_t a, uint32_t b, uint32_t c, uint32_t d, LdsDouble* buf) {
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
__device__ __noinline__ double ks_less_w(LdsShared* SS, uint32_t nx, uint32_t ny) {
  const int lane = threadIdx.x & 63;
  const LdsU32* hx = SS->qh[0];
  const LdsU32* hy = SS->qh[1];
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
  // m * n < 10000 bounds the lattice: min(m, n) <= 99 cells per diagonal, D <= 10000 boundary bits
  if (min(m, n) > (uint32_t)kKsDiag || D > 10000u) return __longlong_as_double(0x7ff8000000000000LL);
  // boundary(d): d is 0 or the pooled count at the end of a tie group
  const uint32_t nw = D / 32 + 1;
  for (uint32_t k = lane; k < nw; k += 64) SS->ks_bnd[k] = 0;
  wave_sync();
  if (lane == 0) {
    SS->ks_bnd[0] |= 1u;
    uint32_t cum = 0;
#pragma unroll 1
    for (int v = 0; v < 128; v++) {
      const uint32_t g = hx[v] + hy[v];
      if (!g) continue;
      cum += g;
      SS->ks_bnd[cum >> 5] |= 1u << (cum & 31);
    }
  }
  wave_sync();
  // the diagonal index runs over the smaller dimension
  const bool by_i = (m <= n);
  const uint32_t small = by_i ? m : n, large = by_i ? n : m;
  LdsDouble* A = SS->ks_a;   // diagonal d-1
  LdsDouble* B = SS->ks_b;   // diagonal d
  for (uint32_t d = 0; d <= D; d++) {
    const bool bnd = (SS->ks_bnd[d >> 5] >> (d & 31)) & 1u;
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
    LdsDouble* t = A;
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

// Relative entropy KL(fh || f) of two-allele frequencies: the drop of a binomial log likelihood
// from its maximum at fh, per read and in natural-log units.
__device__ double kl_two(double fh, double f) {
  double t = 0.0;
  if (fh > 0.0) t += fh * (log(fh) - log(f));
  if (fh < 1.0) t += (1.0 - fh) * (log1p(-fh) - log1p(-f));
  return t;
}

// Predicted crossing point of the profile likelihood with the target inside [lo, hi] (bound
// `side`: 0 lower, 1 upper), used only to choose which midpoints to evaluate.  Model:
// profile(f) = vmax - A * KL(fh || f) + c, fitted to the interval ends whose values are known and
// interior (two ends: A and c; one end: A; none: A = n / ln 10, the binomial curvature of n reads).
__device__ __noinline__ double predict_crossing(int side, double lo, double hi, double vlo, double vhi, double fh,
                                                double vmax, double target, int n) {
  double xs[2], vs[2];
  int k = 0;
  if (lo != fh && lo > 0.0 && lo < 1.0 && isfinite(vlo)) {
    xs[k] = lo;
    vs[k++] = vlo;
  }
  if (hi != fh && hi > 0.0 && hi < 1.0 && isfinite(vhi)) {
    xs[k] = hi;
    vs[k++] = vhi;
  }
  double A = (double)n / 2.302585092994046, c = 0.0;
  if (k == 2) {
    const double d0 = kl_two(fh, xs[0]), d1 = kl_two(fh, xs[1]);
    if (d1 != d0) {
      A = (vs[0] - vs[1]) / (d1 - d0);
      c = vs[0] + A * d0 - vmax;
    }
  } else if (k == 1) {
    const double d = kl_two(fh, xs[0]);
    if (d > 0.0) A = (vmax - vs[0]) / d;
  }
  const double want = (vmax - target + c) / A;
  if (!(A > 0.0) || !isfinite(want)) return 0.5 * (lo + hi);
  // KL grows away from fh: falls across the lower interval, rises across the upper one.  20 halvings
  // place the crossing within 2^-20 of the interval, far finer than the eight levels a round uses.
  double a = lo, b = hi;
  for (int it = 0; it < 20; it++) {
    const double m = 0.5 * (a + b);
    const bool beyond = kl_two(fh, m) > want;   // m lies outside the crossing (farther from fh)
    if ((side == 0) == beyond) a = m;
    else b = m;
  }
  return 0.5 * (a + b);
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

// kKeysGlobal: the candidate's read list is this block's slice of a device scratch list (columns
// too deep for LDS; the table-row cache then takes all of the dynamic LDS), else it is in LDS.  The
// location is a compile-time choice, so that each instantiation's accesses to the list keep their
// address space (LDS instructions in the usual case); one pointer that may be either would make
// them all generic (flat) accesses.
template <bool kKeysGlobal>
__global__ void __launch_bounds__(kSW * 64) __attribute__((amdgpu_waves_per_eu(3))) k_score_candidates_t(ImParams P) {
  extern __shared__ __align__(16) unsigned char dyn[];
  uint32_t* keys;
  size_t key_lds;
  if constexpr (kKeysGlobal) {
    keys = P.key_scratch + (size_t)blockIdx.x * (size_t)P.key_cap;
    key_lds = 0;
  } else {
    keys = (uint32_t*)dyn;
    key_lds = ((size_t)P.key_cap * sizeof(uint32_t) + 15) & ~(size_t)15;
  }
  double* rows = (double*)(dyn + key_lds);
  __shared__ ScoreShared SS;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  LdsTasks* W = (LdsTasks*)&SS.wt[wave];
  LdsDouble* term = &W->term[0][0][0];
  const int S = P.read_set_max, Q = P.quality_max;
  const double tol = P.precision_decimal;
  const double kDrop = 0.587566;   // kProfileLikelihoodLog10Drop (identify_mutations.cpp:3173)
  const int n_cand = min(*P.n_cand, P.cand_cap);
  // With many more candidates than resident blocks the pass is bound by total work, not by one
  // candidate's chain: the null fit then runs alone first, and only candidates that are emitted go on
  // to the profile likelihoods and bisections (one phase more on their chain).  With few, all of
  // phase 1 runs together.
  const bool two_stage = n_cand > 2 * (int)gridDim.x;
  const int prof_ph = two_stage ? 2 : 1;   // the phase of the profile likelihoods and first path levels
  uint32_t err = 0;
  // Blocks take candidates from the pass's work counter, one at a time: a candidate's cost varies
  // several-fold (emitted ones go on to the profiles and bisections), so a fixed stride over the
  // grid would leave blocks idle while others work through a run of costly candidates.
  __shared__ int s_next;
  for (;;) {
    if (tid == 0) s_next = atomicAdd(P.score_next, 1);
    __syncthreads();
    const int ci = __builtin_amdgcn_readfirstlane(s_next);
    __syncthreads();   // every wavefront has read s_next before the next draw overwrites it
    if (ci >= n_cand) break;
    // every index the candidate leads to is checked before it is used: a column inside the
    // sample, a window inside the index, a read range inside the alignments.  The values are read
    // as wave-uniform scalars, so the checks are uniform branches: a skip decided on per-lane
    // copies of the same value is a divergent branch to the compiler, and around the block's
    // barriers that has paired barriers across wavefronts wrongly (results read stale LDS).
    const int2 cd0 = P.cand[ci];
    const int2 cd = make_int2(__builtin_amdgcn_readfirstlane(cd0.x), __builtin_amdgcn_readfirstlane(cd0.y));
    if (cd.x < 0 || cd.x >= P.n_pos || cd.y < 0) {
      err |= kErrIndex;
      continue;
    }
    const int p = cd.x, slot = cd.y;
    const int64_t w = p / 64;
    if (w >= P.n_windows) {
      err |= kErrIndex;
      continue;
    }
    const int rb = __builtin_amdgcn_readfirstlane(P.win_rb[w]), re = __builtin_amdgcn_readfirstlane(P.win_re[w]);
    if (rb < 0 || re < rb || re > P.n_reads) {
      err |= kErrIndex;
      continue;
    }
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
      if (use && key >= (uint32_t)P.tab_keys) {
        err |= kErrIndex;   // a key outside the likelihood tables: counted, never looked up
        key = 0;
      }
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
    if (tid < 5) SS.obs[tid] = SS.cov[tid][0] + SS.cov[tid][2];
    // ---- table rows of the reads, cached in LDS
    const bool cached = n <= P.rows_cap;
    if (cached)
      for (int i = tid; i < n * 6; i += kSW * 64) rows[i] = P.tab_r[(size_t)keys[i / 6] * 6 + i % 6];
    if (tid == 0) {
      SS.done = 0;
      SS.emitted = 0;
      SS.active[0] = SS.active[1] = 0;
      SS.depth[0] = SS.depth[1] = 0;
      SS.path[0] = SS.path[1] = 0;
      SS.tree_only[0] = SS.tree_only[1] = 0;
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
    // ---- phases: each wavefront starts its tasks of the phase, then runs them to completion
    for (int ph = 0;; ph++) {
      {
        // task slots of this wavefront for this phase, set up by every lane with the same values
        // (identical LDS stores).  Kept out of a lane-0 branch on purpose: at the join of such a
        // branch, ahead of the run_tasks call, the compiler has placed a register spill before the
        // exec-mask restore, so the lanes outside the branch reloaded another value's slot; the
        // build rejects that spill placement (build.py, spill_before_exec_restore).
        int k = 0;
        auto start = [&](int mode, uint32_t allowed, int v, double x, int res) {
          task_start(W, k++, mode, allowed, v, x, res, (const LdsU32*)SS.obs, (const LdsDouble*)SS.full_f, n,
                     (LdsDouble*)SS.res);
        };
        if (ph == 0) {
          if (wave == 0) start(1, 31u, -1, 0.0, kResFull);
        } else if (ph == 1 && two_stage) {
          if (variant != 5 && n > 0 && wave == 0) start(1, 31u & ~(1u << variant), -1, 0.0, kResNull);
        } else if (ph == prof_ph) {
          if (variant != 5 && n > 0) {
            // the first path levels of the bounds (lower bound's first) fill the slots that the
            // four phase-1 tasks leave idle; plmax and p1 end after a pass or two and share a wave
            auto spec_start = [&](int j) {
              const int side = j < SS.spec[0] ? 0 : 1, level = side ? j - SS.spec[0] : j;
              if (level < SS.spec[side]) start(2, 31u, variant, SS.pts[side][level], kResBisect + side * kBisectPts + level);
            };
            if (wave == 0) {
              if (!two_stage) start(1, 31u & ~(1u << variant), -1, 0.0, kResNull);
              spec_start(0);
            } else if (wave == 1) {
              start(2, 31u, variant, SS.full_f[variant], kResPlmax);
              start(2, 31u, variant, 1.0, kResP1);
            } else if (wave == 2) {
              start(2, 31u, variant, 0.0, kResP0);
              spec_start(1);
            } else {
              spec_start(2);
              spec_start(3);
            }
          }
        } else {
          // one bound open: its midpoints on nodes 2w, 2w+1; both open: the lower bound's on
          // wavefronts 0-1 and the upper bound's on wavefronts 2-3, nodes 2(w%2), 2(w%2)+1
          const bool a0 = SS.active[0] != 0, a1 = SS.active[1] != 0;
          if (a0 || a1) {
            const int side = (a0 && a1) ? wave / 2 : (a0 ? 0 : 1);
            const int n0 = (a0 && a1) ? 2 * (wave % 2) : 2 * wave;
            const int npts = SS.path[side] ? SS.depth[side] : (1 << SS.depth[side]) - 1;
            for (int node = n0; node < min(n0 + 2, npts); node++)
              start(2, 31u, variant, SS.pts[side][node], kResBisect + side * kBisectPts + node);
          }
        }
        for (; k < kTW; k++) W->state[k] = kTaskIdle;
      }
      wave_sync();
      run_tasks(RR, n, W, tol, (LdsDouble*)SS.res);
      if (ph == 0 && wave == 0 && lane < 5) SS.full_f[lane] = W->f[0][lane];
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
      if (ph == prof_ph && wave == 1) {
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
        const double ks = (nx > 0 && ny > 0) ? ks_less_w((LdsShared*)&SS, nx, ny) : 1.0;
        const uint32_t mjt = (major < 5) ? SS.cov[major][2] : 0, mjb = (major < 5) ? SS.cov[major][0] : 0;
        const uint32_t mnt = (minor < 5) ? SS.cov[minor][2] : 0, mnb = (minor < 5) ? SS.cov[minor][0] : 0;
        const double fs = fisher_2x2_w(mnt, mnb, mjt, mjb, (LdsDouble*)SS.fisher);
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
        if (tid == 0) {
          // the first levels of each bound's predicted path, evaluated speculatively in phase 1:
          // four for a bound whose interval alone is wide enough to be searched, else two each
          // (the model needs only the ML frequency and n until profile values are known)
          SS.spec[0] = SS.spec[1] = 0;
          if (variant != 5 && n > 0) {
            const double fh = SS.full_f[variant];
            const bool open0 = fh > tol, open1 = (1.0 - fh) > tol;
            SS.spec[0] = open0 ? (open1 ? 2 : 4) : 0;
            SS.spec[1] = open1 ? (open0 ? 2 : 4) : 0;
            for (int side = 0; side < 2; side++) {
              const double lo = side ? fh : 0.0, hi = side ? 1.0 : fh;
              const double fs = predict_crossing(side, lo, hi, 0.0, 0.0, fh, kDrop, 0.0, n);
              double a = lo, b = hi;
              for (int d = 0; d < SS.spec[side]; d++) {
                const double m = 0.5 * (a + b);
                SS.pts[side][d] = m;
                const bool left = (side == 0) ? (fs <= m) : (fs < m);
                if (left) b = m;
                else a = m;
              }
            }
          }
        }
        __syncthreads();
        continue;   // phase 1 reads the derived calls and the speculative midpoints
      }
      if (tid == 0) {
        if (ph == 1) {
          bool poly = false;
          if (variant != 5 && n > 0) poly = ((SS.res[kResFull] - SS.res[kResNull]) - P.log10_ref_length) >= P.polymorphism_cutoff;
          SS.emitted = (cons || poly) ? 1 : 0;
        }
        if (ph == 1 && two_stage) {
          // the profile phase runs only for an emitted entry with a variant allele
          SS.done = (SS.emitted && variant < 5 && n > 0) ? 0 : 1;
        } else if (ph == prof_ph) {
          // write_RA_frequency_bounds (identify_mutations.cpp:3175-3217)
          if (SS.emitted && variant < 5 && n > 0) {
            const double fh = SS.full_f[variant];
            const double target = SS.res[kResPlmax] - kDrop;
            SS.target = target;
            if (SS.res[kResP0] < target) {
              SS.lo[0] = 0.0;
              SS.hi[0] = fh;
              SS.vlo[0] = SS.res[kResP0];
              SS.vhi[0] = SS.res[kResPlmax];
              SS.iter[0] = 0;
              SS.active[0] = (fh - 0.0) > tol ? 1 : 0;
              SS.lower = fh;
            }
            if (SS.res[kResP1] < target) {
              SS.lo[1] = fh;
              SS.hi[1] = 1.0;
              SS.vlo[1] = SS.res[kResPlmax];
              SS.vhi[1] = SS.res[kResP1];
              SS.iter[1] = 0;
              SS.active[1] = (1.0 - fh) > tol ? 1 : 0;
              SS.upper = fh;
            }
          }
          // the speculative levels evaluated with the profiles are walked like a round's path
          for (int side = 0; side < 2; side++) {
            SS.path[side] = 1;
            SS.depth[side] = SS.spec[side];
          }
        }
        if (ph >= prof_ph) {
          // walk each open bound down the evaluated levels: every level is one sequential
          // bisection step (interval halved at 0.5 * (lo + hi), ≤ 40 steps, stop at width ≤ tol)
          for (int side = 0; side < 2; side++) {
            if (!SS.active[side]) continue;
            const double* r = &SS.res[kResBisect + side * kBisectPts];
            const bool path = SS.path[side] != 0;
            double l = SS.lo[side], h = SS.hi[side];
            double vl = SS.vlo[side], vh = SS.vhi[side];
            int i = SS.iter[side];
            int node = 0, advanced = 0;
            for (int level = 0; level < SS.depth[side]; level++) {
              if (level > 0 && !(i < 40 && (h - l) > tol)) break;
              const double m = 0.5 * (l + h);
              // a path's midpoint at this level is the sequential one only while every earlier
              // decision followed the path (off it, the interval and so its midpoint differ)
              if (path && m != SS.pts[side][level]) break;
              // lower bound: a profile >= target moves hi down; upper bound: moves lo up
              const double v = r[node];
              const bool ge = v >= SS.target;
              const bool left = (side == 0) ? ge : !ge;   // the interval kept is [l, m]
              if (left) {
                h = m;
                vh = v;
              } else {
                l = m;
                vl = v;
              }
              i++;
              advanced++;
              node = path ? level + 1 : 2 * node + (left ? 1 : 2);
            }
            if (ph > prof_ph && path && advanced < 3) SS.tree_only[side] = 1;
            SS.lo[side] = l;
            SS.hi[side] = h;
            SS.vlo[side] = vl;
            SS.vhi[side] = vh;
            SS.iter[side] = i;
            SS.active[side] = (i < 40 && (h - l) > tol) ? 1 : 0;
            if (side == 0) SS.lower = h;
            else SS.upper = l;
          }
        }
        // midpoints of the next round, each the midpoint of the interval a sequential bisection
        // would hold there: along the predicted path, or breadth-first over a full tree
        const bool both = SS.active[0] && SS.active[1];
        for (int side = 0; side < 2 && ph >= prof_ph; side++) {
          if (!SS.active[side]) {
            SS.depth[side] = 0;
            continue;
          }
          const double lo = SS.lo[side], hi = SS.hi[side];
          double* pt = SS.pts[side];
          if (!SS.tree_only[side]) {
            const int depth = both ? kBisectPts / 2 : kBisectPts;
            const double fs = predict_crossing(side, lo, hi, SS.vlo[side], SS.vhi[side], SS.full_f[variant],
                                               SS.res[kResPlmax], SS.target, n);
            double a = lo, b = hi;
            for (int d = 0; d < depth; d++) {
              const double m = 0.5 * (a + b);
              pt[d] = m;
              const bool left = (side == 0) ? (fs <= m) : (fs < m);
              if (left) b = m;
              else a = m;
            }
            SS.path[side] = 1;
            SS.depth[side] = depth;
            continue;
          }
          SS.path[side] = 0;
          SS.depth[side] = both ? 2 : 3;
          const double m0 = 0.5 * (lo + hi);
          const double m1 = 0.5 * (lo + m0), m2 = 0.5 * (m0 + hi);
          pt[0] = m0;
          pt[1] = m1;
          pt[2] = m2;
          pt[3] = 0.5 * (lo + m1);
          pt[4] = 0.5 * (m1 + m0);
          pt[5] = 0.5 * (m0 + m2);
          pt[6] = 0.5 * (m2 + hi);
        }
        if (ph >= prof_ph) SS.done = (SS.active[0] || SS.active[1]) ? 0 : 1;
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

template __global__ void k_score_candidates_t<false>(ImParams P);
template __global__ void k_score_candidates_t<true>(ImParams P);

}  // namespace bsq
