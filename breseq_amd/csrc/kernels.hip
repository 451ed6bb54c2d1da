// kernels.hip -- CDNA4 (gfx950) kernels of pass 2 that follow the column kernel.
//
//  k_insert_slots     enumerates insert slots >= 1 of columns whose unique reads carry insertions
//  k_runs             MC and UN interval detection (identify_mutations.cpp:2262-2344, 2972-3007)
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
    if (p < 0 || p >= P.n_pos || w >= P.n_windows) {
      err |= kErrIndex;
      continue;
    }
    const int rb = P.win_rb[w], re = P.win_re[w];
    if (rb < 0 || re < rb || re > P.n_reads) {
      err |= kErrIndex;
      continue;
    }
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
/* MC / UN runs                                                                                */

// Runs are found by the columns that open them and walked by a whole wavefront, 256 columns a step
// (four coalesced loads per lane), so a run of L columns costs L / 256 dependent steps:
//   MC: maximal runs with unique coverage <= the propagation cutoff; the seed is reached if any
//       column of the run, or the column closing it, has total coverage <= the seed cutoff
//       (identify_mutations.cpp:2270-2281); the redundant sub-range follows :2326-2343.
//   UN: maximal runs of columns without a predicted base (:2972-3007).
// The reference's walk over an MC run keeps red_start at the column before the first column without
// redundant coverage and red_end at the column after the last one, each -1 when the run starts
// (ends) there; a run with redundant coverage throughout has red_start at its last column.  The
// walk below gets both from the first and last such column of the run.
constexpr int kRunSpan = 4 * WAVE;

__device__ __forceinline__ int first_in(uint64_t m) { return __builtin_ctzll(m); }
__device__ __forceinline__ int last_in(uint64_t m) { return 63 - __builtin_clzll(m); }

__device__ void walk_mc_run(const RunParams& P, int s, int lane) {
  const TargetInfo T = P.targets[P.pos_target[s]];
  const int first = T.off, last = T.off + T.len - 1;
  bool seed = false;
  int fz = -1, lz = -1, e = last;
  for (int q0 = s;; q0 += kRunSpan) {
    bool in[4];
    uint8_t fl[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int q = q0 + j * WAVE + lane;
      const bool valid = q <= last;
      const int32_t u = valid ? P.col_unique[q] : 0;
      fl[j] = valid ? P.col_flags[q] : 0;
      in[j] = valid && (double)u <= T.prop_cutoff;
    }
    int end = kRunSpan;   // first index of this step outside the run
#pragma unroll
    for (int j = 3; j >= 0; j--) {
      const uint64_t out = __ballot(!in[j]);
      if (out) end = j * WAVE + first_in(out);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const bool member = j * WAVE + lane < end;
      seed |= member && (fl[j] & kPosSeed);
      const uint64_t nr = __ballot(member && !(fl[j] & kPosRedundant));
      if (nr) {
        if (fz < 0) fz = q0 + j * WAVE + first_in(nr);
        lz = q0 + j * WAVE + last_in(nr);
      }
    }
    if (end < kRunSpan) {
      e = q0 + end - 1;
      break;
    }
  }
  const bool any_seed = __any(seed);
  if (lane != 0) return;
  RunRec R;
  R.kind = 0;
  R.s = s;
  R.e = e;
  R.seed = (any_seed || (e < last && (P.col_flags[e + 1] & kPosSeed))) ? 1 : 0;
  R.red_start = fz < 0 ? e : (fz > s ? fz - 1 : -1);
  R.red_end = (fz >= 0 && lz < e) ? lz + 1 : -1;
  R.lo = (s > first) ? P.col_unique[s - 1] : -1;
  R.li = P.col_unique[s];
  R.ri = P.col_unique[e];
  R.ro = (e < last) ? P.col_unique[e + 1] : -1;
  const int k = atomicAdd(P.n_runs, 1);
  if (k < P.run_cap) P.runs[k] = R;
  else atomicOr(P.err, kErrCapacity);
}

__device__ void walk_un_run(const RunParams& P, int s, int lane) {
  const TargetInfo T = P.targets[P.pos_target[s]];
  const int last = T.off + T.len - 1;
  int e = last;
  for (int q0 = s;; q0 += kRunSpan) {
    int end = kRunSpan;
#pragma unroll
    for (int j = 3; j >= 0; j--) {
      const int q = q0 + j * WAVE + lane;
      const bool in = q <= last && !(P.col_flags[q] & kPosPredicted);
      const uint64_t out = __ballot(!in);
      if (out) end = j * WAVE + first_in(out);
    }
    if (end < kRunSpan) {
      e = q0 + end - 1;
      break;
    }
  }
  if (lane != 0) return;
  RunRec R;
  R.kind = 1;
  R.s = s;
  R.e = e;
  R.seed = 0;
  R.red_start = R.red_end = -1;
  R.lo = R.li = R.ri = R.ro = 0;
  const int k = atomicAdd(P.n_runs, 1);
  if (k < P.run_cap) P.runs[k] = R;
  else atomicOr(P.err, kErrCapacity);
}

// One lane per column marks the runs that open there; the wavefront then walks each of its marked
// runs in column order.  Every lane stays to the end: the walks need all 64.
__global__ void __launch_bounds__(256) k_runs(RunParams P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (WAVE - 1);
  bool mc_open = false, un_open = false;
  if (p < P.n_pos) {
    const int t = P.pos_target[p];
    if (t >= 0) {
      const TargetInfo T = P.targets[t];
      if (T.scored && T.prop_cutoff >= 0.0) {
        const bool at_first = p == T.off;
        if (P.do_mc && (double)P.col_unique[p] <= T.prop_cutoff)
          mc_open = at_first || !((double)P.col_unique[p - 1] <= T.prop_cutoff);
        if (!(P.col_flags[p] & kPosPredicted)) un_open = at_first || (P.col_flags[p - 1] & kPosPredicted);
      }
    }
  }
  const int base = uniform_i(p - lane);
  for (uint64_t m = __ballot(mc_open); m; m &= m - 1) walk_mc_run(P, base + first_in(m), lane);
  for (uint64_t m = __ballot(un_open); m; m &= m - 1) walk_un_run(P, base + first_in(m), lane);
}

}  // namespace bsq
