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
