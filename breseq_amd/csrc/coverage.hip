// coverage.hip -- BAM2COV coverage tables on CDNA4 (coverage_output::pileup_callback,
// coverage_output.cpp:307-503): per reference column of a region, the unique and redundant
// coverage of each strand, the raw count of redundant reads and the unique reads that begin there.
//
// One wavefront owns 64 consecutive columns (lane = column) and walks the alignments overlapping
// them in file order, so the FP64 sums of 1/X1 are formed in the reference's order.  Deleted and
// reference-skipped positions have no entry (the reference skips is_del entries).
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace bsq {

__global__ void __launch_bounds__(256) k_coverage_columns(CovTabParams C) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t w0 = C.g0 / 64, w1 = C.g1 / 64;
  const int64_t n = (int64_t)C.g1 - C.g0 + 1;
  uint32_t err = 0;
  int32_t last_end = -1;
  for (int64_t w = w0 + wave; w <= w1; w += nwaves) {
    const int p = (int)(w * 64 + lane);
    const bool in = p >= C.g0 && p <= C.g1;
    const int rb = C.win_rb[w], re = C.win_re[w];
    uint32_t u0 = 0, u1 = 0, x0 = 0, x1 = 0, b0 = 0, b1 = 0;
    double r0 = 0.0, r1 = 0.0;
    for (int r = rb; r < re; r++) {
      const DRead R = load_read(C.reads, r);
      if (R.gpos <= C.g1 && R.gend > C.g0) last_end = max(last_end, R.gend);
      if (C.rg_filter >= 0) {
        // read_group_index_map::index (alignment.h:545-559): one group or none, no tag, or an ID
        // not in the header all resolve to group 0
        const int rg = C.rg[r];
        const int g = (C.n_rg <= 1 || rg == 255 || rg >= C.n_rg) ? 0 : rg;
        if (g != C.rg_filter) continue;
      }
      int qpos = 0, indel = 0;
      bool is_del = false;
      const bool cover = entry_at(R, C.cigar, p, qpos, indel, is_del, err);
      if (!in || !cover || is_del) continue;
      const bool rev = (R.flags & kReadReversed) != 0;
      if (R.x1 == 1) {
        // the read's first base in read direction: query position 1 (top) or read length (bottom)
        const bool first = rev ? (qpos == (int)R.l_qseq - 1) : (qpos == 0);
        if (rev) {
          u1++;
          b1 += first ? 1u : 0u;
        } else {
          u0++;
          b0 += first ? 1u : 0u;
        }
      } else {
        const double inv = 1.0 / (double)R.x1;
        if (rev) {
          x1++;
          r1 += inv;
        } else {
          x0++;
          r0 += inv;
        }
      }
    }
    if (in) {
      const int64_t i = p - C.g0;
      C.uniq[i] = u0;
      C.uniq[n + i] = u1;
      C.red[i] = r0;
      C.red[n + i] = r1;
      C.raw[i] = x0;
      C.raw[n + i] = x1;
      C.begin[i] = b0;
      C.begin[n + i] = b1;
    }
  }
  if (lane == 0 && last_end >= 0) atomicMax(C.last_end, last_end);   // the same in every lane
  if (err) atomicOr(C.err, err);
}

}  // namespace bsq
