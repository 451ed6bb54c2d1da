// coverage.hip -- BAM2COV coverage tables on CDNA4 (coverage_output::pileup_callback,
// coverage_output.cpp:307-503): per printed reference column, the unique and redundant coverage of
// each strand, the raw count of redundant reads and the unique reads that begin there.
//
// Only the columns that get a row are computed, packed as rows j = 0..n-1 at columns h0 + j*step.
// The FP64 sums of 1/X1 are formed in read (file) order, as the reference's pileup visits them.
// Deleted and reference-skipped positions count nothing (the reference skips is_del entries).
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace bsq {

// One read at column p: whether it counts, and how.
struct CovHit {
  bool hit;     // covers p with a base (not a deletion) and passes the read-group filter
  bool rev;     // bottom strand
  bool uniq;    // X1 == 1
  bool first;   // its first base in read direction is at p
  double inv;   // 1 / X1
};

__device__ __forceinline__ CovHit cov_hit(const CovTabParams& C, int r, int p, uint32_t& err) {
  CovHit h{false, false, false, false, 0.0};
  const DRead R = load_read(C.reads, r);
  if (C.rg_filter >= 0) {
    // read_group_index_map::index (alignment.h:545-559): one group or none, no tag, or an ID
    // not in the header all resolve to group 0
    const int rg = C.rg[r];
    const int g = (C.n_rg <= 1 || rg == 255 || rg >= C.n_rg) ? 0 : rg;
    if (g != C.rg_filter) return h;
  }
  int qpos = 0, indel = 0;
  bool is_del = false;
  if (!entry_at(R, C.cigar, p, qpos, indel, is_del, err) || is_del) return h;
  h.hit = true;
  h.rev = (R.flags & kReadReversed) != 0;
  h.uniq = R.x1 == 1;
  // query position 1 on the top strand, the read length on the bottom strand
  h.first = h.rev ? (qpos == (int)R.l_qseq - 1) : (qpos == 0);
  h.inv = 1.0 / (double)R.x1;
  return h;
}

__device__ __forceinline__ void put_row(const CovTabParams& C, int64_t j, uint32_t u0, uint32_t u1, double r0,
                                        double r1, uint32_t x0, uint32_t x1, uint32_t b0, uint32_t b1) {
  C.uniq[j] = u0;
  C.uniq[C.n + j] = u1;
  C.red[j] = r0;
  C.red[C.n + j] = r1;
  C.raw[j] = x0;
  C.raw[C.n + j] = x1;
  C.begin[j] = b0;
  C.begin[C.n + j] = b1;
}

// step == 1: wavefront w owns columns [64w, 64w + 64) and walks the reads overlapping them in file
// order; every lane tests its own column, so one read record feeds 64 columns.
__global__ void __launch_bounds__(256) k_coverage_columns(CovTabParams C) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t last = (int64_t)C.h0 + C.n - 1;
  const int64_t w0 = C.h0 / 64, w1 = last / 64;
  uint32_t err = 0;
  for (int64_t w = w0 + wave; w <= w1; w += nwaves) {
    const int p = (int)(w * 64 + lane);
    const bool in = p >= C.h0 && p <= last;
    const int rb = C.win_rb[w], re = C.win_re[w];
    uint32_t u0 = 0, u1 = 0, x0 = 0, x1 = 0, b0 = 0, b1 = 0;
    double r0 = 0.0, r1 = 0.0;
    for (int r = rb; r < re; r++) {
      const CovHit h = cov_hit(C, r, p, err);
      if (!in || !h.hit) continue;
      if (h.uniq) {
        if (h.rev) {
          u1++;
          b1 += h.first ? 1u : 0u;
        } else {
          u0++;
          b0 += h.first ? 1u : 0u;
        }
      } else if (h.rev) {
        x1++;
        r1 += h.inv;
      } else {
        x0++;
        r0 += h.inv;
      }
    }
    if (in) put_row(C, p - C.h0, u0, u1, r0, r1, x0, x1, b0, b1);
  }
  if (err) atomicOr(C.err, err);
}

// step > 1: the rows are far apart, so wavefront j takes row j alone and its lanes take the reads
// of the row's window 64 at a time.  The integer counts are reduced across lanes at the end; the
// redundant sums take the chunk's hits lane by lane, which is read order.
__global__ void __launch_bounds__(256) k_coverage_rows(CovTabParams C) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  uint32_t err = 0;
  for (int64_t j = wave; j < C.n; j += nwaves) {
    const int p = (int)(C.h0 + j * C.step);
    const int w = p >> 6;
    const int rb = C.win_rb[w], re = C.win_re[w];
    uint32_t u0 = 0, u1 = 0, x0 = 0, x1 = 0, b0 = 0, b1 = 0;
    double r0 = 0.0, r1 = 0.0;
    for (int base = rb; base < re; base += 64) {
      const int r = base + lane;
      CovHit h{false, false, false, false, 0.0};
      if (r < re) h = cov_hit(C, r, p, err);
      if (h.hit && h.uniq) {
        if (h.rev) {
          u1++;
          b1 += h.first ? 1u : 0u;
        } else {
          u0++;
          b0 += h.first ? 1u : 0u;
        }
      }
      if (h.hit && !h.uniq) {
        if (h.rev) x1++;
        else x0++;
      }
      unsigned long long m0 = __ballot(h.hit && !h.uniq && !h.rev);
      unsigned long long m1 = __ballot(h.hit && !h.uniq && h.rev);
      while (m0) {
        const int l = __builtin_ctzll(m0);
        m0 &= m0 - 1;
        r0 += __shfl(h.inv, l, 64);
      }
      while (m1) {
        const int l = __builtin_ctzll(m1);
        m1 &= m1 - 1;
        r1 += __shfl(h.inv, l, 64);
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      u0 += __shfl_xor(u0, o, 64);
      u1 += __shfl_xor(u1, o, 64);
      x0 += __shfl_xor(x0, o, 64);
      x1 += __shfl_xor(x1, o, 64);
      b0 += __shfl_xor(b0, o, 64);
      b1 += __shfl_xor(b1, o, 64);
    }
    if (lane == 0) put_row(C, j, u0, u1, r0, r1, x0, x1, b0, b1);
  }
  if (err) atomicOr(C.err, err);
}

}  // namespace bsq
