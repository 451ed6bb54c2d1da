// error_count.hip -- pass 1 of the pileup path on CDNA4 (gfx950): the covariate count table and the
// unique-only coverage histogram (error_count_pileup::pileup_callback, error_count.cpp:125-199, with
// cErrorTable::count_alignment_position, :854-986).
//
// Both outputs are integer sums, so they do not depend on the order in which pileup entries are
// visited.  Pass 1 therefore walks each alignment once (read-parallel) instead of walking columns:
//
//   k_ec_reads       32 consecutive alignments per work item.  First the item's counted
//                    single-run alignments of at most 168 bases, three to a wavefront (21 lanes
//                    each, eight query positions per lane); then its other alignments one per
//                    wavefront, each lane four consecutive query positions of an M/=/X run.  Bases
//                    and qualities are dword loads from 8-byte aligned read starts, reference codes
//                    aligned dword loads.  Covariates go to an LDS-privatized count table.  The
//                    first lane of an alignment adds +1/-1 at the ends of each aligned run to a
//                    per-column difference array (unique or redundant).
//   k_cov_tile_sums  phase 1 of the prefix scan of the difference arrays (per 4096-column tile)
//   k_cov_tile_scan  phase 2: exclusive scan of the tile sums (one block)
//   k_cov_hist       phase 3: per-column unique and redundant coverage -> histogram of the unique
//                    coverage of columns without redundant reads (LDS-privatized)
//
// Deleted reference bases (is_del entries) never count, so only M/=/X runs contribute to either
// output; the entry of the last base of a run carries the run's indel (next CIGAR operations), as
// htslib's resolve_cigar2 reports it.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace bsq {

// Copies each alignment's bases and qualities to its 8-byte aligned offset (DRead::seq_off) in the
// device layout; one wavefront per alignment.
__global__ void k_repack_reads(const DRead* reads, const int64_t* raw_off, int64_t n_reads, const uint8_t* src_b,
                               const uint8_t* src_q, uint8_t* dst_b, uint8_t* dst_q) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < n_reads; r += nw) {
    const DRead R = reads[r];
    const int64_t s = raw_off[r];
    for (int j = lane; j < (int)R.l_qseq; j += 64) {
      dst_b[R.seq_off + j] = src_b[s + j];
      dst_q[R.seq_off + j] = src_q[s + j];
    }
  }
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// BAM 4-bit base code -> A0 C1 G2 T3; every other code -> 15.
__device__ __forceinline__ uint32_t base_idx(uint32_t b) {
  const uint32_t i = (uint32_t)(0xFFFFFFF3FFF2F10FULL >> ((b & 15u) * 4u)) & 15u;
  return b > 15u ? 15u : i;
}

// Complement of a base index when rv == 3 (A<->T, C<->G); other codes are unchanged.
__device__ __forceinline__ uint32_t comp_if(uint32_t i, uint32_t rv) { return i < 4u ? (i ^ rv) : i; }

typedef __attribute__((address_space(4))) const uint32_t ConstU32;
typedef __attribute__((address_space(4))) const int32_t ConstI32;

// Indel carried by the entry of the last base of the reference-consuming operation at index k
// (htslib resolve_cigar2 as breseq reads it; same rule as entry_at).
__device__ __forceinline__ int run_end_indel(const ConstU32* c, int k, int n) {
  if (k + 1 >= n) return 0;
  const uint32_t op2 = c[k + 1] & 15u;
  int indel = 0;
  if (op2 == OP_D) {
    indel = -(int)(c[k + 1] >> 4);
    for (int kk = k + 2; kk < n; ++kk) {
      if ((c[kk] & 15u) != OP_D) break;
      indel -= (int)(c[kk] >> 4);
    }
  } else if (op2 == OP_I) {
    indel = (int)(c[k + 1] >> 4);
    for (int kk = k + 2; kk < n; ++kk) {
      const uint32_t o3 = c[kk] & 15u;
      if (o3 == OP_I) indel += (int)(c[kk] >> 4);
      else if (o3 != OP_P) break;
    }
  } else if (op2 == OP_P && k + 2 < n) {
    int l3 = 0;
    for (int kk = k + 2; kk < n; ++kk) {
      const uint32_t o3 = c[kk] & 15u;
      if (o3 == OP_I) l3 += (int)(c[kk] >> 4);
      else if (is_ref_op(o3)) break;
    }
    if (l3 > 0) indel = l3;
  }
  return indel;
}

__device__ __forceinline__ uint32_t byte_of(uint32_t w, int k) { return (w >> (8 * k)) & 0xFFu; }

// Bin of (read set rs, reference class ri, observed class oi, quality q) in the count table,
// rs + S*(ri + 5*oi + 25*q), as full-rate 24-bit multiply-adds: every operand is below 2^16
// (classes below 16, qualities below 256, S at most 256).
__device__ __forceinline__ uint32_t ec_bin(uint32_t rs, uint32_t S, uint32_t ri, uint32_t oi, uint32_t q) {
  return __umul24(__umul24(q, 25u) + __umul24(oi, 5u) + ri, S) + rs;
}

// Strand-folded base indices as 4-bit lookup tables: read bases by BAM code (A1 C2 G4 T8), reference
// bases by kRef* code (A0 C1 G2 T3 '.'4 N5).  The reverse-strand tables complement A<->T, C<->G, as
// count_alignment_position does before indexing (error_count.cpp:887-903).  15 = not a base.
constexpr uint64_t kReadLutFwd = 0xFFFFFFF3FFF2F10FULL;
constexpr uint64_t kReadLutRev = 0xFFFFFFF0FFF1F23FULL;
constexpr uint32_t kRefLutFwd = 0xFF543210u;
constexpr uint32_t kRefLutRev = 0xFF540123u;

// count_alignment_position (error_count.cpp:854-986) for the aligned bases [x, x+len) <-> query
// [y, y+len) of one unique read; the last base carries `last_indel`.  Each lane takes four
// consecutive query positions of a 4-aligned group: reads start at 8-byte aligned offsets, so the
// bases and qualities of the group and of the next four positions are two dword loads each; the
// five reference codes the group needs come from two aligned dwords of ref_idx.  Per base: the
// base/reference observation (1) and the '..' pair observation (2) are one LDS atomic each; the
// one-base deletion (3) and insertion (4) observations can only occur at the run's last base and
// are handled there.
template <bool kRev>
__device__ __forceinline__ void ec_count_run(const EcParams& P, const uint8_t* __restrict__ sq,
                                             const uint8_t* __restrict__ ql, const DRead& R, int x, int y, int len,
                                             int last_indel, uint32_t* table, uint32_t& err) {
  const uint32_t S = (uint32_t)P.read_set_max, Q = (uint32_t)P.quality_max;
  const uint64_t rlut = kRev ? kReadLutRev : kReadLutFwd;
  const uint32_t flut = kRev ? kRefLutRev : kRefLutFwd;
  const uint32_t rs = R.rfile;
  const uint32_t pair_base = rs + 24u * S;   // ref '.', obs '.'
  const int lane = threadIdx.x & 63;
  const int yend = y + len;
  const int last = yend - 1;
  const int qe0 = (int)R.qe0;
  // pair observations end at q_end, and at the last base when a deletion or insertion follows it
  const int pair_end = (last_indel != 0) ? min(qe0, last) : qe0;
  for (int q0 = (y & ~3) + 4 * lane; q0 < yend; q0 += 256) {
    const uint32_t bw0 = *(const uint32_t*)(sq + q0), bw1 = *(const uint32_t*)(sq + q0 + 4);
    const uint32_t qw0 = *(const uint32_t*)(ql + q0), qw1 = *(const uint32_t*)(ql + q0 + 4);
    const int c0 = x + (q0 - y);   // column of query position q0 (>= x - 3; ref_idx is front-padded)
    const uint32_t* rp = (const uint32_t*)(P.ref_idx + (c0 & ~3));
    const uint32_t rw0 = rp[0], rw1 = rp[1];
    const int sh = c0 & 3;
    const uint32_t rlo = __builtin_amdgcn_alignbyte(rw1, rw0, sh);   // codes of columns c0 .. c0+3
    const uint32_t r4 = (rw1 >> (8 * sh)) & 0xFFu;                   // code of column c0+4
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int qpos = q0 + k;
      const bool in = qpos >= y && qpos <= last;
      const uint32_t b = byte_of(bw0, k), q = byte_of(qw0, k);
      const uint32_t r0 = byte_of(rlo, k);
      // (1) base against reference (:877-903)
      const uint32_t oi = (b > 15u) ? 15u : (uint32_t)(rlut >> (4u * b)) & 15u;
      const uint32_t ri = (flut >> (4u * r0)) & 15u;
      const bool v1 = in && b != 15u && r0 != kRefN;
      const bool good1 = v1 && oi < 4u && ri < 4u && q < Q;
      if (v1 && !good1) err |= (oi > 3u) ? kErrBaseCode : (ri > 3u) ? kErrRefChar : kErrQuality;
      if (good1) atomicAdd(&table[ec_bin(rs, S, ri, oi, q)], 1u);
      // (2) '..' pair (:909-928): the base after qpos in read direction (qpos+1 forward, qpos
      // itself reverse) against the reference after it in reference direction
      const uint32_t b2 = kRev ? b : ((k < 3) ? byte_of(bw0, k + 1) : byte_of(bw1, 0));
      const uint32_t q2 = kRev ? q : ((k < 3) ? byte_of(qw0, k + 1) : byte_of(qw1, 0));
      const uint32_t rp2 = kRev ? r0 : ((k < 3) ? byte_of(rlo, k + 1) : r4);
      const bool v2 = in && qpos < pair_end && b2 != 15u && rp2 != kRefN;
      if (v2 && q2 >= Q) err |= kErrQuality;
      if (v2 && q2 < Q) atomicAdd(&table[__umul24(q2, 25u * S) + pair_base], 1u);
    }
  }
  if (last_indel != -1 && last_indel != 1) return;
  // (3) one-base deletion (:934-956) / (4) one-base insertion (:963-982) after the last base
  if (lane != 0) return;
  const int qpos = last, p = x + len - 1;
  const uint32_t b = sq[qpos], q = ql[qpos];
  const uint32_t r1 = P.ref_idx[p + 1];
  if (last_indel == -1) {
    if (!kRev && qpos + 1 >= (int)R.l_qseq) {
      err |= kErrQueryIndex;
      return;
    }
    const uint32_t b2 = kRev ? b : (uint32_t)sq[qpos + 1];
    const uint32_t q2 = kRev ? q : (uint32_t)ql[qpos + 1];
    if (b2 == 15u || r1 == kRefN) return;
    const uint32_t rd = (flut >> (4u * r1)) & 15u;
    if (rd > 4u) err |= kErrRefChar;
    else if (q2 >= Q) err |= kErrQuality;
    else atomicAdd(&table[rs + S * rd + 20u * S + 25u * S * q2], 1u);
  } else {
    const int mq = qpos + 1;
    if (mq >= (int)R.l_qseq) {
      err |= kErrQueryIndex;
      return;
    }
    if (mq > qe0 || mq < (int)R.qs0) return;
    const uint32_t b3 = sq[mq];
    if (b3 == 15u) return;
    const uint32_t oi = (b3 > 15u) ? 15u : (uint32_t)(rlut >> (4u * b3)) & 15u;
    const uint32_t q3 = ql[mq];
    if (oi > 3u) err |= kErrBaseCode;
    else if (q3 >= Q) err |= kErrQuality;
    else atomicAdd(&table[rs + 4u * S + 5u * S * oi + 25u * S * q3], 1u);
  }
}

// Observations (1) and (2) of ec_count_run for one lane's four query positions [q0, q0 + 4) of an
// alignment whose M/=/X run is [x, x + len) <-> [y, y + len) and whose last base carries no indel,
// with the strand a run-time value (the three alignments of a shared wavefront differ in strand).
__device__ __forceinline__ void ec_group4(const EcParams& P, const uint8_t* __restrict__ sq,
                                          const uint8_t* __restrict__ ql, uint32_t rs, bool rev, int x, int y,
                                          int last, int pair_end, int q0, uint32_t* table, uint32_t& err) {
  const uint32_t S = (uint32_t)P.read_set_max, Q = (uint32_t)P.quality_max;
  const uint64_t rlut = rev ? kReadLutRev : kReadLutFwd;
  const uint32_t flut = rev ? kRefLutRev : kRefLutFwd;
  const uint32_t pair_base = rs + 24u * S;   // ref '.', obs '.'
  const uint32_t bw0 = *(const uint32_t*)(sq + q0), bw1 = *(const uint32_t*)(sq + q0 + 4);
  const uint32_t qw0 = *(const uint32_t*)(ql + q0), qw1 = *(const uint32_t*)(ql + q0 + 4);
  const int c0 = x + (q0 - y);   // column of query position q0 (>= x - 3; ref_idx is front-padded)
  const uint32_t* rp = (const uint32_t*)(P.ref_idx + (c0 & ~3));
  const uint32_t rw0 = rp[0], rw1 = rp[1];
  const int sh = c0 & 3;
  const uint32_t rlo = __builtin_amdgcn_alignbyte(rw1, rw0, sh);   // codes of columns c0 .. c0+3
  const uint32_t r4 = (rw1 >> (8 * sh)) & 0xFFu;                   // code of column c0+4
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int qpos = q0 + k;
    const bool in = qpos >= y && qpos <= last;
    const uint32_t b = byte_of(bw0, k), q = byte_of(qw0, k);
    const uint32_t r0 = byte_of(rlo, k);
    const uint32_t oi = (b > 15u) ? 15u : (uint32_t)(rlut >> (4u * b)) & 15u;
    const uint32_t ri = (flut >> (4u * r0)) & 15u;
    const bool v1 = in && b != 15u && r0 != kRefN;
    const bool good1 = v1 && oi < 4u && ri < 4u && q < Q;
    if (v1 && !good1) err |= (oi > 3u) ? kErrBaseCode : (ri > 3u) ? kErrRefChar : kErrQuality;
    if (good1) atomicAdd(&table[ec_bin(rs, S, ri, oi, q)], 1u);
    const uint32_t b2 = rev ? b : ((k < 3) ? byte_of(bw0, k + 1) : byte_of(bw1, 0));
    const uint32_t q2 = rev ? q : ((k < 3) ? byte_of(qw0, k + 1) : byte_of(qw1, 0));
    const uint32_t rp2 = rev ? r0 : ((k < 3) ? byte_of(rlo, k + 1) : r4);
    const bool v2 = in && qpos < pair_end && b2 != 15u && rp2 != kRefN;
    if (v2 && q2 >= Q) err |= kErrQuality;
    if (v2 && q2 < Q) atomicAdd(&table[__umul24(q2, 25u * S) + pair_base], 1u);
  }
}

constexpr int kEcGroup = 21;   // lanes per alignment when three share a wavefront (8 positions each)

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) k_ec_reads(EcParams P) {
  extern __shared__ uint32_t table[];
  for (int i = threadIdx.x; i < P.n_bins; i += blockDim.x) table[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t gw = ((int64_t)xcd_major_block() * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const ConstI32* pos_target = (const ConstI32*)P.pos_target;
  const ConstI32* scored = (const ConstI32*)P.target_scored;
  int32_t* diff_u = P.cov_diff;
  int32_t* diff_r = P.cov_diff + P.n_pos + 1;
  uint32_t err = 0;
  // one alignment at a time, wave-uniform (scalar loads; any CIGAR, redundant alignments)
  auto one = [&](int64_t r) {
    const DRead R = load_read(P.reads, (int)r);
    const int t = pos_target[R.gpos];
    if (t < 0 || !scored[t]) return;
    const bool red = R.x1 > 1;
    int32_t* diff = red ? diff_r : diff_u;
    bool count = !red && P.do_errors;
    if (count && (int)R.rfile >= P.read_set_max) {   // covariate read_set out of range: ASSERT in the reference
      err |= kErrReadSet;
      count = false;
    }
    const uint8_t* sq = P.bases + R.seq_off;
    const uint8_t* ql = P.quals + R.seq_off;
    if (R.flags & kReadSimple) {
      if (lane == 0) {
        atomicAdd(&diff[R.gpos], 1);
        atomicAdd(&diff[R.gend], -1);
      }
      if (count) {
        if (R.flags & kReadReversed) ec_count_run<true>(P, sq, ql, R, R.gpos, R.y0, R.gend - R.gpos, 0, table, err);
        else ec_count_run<false>(P, sq, ql, R, R.gpos, R.y0, R.gend - R.gpos, 0, table, err);
      }
      return;
    }
    const ConstU32* c = (const ConstU32*)(P.cigar + R.cig_off);
    const int n = R.n_cigar;
    int x = R.gpos, y = 0;
    for (int k = 0; k < n; k++) {
      const uint32_t w = c[k];
      const uint32_t op = w & 15u;
      const int l = (int)(w >> 4);
      if (op == OP_M || op == OP_EQ || op == OP_X) {
        if (l > 0) {
          if (lane == 0) {
            atomicAdd(&diff[x], 1);
            atomicAdd(&diff[x + l], -1);
          }
          if (count) {
            const int li = run_end_indel(c, k, n);
            if (R.flags & kReadReversed) ec_count_run<true>(P, sq, ql, R, x, y, l, li, table, err);
            else ec_count_run<false>(P, sq, ql, R, x, y, l, li, table, err);
          }
        }
        x += l;
        y += l;
      } else if (op == OP_I || op == OP_S) {
        y += l;
      } else if (op == OP_D || op == OP_N) {
        x += l;
      }
    }
    if (x != R.gend) err |= kErrCigar;
  };
  const int grp = lane / kEcGroup, gl = lane - kEcGroup * grp;
  for (int64_t c0 = gw * kEcChunk; c0 < P.n_reads; c0 += nw * kEcChunk) {
    const int n = (int)min<int64_t>(kEcChunk, P.n_reads - c0);
    // Counted single-run alignments of at most 168 bases go three to a wavefront (lanes
    // 21g .. 21g+20 take one, eight query positions per lane), which fills the lanes that one
    // 150-base alignment leaves idle; the chunk's other alignments then go one at a time.  The
    // counts are integers, so the order does not matter.
    bool fit = false;
    if (lane < n) {
      const DRead R = P.reads[c0 + lane];
      const int t = pos_target[R.gpos];
      fit = t >= 0 && scored[t] && (R.flags & kReadSimple) && R.x1 <= 1 && P.do_errors &&
            (int)R.rfile < P.read_set_max && (R.y0 & 3) + (R.gend - R.gpos) <= 8 * kEcGroup;
    }
    unsigned long long fm = __ballot(fit);
    unsigned long long sm = ((n == 64) ? ~0ull : ((1ull << n) - 1ull)) & ~fm;
    while (fm) {
      const int ja = __builtin_ctzll(fm);
      fm &= fm - 1ull;
      const int jb = fm ? __builtin_ctzll(fm) : -1;
      if (jb >= 0) fm &= fm - 1ull;
      const int jc = fm ? __builtin_ctzll(fm) : -1;
      if (jc >= 0) fm &= fm - 1ull;
      const int mine = (grp == 0) ? ja : (grp == 1) ? jb : (grp == 2) ? jc : -1;
      if (mine < 0) continue;
      const DRead R = P.reads[c0 + mine];
      if (gl == 0) {
        atomicAdd(&diff_u[R.gpos], 1);
        atomicAdd(&diff_u[R.gend], -1);
      }
      const uint8_t* sq = P.bases + R.seq_off;
      const uint8_t* ql = P.quals + R.seq_off;
      const int y = R.y0, yend = y + (R.gend - R.gpos), last = yend - 1;
      const bool rev = (R.flags & kReadReversed) != 0;
      const int q0 = (y & ~3) + 8 * gl;
      if (q0 < yend) ec_group4(P, sq, ql, R.rfile, rev, R.gpos, y, last, (int)R.qe0, q0, table, err);
      if (q0 + 4 < yend) ec_group4(P, sq, ql, R.rfile, rev, R.gpos, y, last, (int)R.qe0, q0 + 4, table, err);
    }
    while (sm) {
      const int j = __builtin_ctzll(sm);
      sm &= sm - 1ull;
      one(c0 + j);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P.n_bins; i += blockDim.x)
    if (table[i]) atomicAdd(&P.counts[i], (unsigned long long)table[i]);
  if (err) atomicOr(P.err, err);
}

/* ------------------------------------------------------------------------------------------ */
/* Preprocess mode: where qualifying reads start                                               */

// One thread per alignment: a unique alignment of a scored target whose first query base is
// aligned (nothing but deletions, skips, hard clips or pads before its first M/=/X run) marks its
// strand's bit at that column when its stranded aligned end (query_stranded_end_1 over
// query_bounds_1: soft clips, hard clips and skips at either end excluded; alignment.cpp:120-239)
// reaches required_junction_read_end_min_coordinate(read length) (settings.h:345-354).  The
// pileup entry there has query_position_1 == 1 (error_count.cpp:159-166).
__global__ void __launch_bounds__(256) k_read_starts(EcParams P) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n_reads) return;
  const DRead R = P.reads[i];
  const int t = P.pos_target[R.gpos];
  if (t < 0 || !P.target_scored[t] || R.x1 > 1) return;
  const uint32_t* c = P.cigar + R.cig_off;
  const int n = R.n_cigar;
  int x = R.gpos, col = -1;
  for (int k = 0; k < n; k++) {
    const uint32_t op = c[k] & 15u;
    const int l = (int)(c[k] >> 4);
    if (op == OP_M || op == OP_EQ || op == OP_X) {
      if (l > 0) col = x;
      break;
    }
    if (op == OP_I || op == OP_S) break;   // a query base comes first
    if (op == OP_D || op == OP_N) x += l;
  }
  if (col < 0) return;
  uint32_t start1 = 1, qlen = 0;
  for (int k = 0; k < n; k++) {
    const uint32_t op = c[k] & 15u;
    if (op == OP_M || op == OP_I || op == OP_S || op == OP_EQ || op == OP_X) qlen += c[k] >> 4;
  }
  for (int k = 0; k < n; k++) {
    const uint32_t op = c[k] & 15u;
    if (op != OP_S && op != OP_H && op != OP_N) break;
    if (op == OP_S) start1 += c[k] >> 4;
  }
  uint32_t end1 = qlen;
  for (int k = n - 1; k > 0; k--) {
    const uint32_t op = c[k] & 15u;
    if (op != OP_S && op != OP_H && op != OP_N) break;
    if (op == OP_S) end1 -= c[k] >> 4;
  }
  const bool rev = (R.flags & kReadReversed) != 0;
  const uint32_t rl = R.l_qseq;
  const uint32_t stranded_end = rev ? rl - start1 + 1u : end1;
  const int32_t m = (int32_t)floor((double)((int32_t)rl - (int32_t)P.end_min_len) * P.end_factor);
  const uint32_t need = (m <= 0) ? rl : rl - (uint32_t)m;
  if (stranded_end >= need) atomicOr(&P.start_bits[col >> 4], 1u << ((col & 15) * 2 + (rev ? 1 : 0)));
}

/* ------------------------------------------------------------------------------------------ */
/* Coverage histogram: prefix scan of the difference arrays                                    */

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Inclusive scan across the 64 lanes of a wavefront.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

__global__ void __launch_bounds__(256) k_cov_tile_sums(CovParams C) {
  __shared__ int red[2][4];
  const int64_t n = (int64_t)C.n_pos + 1;
  const int64_t base = (int64_t)blockIdx.x * kCovTile;
  const int32_t* du = C.cov_diff;
  const int32_t* dr = C.cov_diff + n;
  int su = 0, sr = 0;
#pragma unroll 4
  for (int i = 0; i < kCovTile / 256; i++) {
    const int64_t idx = base + i * 256 + threadIdx.x;
    if (idx < n) {
      su += du[idx];
      sr += dr[idx];
    }
  }
  su = wave_sum(su);
  sr = wave_sum(sr);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][wv] = su;
    red[1][wv] = sr;
  }
  __syncthreads();
  if (threadIdx.x == 0) C.tile_sum[blockIdx.x] = make_int2(red[0][0] + red[0][1] + red[0][2] + red[0][3],
                                                           red[1][0] + red[1][1] + red[1][2] + red[1][3]);
}

// One block of 1024 threads: tile_sum becomes the exclusive prefix of the tile sums.
__global__ void __launch_bounds__(1024) k_cov_tile_scan(CovParams C) {
  __shared__ int wsum[2][16];
  __shared__ int carry[2];
  if (threadIdx.x == 0) carry[0] = carry[1] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int b0 = 0; b0 < C.n_tiles; b0 += 1024) {
    const int b = b0 + threadIdx.x;
    const int2 v = (b < C.n_tiles) ? C.tile_sum[b] : make_int2(0, 0);
    const int iu = wave_incl_scan(v.x), ir = wave_incl_scan(v.y);
    if (lane == 63) {
      wsum[0][wv] = iu;
      wsum[1][wv] = ir;
    }
    __syncthreads();
    int ou = carry[0], orr = carry[1];
    for (int k = 0; k < wv; k++) {
      ou += wsum[0][k];
      orr += wsum[1][k];
    }
    if (b < C.n_tiles) C.tile_sum[b] = make_int2(ou + iu - v.x, orr + ir - v.y);
    __syncthreads();
    if (threadIdx.x == 1023) {
      carry[0] = ou + iu;
      carry[1] = orr + ir;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) k_cov_hist(CovParams C) {
  extern __shared__ uint32_t hist[];   // n_groups_lds * hist_lds
  __shared__ int wtot[2][4];
  const int nh = C.n_groups_lds * C.hist_lds;
  for (int i = threadIdx.x; i < nh; i += blockDim.x) hist[i] = 0;
  const int64_t n = (int64_t)C.n_pos + 1;
  const int64_t c0 = (int64_t)blockIdx.x * kCovTile + (int64_t)threadIdx.x * 16;
  const int32_t* du = C.cov_diff;
  const int32_t* dr = C.cov_diff + n;
  int pu[16], pr[16];
  int su = 0, sr = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int64_t idx = c0 + i;
    const int a = (idx < n) ? du[idx] : 0;
    const int b = (idx < n) ? dr[idx] : 0;
    su += a;
    sr += b;
    pu[i] = su;
    pr[i] = sr;
  }
  // exclusive prefix of the thread totals within the block
  const int iu = wave_incl_scan(su), ir = wave_incl_scan(sr);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) {
    wtot[0][wv] = iu;
    wtot[1][wv] = ir;
  }
  __syncthreads();
  const int2 tile = C.tile_sum[blockIdx.x];
  int ou = tile.x + iu - su, orr = tile.y + ir - sr;
  for (int k = 0; k < wv; k++) {
    ou += wtot[0][k];
    orr += wtot[1][k];
  }
  uint32_t err = 0;
  // preprocess mode: (position, strand) pairs without / with a qualifying start, per target
  int st_t = -1;
  uint32_t st_none = 0, st_some = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int64_t col = c0 + i;
    if (col >= C.n_pos) break;
    const int t = C.pos_target[col];
    if (t < 0 || !C.target_scored[t]) continue;
    const int cu = ou + pu[i], cr = orr + pr[i];
    if (cr != 0) continue;   // a redundant read covers the column (error_count.cpp:172-175)
    const int g = C.target_group[t];
    if (g < C.n_groups_lds && cu < C.hist_lds) atomicAdd(&hist[g * C.hist_lds + cu], 1u);
    else if (cu >= 0 && cu < C.hist_len) atomicAdd(&C.hist[(int64_t)g * C.hist_len + cu], 1ull);
    else err |= kErrCapacity;
    if (C.start_bits) {
      if (t != st_t) {
        if (st_t >= 0) {
          atomicAdd(&C.start_counts[2 * st_t], (unsigned long long)st_none);
          atomicAdd(&C.start_counts[2 * st_t + 1], (unsigned long long)st_some);
        }
        st_t = t;
        st_none = st_some = 0;
      }
      const uint32_t b = (C.start_bits[col >> 4] >> ((col & 15) * 2)) & 3u;
      const uint32_t some = (b & 1u) + (b >> 1);
      st_some += some;
      st_none += 2u - some;
    }
  }
  if (st_t >= 0) {
    atomicAdd(&C.start_counts[2 * st_t], (unsigned long long)st_none);
    atomicAdd(&C.start_counts[2 * st_t + 1], (unsigned long long)st_some);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nh; i += blockDim.x) {
    if (hist[i]) {
      const int g = i / C.hist_lds, c = i % C.hist_lds;
      atomicAdd(&C.hist[(int64_t)g * C.hist_len + c], (unsigned long long)hist[i]);
    }
  }
  if (err) atomicOr(C.err, err);
}

}  // namespace bsq
