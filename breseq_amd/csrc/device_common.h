// device_common.h -- device-side building blocks shared by the pileup kernels: base encodings
// (common.h:202-316), the pileup entry of one alignment at one reference column (htslib
// resolve_cigar2 as breseq consumes it, pileup_base.cpp:141-210), and the per-slot read filters
// of identify_mutations_pileup::pileup_callback (identify_mutations.cpp:1561-1657).
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace bsq {

enum { OP_M = 0, OP_I = 1, OP_D = 2, OP_N = 3, OP_S = 4, OP_H = 5, OP_P = 6, OP_EQ = 7, OP_X = 8 };

// error bits reported back to the host (each is an ASSERT in the reference)
enum : uint32_t {
  kErrBaseCode = 1,      // read base is not A/C/G/T/N (basebam2index, common.h:271-283)
  kErrRefChar = 2,       // reference base is not A/C/G/T/N (basechar2index)
  kErrQuality = 4,       // quality >= quality covariate maximum (error_count.cpp:484-491)
  kErrReadSet = 8,       // read file index >= read_set covariate maximum
  kErrQueryIndex = 16,   // quality index outside the read (error_count.cpp:914,940,967)
  kErrCigar = 32,        // CIGAR walk did not reach the column
  kErrCapacity = 64,     // a fixed device buffer was too small (host retries or reports)
  kErrIndex = 128        // an index read from a device list fell outside the buffer it indexes;
                         // the access was skipped (a library fault, reported rather than taken)
};

// Alignment records are read-only for a kernel's lifetime; reading them through the constant
// address space lets the compiler use scalar (SMEM) loads when the index is wave-uniform.
typedef __attribute__((address_space(4))) const uint64_t ConstU64;
__device__ __forceinline__ DRead load_read(const DRead* reads, int r) {
  const ConstU64* c = (const ConstU64*)(reads + r);
  uint64_t w[5];
#pragma unroll
  for (int i = 0; i < 5; i++) w[i] = c[i];
  DRead d;
  __builtin_memcpy(&d, w, sizeof(DRead));
  return d;
}

__device__ __forceinline__ bool is_ref_op(uint32_t op) {
  return op == OP_M || op == OP_D || op == OP_N || op == OP_EQ || op == OP_X;
}

__device__ __forceinline__ uint8_t comp_bam(uint8_t b) {
  // 1<->8, 2<->4, 15->15, '.'->'.'
  switch (b) {
    case 1: return 8;
    case 2: return 4;
    case 4: return 2;
    case 8: return 1;
    default: return b;
  }
}

__device__ __forceinline__ char comp_char(char c) {
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    default: return c;
  }
}

// BAM 4-bit code -> index A0 C1 G2 T3; '.' -> 4; anything else -> -1
__device__ __forceinline__ int bam_index(uint8_t b) {
  switch (b) {
    case 1: return 0;
    case 2: return 1;
    case 4: return 2;
    case 8: return 3;
    case '.': return 4;
    default: return -1;
  }
}

__device__ __forceinline__ int char_index(char c) {
  switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    case '.': return 4;
    case 'N': return 5;
    default: return -1;
  }
}

__device__ __forceinline__ uint8_t comp_index(uint8_t i) { return i < 4 ? (uint8_t)(3 - i) : (uint8_t)4; }

// Pileup entry of read R at global column p: returns false if R does not cover p.
__device__ __forceinline__ bool entry_at(const DRead& R, const uint32_t* __restrict__ cigar, int p, int& qpos,
                                         int& indel, bool& is_del, uint32_t& err) {
  if (p < R.gpos || p >= R.gend) return false;
  if (R.flags & kReadSimple) {
    qpos = (int)R.y0 + (p - R.gpos);
    indel = 0;
    is_del = false;
    return true;
  }
  const uint32_t* c = cigar + R.cig_off;
  const int n = R.n_cigar;
  int x = R.gpos, y = 0, k = 0;
  for (; k < n; ++k) {
    uint32_t op = c[k] & 15;
    if (is_ref_op(op)) break;
    if (op == OP_I || op == OP_S) y += (int)(c[k] >> 4);
  }
  for (;;) {
    if (k >= n) {
      err |= kErrCigar;
      return false;
    }
    uint32_t w = c[k];
    int l = (int)(w >> 4);
    if (p - x < l) break;
    uint32_t op = w & 15;
    if (op == OP_M || op == OP_EQ || op == OP_X) y += l;
    x += l;
    ++k;
    for (; k < n; ++k) {
      uint32_t op2 = c[k] & 15;
      if (is_ref_op(op2)) break;
      if (op2 == OP_I || op2 == OP_S) y += (int)(c[k] >> 4);
    }
  }
  uint32_t op = c[k] & 15;
  int l = (int)(c[k] >> 4);
  indel = 0;
  if (x + l - 1 == p && k + 1 < n) {
    uint32_t op2 = c[k + 1] & 15;
    int l2 = (int)(c[k + 1] >> 4);
    if (op2 == OP_D && op != OP_D) {
      indel = -l2;
      for (int kk = k + 2; kk < n; ++kk) {
        if ((c[kk] & 15) == OP_D) indel -= (int)(c[kk] >> 4);
        else break;
      }
    } else if (op2 == OP_I) {
      indel = l2;
      for (int kk = k + 2; kk < n; ++kk) {
        uint32_t o3 = c[kk] & 15;
        if (o3 == OP_I) indel += (int)(c[kk] >> 4);
        else if (o3 != OP_P) break;
      }
    } else if (op2 == OP_P && k + 2 < n) {
      int l3 = 0;
      for (int kk = k + 2; kk < n; ++kk) {
        uint32_t o3 = c[kk] & 15;
        if (o3 == OP_I) l3 += (int)(c[kk] >> 4);
        else if (is_ref_op(o3)) break;
      }
      if (l3 > 0) indel = l3;
    }
  }
  if (op == OP_M || op == OP_EQ || op == OP_X) {
    qpos = y + (p - x);
    is_del = false;
  } else {
    qpos = y;
    is_del = true;
  }
  return true;
}

// is_trimmed (alignment.h:389-410)
__device__ __forceinline__ bool is_trimmed(const DRead& R, int qpos, bool past) {
  uint32_t q1 = (uint32_t)qpos + 1;
  if (R.xl >= 0 && q1 <= (uint32_t)R.xl) return true;
  if (R.xr >= 0) {
    uint32_t rem = (uint32_t)R.l_qseq - q1;
    if (rem + 1 <= (uint32_t)R.xr) return true;
    if (past && rem == (uint32_t)R.xr) return true;
  }
  return false;
}

// Observation a unique, untrimmed read contributes at (column, insert slot), after
// alignment_position_to_covariates (error_count.cpp:1049-1105).  Returns false when the read is
// skipped there.  obs: 0..4; q: quality of the base the reference reads it from.
__device__ __forceinline__ bool slot_observation(const DRead& R, const uint8_t* __restrict__ bases,
                                                 const uint8_t* __restrict__ quals, int qpos, int ondel,
                                                 int slot, uint8_t read_base, int& obs, int& q, uint32_t& err) {
  const int rev = (R.flags & kReadReversed) ? 1 : 0;
  int qq = qpos;
  if (ondel == -1) {
    qq += 1 - rev;
    if (qq < 0 || qq >= (int)R.l_qseq) {
      err |= kErrQueryIndex;
      return false;
    }
    if (bases[R.seq_off + qq] == 15) return false;
  } else if (slot > 0) {
    int mo = slot;
    if (ondel < mo) mo = ondel;
    qq += mo + 1 - rev;
    if ((uint32_t)qq > (uint32_t)R.qbe0) return false;
    if (qq < 0 || qq >= (int)R.l_qseq) {
      err |= kErrQueryIndex;
      return false;
    }
    if (bases[R.seq_off + qq] == 15) return false;
  }
  obs = bam_index(read_base);
  if (obs < 0) {
    err |= kErrBaseCode;
    return false;
  }
  q = quals[R.seq_off + qq];
  return true;
}

// Workgroups are dispatched to the eight XCDs round robin (block b runs on XCD b % 8).  This
// renumbering gives each XCD one contiguous range of logical block indices, so work assigned by
// logical index in contiguous chunks keeps neighbouring chunks in the same L2.  Grids are launched
// with a multiple of 8 blocks; any remainder keeps its index.
__device__ __forceinline__ int xcd_major_block() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int per = nb / 8;
  if (b >= per * 8) return b;
  return (b % 8) * per + b / 8;
}

// pdata entry of alignment R at (global column p, insert slot): the per-read filters of
// identify_mutations_pileup::pileup_callback (identify_mutations.cpp:1561-1657).  On success key is
// ((mslot*S + read file)*Q + quality)*10 + obs*2 + (reverse strand), the row of the likelihood
// tables; strand_idx is 2 for the forward strand and 0 for the reverse strand.
__device__ __forceinline__ bool pdata_key(const DRead& R, const uint32_t* __restrict__ cigar,
                                          const uint8_t* __restrict__ bases, const uint8_t* __restrict__ quals,
                                          int p, int slot, int S, int Q, int qcut, uint32_t& key, int& obs,
                                          int& strand_idx, uint32_t& err) {
  int qpos, indel;
  bool is_del;
  if (!entry_at(R, cigar, p, qpos, indel, is_del, err)) return false;
  const int ondel = is_del ? -1 : (indel < 0 ? 0 : indel);
  uint8_t read_base = '.';
  bool past = true;
  if (ondel >= slot) {
    read_base = bases[R.seq_off + qpos + slot];
    past = false;
  }
  // X1 > 1: redundant, no pdata
  if (read_base == 15 || R.x1 > 1 || is_trimmed(R, qpos, past)) return false;
  int q;
  if (!slot_observation(R, bases, quals, qpos, ondel, slot, read_base, obs, q, err)) return false;
  if (q < qcut) return false;
  if ((int)R.rfile >= S) {
    err |= kErrReadSet;
    return false;
  }
  if (q >= Q) {
    err |= kErrQuality;
    return false;
  }
  const bool top = !(R.flags & kReadReversed);
  key = ((((uint32_t)R.mslot * S + R.rfile) * Q + q) * 10u) + obs * 2u + (top ? 0u : 1u);
  strand_idx = top ? 2 : 0;
  return true;
}

// Correctly rounded quotients x / s for several x over one denominator s.  The reciprocal y = 1/s
// is formed once by the ordinary (correctly rounded) division; each quotient is then
// q0 = RN(x*y), r = x - s*q0 (exact through the FMA) and RN(q0 + r*y): the final correction step
// of the hardware's own division sequence, which with y within half an ulp of 1/s gives RN(x/s)
// (Markstein).  It needs neither operand near the ends of the exponent range, so a denominator or
// numerator outside [2^-500, 2^500] (zero numerators aside) takes the ordinary division.  The
// five quotients of an EM term then cost one division instead of five serialized ones.
struct SharedDiv {
  double s, y;
  bool ok;
  __device__ __forceinline__ explicit SharedDiv(double den) : s(den) {
    const double a = fabs(den);
    ok = a >= 0x1p-500 && a <= 0x1p500;
    y = ok ? 1.0 / den : 0.0;
  }
  __device__ __forceinline__ double operator()(double x) const {
    const double a = fabs(x);
    if (!ok || !(a == 0.0 || (a >= 0x1p-500 && a <= 0x1p500))) return x / s;
    const double q0 = x * y;
    const double r = fma(-s, q0, x);
    return fma(r, y, q0);
  }
};

// pure_genotype_call (identify_mutations.cpp:3398-3433)
__device__ __forceinline__ void pure_genotype(const double lps[5], uint32_t n, int& best, double& score) {
  best = 5;
  score = __longlong_as_double(0x7ff8000000000000LL);
  if (n == 0) return;
  best = 0;
  for (int b = 1; b < 5; b++)
    if (lps[b] > lps[best]) best = b;
  double off = -1.7976931348623157e308;
  for (int b = 0; b < 5; b++)
    if (b != best) off = fmax(off, lps[b]);
  double tot = 0.0;
  // kept rolled: one pow at a time keeps the register footprint of the callers' loops small
#pragma unroll 1
  for (int b = 0; b < 5; b++)
    if (b != best) tot += pow(10.0, lps[b] - off);
  double lt = log10(tot);
  lt += off;
  score = lps[best] - lt;
}

}  // namespace bsq
