// gd_emit.hip -- ra_mc_evidence.gd written on the device from the identify pass's results: the RA
// entries of the exact scorer, the MC / UN runs and the whole-target MC entries of failed coverage
// fits (identify_mutations.cpp:1836-1910, 2262-2344, 2972-3007; cGenomeDiff::write and sort,
// genome_diff.cpp:685-830, genome_diff_entry.cpp:566-700).
//
//   k_gd_collect   one thread per candidate / run / failed target: the entry, its numbering key
//                  (target in traversal order, emission column, kind: the order add() numbers
//                  entries in) and its output key (evidence type, seq_id, position, then the
//                  type's own fields);
//   k_gd_rank      the rank of every key among all keys (keys are unique), tile by tile;
//   k_gd_number    the ID (numbering rank + 1) joins the output key;
//   k_gd_lines     the text of each entry into a fixed slot at its output rank, with its length
//                  (the numbers of a line printed by sixteen lanes side by side);
//   k_gd_offsets   exclusive prefix sum of the line lengths (one block);
//   k_gd_gather    the lines packed into one text, one wavefront per line.
// Numbers are printed by the exact formatter of dec_format.h (printf's %.*e and %.*f), so the text
// is byte-identical to the host writer's.
#include <hip/hip_runtime.h>

#include "dec_format.h"
#include "gd_emit.h"

namespace bsq {

namespace {

__device__ const char kIdxChar[6] = {'A', 'C', 'G', 'T', '.', 'N'};

struct Sink {
  char* p;
  int n;
  int cap;
  __device__ void c(char ch) {
    if (n < cap) p[n] = ch;
    n++;
  }
  __device__ void s(const char* x) {
    while (*x) c(*x++);
  }
  __device__ void mem(const char* x, int k) {
    for (int i = 0; i < k; i++) c(x[i]);
  }
  __device__ void i64(int64_t v) {
    char b[24];
    int k = 0;
    uint64_t u = v < 0 ? (uint64_t)0 - (uint64_t)v : (uint64_t)v;
    do {
      b[k++] = (char)('0' + u % 10u);
      u /= 10u;
    } while (u);
    if (v < 0) c('-');
    while (k) c(b[--k]);
  }
  __device__ void key(const char* k) {
    c('\t');
    s(k);
    c('=');
  }
};

__device__ __forceinline__ bool key_less(const GdEnt& a, const GdEnt& b, int which) {
  const uint64_t a0 = which ? a.o0 : a.k0, a1 = which ? a.o1 : a.k1, a2 = which ? a.o2 : a.k2;
  const uint64_t b0 = which ? b.o0 : b.k0, b1 = which ? b.o1 : b.k1, b2 = which ? b.o2 : b.k2;
  if (a0 != b0) return a0 < b0;
  if (a1 != b1) return a1 < b1;
  return a2 < b2;
}

// allele_model::reported_frequency (identify_mutations.cpp:3042-3086): below half a read, 0
__device__ __forceinline__ double reported(const RaOut& r, int b) {
  if (r.n == 0 || b >= 5) return 0.0;
  return (r.f[b] < 0.5 / (double)r.n) ? 0.0 : r.f[b];
}

}  // namespace

__global__ void __launch_bounds__(256) k_gd_collect(GdParams P) {
  const int64_t n_cand = min(*P.n_cand, P.cand_cap), n_runs = min(*P.n_runs, P.run_cap);
  const int64_t total = n_cand + n_runs + P.n_whole;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.ent_cap; i += (int64_t)gridDim.x * blockDim.x) {
    P.rank_id[i] = 0;
    P.rank_out[i] = 0;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    GdEnt e;
    e.o2 = 0;
    e.id = 0;
    if (i < n_cand) {
      e.src = (int32_t)i;
      const RaOut& r = P.ra[i];
      if (!r.emitted) continue;
      if (r.gpos < 0 || r.gpos >= P.n_pos) {
        atomicOr(P.err, 1u);
        continue;
      }
      const int t = P.pos_target[r.gpos];
      if (t < 0 || t >= P.n_targets) {
        atomicOr(P.err, 1u);
        continue;
      }
      const uint64_t position = (uint64_t)(r.gpos - P.targets[t].off + 1);
      e.type = 0;
      e.t = t;
      e.k0 = (uint64_t)P.rank[t];
      e.k1 = position;
      e.k2 = 2ull + (uint64_t)(uint32_t)r.slot;
      e.o0 = (0ull << 56) | ((uint64_t)P.lexrank[t] << 32) | position;
      e.o1 = ((uint64_t)(uint32_t)r.slot << 32) | ((uint64_t)(uint8_t)kIdxChar[min((int)r.ref_idx, 5)] << 8) |
             (uint64_t)(uint8_t)kIdxChar[min((int)r.variant, 5)];
      atomicAdd(P.n_type + 0, 1);
    } else if (i < n_cand + n_runs) {
      e.src = (int32_t)(i - n_cand);
      const RunRec& rr = P.runs[e.src];
      if (rr.kind == 0 && !rr.seed) continue;   // an MC run whose seed was never reached
      if (rr.s < 0 || rr.s >= P.n_pos || rr.e < rr.s || rr.e >= P.n_pos) {
        atomicOr(P.err, 1u);
        continue;
      }
      const int t = P.pos_target[rr.s];
      if (t < 0 || t >= P.n_targets) {
        atomicOr(P.err, 1u);
        continue;
      }
      const int64_t base = P.targets[t].off;
      const uint64_t pos = (uint64_t)(rr.s - base + 1), end = (uint64_t)(rr.e - base + 1);
      e.t = t;
      e.k0 = (uint64_t)P.rank[t];
      e.k1 = end + 1;
      if (rr.kind == 0) {
        const uint64_t red_start = rr.red_start >= 0 ? (uint64_t)(rr.red_start - base + 1) : pos;
        const uint64_t red_end = rr.red_end >= 0 ? (uint64_t)(rr.red_end - base + 1) : end;
        e.type = 1;
        e.k2 = 0;
        e.o0 = (1ull << 56) | ((uint64_t)P.lexrank[t] << 32) | pos;
        e.o1 = (end << 32) | (uint32_t)(red_start - pos);
        e.o2 = (uint64_t)(uint32_t)(end - red_end) << 32;
        atomicAdd(P.n_type + 1, 1);
      } else {
        e.type = 2;
        e.k2 = 1;
        e.o0 = (2ull << 56) | ((uint64_t)P.lexrank[t] << 32) | pos;
        e.o1 = end << 32;
        atomicAdd(P.n_type + 2, 1);
      }
    } else {
      e.src = (int32_t)(i - n_cand - n_runs);
      const int t = P.whole[e.src];
      if (t < 0 || t >= P.n_targets) {
        atomicOr(P.err, 1u);
        continue;
      }
      const uint64_t len = (uint64_t)P.targets[t].len;
      e.type = 3;   // MC of a whole target whose coverage fit failed
      e.t = t;
      e.k0 = (uint64_t)P.rank[t];
      e.k1 = len + 2;
      e.k2 = 0;
      e.o0 = (1ull << 56) | ((uint64_t)P.lexrank[t] << 32) | 1ull;
      e.o1 = len << 32;
      atomicAdd(P.n_type + 1, 1);
    }
    const int k = atomicAdd(P.n_ents, 1);
    if (k < P.ent_cap) P.ents[k] = e;
    else atomicOr(P.err, 2u);
  }
}

// rank of every entry's key (numbering key, or output key) among all n keys: block (x, y) takes
// the tiles y, y + gridDim.y, ... (kGdRankTile keys in LDS each) and, for each, counts the keys of
// the tile below each of the entries x * 256 + tid, stepping by gridDim.x * 256, adding the count
// to the entry's rank (zeroed by k_gd_collect).  n is read on the device, so the grid is fixed.
__global__ void __launch_bounds__(256) k_gd_rank(const GdEnt* ents, const int32_t* n_ents, int cap, int which,
                                                 int32_t* rank) {
  __shared__ GdEnt tile[kGdRankTile];
  // the count read as a wave-uniform scalar: every thread of the block runs the same tile loop and
  // reaches its barriers
  const int n = __builtin_amdgcn_readfirstlane(min(*n_ents, cap));
  for (int j0 = blockIdx.y * kGdRankTile; j0 < n; j0 += gridDim.y * kGdRankTile) {
    __syncthreads();   // the previous tile is no longer read
    if (threadIdx.x < kGdRankTile && j0 + (int)threadIdx.x < n) tile[threadIdx.x] = ents[j0 + threadIdx.x];
    __syncthreads();
    const int m = min(kGdRankTile, n - j0);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      const GdEnt me = ents[i];
      int r = 0;
      for (int j = 0; j < m; j++) r += key_less(tile[j], me, which) ? 1 : 0;
      if (r) atomicAdd(rank + i, r);
    }
  }
}

// the ID (numbering rank + 1) becomes the last output key, as cGenomeDiff::sort breaks ties by it
__global__ void __launch_bounds__(256) k_gd_number(GdEnt* ents, const int32_t* n_ents, int cap, const int32_t* rank) {
  const int n = min(*n_ents, cap);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t id = (uint32_t)rank[i] + 1u;
    ents[i].id = id;
    ents[i].o2 |= id;
  }
}

// The line of each entry into slot out_rank (cGenomeDiff::write: type, ID, evidence '.', the
// positional fields, then key=value fields in name order).  A line is cut into kGdPieces pieces,
// one per lane of a 16-lane group (four entries per wavefront): the positional header, each
// allele frequency, each printed number with its key, the coverage pairs.  Every lane prints its
// piece into LDS, the group's lengths give each piece its offset, and the pieces are copied into
// the slot side by side, so the numbers of one line are formatted in parallel.
__global__ void __launch_bounds__(64) k_gd_lines(GdParams P, const int32_t* out_rank) {
  __shared__ char piece[64][kGdPiece];
  __shared__ int32_t plen[64];
  const int lane = threadIdx.x, grp = lane / kGdPieces, k = lane % kGdPieces;
  const int n = __builtin_amdgcn_readfirstlane(min(*P.n_ents, P.ent_cap));
  const int step = gridDim.x * (64 / kGdPieces);
  for (int i0 = blockIdx.x * (64 / kGdPieces); i0 < n; i0 += step) {
    const int i = i0 + grp;
    int slot = -1;
    Sink o{piece[lane], 0, kGdPiece};
    bool num = false, nsci = true;   // the piece ends with a number: nv printed with nprec digits
    double nv = 0.0;
    int nprec = 0;
    if (i < n) {
      const GdEnt e = P.ents[i];
      slot = out_rank[i];
      if (slot < 0 || slot >= n) {
        atomicOr(P.err, 1u);
        slot = -1;
      } else if (k == 0) {
        // type, ID, evidence, seq_id and the positional fields
        const int t = e.t;
        o.s(e.type == 0 ? "RA" : (e.type == 2 ? "UN" : "MC"));
        o.c('\t');
        o.i64(e.id);
        o.s("\t.\t");
        o.mem(P.names + P.name_off[t], P.name_off[t + 1] - P.name_off[t]);
        o.c('\t');
        if (e.type == 0) {
          const RaOut& r = P.ra[e.src];
          o.i64(r.gpos - P.targets[t].off + 1);
          o.c('\t');
          o.i64(r.slot);
          o.c('\t');
          o.c(kIdxChar[min((int)r.ref_idx, 5)]);
          o.c('\t');
          o.c(kIdxChar[min((int)r.variant, 5)]);
        } else {
          o.i64((int64_t)(e.o0 & 0xffffffffull));
          o.c('\t');
          o.i64((int64_t)(e.o1 >> 32));
          if (e.type != 2) {
            o.c('\t');
            o.i64((int64_t)(e.o1 & 0xffffffffull));
            o.c('\t');
            o.i64((int64_t)(e.o2 >> 32));
          }
        }
      } else if (e.type == 0 || e.type == 1 || e.type == 3) {
        // the piece's text before its number (if any) is written here; the numbers of all lanes
        // are then printed by one call below, so that the lanes run the formatter together
        if (e.type == 0) {
          // RA fields (identify_mutations.cpp:1836-1910), in the entry's name order
          const RaOut& r = P.ra[e.src];
          const int places = P.places;
          auto cov = [&](const char* key, int b) {
            o.key(key);
            if (b >= 5) {
              o.s("0/0");
              return;
            }
            o.i64((int32_t)r.cov[b][0]);
            o.c('/');
            o.i64((int32_t)r.cov[b][1]);
          };
          if (k <= 5) {
            const int b = k - 1;
            const double fr = reported(r, b);
            if (fr > 0.0) {
              bool first = true;
              for (int q = 0; q < b; q++) first = first && !(reported(r, q) > 0.0);
              if (first) o.key("allele_frequencies");
              else o.c(',');
              o.c(kIdxChar[b]);
              o.c(':');
              num = true;
              nv = fr;
              nprec = places;
            }
          } else if (k == 11) {
            o.key("major_base");
            o.c(kIdxChar[min((int)r.major, 5)]);
            cov("major_cov", r.major);
          } else if (k == 13) {
            o.key("minor_base");
            o.c(kIdxChar[min((int)r.minor, 5)]);
            cov("minor_cov", r.minor);
            cov("new_cov", r.variant);
            cov("ref_cov", r.ref_idx);
          } else if (k == 15) {
            o.key("total_cov");
            o.i64(r.total_top);
            o.c('/');
            o.i64(r.total_bot);
            o.c('\n');
          } else {
            num = true;
            const char* key;
            if (k == 6) {
              key = "fisher_strand_p_value";
              nv = r.fisher_p;
              nprec = 5;
            } else if (k == 7) {
              key = "frequency";
              nv = reported(r, r.variant);
              nprec = places;
            } else if (k == 8) {
              key = "frequency_lower";
              nv = r.lower;
              nprec = places;
            } else if (k == 9) {
              key = "frequency_upper";
              nv = r.upper;
              nprec = places;
            } else if (k == 10) {
              key = "ks_quality_p_value";
              nv = r.ks_p;
              nprec = 5;
            } else if (k == 12) {
              key = "major_frequency";
              nv = reported(r, r.major);
              nprec = places;
            } else {
              key = "score";
              nv = r.variant_score;
              nprec = 1;
              nsci = false;
            }
            o.key(key);
          }
        } else {
          // MC: left_inside_cov, left_outside_cov, right_inside_cov, right_outside_cov (%.0f, NA
          // when unset); a whole-target MC has 0 inside and nothing outside
          int32_t li = 0, lo = -1, ri = 0, ro = -1;
          if (e.type == 1) {
            const RunRec& rr = P.runs[e.src];
            li = rr.li;
            lo = rr.lo;
            ri = rr.ri;
            ro = rr.ro;
          }
          const double nan = __longlong_as_double(0x7ff8000000000000LL);
          if (k >= 1 && k <= 4) {
            num = true;
            nprec = 0;
            nsci = false;
            if (k == 1) {
              o.key("left_inside_cov");
              nv = (double)li;
            } else if (k == 2) {
              o.key("left_outside_cov");
              nv = lo < 0 ? nan : (double)lo;
            } else if (k == 3) {
              o.key("right_inside_cov");
              nv = (double)ri;
            } else {
              o.key("right_outside_cov");
              nv = ro < 0 ? nan : (double)ro;
            }
          } else if (k == 5) {
            o.c('\n');
          }
        }
      } else if (k == 1) {
        o.c('\n');
      }
    }
    if (num) {
      char buf[dec::kMaxText];
      const int m = dec::format_number(nv, nprec, nsci, buf);
      if (m > 0) o.mem(buf, m);
      else atomicOr(P.err, 8u);
    }
    if (o.n > kGdPiece) atomicOr(P.err, 4u);
    plen[lane] = min(o.n, kGdPiece);
    __syncthreads();
    int off = 0, total = 0;
    for (int q = 0; q < kGdPieces; q++) {
      const int l = plen[grp * kGdPieces + q];
      if (q < k) off += l;
      total += l;
    }
    if (slot >= 0) {
      if (total > kGdSlot) {
        atomicOr(P.err, 4u);
      } else {
        char* dst = P.slots + (size_t)slot * kGdSlot + off;
        for (int q = 0; q < plen[lane]; q++) dst[q] = piece[lane][q];
        if (k == 0) P.lens[slot] = total;
      }
    }
    __syncthreads();
  }
}

// exclusive prefix sum of the n line lengths, offsets[n] = the text length (one block)
__global__ void __launch_bounds__(1024) k_gd_offsets(const int32_t* lens, const int32_t* n_ents, int cap, int64_t* offsets,
                                                     int64_t* total) {
  __shared__ int64_t part[1024];
  const int n = __builtin_amdgcn_readfirstlane(min(*n_ents, cap));
  const int per = (n + 1023) / 1024;
  const int b = threadIdx.x * per, e = min(n, b + per);
  int64_t s = 0;
  for (int i = b; i < e; i++) s += lens[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int64_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int64_t run = part[threadIdx.x] - s;
  for (int i = b; i < e; i++) {
    offsets[i] = run;
    run += lens[i];
  }
  if (threadIdx.x == 1023) {
    offsets[n] = part[1023];
    *total = part[1023];
  }
}

// line k into the text at offsets[k], one wavefront per line
__global__ void __launch_bounds__(256) k_gd_gather(const char* slots, const int32_t* lens, const int64_t* offsets,
                                                   const int32_t* n_ents, int cap, char* text) {
  const int n = min(*n_ents, cap);
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64, lane = threadIdx.x % 64;
  const int waves = gridDim.x * blockDim.x / 64;
  for (int k = wave; k < n; k += waves) {
    const int len = lens[k];
    const char* src = slots + (size_t)k * kGdSlot;
    char* dst = text + offsets[k];
    for (int i = lane; i < len; i += 64) dst[i] = src[i];
  }
}

// Exact decimal text of v (the formatter the writer uses), one value per thread; out holds
// kGdFmtStride bytes per value, lens the text lengths.
__global__ void __launch_bounds__(64) k_format_doubles(const double* v, int64_t n, int32_t precision, int32_t scientific,
                                                       char* out, int32_t* lens) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    char b[dec::kMaxText];
    const int k = scientific ? dec::format_sci(v[i], precision, b) : dec::format_fixed(v[i], precision, b);
    lens[i] = k;
    for (int j = 0; j < k && j < kGdFmtStride; j++) out[i * kGdFmtStride + j] = b[j];
  }
}

}  // namespace bsq

#include <string>
#include <vector>

#include "host_util.h"

// The device formatter over n values (test entry: the device's text is compared with printf's).
extern "C" int bsq_format_doubles_device(const double* v, int64_t n, int32_t precision, int32_t scientific, char* out,
                                         int32_t* lens, bsq_status* status) {
  using namespace bsq;
  double* d_v = nullptr;
  char* d_out = nullptr;
  int32_t* d_len = nullptr;
  auto cleanup = [&]() {
    if (d_v) (void)hipFree(d_v);
    if (d_out) (void)hipFree(d_out);
    if (d_len) (void)hipFree(d_len);
  };
  try {
    if (n < 0 || (n > 0 && (!v || !out || !lens))) throw BsqError(BSQ_ERR_ARG, "bad arguments");
    if (n == 0) {
      set_status(status, BSQ_OK, "");
      return BSQ_OK;
    }
    if (hipMalloc((void**)&d_v, (size_t)n * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&d_out, (size_t)n * kGdFmtStride) != hipSuccess ||
        hipMalloc((void**)&d_len, (size_t)n * sizeof(int32_t)) != hipSuccess)
      throw BsqError(BSQ_ERR_DEVICE, "device allocation failed");
    if (hipMemcpy(d_v, v, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      throw BsqError(BSQ_ERR_DEVICE, "copy failed");
    const int64_t blocks = std::min<int64_t>(4096, (n + 63) / 64);
    hipLaunchKernelGGL(k_format_doubles, dim3((unsigned)blocks), dim3(64), 0, 0, d_v, n, precision, scientific, d_out, d_len);
    if (hipGetLastError() != hipSuccess) throw BsqError(BSQ_ERR_DEVICE, "kernel launch failed");
    if (hipMemcpy(out, d_out, (size_t)n * kGdFmtStride, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(lens, d_len, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
      throw BsqError(BSQ_ERR_DEVICE, "copy failed");
    cleanup();
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    cleanup();
    set_status(status, e.code, e.what());
    return e.code;
  }
}

