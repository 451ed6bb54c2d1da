// extra.hip -- per-position reference trims on the GPU (SequenceTrims, calculate_trims.cpp:67-138).
//
// The reference scans each position for perfect tandem repeats of unit 1..18 and max-scatters the
// resulting trim lengths forward.  Here the scatter is turned into a gather so every output byte is
// written by exactly one thread: k_repeat_extent computes each position's maximal repeat extent m[p]
// (capped at 255), and k_trim_gather computes for each x
//   right[x] = max over p in (x-255, x] with p + m[p] > x of (x - p + 1)
//   left[x]  = max over the same p of (m[p] - (x - p)).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "engine.h"
#include "host_util.h"

namespace bsq {

// Device memory of bsq_sequence_trims_static, for callers that cannot allocate: the sequence, its
// repeat extents and the two trim rows of a sequence of up to kStaticTrimsMax bases (4 bytes per
// base, 32 MiB).  Every GPU the library runs this entry point on holds its own copy until the
// process ends.  Host code never uses the array's name as an address: each device's copy is found
// once with hipGetSymbolAddress (scratch_on_device) and only that pointer reaches kernels and copies.
constexpr int64_t kStaticTrimsMax = (int64_t)8 << 20;
static __device__ __align__(256) uint8_t scratch[4 * kStaticTrimsMax];

static std::mutex g_scratch_mu;
static std::vector<uint8_t*> g_scratch_ptr;   // per device id

// Address of `scratch` on the current device, looked up on the first call on that device.
static uint8_t* scratch_on_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) throw BsqError(BSQ_ERR_DEVICE, "no current HIP device");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_ptr[(size_t)dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw BsqError(BSQ_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[(size_t)dev] = (uint8_t*)p;
  }
  return g_scratch_ptr[(size_t)dev];
}

__global__ void __launch_bounds__(256) k_repeat_extent(const char* __restrict__ s, int64_t len, uint8_t* __restrict__ m) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= len) return;
  uint32_t mx = 1;
  for (int size = 1; size <= 18; size++) {
    // repeat_match(pos, size, num): s[p+i] == s[p+i+size*num] for i < size, all indexes < len
    uint32_t num = 1;
    for (;;) {
      const int64_t off = (int64_t)size * num;
      if (p + size - 1 + off >= len) break;
      bool ok = true;
      for (int i = 0; i < size; i++)
        if (s[p + i] != s[p + i + off]) {
          ok = false;
          break;
        }
      if (!ok) break;
      num++;
      if ((uint64_t)num * size > 255) break;   // capped below; longer runs change nothing
    }
    if (num > 1 && num * size > mx) mx = num * size;
  }
  if (mx > 255) mx = 255;
  m[p] = (uint8_t)mx;
}

__global__ void __launch_bounds__(256) k_trim_gather(const uint8_t* __restrict__ m, int64_t len, uint8_t* __restrict__ left_out,
                                                     uint8_t* __restrict__ right_out) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= len) return;
  uint32_t right = 0, left = 0;
  const int64_t p0 = (x >= 254) ? x - 254 : 0;
  for (int64_t p = p0; p <= x; p++) {
    const uint32_t mp = m[p];
    const uint32_t o = (uint32_t)(x - p);
    if (o < mp) {
      right = max(right, o + 1);
      left = max(left, mp - o);
    }
  }
  left_out[x] = (uint8_t)left;
  right_out[x] = (uint8_t)right;
}

// BAM code -> the character breseq compares (basebam2char, common.h:263-275): A C G T, else N
__device__ __forceinline__ char trim_char(uint8_t b) {
  switch (b) {
    case 1: return 'A';
    case 2: return 'C';
    case 4: return 'G';
    case 8: return 'T';
    default: return 'N';
  }
}

// get_alignment_trims (calculate_trims.cpp:239-285) of every packed alignment, one thread each:
//   the genomic trims at the alignment's first and last reference positions, the read-edge trims
//   of its aligned sequence (edge_trims_for_sequence, :183-230: the longest perfect tandem repeat of
//   unit 1..18 starting at either end, at least 1), the larger of each pair, plus the bases outside
//   the match (query_start_0 on the left, read length - query_end_1 on the right; soft clips only,
//   alignment.cpp:248-289).  A repeat compared on the reverse complement matches exactly where the
//   reversed sequence does (complementing is a bijection on A, C, G, T, N), so the right edge reads
//   the sequence backwards.  left/right: the sample's per-column trims (global columns).
__global__ void __launch_bounds__(256) k_alignment_trims(const DRead* __restrict__ reads, int64_t n,
                                                         const uint32_t* __restrict__ cigar,
                                                         const uint8_t* __restrict__ bases,
                                                         const uint8_t* __restrict__ left, const uint8_t* __restrict__ right,
                                                         int32_t* __restrict__ xl, int32_t* __restrict__ xr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DRead R = reads[i];
  const uint32_t* c = cigar + R.cig_off;
  const int nc = R.n_cigar;
  uint32_t qs1 = 1, qlen = 0;
  for (int k = 0; k < nc; k++) {
    if ((c[k] & 15u) != 4u) break;
    qs1 += c[k] >> 4;
  }
  for (int k = 0; k < nc; k++) {
    const uint32_t op = c[k] & 15u;
    if (op == 0 || op == 1 || op == 4 || op == 7 || op == 8) qlen += c[k] >> 4;
  }
  uint32_t qe1 = qlen;
  for (int k = nc - 1; k > 0; k--) {
    if ((c[k] & 15u) != 4u) break;
    qe1 -= c[k] >> 4;
  }
  const uint8_t* s = bases + R.seq_off + (qs1 - 1);
  const int64_t len = (qe1 >= qs1) ? (int64_t)(qe1 - qs1 + 1) : 0;
  uint32_t eL = 1, eR = 1;
  for (int size = 1; size <= 18; size++) {
    // repeat_match(seq, 0, size, num) (calculate_trims.cpp:44-56), from each end
    int64_t num = 1;
    for (;;) {
      const int64_t off = (int64_t)size * num;
      if (size - 1 + off >= len) break;
      bool ok = true;
      for (int k = 0; k < size && ok; k++) ok = trim_char(s[k]) == trim_char(s[k + off]);
      if (!ok) break;
      num++;
    }
    if (num > 1) eL = max(eL, (uint32_t)(num * size));
    num = 1;
    for (;;) {
      const int64_t off = (int64_t)size * num;
      if (size - 1 + off >= len) break;
      bool ok = true;
      for (int k = 0; k < size && ok; k++) ok = trim_char(s[len - 1 - k]) == trim_char(s[len - 1 - k - off]);
      if (!ok) break;
      num++;
    }
    if (num > 1) eR = max(eR, (uint32_t)(num * size));
  }
  const uint32_t gL = left[R.gpos], gR = right[R.gend - 1];
  xl[i] = (int32_t)(max(gL, eL) + (qs1 - 1));
  xr[i] = (int32_t)(max(gR, eR) + (R.l_qseq - qe1));
}

}  // namespace bsq

using namespace bsq;

extern "C" int bsq_sequence_trims(const char* seq, int64_t len, uint8_t* out, bsq_status* status) {
  char* d_s = nullptr;
  uint8_t *d_m = nullptr, *d_o = nullptr;
  auto cleanup = [&]() {
    if (d_s) (void)hipFree(d_s);
    if (d_m) (void)hipFree(d_m);
    if (d_o) (void)hipFree(d_o);
  };
  try {
    if (!seq || !out || len < 0) throw BsqError(BSQ_ERR_ARG, "bad arguments");
    if (len == 0) {
      set_status(status, BSQ_OK, "");
      return BSQ_OK;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw BsqError(BSQ_ERR_DEVICE, "no HIP device available");
    if (hipMalloc((void**)&d_s, (size_t)len) != hipSuccess || hipMalloc((void**)&d_m, (size_t)len) != hipSuccess ||
        hipMalloc((void**)&d_o, 2 * (size_t)len) != hipSuccess)
      throw BsqError(BSQ_ERR_DEVICE, "hipMalloc failed");
    if (hipMemcpy(d_s, seq, (size_t)len, hipMemcpyHostToDevice) != hipSuccess) throw BsqError(BSQ_ERR_DEVICE, "copy failed");
    const int64_t blocks = (len + 255) / 256;
    hipLaunchKernelGGL(k_repeat_extent, dim3((unsigned)blocks), dim3(256), 0, 0, d_s, len, d_m);
    hipLaunchKernelGGL(k_trim_gather, dim3((unsigned)blocks), dim3(256), 0, 0, d_m, len, d_o, d_o + len);
    if (hipGetLastError() != hipSuccess) throw BsqError(BSQ_ERR_DEVICE, "kernel launch failed");
    if (hipMemcpy(out, d_o, 2 * (size_t)len, hipMemcpyDeviceToHost) != hipSuccess) throw BsqError(BSQ_ERR_DEVICE, "copy failed");
    cleanup();
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    cleanup();
    set_status(status, e.code, e.what());
    return e.code;
  }
}

// The trims of bsq_sequence_trims with no device allocation: input, repeat extents and output live
// in this device's `scratch` (sequences of at most kStaticTrimsMax bases).  Both kernels write
// every byte they later read, so nothing needs clearing between calls.  Calls that overlap on one
// GPU share the array and race.
extern "C" int bsq_sequence_trims_static(const char* seq, int64_t len, uint8_t* out, bsq_status* status) {
  try {
    if ((!seq || !out) && len > 0) throw BsqError(BSQ_ERR_ARG, "bad arguments");
    if (len < 0) throw BsqError(BSQ_ERR_ARG, "negative length");
    if (len > kStaticTrimsMax)
      throw BsqError(BSQ_ERR_LIMIT, "sequence of " + std::to_string(len) + " bases exceeds the " +
                                        std::to_string(kStaticTrimsMax) + "-base capacity of the allocation-free trims");
    if (len == 0) {
      set_status(status, BSQ_OK, "");
      return BSQ_OK;
    }
    uint8_t* base = scratch_on_device();
    char* d_s = (char*)base;
    uint8_t* d_m = base + kStaticTrimsMax;
    uint8_t* d_o = base + 2 * kStaticTrimsMax;   // 2 * len bytes: left row, then right row
    if (hipMemcpy(d_s, seq, (size_t)len, hipMemcpyHostToDevice) != hipSuccess) throw BsqError(BSQ_ERR_DEVICE, "copy failed");
    const int64_t blocks = (len + 255) / 256;
    hipLaunchKernelGGL(k_repeat_extent, dim3((unsigned)blocks), dim3(256), 0, 0, d_s, len, d_m);
    hipLaunchKernelGGL(k_trim_gather, dim3((unsigned)blocks), dim3(256), 0, 0, d_m, len, d_o, d_o + len);
    if (hipGetLastError() != hipSuccess) throw BsqError(BSQ_ERR_DEVICE, "kernel launch failed");
    if (hipMemcpy(out, d_o, 2 * (size_t)len, hipMemcpyDeviceToHost) != hipSuccess) throw BsqError(BSQ_ERR_DEVICE, "copy failed");
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& e) {
    set_status(status, e.code, e.what());
    return e.code;
  }
}
