// selftest.hip -- on-device check of SharedDiv (device_common.h) against the ordinary correctly
// rounded division, over random and adversarial operands.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "host_util.h"
#include "kernels.h"

namespace bsq {

namespace {

__device__ __forceinline__ uint64_t splitmix(uint64_t& st) {
  uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// A double with the given exponent (unbiased) and random significand; some significands all
// ones or all zeros, the cases nearest to a rounding boundary of a reciprocal.
__device__ __forceinline__ double make_double(uint64_t bits, int e) {
  uint64_t mant = bits & 0xFFFFFFFFFFFFFULL;
  const uint32_t kind = (uint32_t)(bits >> 60);
  if (kind == 0) mant = 0xFFFFFFFFFFFFFULL;
  else if (kind == 1) mant = 0;
  else if (kind == 2) mant |= 0xFFFFFFFF00000ULL;
  return __longlong_as_double((long long)(((uint64_t)(e + 1023) << 52) | mant));
}

__global__ void k_selftest_division(uint64_t n, uint64_t seed, unsigned long long* mismatches,
                                    unsigned long long* fast_path) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long bad = 0, fast = 0;
  for (uint64_t i = tid; i < n; i += stride) {
    uint64_t st = seed ^ (i * 0xD1B54A32D192ED03ULL);
    const uint64_t a = splitmix(st), b = splitmix(st), c = splitmix(st);
    double s, x;
    switch ((uint32_t)(c & 3u)) {
      case 0: {   // EM terms: f, r in (0, 1], s = f r + rest, x = f r
        const double f = make_double(a, -(int)((c >> 8) % 60u) - 1);
        const double r = make_double(b, -(int)((c >> 16) % 40u));
        x = f * r;
        s = x + make_double(a ^ b, -(int)((c >> 24) % 30u));
        break;
      }
      case 1:   // arbitrary operands across the whole guarded range and beyond it
        s = make_double(a, (int)((c >> 8) % 1200u) - 600);
        x = make_double(b, (int)((c >> 20) % 1200u) - 600);
        break;
      case 2: {   // frequency updates: sums / n and (1 - x) sums / others
        const double nn = (double)(1 + ((c >> 8) % 2000u));
        x = make_double(a, (int)((c >> 24) % 12u) - 1) * nn * 0.25;
        s = ((c >> 40) & 1u) ? nn : make_double(b, (int)((c >> 32) % 12u) - 11);
        break;
      }
      default:   // quotients near 1 and exact ones
        s = make_double(a, (int)((c >> 8) % 16u) - 8);
        x = ((c >> 20) & 1u) ? s : s * make_double(b, 0);
        break;
    }
    const SharedDiv d(s);
    const double q = d(x), want = x / s;
    const double ax = fabs(x);
    fast += (d.ok && (ax == 0.0 || (ax >= 0x1p-500 && ax <= 0x1p500))) ? 1u : 0u;
    bad += (__double_as_longlong(q) != __double_as_longlong(want)) ? 1u : 0u;
  }
  atomicAdd(mismatches, bad);
  atomicAdd(fast_path, fast);
}

}  // namespace

}  // namespace bsq

using namespace bsq;

extern "C" int bsq_selftest_division(int32_t device, uint64_t n, uint64_t seed, uint64_t* mismatches,
                                     uint64_t* fast_path, bsq_status* status) {
  try {
    if (!mismatches || !fast_path) throw BsqError(BSQ_ERR_ARG, "null argument");
    if (hipSetDevice(device) != hipSuccess) throw BsqError(BSQ_ERR_DEVICE, "no such HIP device");
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, 2 * sizeof(unsigned long long)) != hipSuccess) throw BsqError(BSQ_ERR_DEVICE, "hipMalloc failed");
    unsigned long long h[2] = {0, 0};
    hipError_t e = hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_selftest_division, dim3(4096), dim3(256), 0, 0, n, seed, d, d + 1);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) throw BsqError(BSQ_ERR_DEVICE, hipGetErrorString(e));
    *mismatches = h[0];
    *fast_path = h[1];
    set_status(status, BSQ_OK, "");
    return BSQ_OK;
  } catch (const BsqError& ex) {
    set_status(status, ex.code, ex.what());
    return ex.code;
  }
}
