// noalloc.hip -- lcr_depth_to_cloud_noalloc: the depth back-projection (K13) for
// callers that cannot allocate device memory (code under HIP graph capture, a
// per-frame capture loop).  The caller passes device buffers for the inputs and
// outputs; the kernels' scratch -- per-pixel validity flags, their exclusive scan
// (the output slot of every point) and the scan's tile sums -- lives in one fixed
// file-scope device array, so the call allocates nothing.
//
// Each device holds its own copy of `scratch` at its own address.  The address is
// taken once per device with hipGetSymbolAddress (in lcr_init for the device it
// selects, or on the entry point's first call on a device) and kept in
// g_scratch[device]; host code never passes the array's name to a kernel.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/lcr.h"
#include "lcr_kernels.hpp"
#include "runtime.hpp"

namespace lcr {

namespace {

constexpr int64_t kMaxPixels = LCR_NOALLOC_MAX_PIXELS;
// flags [0, kMaxPixels), scan [kMaxPixels, 2 kMaxPixels), tile sums after them
// (the scan needs scan_tmp_ints(n) <= n / 1024 + 2 of them)
constexpr int64_t kTmpInts = kMaxPixels / 1024 + 2;
constexpr int64_t kScratchInts = 2 * kMaxPixels + kTmpInts;

static __device__ __align__(256) int scratch[kScratchInts];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
int* g_scratch[kMaxDevices] = {};

__global__ void k_zero_count(int64_t* out_n) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *out_n = 0;
}

// The current device's scratch address (taken once per device).
int* scratch_here() {
  int dev = 0;
  LCR_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error{LCR_E_LIMIT, "device index beyond 63"};
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    int* p = nullptr;
    LCR_HIP(hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch)));
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

}  // namespace

// Called by lcr_init with `device` current.
void noalloc_init(int device) {
  (void)device;
  (void)scratch_here();
}

}  // namespace lcr

using namespace lcr;

extern "C" int lcr_depth_to_cloud_noalloc(const uint16_t* depth, int32_t width, int32_t height,
                                          const uint8_t* color, const double intrinsics[4],
                                          int32_t depth_min, int32_t depth_max,
                                          const int32_t crop[4], float* out_xyz, float* out_nrm,
                                          uint8_t* out_rgb, int64_t* out_n, void* stream) {
  try {
    const int64_t np = (int64_t)width * height;
    if (width < 0 || height < 0 || !intrinsics || !out_n || (np > 0 && (!depth || !out_xyz)))
      return set_last_error(LCR_E_ARG, "invalid arguments");
    if (np > kMaxPixels)
      return set_last_error(LCR_E_LIMIT,
                            "depth map larger than LCR_NOALLOC_MAX_PIXELS (2^20) pixels: the "
                            "fixed scratch of lcr_depth_to_cloud_noalloc holds no more");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int* base = scratch_here();
    if (np == 0) {
      hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, s, out_n);
      LCR_HIP(hipGetLastError());
      return LCR_OK;
    }
    if (scan_tmp_ints(np) > kTmpInts) return set_last_error(LCR_E_LIMIT, "scan scratch too small");
    int* flags = base;
    int* pos = base + kMaxPixels;
    int* tmp = base + 2 * kMaxPixels;
    DepthParams P;
    P.fx = intrinsics[0];
    P.fy = intrinsics[1];
    P.cx = intrinsics[2];
    P.cy = intrinsics[3];
    P.depth_min = depth_min;
    P.depth_max = depth_max;
    P.top = crop ? crop[0] : 0;
    P.bottom = crop ? crop[1] : height - 1;
    P.left = crop ? crop[2] : 0;
    P.right = crop ? crop[3] : width - 1;
    // every scratch word a kernel reads is written earlier in the same call (flags
    // for every pixel, the scan and its tile sums in full), so nothing is cleared
    launch_depth_flags(depth, width, height, P, flags, s);
    launch_exclusive_scan(flags, pos, np, tmp, s);
    launch_depth_points(depth, width, height, color, P, pos, out_xyz, out_nrm,
                        color ? out_rgb : nullptr, s);
    launch_depth_count(pos, flags, np, out_n, s);
    LCR_HIP(hipGetLastError());
    return LCR_OK;
  } catch (const Error& e) {
    return set_last_error(e.code, e.msg);
  } catch (const std::exception& e) {
    return set_last_error(LCR_E_ARG, e.what());
  }
}
