// fuse.hip -- K14: burst depth fusion on the device, and its C ABI entry point.
//
// Replaces the reconstruction half of capture/super_resolution/super_resolution.cpp
// (:283-314) for a burst of depth frames that are already aligned (the ECC
// registration and warp of :255-279 are the identity for a static rig):
//   * each frame upsampled by `factor` with nearest neighbours (:283, INTER_NEAREST):
//     the upsampled image repeats every pixel factor x factor times, so it is
//     never stored -- the resize samples pixel (r / factor, c / factor);
//   * the zero-eliminating mean per pixel over the frames (:295-306): the mean of
//     the positive depths, accumulated in double as std::accumulate(.., 0.0) does,
//     NaN when no frame has one;
//   * the Lanczos-4 resize back to the frame size (:311, INTER_LANCZOS4): output
//     pixel x samples source position (x + 0.5) factor - 0.5 with 8 taps (clamped
//     to the image) per axis, the horizontal pass first, then the vertical, in
//     float; a NaN anywhere under the taps gives NaN;
//   * the 16-bit conversion (:312, convertTo CV_16UC1): round to nearest even,
//     clamped to [0, 65535], NaN -> 0.
// One thread per frame pixel for the mean (frames read once, coalesced); one per
// output pixel for the resize (64 float taps from the mean image, cached).  HBM
// bound: 2 B per frame pixel read, 4 + 4 B for the mean, 2 B written.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <memory>

#include "../../include/lcr.h"
#include "runtime.hpp"

namespace lcr {

#pragma clang fp contract(off)

namespace {

struct Taps8 {
  float w[8];
};

// Lanczos-4 weights of the 8 taps around a sample at fractional offset fr: the
// windowed sinc sin(pi d) sin(pi d / 4) / (pi^2 d^2 / 4), d = fr + 3 - i, normalised
// to sum 1 in double and rounded to float; fr == 0 is the unit tap.
Taps8 lanczos4_taps(double fr) {
  Taps8 t;
  const double pi = 3.14159265358979323846;
  if (fr == 0.0) {
    for (int i = 0; i < 8; ++i) t.w[i] = i == 3 ? 1.0f : 0.0f;
    return t;
  }
  double c[8], s = 0.0;
  for (int i = 0; i < 8; ++i) {
    const double d = fr + 3.0 - i;
    c[i] = sin(pi * d) * sin(pi * d / 4.0) / (pi * pi * d * d / 4.0);
    s += c[i];
  }
  for (int i = 0; i < 8; ++i) t.w[i] = (float)(c[i] / s);
  return t;
}

__global__ void k_fuse_mean(const uint16_t* __restrict__ frames, int nf, int64_t np,
                            float* __restrict__ mean) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double s = 0.0;
  int c = 0;
  for (int f = 0; f < nf; ++f) {
    const uint16_t v = frames[(int64_t)f * np + p];
    if (v > 0) {
      s += (double)(float)v;
      ++c;
    }
  }
  mean[p] = c > 0 ? (float)(s / c) : NAN;
}

__global__ void k_fuse_resize(const float* __restrict__ mean, int w, int h, int factor, int base,
                              Taps8 taps, uint16_t* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= w || y >= h) return;
  const int W = w * factor, H = h * factor;
  const int sx = factor * x + base, sy = factor * y + base;
  float acc = 0.0f;
#pragma unroll
  for (int ky = 0; ky < 8; ++ky) {
    const int r = min(max(sy - 3 + ky, 0), H - 1);
    const float* mr = mean + (int64_t)(r / factor) * w;
    float hs = 0.0f;
#pragma unroll
    for (int kx = 0; kx < 8; ++kx) {
      const int q = min(max(sx - 3 + kx, 0), W - 1);
      hs = hs + taps.w[kx] * mr[q / factor];
    }
    acc = acc + taps.w[ky] * hs;
  }
  uint16_t v = 0;
  if (!isnan(acc)) {
    const float rv = rintf(acc);
    v = rv <= 0.0f ? (uint16_t)0 : (rv >= 65535.0f ? (uint16_t)65535 : (uint16_t)rv);
  }
  out[(int64_t)y * w + x] = v;
}

}  // namespace

}  // namespace lcr

using namespace lcr;

extern "C" int lcr_depth_fuse(const uint16_t* frames, int32_t n_frames, int32_t width,
                              int32_t height, int32_t factor, uint16_t* out) {
  try {
    const int64_t np = (int64_t)width * height;
    if (width < 0 || height < 0 || n_frames < 1 || (np > 0 && (!frames || !out)))
      return set_last_error(LCR_E_ARG, "invalid arguments");
    if (factor < 1 || factor > 16) return set_last_error(LCR_E_ARG, "factor must be in 1..16");
    if (np > ((int64_t)1 << 26) || (int64_t)width * factor > INT32_MAX / 2 ||
        (int64_t)height * factor > INT32_MAX / 2 || (int64_t)n_frames * np > ((int64_t)1 << 31))
      return set_last_error(LCR_E_LIMIT, "burst larger than 2^26 pixels per frame or 2^31 in all");
    if (np == 0) return LCR_OK;
    int dev = 0;
    LCR_HIP(hipGetDevice(&dev));
    hipStream_t s = nullptr;
    LCR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::unique_ptr<ihipStream_t, void (*)(hipStream_t)> guard(s, [](hipStream_t q) {
      (void)hipStreamDestroy(q);
    });
    DevBuf d_frames, d_mean, d_out;
    d_frames.reserve((size_t)(n_frames * np) * sizeof(uint16_t));
    d_mean.reserve((size_t)np * sizeof(float));
    d_out.reserve((size_t)np * sizeof(uint16_t));
    LCR_HIP(hipMemcpyAsync(d_frames.p, frames, (size_t)(n_frames * np) * sizeof(uint16_t),
                           hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fuse_mean, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s,
                       d_frames.as<uint16_t>(), (int)n_frames, np, d_mean.as<float>());
    // output x samples (x + 0.5) factor - 0.5 = factor x + (factor - 1) / 2
    const double pos = (factor - 1) / 2.0;
    const int base = (int)floor(pos);
    const Taps8 taps = lanczos4_taps(pos - base);
    hipLaunchKernelGGL(k_fuse_resize, dim3((unsigned)((width + 127) / 128), (unsigned)height),
                       dim3(128), 0, s, d_mean.as<float>(), (int)width, (int)height, (int)factor,
                       base, taps, d_out.as<uint16_t>());
    LCR_HIP(hipGetLastError());
    LCR_HIP(hipMemcpyAsync(out, d_out.p, (size_t)np * sizeof(uint16_t), hipMemcpyDeviceToHost, s));
    LCR_HIP(hipStreamSynchronize(s));
    return LCR_OK;
  } catch (const Error& e) {
    return set_last_error(e.code, e.msg);
  } catch (const std::exception& e) {
    return set_last_error(LCR_E_ARG, e.what());
  }
}
