// filters.hip -- the steps before ICP: K11 VoxelGrid centroids, K12 statistical
// outlier removal statistics, K13 depth-map back-projection.
//
// K11 replaces pcl::VoxelGrid<PointXYZRGBNormal>::filter with
// setLeafSize(l, l, l) (pcl_tools/cloud_downsampling.cpp:73-75): voxel
// coordinates floor(p * (1/l)) relative to the minimum's, one centroid per
// occupied voxel in increasing voxel index (x fastest).  Counting and prefix
// sums give every voxel its range; a stable radix sort of (voxel, input index)
// gives every voxel its points in input order (linear in the points, whatever
// the voxel sizes: with the CLI's default 1 m leaf a whole capture falls into a
// few voxels); one thread per voxel then sums them in that order (FP32, as PCL's
// CentroidPoint), so the output does not depend on the order in which atomics
// arrived.
//
// K12 finishes pcl::StatisticalOutlierRemoval (pcl_tools/outlier_removal.cpp:80-84):
// given each point's mean distance to its k nearest neighbours (normals.hip), the
// mean and sample standard deviation over the finite points (FP64, fixed-order
// two-level reduction) and the keep flags.
//
// K13 replaces the depth tool's cross-product normals and back-projection
// (capture/depth_capture/depth_capture.cpp:230-305) for a 16-bit depth map:
// per-pixel validity, a prefix sum for row-major output positions, then the
// points, normals and colours.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <hipcub/hipcub.hpp>

#include "lcr_kernels.hpp"

namespace lcr {

#pragma clang fp contract(off)

namespace {

inline unsigned blocks_of(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

__device__ __forceinline__ bool finite3(const float4& p) {
  return isfinite(p.x) && isfinite(p.y) && isfinite(p.z);
}

// Per point its voxel (skey: nvox for a non-finite point, so that those sort
// last) and its input index (the sort's payload).
__global__ void k_vox_key(const float4* __restrict__ raw, int64_t n, VoxLattice L, int64_t nvox,
                          uint32_t* __restrict__ skey, int* __restrict__ idx,
                          int* __restrict__ counts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  idx[i] = (int)i;
  const float4 p = raw[i];
  if (!finite3(p)) {
    skey[i] = (uint32_t)nvox;
    return;
  }
  const int ix = (int)(floorf(p.x * L.inv) - (float)L.minb[0]);
  const int iy = (int)(floorf(p.y * L.inv) - (float)L.minb[1]);
  const int iz = (int)(floorf(p.z * L.inv) - (float)L.minb[2]);
  const int64_t k = (int64_t)ix + (int64_t)iy * L.div[0] + (int64_t)iz * L.div[0] * L.div[1];
  skey[i] = (uint32_t)k;
  atomicAdd(&counts[k], 1);
}

__global__ void k_vox_flags(const int* __restrict__ counts, int64_t nvox, int* __restrict__ flags) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nvox) flags[v] = counts[v] > 0 ? 1 : 0;
}

__global__ void k_vox_centroid(const int* __restrict__ counts, const int* __restrict__ start,
                               const int* __restrict__ outidx, int64_t nvox,
                               const int* __restrict__ sorted_raw, const float4* __restrict__ raw,
                               const float4* __restrict__ nrm, const float* __restrict__ curv,
                               const uint8_t* __restrict__ rgb, float* __restrict__ out_xyz,
                               float* __restrict__ out_nrm, float* __restrict__ out_curv,
                               uint8_t* __restrict__ out_rgb) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  const int c = counts[v];
  if (c == 0) return;
  const int b = start[v];
  float sx = 0.f, sy = 0.f, sz = 0.f, snx = 0.f, sny = 0.f, snz = 0.f, sc = 0.f;
  float sr = 0.f, sg = 0.f, sb = 0.f;
  for (int t = b; t < b + c; ++t) {
    const int i = sorted_raw[t];
    const float4 p = raw[i];
    sx = sx + p.x;
    sy = sy + p.y;
    sz = sz + p.z;
    if (nrm) {
      const float4 m = nrm[i];
      snx = snx + m.x;
      sny = sny + m.y;
      snz = snz + m.z;
    }
    if (curv) sc = sc + curv[i];
    if (rgb) {
      sr = sr + (float)rgb[3 * (int64_t)i];
      sg = sg + (float)rgb[3 * (int64_t)i + 1];
      sb = sb + (float)rgb[3 * (int64_t)i + 2];
    }
  }
  const int64_t o = outidx[v];
  const float cnt = (float)c;
  out_xyz[3 * o] = sx / cnt;
  out_xyz[3 * o + 1] = sy / cnt;
  out_xyz[3 * o + 2] = sz / cnt;
  if (out_nrm && nrm) {
    // Eigen normalized() of the summed normal: v / sqrt(v . v)
    float nn = snx * snx;
    nn = nn + sny * sny;
    nn = nn + snz * snz;
    const float len = sqrtf(nn);
    out_nrm[3 * o] = len > 0.f ? snx / len : snx;
    out_nrm[3 * o + 1] = len > 0.f ? sny / len : sny;
    out_nrm[3 * o + 2] = len > 0.f ? snz / len : snz;
  }
  if (out_curv && curv) out_curv[o] = sc / cnt;
  if (out_rgb && rgb) {
    out_rgb[3 * o] = (uint8_t)(uint32_t)(sr / cnt);
    out_rgb[3 * o + 1] = (uint8_t)(uint32_t)(sg / cnt);
    out_rgb[3 * o + 2] = (uint8_t)(uint32_t)(sb / cnt);
  }
}

// K12: block partials of (sum, sum of float squares, finite count) in FP64
constexpr int kSorThreads = 256;

__global__ void __launch_bounds__(kSorThreads)
k_sor_partials(const float* __restrict__ dist, const uint8_t* __restrict__ finite, int64_t n,
               double* __restrict__ part) {
  __shared__ double s0[kSorThreads], s1[kSorThreads], s2[kSorThreads];
  double a = 0.0, b = 0.0, c = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kSorThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kSorThreads) {
    const float d = dist[i];
    a += d;
    b += (double)(d * d);
    c += finite[i] ? 1.0 : 0.0;
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  s2[threadIdx.x] = c;
  __syncthreads();
  for (int o = kSorThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s0[threadIdx.x] += s0[threadIdx.x + o];
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = s0[0];
    part[3 * blockIdx.x + 1] = s1[0];
    part[3 * blockIdx.x + 2] = s2[0];
  }
}

// one thread: fold the partials in block order; stats = mean, stddev, threshold
__global__ void k_sor_stats(const double* __restrict__ part, int nparts, double mul,
                            double* __restrict__ stats) {
  double sum = 0.0, sq = 0.0, valid = 0.0;
  for (int k = 0; k < nparts; ++k) {
    sum += part[3 * k];
    sq += part[3 * k + 1];
    valid += part[3 * k + 2];
  }
  const double mean = valid > 0.0 ? sum / valid : 0.0;
  const double var = valid > 1.0 ? (sq - sum * sum / valid) / (valid - 1.0) : 0.0;
  const double sd = sqrt(var);
  stats[0] = mean;
  stats[1] = sd;
  stats[2] = mean + mul * sd;
}

__global__ void k_sor_keep(const float* __restrict__ dist, int64_t n, const double* __restrict__ stats,
                           uint8_t* __restrict__ keep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keep[i] = (double)dist[i] > stats[2] ? 0 : 1;
}

// K13
__device__ __forceinline__ bool depth_valid(const DepthParams& P, int z, int i, int j) {
  return z != 0 && z >= P.depth_min && z <= P.depth_max && i >= P.top && i <= P.bottom &&
         j >= P.left && j <= P.right;
}

__global__ void k_depth_flags(const uint16_t* __restrict__ depth, int w, int h, DepthParams P,
                              int* __restrict__ flags) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)w * h) return;
  const int i = (int)(p / w), j = (int)(p % w);
  flags[p] = depth_valid(P, depth[p], i, j) ? 1 : 0;
}

__global__ void k_depth_points(const uint16_t* __restrict__ depth, int w, int h,
                               const uint8_t* __restrict__ color, DepthParams P,
                               const int* __restrict__ pos, float* __restrict__ out_xyz,
                               float* __restrict__ out_nrm, uint8_t* __restrict__ out_rgb) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)w * h) return;
  const int i = (int)(p / w), j = (int)(p % w);
  const int z = depth[p];
  if (!depth_valid(P, z, i, j)) return;
  int rt = 0, lf = 0, dn = 0, up = 0;
  if (i > 0 && j > 0 && i < h - 1 && j < w - 1) {
    rt = depth[p + w];
    lf = depth[p - w];
    dn = depth[p + 1];
    up = depth[p - 1];
  }
  if (abs(z - rt) > 20) rt = 0;
  if (abs(z - lf) > 20) lf = 0;
  if (abs(z - dn) > 20) dn = 0;
  if (abs(z - up) > 20) up = 0;
  const double dzdx = (rt - lf) / 2.0, dzdy = (dn - up) / 2.0;
  const double vx = -dzdx, vy = -dzdy, vz = 1.0;
  const double inv = 1.0 / sqrt(vx * vx + vy * vy + vz * vz);
  const double px = (double)z * (j - P.cx) * (1 / P.fx);
  const double py = -(double)z * (i - P.cy) * (1 / P.fy);
  const int64_t m = pos[p];
  out_xyz[3 * m] = (float)px;
  out_xyz[3 * m + 1] = (float)py;
  out_xyz[3 * m + 2] = (float)(-z);
  if (out_nrm) {
    out_nrm[3 * m] = (float)(vx * inv);
    out_nrm[3 * m + 1] = (float)(vy * inv);
    out_nrm[3 * m + 2] = (float)(vz * inv);
  }
  if (out_rgb && color) {
    out_rgb[3 * m] = color[3 * p];
    out_rgb[3 * m + 1] = color[3 * p + 1];
    out_rgb[3 * m + 2] = color[3 * p + 2];
  }
}

__global__ void k_depth_count(const int* __restrict__ pos, const int* __restrict__ flags, int64_t n,
                              int64_t* __restrict__ out_n) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *out_n = (int64_t)pos[n - 1] + flags[n - 1];
}

}  // namespace

// Bits of the voxel keys the sort looks at: up to nvox (the non-finite marker).
static int vox_key_bits(int64_t nvox) {
  int b = 1;
  while (((int64_t)1 << b) <= nvox) ++b;
  return b;
}

size_t voxel_sort_tmp_bytes(int64_t n, int64_t nvox) {
  size_t bytes = 0;
  if (n <= 0) return 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                         (const int*)nullptr, (int*)nullptr, (int)n, 0,
                                         vox_key_bits(nvox)) != hipSuccess)
    return 0;
  return bytes;
}

hipError_t launch_voxel_grid(const float4* raw, const float4* raw_nrm, const float* curv,
                             const uint8_t* rgb, int64_t n, const VoxLattice& L, int64_t nvox,
                             VoxScratch w, float* out_xyz, float* out_nrm, float* out_curv,
                             uint8_t* out_rgb, hipStream_t s) {
  if (n <= 0 || nvox <= 0) return hipSuccess;
  launch_fill_i32(w.counts, nvox, 0, s);
  hipLaunchKernelGGL(k_vox_key, dim3(blocks_of(n, 256)), dim3(256), 0, s, raw, n, L, nvox, w.skey,
                     w.idx, w.counts);
  launch_exclusive_scan(w.counts, w.start, nvox, w.scan_tmp, s);
  hipLaunchKernelGGL(k_vox_flags, dim3(blocks_of(nvox, 256)), dim3(256), 0, s, w.counts, nvox,
                     w.flags);
  launch_exclusive_scan(w.flags, w.outidx, nvox, w.scan_tmp, s);
  // stable: within a voxel the points keep increasing input index
  size_t bytes = w.sort_tmp_bytes;
  const hipError_t e = hipcub::DeviceRadixSort::SortPairs(w.sort_tmp, bytes, w.skey, w.skey_sorted,
                                                          w.idx, w.sorted_raw, (int)n, 0,
                                                          vox_key_bits(nvox), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_vox_centroid, dim3(blocks_of(nvox, 256)), dim3(256), 0, s, w.counts,
                     w.start, w.outidx, nvox, w.sorted_raw, raw, raw_nrm, curv, rgb, out_xyz,
                     out_nrm, out_curv, out_rgb);
  return hipSuccess;
}

int sor_stat_blocks(int64_t n) {
  const int64_t b = (n + kSorThreads - 1) / kSorThreads;
  return (int)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

void launch_sor_stats(const float* dist, const uint8_t* finite, int64_t n, double mul, double* part,
                      double* stats, uint8_t* keep, hipStream_t s) {
  const int nb = sor_stat_blocks(n);
  hipLaunchKernelGGL(k_sor_partials, dim3(nb), dim3(kSorThreads), 0, s, dist, finite, n, part);
  hipLaunchKernelGGL(k_sor_stats, dim3(1), dim3(1), 0, s, part, nb, mul, stats);
  if (n > 0)
    hipLaunchKernelGGL(k_sor_keep, dim3(blocks_of(n, 256)), dim3(256), 0, s, dist, n, stats, keep);
}

void launch_depth_flags(const uint16_t* depth, int w, int h, const DepthParams& P, int* flags,
                        hipStream_t s) {
  const int64_t np = (int64_t)w * h;
  if (np <= 0) return;
  hipLaunchKernelGGL(k_depth_flags, dim3(blocks_of(np, 256)), dim3(256), 0, s, depth, w, h, P,
                     flags);
}

void launch_depth_points(const uint16_t* depth, int w, int h, const uint8_t* color,
                         const DepthParams& P, const int* pos, float* out_xyz, float* out_nrm,
                         uint8_t* out_rgb, hipStream_t s) {
  const int64_t np = (int64_t)w * h;
  if (np <= 0) return;
  hipLaunchKernelGGL(k_depth_points, dim3(blocks_of(np, 256)), dim3(256), 0, s, depth, w, h, color,
                     P, pos, out_xyz, out_nrm, out_rgb);
}

void launch_depth_count(const int* pos, const int* flags, int64_t n, int64_t* out_n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_depth_count, dim3(1), dim3(64), 0, s, pos, flags, n, out_n);
}

}  // namespace lcr
